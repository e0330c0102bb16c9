// RT-RCS — the drop-in command line of the reference
// (src/RT-RCS.cpp:12-31, src/app/rtRcsEntry.cu:59-352), running the sweep
// through libsbr on one MI355X.
//
//   RT-RCS --model <file.obj|.gltf|.glb> [--model-scale S] [--config P] [freq]
//          [--precision fp32|fp64|mixed] [--method per_angle|batched]
//          [--device N] [--print-config]
//
// Same config.txt keys and defaults, same CLI rules (the first bare token is
// the frequency in Hz; unknown flags warn), same ./rcs_results.csv header
// and "theta,phi,rcs_m2,rcs_dbsm" rows in the working precision with default
// ostream formatting, same exit codes (0, or 1 on error). The precision /
// method / device flags are additions; fp32 + batched is the default.
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "mesh_io.hpp"
#include "sbr.h"

namespace {

// simulationConfig defaults (include/app/config.hpp:8-25); values stay in
// double until narrowed to the working precision like static_cast<Real>.
struct Config {
    double phiStart = 0.0, phiEnd = 180.0;
    int phiSamples = 181;
    double thetaStart = 90.0, thetaEnd = 90.0;
    int thetaSamples = 1;
    double gridSize = 10.0;
    int nx = 5000, ny = 5000, tpbx = 16, tpby = 16, maxBounces = 20;
    double reflectionConst = 1.0, freq = 10.0e9;
    bool showInfoGPU = true, showHitStats = true;
    std::string modelPath;
    double modelScale = 1.0;
    int precision = SBR_FP32;
    int method = SBR_BATCHED;
    int device = 0;
    bool printConfig = false;
};

// config.cpp:11-36: "key value" lines, '#' starts a comment
std::map<std::string, double> read_config_file(const std::string& path) {
    std::map<std::string, double> kv;
    std::ifstream f(path);
    if (!f.is_open()) {
        std::cerr << "Warning: Could not open " << path << ". Using defaults." << std::endl;
        return kv;
    }
    std::string line;
    while (std::getline(f, line)) {
        size_t c = line.find('#');
        if (c != std::string::npos) line = line.substr(0, c);
        std::stringstream ss(line);
        std::string key;
        double value;
        if (ss >> key >> value) kv[key] = value;
    }
    return kv;
}

template <typename R> std::string fmt(double v) {
    std::ostringstream o;
    o << static_cast<R>(v);
    return o.str();
}

std::string fmt_real(double v, int precision) {
    return precision == SBR_FP64 ? fmt<double>(v) : fmt<float>(v);
}

// config.cpp:39-155
Config load_config(int argc, char** argv) {
    std::string path = "config.txt";
    for (int i = 1; i < argc; ++i)
        if (std::string(argv[i]) == "--config" && i + 1 < argc) {
            path = argv[i + 1];
            break;
        }
    Config c;
    auto kv = read_config_file(path);
    auto num = [&](const char* k, double& dst) {
        if (kv.count(k)) dst = kv[k];
    };
    auto integer = [&](const char* k, int& dst) {
        if (kv.count(k)) dst = static_cast<int>(kv[k]);
    };
    num("phi_start", c.phiStart);
    num("phi_end", c.phiEnd);
    integer("phi_samples", c.phiSamples);
    num("theta_start", c.thetaStart);
    num("theta_end", c.thetaEnd);
    integer("theta_samples", c.thetaSamples);
    num("grid_size", c.gridSize);
    integer("n_x", c.nx);
    integer("n_y", c.ny);
    integer("tpb_x", c.tpbx);  // accepted and ignored, as in the reference
    integer("tpb_y", c.tpby);
    integer("max_bounces", c.maxBounces);
    num("reflection_const", c.reflectionConst);
    if (kv.count("show_info_GPU")) c.showInfoGPU = static_cast<bool>(kv["show_info_GPU"]);
    if (kv.count("show_hit_stats")) c.showHitStats = static_cast<bool>(kv["show_hit_stats"]);
    bool freqFromConfig = false, freqFromCli = false;
    if (kv.count("freq")) {
        c.freq = kv["freq"];
        freqFromConfig = true;
    }
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--model") {
            if (i + 1 >= argc) {
                std::cerr << "Error: --model requires a file path.\n";
                std::exit(1);
            }
            c.modelPath = argv[++i];
            continue;
        }
        if (a == "--model-scale") {
            if (i + 1 >= argc) {
                std::cerr << "Error: --model-scale requires a numeric value.\n";
                std::exit(1);
            }
            c.modelScale = std::atof(argv[++i]);
            continue;
        }
        if (a == "--config") {
            ++i;
            continue;
        }
        if (a == "--precision" && i + 1 < argc) {
            std::string v = argv[++i];
            c.precision = v == "fp64" ? SBR_FP64 : (v == "mixed" ? SBR_MIXED : SBR_FP32);
            if (v != "fp64" && v != "mixed" && v != "fp32")
                std::cerr << "Warning: unknown precision '" << v << "', using fp32.\n";
            continue;
        }
        if (a == "--method" && i + 1 < argc) {
            std::string v = argv[++i];
            c.method = v == "per_angle" ? SBR_PER_ANGLE : SBR_BATCHED;
            continue;
        }
        if (a == "--device" && i + 1 < argc) {
            c.device = std::atoi(argv[++i]);
            continue;
        }
        if (a == "--print-config") {
            c.printConfig = true;
            continue;
        }
        if (!a.empty() && a[0] != '-') {
            if (!freqFromCli) {
                c.freq = std::atof(a.c_str());
                freqFromCli = true;
            }
            continue;
        }
        std::cerr << "Warning: Unrecognized argument '" << a << "'.\n";
    }
    const std::string f = fmt_real(c.freq, c.precision);
    if (freqFromCli) std::cout << "Using frequency from command line: " << f << " Hz\n";
    else if (freqFromConfig) std::cout << "Using frequency from config file: " << f << " Hz\n";
    else std::cout << "Using default frequency: " << f << " Hz\n";
    return c;
}

void sep(const std::string& title) {
    std::cerr << "\n┌── " << title << "\n";
}
void kv(const std::string& k, const std::string& v) { std::cerr << "│ " << std::left << std::setw(26) << k << ": " << v << "\n"; }
void endsep() { std::cerr << "└──────────────────────────────────────────────────┘\n"; }

int run(int argc, char** argv) {
    std::cerr << "╔═══════════════════════════════════════════════════╗\n";
    std::cerr << "║  SagittaSBR-MI355X  >>---->  MONOSTATIC RCS       ║\n";
    std::cerr << "╚═══════════════════════════════════════════════════╝\n";
    auto tApp = std::chrono::steady_clock::now();
    Config c = load_config(argc, argv);
    if (c.printConfig) {
        std::cout << "phi_start=" << c.phiStart << " phi_end=" << c.phiEnd << " phi_samples=" << c.phiSamples
                  << " theta_start=" << c.thetaStart << " theta_end=" << c.thetaEnd
                  << " theta_samples=" << c.thetaSamples << " grid_size=" << c.gridSize << " n_x=" << c.nx
                  << " n_y=" << c.ny << " max_bounces=" << c.maxBounces << " reflection_const=" << c.reflectionConst
                  << " freq=" << c.freq << " model=" << c.modelPath << " model_scale=" << c.modelScale
                  << " precision=" << c.precision << " method=" << c.method << "\n";
        return 0;
    }
    if (c.modelPath.empty()) {
        std::cerr << "Error: --model is required for simulation runs.\n";
        return 1;
    }
    {
        std::string l = c.modelPath;
        for (auto& ch : l) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
        auto ends = [&](const char* s) {
            std::string t(s);
            return l.size() >= t.size() && l.compare(l.size() - t.size(), t.size(), t) == 0;
        };
        if (!ends(".obj") && !ends(".gltf") && !ends(".glb")) {
            std::cerr << "Error: Unsupported model extension for '" << c.modelPath << "'.\n";
            return 1;
        }
    }
    auto t0 = std::chrono::steady_clock::now();
    sbr::MeshData mesh;
    std::string err;
    if (!sbr::load_model(c.modelPath, mesh, err)) {
        std::cerr << "Error: Failed to load model '" << c.modelPath << "'.\n";
        if (!err.empty()) std::cerr << err << "\n";
        return 1;
    }
    if (mesh.indices.empty() || mesh.positions.empty() || mesh.indices.size() % 3 != 0) {
        std::cerr << "Error: Model '" << c.modelPath << "' contains no valid triangles.\n";
        return 1;
    }
    const size_t nv = mesh.positions.size() / 3, nt = mesh.indices.size() / 3;
    for (uint32_t i : mesh.indices)
        if (i >= nv) {
            std::cerr << "Error: Model '" << c.modelPath << "' contains out-of-range indices.\n";
            return 1;
        }
    double loadMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();

    sbr_ctx* ctx = nullptr;
    if (sbr_ctx_create(c.device, &ctx)) {
        std::cerr << "Error: " << sbr_last_error(nullptr) << "\n";
        return 1;
    }
    if (c.showInfoGPU) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, c.device) == hipSuccess) {
            sep("GPU INFORMATION");
            kv("Device", prop.name);
            kv("Arch", prop.gcnArchName);
            kv("Total Global Memory", std::to_string(prop.totalGlobalMem / double(1 << 30)) + " GB");
            kv("Compute Units", std::to_string(prop.multiProcessorCount));
            endsep();
        }
    }
    t0 = std::chrono::steady_clock::now();
    int mesh_id = -1;
    if (sbr_mesh_create(ctx, mesh.positions.data(), int(nv), mesh.indices.data(), int(nt), c.modelScale, nullptr,
                        &mesh_id)) {
        std::cerr << "Error: " << sbr_last_error(ctx) << "\n";
        sbr_ctx_destroy(ctx);
        return 1;
    }
    sbr_mesh_info mi{};
    sbr_mesh_info_get(ctx, mesh_id, c.precision == SBR_FP64 ? SBR_FP64 : SBR_FP32, &mi);
    double bvhMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    sep("INITIALIZATION timing");
    kv("Model Loading", std::to_string(loadMs) + " ms");
    kv("BVH Construction + upload", std::to_string(bvhMs) + " ms");
    kv("Triangle Count", std::to_string(nt));
    kv("BVH Node Count", std::to_string(mi.n_nodes));
    kv("BVH Depth", std::to_string(mi.depth));
    endsep();

    const double c0 = 299792458.0;
    const double lambda = c0 / c.freq;
    const double rayArea = c.gridSize * c.gridSize / (double(c.nx) * c.ny);
    sep("SIMULATION PARAMETERS");
    kv("Frequency", std::to_string(c.freq / 1e9) + " GHz");
    kv("Wavelength", std::to_string(lambda * 1000) + " mm");
    kv("Grid Size", std::to_string(c.nx) + "x" + std::to_string(c.ny));
    kv("Ray Area", fmt<double>(rayArea) + " m²");
    kv("Angles", std::to_string(c.thetaSamples) + " x " + std::to_string(c.phiSamples));
    kv("Precision", c.precision == SBR_FP64 ? "fp64" : (c.precision == SBR_MIXED ? "mixed" : "fp32"));
    kv("Method", c.method == SBR_PER_ANGLE ? "per-angle" : "batched (persistent)");
    endsep();
    if (std::sqrt(rayArea) / lambda > 0.2) {
        sep("WARNING");
        std::cerr << "│ Physics is under-resolved at your frequency!\n";
        kv("Ray spacing / λ", std::to_string(std::sqrt(rayArea) / lambda) + " (< 0.1 needed)");
        endsep();
    }

    // header (rtRcsEntry.cu:298-302)
    std::ofstream out("rcs_results.csv");
    out << "# Frequency: " << fmt_real(c.freq, c.precision) << "\n# Grid: " << c.nx << "x" << c.ny << "\n";
    out << "# GridSize: " << fmt_real(c.gridSize, c.precision) << "\n# ThetaSamples: " << c.thetaSamples << "\n";
    out << "# PhiSamples: " << c.phiSamples << "\n"
        << "theta,phi,rcs_m2,rcs_dbsm\n";

    sbr_sweep_params p{};
    p.phi_start = c.phiStart;
    p.phi_end = c.phiEnd;
    p.phi_samples = c.phiSamples;
    p.theta_start = c.thetaStart;
    p.theta_end = c.thetaEnd;
    p.theta_samples = c.thetaSamples;
    p.grid_size = c.gridSize;
    p.n_x = c.nx;
    p.n_y = c.ny;
    p.max_bounces = c.maxBounces;
    p.reflection_const = c.reflectionConst;
    p.freq = c.freq;
    const int n = c.thetaSamples * c.phiSamples;
    if (n <= 0) {
        std::cerr << "Error: theta_samples and phi_samples must be positive.\n";
        sbr_ctx_destroy(ctx);
        return 1;
    }
    std::vector<sbr_job> jobs(n);
    sbr_plan* plan = nullptr;
    if (sbr_sweep_jobs(&p, mesh_id, c.precision, jobs.data(), n) ||
        sbr_plan_create(ctx, jobs.data(), n, c.precision, c.method, &plan)) {
        std::cerr << "Error: " << sbr_last_error(ctx) << "\n";
        sbr_ctx_destroy(ctx);
        return 1;
    }
    sep("EXECUTING SWEEP");
    std::cerr << "│ Total iterations: " << n << "\n";
    endsep();
    auto ts = std::chrono::steady_clock::now();
    std::vector<double> reim(2 * size_t(n)), sigma(n), dbsm(n);
    double kernelMs = 0;
    if (sbr_plan_run(plan) || sbr_plan_wait(plan) || sbr_plan_last_ms(plan, &kernelMs) ||
        sbr_plan_results(plan, reim.data(), sigma.data(), dbsm.data())) {
        std::cerr << "Error: " << sbr_last_error(ctx) << "\n";
        sbr_plan_destroy(plan);
        sbr_ctx_destroy(ctx);
        return 1;
    }
    double sweepMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ts).count();
    // progress lines of every 100th angle and the last (sweep.cu:197,236-303),
    // after the fact: the ray and PO work of an angle is one fused launch
    // (and the batched method sweeps every angle in one launch), so the time
    // shown is the sweep's kernel time per angle
    for (int i = 0; i < n; ++i) {
        if ((i + 1) % 100 != 0 && i != n - 1) continue;
        std::cerr << "[R0:" << std::setw(3) << i + 1 << "/" << n << "]"
                  << " │ Θ: " << std::setw(3) << static_cast<int>(jobs[i].theta_deg) << "°"
                  << " │ Φ: " << std::setw(3) << static_cast<int>(jobs[i].phi_deg) << "°"
                  << " │ RT+PO: " << std::fixed << std::setprecision(2) << kernelMs / n << "ms (sweep mean)"
                  << " │ σ: " << std::setprecision(2) << dbsm[i] << " dBsm" << std::defaultfloat << "\n";
    }
    // rows sorted by global index (sweep.cu:529-545)
    for (int i = 0; i < n; ++i)
        out << fmt_real(jobs[i].theta_deg, c.precision) << "," << fmt_real(jobs[i].phi_deg, c.precision) << ","
            << fmt_real(sigma[i], c.precision) << "," << fmt_real(dbsm[i], c.precision) << "\n";
    out.close();
    if (c.showHitStats) {
        sbr_counters cnt{};
        if (sbr_plan_count(plan, &cnt) == 0) {
            sep("HIT STATISTICS");
            kv("Rays traced (after culling)", std::to_string(cnt.rays));
            kv("Ray-bounces", std::to_string(cnt.ray_bounces));
            kv("Node visits", std::to_string(cnt.node_visits));
            kv("Triangle tests", std::to_string(cnt.tri_tests));
            endsep();
        }
    }
    sbr_plan_destroy(plan);
    sbr_ctx_destroy(ctx);
    double appMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tApp).count();
    sep("SWEEP PERFORMANCE");
    kv("GPU Time (kernels only)", std::to_string(kernelMs) + " ms");
    kv("Total Sweep Wall Time", std::to_string(sweepMs) + " ms");
    kv("Average Time/Point", std::to_string(sweepMs / n) + " ms");
    kv("Angles/s", std::to_string(n / (sweepMs / 1000.0)));
    kv("Total App Time", std::to_string(appMs) + " ms");
    endsep();
    return 0;
}

}  // namespace

int main(int argc, char** argv) { return run(argc, argv); }
