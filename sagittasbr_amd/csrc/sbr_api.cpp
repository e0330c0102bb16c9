// C ABI of libsbr (declared in include/sbr.h): context, meshes, plans.
//
// Host-side arithmetic (launch frames, physics constants) restates
// src/app/sweep.cu:50-55,336-364 in the working precision and is compiled
// with -ffp-contract=off, so the frames are bit-identical to a literal
// reading of the reference source.
#include "sbr.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <set>
#include <string>
#include <type_traits>
#include <vector>

#include "bvh_build.hpp"
#include "mesh_io.hpp"
#include "sbr_device.hpp"

namespace sbr {
template <typename R>
hipError_t launch_trace(const TraceArgs<R>& a, int mode, int sd, bool rec, bool cnt, bool persist, bool pool, bool ww,
                        bool top, int width, int grid, hipStream_t s);
hipError_t launch_zero(unsigned* p, size_t n, hipStream_t s);
template <typename R>
hipError_t launch_coverage(const Job<R>* jobs, const Mesh<R>* meshes, int nJobs, int maxChunks,
                           unsigned long long* mask, hipStream_t s);
template <typename R>
hipError_t launch_finalize(const Job<R>* jobs, int nJobs, int mode, const void* partials, double* jobOut,
                           hipStream_t s);
size_t partial_bytes(bool fp64, int mode);
hipError_t launch_adjoint(const AdjArgs& a, int blocks, hipStream_t s);
template <typename R> int trace_occupancy(int mode, int sd, bool pool, bool top, int width);
hipError_t scratch_init(int device);
void* scratch_ptr(int device);
}  // namespace sbr

using namespace sbr;

static_assert(kScratchBytes == SBR_NOALLOC_SCRATCH_BYTES, "scratch size of the header and the kernels differ");

namespace {

thread_local std::string g_err;

int fail(std::string* slot, const std::string& msg) {
    if (slot) *slot = msg;
    g_err = msg;
    return 1;
}

#define HIP_TRY(slot, expr)                                                                       \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail(slot, std::string(#expr) + " failed: " + hipGetErrorString(e_));          \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t ensure(size_t n) {
        if (n <= bytes && p) return hipSuccess;
        if (p) {
            (void)hipFree(p);
            p = nullptr;
            bytes = 0;
        }
        hipError_t e = hipMalloc(&p, std::max<size_t>(n, 256));
        if (e == hipSuccess) bytes = std::max<size_t>(n, 256);
        return e;
    }
};

template <typename R> struct PrecMesh {
    bool built = false;
    Bvh<R> bvh;
    GpuBvh<R> gpu;
    DevBuf nodes, tris;
    DevBuf tris64;  // FP32 tree: FP64 vertex records in tree order (80 B each), unless vertsExact
    bool vertsExact = false;  // FP32 tree: every FP64 vertex is its FP32 vertex widened
    GpuBvhW gpuW[3];  // FP32 tree: 2-, 4- and 8-wide collapses (mixed mode; width_slot order)
    DevBuf nodesW[3];
    bool builtW[3] = {false, false, false};
    DevBuf cull;    // boxes of the per-tile emptiness test
    int nCull = 0;
    DevBuf cov;     // boxes of the batched runs' coverage masks
    int nCov = 0;
    double buildMs = 0;
};

struct MeshEntry {
    std::vector<float> xyz;
    std::vector<uint32_t> idx;
    double scale = 1.0;
    double radius = 0.0;  // largest |vertex| after scaling (launch grid is centred on the origin)
    BuildOptions opt;
    PrecMesh<float> f32;
    PrecMesh<double> f64;
    DevBuf adjVerts, adjIdx;  // adjoint: reference FP64 vertices (scaled) and indices, original order
};

}  // namespace

struct sbr_ctx {
    int device = 0;
    int cuCount = 0;
    hipStream_t stream = nullptr;     // the stream work is issued on
    hipStream_t ownStream = nullptr;  // the context's own stream
    std::string err;
    std::vector<std::unique_ptr<MeshEntry>> meshes;
    DevBuf meshTable32, meshTable64;
    bool tableDirty32 = true, tableDirty64 = true;
    // bumped whenever a mesh table moves to a new device address: a plan's
    // captured hipGraph holds the address it was captured with
    uint64_t tableGen = 0;
    std::set<sbr_plan*> plans;  // live plans, orphaned when the context goes
    std::string scratchErr;     // why the fixed scratch array has no address
};

struct sbr_plan {
    sbr_ctx* ctx = nullptr;
    int precision = 0;
    int method = 0;
    int nJobs = 0;
    std::vector<sbr_job> jobs;
    std::vector<int64_t> tileBase;
    std::vector<int> nTiles;
    std::vector<int> meshIds;  // distinct meshes of the jobs
    int64_t totalTiles = 0;
    DevBuf dChunkJob;  // job of the first tile of each kChunkTiles chunk
    int64_t nChunks = 0;
    int maxDepth = 0;
    sbr_tuning tune{0, 0, 0, 0, 0, 0, 0, 0, 0};
    int tileSide = 32;
    DevBuf dJobs, dPartials, dJobOut, dCounter, dCounters;
    // mixed mode
    DevBuf dJobs64, dRecFlag;
    // batched runs: the jobs' ray-patch coverage masks (maskWords 64-bit
    // words), rebuilt by every run from the coverage boxes of its meshes in
    // LDS chunks (at most maxChunks per job)
    DevBuf dMask;
    int64_t maskWords = 0;
    int maxChunks = 0;
    // a counting run skips what the coverage masks rule out (sbr_plan_count_traced)
    bool countTraced = false;
    std::vector<double> out;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    uint64_t graphTableGen = 0;  // ctx->tableGen when the graph was captured
    bool ran = false;
};

namespace {

template <typename R> PrecMesh<R>& prec_of(MeshEntry& m);
template <> PrecMesh<float>& prec_of<float>(MeshEntry& m) { return m.f32; }
template <> PrecMesh<double>& prec_of<double>(MeshEntry& m) { return m.f64; }

// A model scale that rounds FP32 vertices leaves each FP32 vertex within 2
// ulps of the reference FP64 one (the FP32 scale's rounding plus the
// product's): the mixed mode's FP32 boxes are then widened by kWidenUlps
// outwards so that they bound the FP64 triangles it tests.
constexpr int kWidenUlps = 3;
float widen_ulps(float v, int n) {
    for (int k = 0; k < (n < 0 ? -n : n); ++k) v = std::nextafter(v, n < 0 ? -INFINITY : INFINITY);
    return v;
}
void widen_boxes(GpuBvhW& g) {
    const int W = g.width;
    const size_t rb = size_t(W * 28 + 63) / 64 * 64;
    const size_t nRec = g.nodes.size() / rb;
    std::vector<float> f(rb / 4);
    for (size_t i = 0; i < nRec; ++i) {
        std::memcpy(f.data(), g.nodes.data() + i * rb, rb);
        for (int q = 0; q < W; ++q) {
            int ref;
            std::memcpy(&ref, &f[6 * W + q], sizeof(int));
            if (ref == kEmptySlot) continue;
            for (int k = 0; k < 3; ++k) {
                f[k * W + q] = widen_ulps(f[k * W + q], -kWidenUlps);
                f[(3 + k) * W + q] = widen_ulps(f[(3 + k) * W + q], kWidenUlps);
            }
        }
        std::memcpy(g.nodes.data() + i * rb, f.data(), rb);
    }
}

// The mixed mode's wide collapse of the mesh's FP32 tree at this width,
// built (on the host, then uploaded) on first use.
// Every node and triangle array of a built tree below 4 GiB (the kernels
// address records with 32-bit byte offsets); an unbuilt tree passes.
constexpr uint64_t kRecordSpan = uint64_t(1) << 32;
template <typename R> bool records_fit(const PrecMesh<R>& pm) {
    if (!pm.built) return true;
    if (pm.gpu.nodes.size() >= kRecordSpan || pm.gpu.tris.size() >= kRecordSpan) return false;
    if constexpr (std::is_same_v<R, float>) {
        if (!pm.vertsExact && uint64_t(pm.bvh.perm.size()) * 80u >= kRecordSpan) return false;
    }
    return true;
}

int ensure_width(sbr_ctx* ctx, PrecMesh<float>& pm, int width) {
    const int slot = width_slot(width);
    if (pm.builtW[slot]) return 0;
    std::string err;
    if (!to_gpu_layout_wide(pm.bvh, width, pm.gpuW[slot], err)) return fail(&ctx->err, err);
    // the kernels address wide records with 32-bit byte offsets
    if (uint64_t(pm.gpuW[slot].nodes.size()) >= (uint64_t(1) << 32)) {
        pm.gpuW[slot] = GpuBvhW{};
        return fail(&ctx->err, "the " + std::to_string(width) + "-wide tree exceeds 4 GiB");
    }
    if (!pm.vertsExact) widen_boxes(pm.gpuW[slot]);
    HIP_TRY(&ctx->err, pm.nodesW[slot].ensure(pm.gpuW[slot].nodes.size()));
    HIP_TRY(&ctx->err, hipMemcpy(pm.nodesW[slot].p, pm.gpuW[slot].nodes.data(), pm.gpuW[slot].nodes.size(),
                                 hipMemcpyHostToDevice));
    pm.builtW[slot] = true;
    ctx->tableDirty32 = true;
    return 0;
}

template <typename R> int upload_table(sbr_ctx* ctx);

// Triangle soup in R (src/app/rtRcsEntry.cu:118-148): float position
// converted to R, times the model scale in R. The context's mesh table of
// that precision is uploaded again at once, so it is current after every
// build.
template <typename R> int ensure_built(sbr_ctx* ctx, MeshEntry& m) {
    PrecMesh<R>& pm = prec_of<R>(m);
    if (pm.built) return 0;
    auto t0 = std::chrono::steady_clock::now();
    const size_t nt = m.idx.size() / 3;
    std::vector<R> soup(nt * 9);
    const R s = static_cast<R>(m.scale);
    for (size_t f = 0; f < nt; ++f)
        for (int c = 0; c < 3; ++c) {
            uint32_t vi = m.idx[3 * f + c];
            for (int k = 0; k < 3; ++k) soup[9 * f + 3 * c + k] = static_cast<R>(m.xyz[3 * size_t(vi) + k]) * s;
        }
    pm.bvh = build_bvh<R>(soup, m.opt);
    if (pm.bvh.depth > kMaxDepth)
        return fail(&ctx->err, "BVH depth " + std::to_string(pm.bvh.depth) +
                                   " exceeds the 64-entry traversal stack of the reference traversal");
    std::string err;
    if (!to_gpu_layout<R>(pm.bvh, pm.gpu, err)) return fail(&ctx->err, err);
    HIP_TRY(&ctx->err, pm.nodes.ensure(pm.gpu.nodes.size()));
    HIP_TRY(&ctx->err, pm.tris.ensure(pm.gpu.tris.size()));
    HIP_TRY(&ctx->err, hipMemcpy(pm.nodes.p, pm.gpu.nodes.data(), pm.gpu.nodes.size(), hipMemcpyHostToDevice));
    HIP_TRY(&ctx->err, hipMemcpy(pm.tris.p, pm.gpu.tris.data(), pm.gpu.tris.size(), hipMemcpyHostToDevice));
    if constexpr (std::is_same_v<R, float>) {
        // reference FP64 vertices (rtRcsEntry.cu:136-144 with Real = double)
        // of every triangle in this tree's order, for the mixed mode
        const double sd = m.scale;
        std::vector<double> t64(nt * 10);
        bool exact = true;
        for (size_t h = 0; h < nt; ++h) {
            const size_t f = size_t(pm.bvh.perm[h]);
            for (int c = 0; c < 3; ++c) {
                uint32_t vi = m.idx[3 * f + c];
                for (int k = 0; k < 3; ++k) {
                    const float x = m.xyz[3 * size_t(vi) + k];
                    t64[10 * h + 3 * c + k] = double(x) * sd;
                    // the FP32 soup's value of the same vertex
                    exact = exact && double(x * s) == t64[10 * h + 3 * c + k];
                }
            }
            long long fid = static_cast<long long>(f);
            std::memcpy(&t64[10 * h + 9], &fid, sizeof(double));
        }
        pm.vertsExact = exact;
        // the default four-wide collapse now; other widths when a plan asks
        if (ensure_width(ctx, pm, kWide)) return 1;
        if (!exact) {
            HIP_TRY(&ctx->err, pm.tris64.ensure(t64.size() * sizeof(double)));
            HIP_TRY(&ctx->err,
                    hipMemcpy(pm.tris64.p, t64.data(), t64.size() * sizeof(double), hipMemcpyHostToDevice));
        }
    }
    {
        std::vector<float> cb = cull_boxes<R>(pm.bvh, kCullBoxes);
        // (FP32 tree whose vertices round: the boxes must bound the FP64
        // triangles of the mixed mode too)
        if constexpr (std::is_same_v<R, float>) {
            if (!pm.vertsExact)
                for (size_t b = 0; b < cb.size() / 6; ++b)
                    for (int k = 0; k < 3; ++k) {
                        cb[6 * b + k] = widen_ulps(cb[6 * b + k], -kWidenUlps);
                        cb[6 * b + 3 + k] = widen_ulps(cb[6 * b + 3 + k], kWidenUlps);
                    }
        }
        pm.nCull = int(cb.size() / 6);
        HIP_TRY(&ctx->err, pm.cull.ensure(cb.size() * sizeof(float)));
        HIP_TRY(&ctx->err, hipMemcpy(pm.cull.p, cb.data(), cb.size() * sizeof(float), hipMemcpyHostToDevice));
        std::vector<float> vb = cover_boxes<R>(pm.bvh, kCovBoxes);
        if constexpr (std::is_same_v<R, float>) {
            if (!pm.vertsExact)
                for (size_t b = 0; b < vb.size() / 6; ++b)
                    for (int k = 0; k < 3; ++k) {
                        vb[6 * b + k] = widen_ulps(vb[6 * b + k], -kWidenUlps);
                        vb[6 * b + 3 + k] = widen_ulps(vb[6 * b + 3 + k], kWidenUlps);
                    }
        }
        pm.nCov = int(vb.size() / 6);
        HIP_TRY(&ctx->err, pm.cov.ensure(vb.size() * sizeof(float)));
        HIP_TRY(&ctx->err, hipMemcpy(pm.cov.p, vb.data(), vb.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    pm.built = true;
    pm.buildMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (sizeof(R) == 4) ctx->tableDirty32 = true;
    else ctx->tableDirty64 = true;
    return upload_table<R>(ctx);
}

template <typename R> int upload_table(sbr_ctx* ctx) {
    bool& dirty = sizeof(R) == 4 ? ctx->tableDirty32 : ctx->tableDirty64;
    DevBuf& buf = sizeof(R) == 4 ? ctx->meshTable32 : ctx->meshTable64;
    if (!dirty && buf.p) return 0;
    std::vector<Mesh<R>> tab(ctx->meshes.size());
    for (size_t i = 0; i < ctx->meshes.size(); ++i) {
        Mesh<R> mm{};
        if (!ctx->meshes[i]) {
            tab[i] = mm;  // destroyed
            continue;
        }
        PrecMesh<R>& pm = prec_of<R>(*ctx->meshes[i]);
        if (pm.built) {
            mm.nodes = pm.nodes.p;
            mm.tris = pm.tris.p;
            mm.tris64 = pm.tris64.p;
            mm.cull = static_cast<const float*>(pm.cull.p);
            mm.nCull = pm.nCull;
            mm.cov = static_cast<const float*>(pm.cov.p);
            mm.nCov = pm.nCov;
            if constexpr (std::is_same_v<R, float>) {
                for (int w = 0; w < 3; ++w) {
                    mm.nodesW[w] = pm.builtW[w] ? pm.nodesW[w].p : nullptr;
                    mm.rootRefW[w] = pm.gpuW[w].rootRef;
                    for (int k = 0; k <= kTopLevels; ++k) mm.topRecsW[w][k] = pm.gpuW[w].topRecs[k];
                }
                mm.vertsExact = pm.vertsExact ? 1 : 0;
                for (int k = 0; k < 3; ++k) {
                    mm.rootLo4[k] = pm.vertsExact ? pm.gpu.rootLo[k] : widen_ulps(pm.gpu.rootLo[k], -kWidenUlps);
                    mm.rootHi4[k] = pm.vertsExact ? pm.gpu.rootHi[k] : widen_ulps(pm.gpu.rootHi[k], kWidenUlps);
                }
            }
            for (int k = 0; k < 3; ++k) {
                mm.rootLo[k] = pm.gpu.rootLo[k];
                mm.rootHi[k] = pm.gpu.rootHi[k];
            }
            mm.tieScale = 1.0;
            for (int k = 0; k < 3; ++k)
                mm.tieScale += std::fmax(std::fabs(double(mm.rootLo[k])), std::fabs(double(mm.rootHi[k])));
            mm.rootRef = pm.gpu.rootRef;
            mm.nTris = static_cast<int>(pm.bvh.perm.size());
        }
        tab[i] = mm;
    }
    const void* before = buf.p;
    const size_t need = sizeof(Mesh<R>) * std::max<size_t>(tab.size(), 1);
    // a larger table replaces the old one: let runs still reading it finish
    if (buf.p && need > buf.bytes) HIP_TRY(&ctx->err, hipStreamSynchronize(ctx->stream));
    HIP_TRY(&ctx->err, buf.ensure(need));
    if (buf.p != before) ++ctx->tableGen;
    HIP_TRY(&ctx->err, hipMemcpy(buf.p, tab.data(), sizeof(Mesh<R>) * tab.size(), hipMemcpyHostToDevice));
    dirty = false;
    return 0;
}

template <typename R> struct V {
    R x, y, z;
};
template <typename R> V<R> crossp(V<R> a, V<R> b) {
    return {a.y * b.z - a.z * b.y, -(a.x * b.z - a.z * b.x), a.x * b.y - a.y * b.x};
}
template <typename R> V<R> unitp(V<R> v) {
    R len = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    return {v.x / len, v.y / len, v.z / len};
}

// Angle of sweep index -> degrees in R (src/app/sweep.cu:336-347).
template <typename R> void sweep_degrees(const sbr_sweep_params* p, int idx, R& th, R& ph) {
    const int t = idx / p->phi_samples, q = idx % p->phi_samples;
    const R ts = R(p->theta_start), te = R(p->theta_end), ps = R(p->phi_start), pe = R(p->phi_end);
    th = ts + (p->theta_samples > 1 ? t * (te - ts) / (p->theta_samples - 1) : R(0));
    ph = ps + (p->phi_samples > 1 ? q * (pe - ps) / (p->phi_samples - 1) : R(0));
}

// Launch frame (src/app/sweep.cu:349-364).
template <typename R> void make_frame(R thDeg, R phDeg, R grid, R llc[3], R uv[3], R vv[3], R dir[3]) {
    const R pi = R(M_PI);
    R th = thDeg * pi / R(180.0);
    R ph = phDeg * pi / R(180.0);
    V<R> d{std::sin(th) * std::cos(ph), std::sin(th) * std::sin(ph), std::cos(th)};
    V<R> up{R(0), R(1), R(0)};
    if (std::fabs(d.y) > R(0.99)) up = {R(1), R(0), R(0)};
    V<R> u = unitp(crossp(up, d));
    u = {u.x * grid, u.y * grid, u.z * grid};
    V<R> v = unitp(crossp(d, u));
    v = {v.x * grid, v.y * grid, v.z * grid};
    const R off = R(10.0) * grid;
    llc[0] = (d.x * off - R(0.5) * u.x) - R(0.5) * v.x;
    llc[1] = (d.y * off - R(0.5) * u.y) - R(0.5) * v.y;
    llc[2] = (d.z * off - R(0.5) * u.z) - R(0.5) * v.z;
    uv[0] = u.x; uv[1] = u.y; uv[2] = u.z;
    vv[0] = v.x; vv[1] = v.y; vv[2] = v.z;
    dir[0] = -d.x; dir[1] = -d.y; dir[2] = -d.z;
}

// Physics constants (src/app/sweep.cu:50-55, poKernels.cu:98-101).
template <typename R> void make_phys(const sbr_job& j, Job<R>& J) {
    const R c0 = R(299792458.0);
    R lambda = c0 / R(j.freq);
    J.k = R(2.0) * R(M_PI) / lambda;
    R g = R(j.grid_size);
    R area = (g * g) / R(j.n_x * j.n_y);
    J.twoPi = R(2.0) * R(M_PI);
    J.invTwoPi = R(1.0) / J.twoPi;
    J.pre = (J.k * area / (R(4.0) * R(M_PI))) * R(2.0);
    J.refl = R(j.reflection_const);
    J.kd = 2.0 * M_PI / (299792458.0 / j.freq);
    J.twoPid = 2.0 * M_PI;
}

// Conservative ray-index rectangle whose rays can reach the root box:
// the 8 corners of the box are projected along the ray direction onto the
// launch plane (exact 3x3 solve in double), plus a 2-ray margin.
template <typename R>
bool cull_rect(const Job<R>& J, const R lo[3], const R hi[3], int& i0, int& i1, int& j0, int& j1) {
    double U[3], W[3], D[3], C[3];
    for (int k = 0; k < 3; ++k) {
        U[k] = J.uv[k];
        W[k] = J.vv[k];
        D[k] = J.dir[k];
        C[k] = J.llc[k];
    }
    // columns U, W, D; solve M [a b c]^T = X - llc by Cramer's rule
    auto det3 = [](const double* a, const double* b, const double* c) {
        return a[0] * (b[1] * c[2] - b[2] * c[1]) - b[0] * (a[1] * c[2] - a[2] * c[1]) +
               c[0] * (a[1] * b[2] - a[2] * b[1]);
    };
    double det = det3(U, W, D);
    if (!(std::fabs(det) > 0)) return false;
    double amin = 1e300, amax = -1e300, bmin = 1e300, bmax = -1e300;
    for (int c = 0; c < 8; ++c) {
        double X[3] = {double((c & 1) ? hi[0] : lo[0]) - C[0], double((c & 2) ? hi[1] : lo[1]) - C[1],
                       double((c & 4) ? hi[2] : lo[2]) - C[2]};
        double a = det3(X, W, D) / det;
        double b = det3(U, X, D) / det;
        amin = std::min(amin, a);
        amax = std::max(amax, a);
        bmin = std::min(bmin, b);
        bmax = std::max(bmax, b);
    }
    // ray (i, j) sits at a = (i + 0.5) / nx
    double fi0 = std::floor(amin * J.nx - 0.5) - 2, fi1 = std::ceil(amax * J.nx - 0.5) + 2;
    double fj0 = std::floor(bmin * J.ny - 0.5) - 2, fj1 = std::ceil(bmax * J.ny - 0.5) + 2;
    fi0 = std::max(fi0, 0.0);
    fj0 = std::max(fj0, 0.0);
    fi1 = std::min(fi1, double(J.nx - 1));
    fj1 = std::min(fj1, double(J.ny - 1));
    if (fi0 > fi1 || fj0 > fj1) return false;
    i0 = int(fi0); i1 = int(fi1); j0 = int(fj0); j1 = int(fj1);
    return true;
}

// Launch-plane coordinates as linear functionals: with M = [U V D] (columns),
// [a b c] = M^-1 (X - llc), so a = X . (V x D) / det - llc . (V x D) / det and
// b = X . (D x U) / det - llc . (D x U) / det (Cramer's rule). Moving llc
// along D changes neither.
template <typename R> void launch_coords(Job<R>& J) {
    double U[3], W[3], D[3], C[3];
    for (int k = 0; k < 3; ++k) {
        U[k] = J.uv[k];
        W[k] = J.vv[k];
        D[k] = J.dir[k];
        C[k] = J.llc[k];
    }
    auto cross = [](const double* a, const double* b, double* o) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
    };
    double wd[3], du[3];
    cross(W, D, wd);
    cross(D, U, du);
    const double det = U[0] * wd[0] + U[1] * wd[1] + U[2] * wd[2];
    if (!(std::fabs(det) > 0)) {
        for (int k = 0; k < 4; ++k) J.cullA[k] = J.cullB[k] = 0.f;
        return;
    }
    double oa = 0, ob = 0;
    for (int k = 0; k < 3; ++k) {
        J.cullA[k] = float(wd[k] / det);
        J.cullB[k] = float(du[k] / det);
        oa += C[k] * wd[k] / det;
        ob += C[k] * du[k] / det;
    }
    J.cullA[3] = float(oa);
    J.cullB[3] = float(ob);
}

constexpr float kErrSafety = 2.0f;  // safety factor on the mixed mode's forward error bounds

// Host side of a plan: device jobs (frames, constants, culled tile
// rectangles, pass grids), the tile -> job table and the partial count.
template <typename R> struct HostPlan {
    std::vector<Job<R>> jobs;
    std::vector<Job<double>> jobs64;  // mixed mode
    std::vector<int> chunkJob;
    int64_t passes = 0;
};

// Device buffers of a plan run (the plan's own, or caller-provided storage).
struct DevPlan {
    void* jobs = nullptr;
    void* partials = nullptr;
    double* jobOut = nullptr;
    unsigned long long* tileCounter = nullptr;
    unsigned long long* counters = nullptr;  // 16 words
    const int* chunkJob = nullptr;
    void* jobs64 = nullptr;
    int8_t* recFlag = nullptr;
    unsigned long long* mask = nullptr;  // batched runs: coverage masks (null: the kCullBoxes test)
};

// Fills hp and the plan's tile bookkeeping. noAlloc: fail instead of
// building a tree that is not built yet (building allocates device memory).
template <typename R> int plan_host(sbr_plan* pl, bool record, bool noAlloc, HostPlan<R>& hp) {
    sbr_ctx* ctx = pl->ctx;
    const bool mixed = pl->precision == SBR_MIXED;
    std::vector<Job<R>>& jobs = hp.jobs;
    std::vector<Job<double>>& jobs64 = hp.jobs64;
    jobs.assign(pl->nJobs, Job<R>{});
    jobs64.assign(mixed ? pl->nJobs : 0, Job<double>{});
    pl->tileBase.assign(pl->nJobs, 0);
    pl->nTiles.assign(pl->nJobs, 0);
    int64_t base = 0;
    int64_t recRays = 0, recIds = 0;
    int64_t rays = 0, passes = 0, maskWords = 0;
    int maxDepth = 0, maxChunks = 0;
    const int ts = pl->tileSide;
    for (int i = 0; i < pl->nJobs; ++i) {
        const sbr_job& j = pl->jobs[i];
        if (j.mesh < 0 || j.mesh >= int(ctx->meshes.size()) || !ctx->meshes[j.mesh])
            return fail(&ctx->err, "job " + std::to_string(i) + ": bad or destroyed mesh id");
        if (j.n_x <= 0 || j.n_y <= 0 || j.max_bounces < 0 || !(j.freq > 0) || !(j.grid_size > 0))
            return fail(&ctx->err, "job " + std::to_string(i) + ": bad grid, bounce count, frequency or size");
        if (int64_t(j.n_x) * j.n_y > 0x7fffffffll)
            return fail(&ctx->err, "job " + std::to_string(i) + ": n_x*n_y exceeds the int ray index");
        MeshEntry& m = *ctx->meshes[j.mesh];
        if (noAlloc && !prec_of<R>(m).built)
            return fail(&ctx->err, std::string("mesh ") + std::to_string(j.mesh) + " has no " +
                                       (sizeof(R) == 8 ? "FP64" : "FP32") + " tree yet (building it allocates)");
        if (ensure_built<R>(ctx, m)) return 1;
        PrecMesh<R>& pm = prec_of<R>(m);
        maxDepth = std::max(maxDepth, pm.bvh.depth);
        Job<R> J{};
        if (!mixed) {
            make_frame<R>(R(j.theta_deg), R(j.phi_deg), R(j.grid_size), J.llc, J.uv, J.vv, J.dir);
            make_phys<R>(j, J);
        } else {
            // reference FP64 job: the ray state and the FP64 answers
            if constexpr (std::is_same_v<R, float>) {
                const int width = pl->tune.node_width ? pl->tune.node_width : kWide;
                if (noAlloc && !pm.builtW[width_slot(width)])
                    return fail(&ctx->err, "mesh " + std::to_string(j.mesh) + " has no " + std::to_string(width) +
                                               "-wide tree yet (building it allocates)");
                if (ensure_width(ctx, pm, width)) return 1;
                const int depth = pm.gpuW[width_slot(width)].depth;
                if (depth > max_depth_wide(width))
                    return fail(&ctx->err, "mixed mode: " + std::to_string(width) + "-wide tree depth " +
                                               std::to_string(depth) + " exceeds " +
                                               std::to_string(max_depth_wide(width)));
            }
            if (noAlloc && !m.f64.built)
                return fail(&ctx->err, "mesh " + std::to_string(j.mesh) +
                                           " has no FP64 tree yet (the mixed mode needs it; building it allocates)");
            if (ensure_built<double>(ctx, m)) return 1;
            Job<double> D{};
            make_frame<double>(j.theta_deg, j.phi_deg, j.grid_size, D.llc, D.uv, D.vv, D.dir);
            make_phys<double>(j, D);
            D.nx = j.n_x;
            D.ny = j.n_y;
            D.maxB = j.max_bounces;
            D.mesh = j.mesh;
            // FP32 job: FP64 frame moved towards the body, rounded once
            const double off = 10.0 * j.grid_size;
            // keep >= 1 cm (> tMin) between the moved plane and any vertex
            double shift = off - m.radius * (1.0 + 1e-6) - std::max(1e-3 * j.grid_size, 0.01);
            if (shift < 0) shift = 0;
            double llc2[3];
            for (int k = 0; k < 3; ++k) {
                llc2[k] = D.llc[k] + shift * D.dir[k];
                J.llc[k] = R(llc2[k]);
                J.uv[k] = R(D.uv[k]);
                J.vv[k] = R(D.vv[k]);
                J.dir[k] = R(D.dir[k]);
            }
            J.k = R(D.k);
            J.pre = R(D.pre);
            J.twoPi = R(D.twoPi);
            J.invTwoPi = R(D.invTwoPi);
            J.refl = R(D.refl);
            J.kd = D.kd;
            J.twoPid = D.twoPid;
            J.L0 = shift;
            auto nrm = [](const double* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); };
            J.ex0 = float(kErrSafety * 5.9604645e-8 * 4.0 * (nrm(llc2) + nrm(D.uv) + nrm(D.vv)));
            jobs64[i] = D;
        }
        // the kernels address node and triangle records with 32-bit byte offsets
        if (!records_fit(pm) || (mixed && !records_fit(m.f64)))
            return fail(&ctx->err, "mesh " + std::to_string(j.mesh) +
                                       ": a node or triangle array of its trees exceeds the 4 GiB the kernels address");
        J.nx = j.n_x;
        J.ny = j.n_y;
        J.maxB = j.max_bounces;
        J.mesh = j.mesh;
        launch_coords<R>(J);
        {
            // FP32 rounding of the tile test's a = X . cullA - cullA[3] over
            // the mesh's boxes: three products and three sums (4 eps of the
            // terms' magnitudes) plus the rounding of cullA itself (1 eps),
            // doubled; 64 eps cover the ray-centre coordinates and the FP32
            // frame of the mixed mode's FP64 rays
            double sa = std::fabs(double(J.cullA[3])), sb = std::fabs(double(J.cullB[3]));
            for (int k = 0; k < 3; ++k) {
                const double mk = std::max(std::fabs(double(pm.gpu.rootLo[k])), std::fabs(double(pm.gpu.rootHi[k])));
                sa += mk * std::fabs(double(J.cullA[k]));
                sb += mk * std::fabs(double(J.cullB[k]));
            }
            constexpr double e32 = 5.9604644775390625e-8;
            J.cullErrA = float(10.0 * e32 * sa + 64.0 * e32);
            J.cullErrB = float(10.0 * e32 * sb + 64.0 * e32);
        }
        int i0 = 0, i1 = j.n_x - 1, j0 = 0, j1 = j.n_y - 1;
        bool any = true;
        if (!record) any = cull_rect<R>(J, pm.gpu.rootLo, pm.gpu.rootHi, i0, i1, j0, j1);
        if (j.max_bounces == 0) any = false;
        J.i0 = i0;
        J.j0 = j0;
        J.tilesX = any ? (i1 - i0 + ts) / ts : 0;
        J.tilesY = any ? (j1 - j0 + ts) / ts : 0;
        J.nTiles = J.tilesX * J.tilesY;
        J.tileBase = base;
        // coverage mask: 64 x 64-ray blocks over the job's tiles
        J.maskX = (J.tilesX * ts + 63) / 64;
        J.maskY = (J.tilesY * ts + 63) / 64;
        J.maskBase = maskWords;
        maskWords += int64_t(J.maskX) * J.maskY;
        maxChunks = std::max(maxChunks, cov_chunks(J.maskX, J.maskY));
        J.passX = any ? (i1 - i0 + 16) / 16 : 0;
        J.passY = any ? (j1 - j0 + 16) / 16 : 0;
        J.passBase = passes;
        J.recBase = recRays;
        J.recIdBase = recIds;
        recRays += int64_t(j.n_x) * j.n_y;
        recIds += int64_t(j.n_x) * j.n_y * j.max_bounces;
        pl->tileBase[i] = base;
        pl->nTiles[i] = J.nTiles;
        base += J.nTiles;
        passes += int64_t(J.passX) * J.passY;
        rays += int64_t(J.nTiles) * ts * ts;
        if (mixed) {
            jobs64[i].tileBase = J.tileBase;
            jobs64[i].recBase = J.recBase;
            jobs64[i].recIdBase = J.recIdBase;
            jobs64[i].nTiles = J.nTiles;
            jobs64[i].passX = J.passX;
            jobs64[i].passY = J.passY;
            jobs64[i].passBase = J.passBase;
        }
        jobs[i] = J;
    }
    pl->totalTiles = base;
    {
        // tile -> job table: the last job whose first tile is <= chunk * kChunkTiles
        pl->nChunks = (base + kChunkTiles - 1) / kChunkTiles;
        std::vector<int>& cj = hp.chunkJob;
        cj.assign(std::max<int64_t>(pl->nChunks, 1), 0);
        int jj = 0;
        for (int64_t c = 0; c < pl->nChunks; ++c) {
            const int64_t t = c * kChunkTiles;
            while (jj + 1 < pl->nJobs && pl->tileBase[jj + 1] <= t) ++jj;
            cj[size_t(c)] = jj;
        }
    }
    pl->maxDepth = maxDepth;
    pl->maxChunks = maxChunks;
    pl->maskWords = maskWords;
    hp.passes = passes;
    return 0;
}

// Host side, then the plan's own device buffers.
template <typename R> int plan_prepare(sbr_plan* pl, bool record) {
    sbr_ctx* ctx = pl->ctx;
    const bool mixed = pl->precision == SBR_MIXED;
    HostPlan<R> hp;
    if (plan_host<R>(pl, record, false, hp)) return 1;
    if (upload_table<R>(ctx)) return 1;
    if (mixed && upload_table<double>(ctx)) return 1;
    const int mode = mixed ? 1 : 0;
    const std::vector<int>& cj = hp.chunkJob;
    HIP_TRY(&ctx->err, pl->dChunkJob.ensure(cj.size() * sizeof(int)));
    HIP_TRY(&ctx->err, hipMemcpy(pl->dChunkJob.p, cj.data(), cj.size() * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(&ctx->err, pl->dJobs.ensure(sizeof(Job<R>) * std::max(pl->nJobs, 1)));
    HIP_TRY(&ctx->err, hipMemcpy(pl->dJobs.p, hp.jobs.data(), sizeof(Job<R>) * pl->nJobs, hipMemcpyHostToDevice));
    HIP_TRY(&ctx->err, pl->dPartials.ensure(partial_bytes(sizeof(R) == 8, mode) * std::max<int64_t>(hp.passes, 1)));
    HIP_TRY(&ctx->err, pl->dJobOut.ensure(sizeof(double) * 2 * std::max(pl->nJobs, 1)));
    HIP_TRY(&ctx->err, pl->dCounter.ensure(sizeof(unsigned long long)));
    HIP_TRY(&ctx->err, pl->dCounters.ensure(sizeof(unsigned long long) * 16));
    if (pl->method == SBR_BATCHED && !record)
        HIP_TRY(&ctx->err, pl->dMask.ensure(sizeof(unsigned long long) * std::max<int64_t>(pl->maskWords, 1)));
    if (mixed) {
        HIP_TRY(&ctx->err, pl->dJobs64.ensure(sizeof(Job<double>) * std::max(pl->nJobs, 1)));
        HIP_TRY(&ctx->err,
                hipMemcpy(pl->dJobs64.p, hp.jobs64.data(), sizeof(Job<double>) * pl->nJobs, hipMemcpyHostToDevice));
    }
    return 0;
}

DevPlan dev_of(sbr_plan* pl) {
    DevPlan d;
    d.jobs = pl->dJobs.p;
    d.partials = pl->dPartials.p;
    d.jobOut = static_cast<double*>(pl->dJobOut.p);
    d.tileCounter = static_cast<unsigned long long*>(pl->dCounter.p);
    d.counters = static_cast<unsigned long long*>(pl->dCounters.p);
    d.chunkJob = static_cast<const int*>(pl->dChunkJob.p);
    d.jobs64 = pl->dJobs64.p;
    d.recFlag = static_cast<int8_t*>(pl->dRecFlag.p);
    d.mask = pl->method == SBR_BATCHED && pl->tune.cull != 2 ? static_cast<unsigned long long*>(pl->dMask.p)
                                                             : nullptr;
    return d;
}

int default_sd(int /*depth*/) { return 8; }

// While-while traversal: measured faster where the BVH stays in cache (the
// trihedral+duct sweep); large meshes can prefer if-if (the autotuner tries both).
int default_traversal() { return 2; }

// One angle launched on its own fills the GPU only with small tiles (a
// 535 x 535 grid gives ~1,100 tiles of 16 x 16 but ~280 of 32 x 32); the
// persistent sweep prefers the larger tile, which halves barrier and
// reduction work per ray.
int default_tile(int method) { return method == SBR_PER_ANGLE ? 16 : 32; }

void drop_graph(sbr_plan* pl) {
    if (pl->graphExec) (void)hipGraphExecDestroy(pl->graphExec);
    if (pl->graph) (void)hipGraphDestroy(pl->graph);
    pl->graphExec = nullptr;
    pl->graph = nullptr;
}

// Free a plan's device resources (graph, events, buffers); its host side
// stays valid for sbr_plan_destroy.
void release_plan_device(sbr_plan* pl) {
    drop_graph(pl);
    if (pl->ev0) (void)hipEventDestroy(pl->ev0);
    if (pl->ev1) (void)hipEventDestroy(pl->ev1);
    pl->ev0 = pl->ev1 = nullptr;
    for (DevBuf* b : {&pl->dJobs, &pl->dPartials, &pl->dJobOut, &pl->dCounter, &pl->dCounters, &pl->dChunkJob,
                      &pl->dJobs64, &pl->dRecFlag, &pl->dMask})
        b->release();
    pl->ctx = nullptr;
}

// Enqueue one run: clear outputs/counters with fill kernels (no memset
// nodes, so the sequence is safe to capture), then the trace launch(es).
template <typename R> int plan_enqueue(sbr_plan* pl, const DevPlan& d, bool record, bool count, int32_t* rIds,
                                       int32_t* rCnt, double* rL, double* rReIm, bool timed) {
    sbr_ctx* ctx = pl->ctx;
    const int mode = pl->precision == SBR_MIXED ? 1 : 0;
    const int sd = pl->tune.lds_stack > 0 ? pl->tune.lds_stack : default_sd(pl->maxDepth);
    // measured on the trihedral+duct sweep: refilling lanes mixes bounce
    // depths within a wave and costs more in traversal coherence than it
    // gains in lane use, so the pool is opt-in (the autotuner tries both)
    const bool pool = pl->tune.ray_pool == 1 && !record;
    TraceArgs<R> a{};
    a.jobs = static_cast<const Job<R>*>(d.jobs);
    a.meshes = static_cast<const Mesh<R>*>(sizeof(R) == 4 ? ctx->meshTable32.p : ctx->meshTable64.p);
    a.nJobs = pl->nJobs;
    a.tileSide = pl->tileSide;
    const bool ww = (pl->tune.traversal ? pl->tune.traversal : default_traversal()) == 2;
    a.totalTiles = pl->totalTiles;
    a.partials = d.partials;
    a.jobOut = d.jobOut;
    a.tileCounter = d.tileCounter;
    a.chunkJob = d.chunkJob;
    a.nChunks = pl->nChunks;
    a.recIds = rIds;
    a.recCount = rCnt;
    a.recL = rL;
    a.recReIm = rReIm;
    a.counters = d.counters;
    const bool mixed = pl->precision == SBR_MIXED;
    // top levels of the wide tree staged in LDS: mixed timed runs, one ray per lane
    const bool top = mixed && !record && !count && !pool && pl->tune.top_levels > 0;
    a.topLevels = top ? pl->tune.top_levels : 0;
    // width of the mixed mode's box tree (the ray pool walks the default one)
    const int width = mixed && !pool && pl->tune.node_width ? pl->tune.node_width : kWide;
    if (mixed) {
        a.jobs64 = static_cast<const Job<double>*>(d.jobs64);
        a.meshes64 = static_cast<const Mesh<double>*>(ctx->meshTable64.p);
        a.errK = kErrSafety;
        a.recFlag = d.recFlag;
    }
    // batched runs cull against the jobs' coverage masks, built first
    // (unless tuned to the per-tile box tests); a counting run culls only
    // when it counts the traced rays
    const bool masked = pl->method == SBR_BATCHED && !record && d.mask && pl->totalTiles > 0 &&
                        (!count || pl->countTraced);
    a.mask = masked ? d.mask : nullptr;
    a.countCulled = count && pl->countTraced ? 1 : 0;
    hipStream_t s = ctx->stream;
    HIP_TRY(&ctx->err, launch_zero(reinterpret_cast<unsigned*>(d.tileCounter), 2, s));
    HIP_TRY(&ctx->err, launch_zero(reinterpret_cast<unsigned*>(d.counters), 32, s));
    if (masked) HIP_TRY(&ctx->err, launch_coverage<R>(a.jobs, a.meshes, pl->nJobs, pl->maxChunks, d.mask, s));
    // the timed region covers the trace launches only
    if (timed) HIP_TRY(&ctx->err, hipEventRecord(pl->ev0, s));
    if (pl->totalTiles > 0) {
        if (record) {
            // every job's tiles in one grid (each block finds its job); job
            // i's records follow those of jobs 0..i-1
            if (pl->totalTiles > INT32_MAX) return fail(&ctx->err, "record run too large for one launch");
            a.tileOffset = 0;
            HIP_TRY(&ctx->err,
                    launch_trace<R>(a, mode, sd, true, count, false, pool, ww, false, width, int(pl->totalTiles), s));
        } else if (pl->method == SBR_BATCHED) {
            int occ = trace_occupancy<R>(mode, sd, pool, top, width);
            int per = pl->tune.wg_per_cu > 0 ? pl->tune.wg_per_cu : std::max(occ, 1);
            int64_t grid = int64_t(ctx->cuCount) * per;
            grid = std::min<int64_t>(grid, pl->totalTiles);
            a.tileOffset = 0;
            HIP_TRY(&ctx->err, launch_trace<R>(a, mode, sd, false, count, true, pool, ww, top, width, int(grid), s));
        } else {
            for (int i = 0; i < pl->nJobs; ++i) {
                if (pl->nTiles[i] == 0) continue;
                a.tileOffset = pl->tileBase[i];
                HIP_TRY(&ctx->err,
                        launch_trace<R>(a, mode, sd, record, count, false, pool, ww, top, width, pl->nTiles[i], s));
            }
        }
    }
    if (timed) HIP_TRY(&ctx->err, hipEventRecord(pl->ev1, s));
    // job sums from the pass partials, in each job's pass-grid order
    HIP_TRY(&ctx->err, launch_finalize<R>(a.jobs, pl->nJobs, mode, d.partials, a.jobOut, s));
    return 0;
}

template <typename R> int plan_launch(sbr_plan* pl, bool record, bool count, int32_t* rIds, int32_t* rCnt, double* rL,
                                      double* rReIm) {
    sbr_ctx* ctx = pl->ctx;
    hipStream_t s = ctx->stream;
    for (int id : pl->meshIds)
        if (!ctx->meshes[id]) return fail(&ctx->err, "a mesh of this plan was destroyed");
    if (pl->tune.use_graph && !record && !count) {
        // the graph holds the mesh table addresses of its capture: capture
        // again after a table moved (a later mesh grew it)
        if (pl->graphExec && pl->graphTableGen != ctx->tableGen) drop_graph(pl);
        if (!pl->graphExec) {
            HIP_TRY(&ctx->err, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            int rc = plan_enqueue<R>(pl, dev_of(pl), false, false, nullptr, nullptr, nullptr, nullptr, false);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(s, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return 1;
            }
            if (e != hipSuccess) return fail(&ctx->err, std::string("graph capture failed: ") + hipGetErrorString(e));
            pl->graph = g;
            HIP_TRY(&ctx->err, hipGraphInstantiate(&pl->graphExec, g, nullptr, nullptr, 0));
            pl->graphTableGen = ctx->tableGen;
        }
        // events bracket the whole replay (fill kernels included)
        HIP_TRY(&ctx->err, hipEventRecord(pl->ev0, s));
        HIP_TRY(&ctx->err, hipGraphLaunch(pl->graphExec, s));
        HIP_TRY(&ctx->err, hipEventRecord(pl->ev1, s));
    } else {
        if (plan_enqueue<R>(pl, dev_of(pl), record, count, rIds, rCnt, rL, rReIm, true)) return 1;
    }
    pl->ran = true;
    return 0;
}

int dispatch_prepare(sbr_plan* pl, bool record) {
    return pl->precision == SBR_FP64 ? plan_prepare<double>(pl, record) : plan_prepare<float>(pl, record);
}
int dispatch_launch(sbr_plan* pl, bool record, bool count, int32_t* a = nullptr, int32_t* b = nullptr,
                    double* c = nullptr, double* d = nullptr) {
    return pl->precision == SBR_FP64 ? plan_launch<double>(pl, record, count, a, b, c, d)
                                     : plan_launch<float>(pl, record, count, a, b, c, d);
}

// bvhBuildOptions defaults (include/scene/bvhBuilder.hpp:13-20) for zero fields
int parse_options(const sbr_bvh_options* o, BuildOptions& opt) {
    if (!o) return 0;
    opt.leafSize = o->leaf_size > 0 ? o->leaf_size : 4;
    opt.binCount = o->bin_count > 0 ? o->bin_count : 16;
    opt.maxParallelDepth = o->max_parallel_depth > 0 ? o->max_parallel_depth : 6;
    opt.minTrianglesForParallel = o->min_triangles_for_parallel > 0 ? o->min_triangles_for_parallel : 4096;
    opt.enableParallel = o->enable_parallel != 0;
    opt.sah = o->algorithm != 0;
    if (opt.leafSize > 16) return fail(nullptr, "leaf_size above 16 is not supported");
    return 0;
}

bool valid_prec(int p) { return p == SBR_FP32 || p == SBR_FP64 || p == SBR_MIXED; }

int create_plan(sbr_ctx* ctx, const sbr_job* jobs, int n, int precision, int method, bool record, sbr_plan** out) {
    if (!ctx) return fail(nullptr, "null context");
    if (!out) return fail(&ctx->err, "null plan output");
    if (n < 0 || (n > 0 && !jobs)) return fail(&ctx->err, "bad job list");
    if (!valid_prec(precision)) return fail(&ctx->err, "unknown precision");
    if (method != SBR_PER_ANGLE && method != SBR_BATCHED) return fail(&ctx->err, "unknown method");
    HIP_TRY(&ctx->err, hipSetDevice(ctx->device));
    auto pl = std::make_unique<sbr_plan>();
    pl->ctx = ctx;
    pl->precision = precision;
    pl->method = method;
    pl->nJobs = n;
    pl->tileSide = default_tile(method);
    for (int i = 0; i < n; ++i) pl->meshIds.push_back(jobs[i].mesh);
    std::sort(pl->meshIds.begin(), pl->meshIds.end());
    pl->meshIds.erase(std::unique(pl->meshIds.begin(), pl->meshIds.end()), pl->meshIds.end());
    pl->jobs.assign(jobs, jobs + n);
    HIP_TRY(&ctx->err, hipEventCreate(&pl->ev0));
    HIP_TRY(&ctx->err, hipEventCreate(&pl->ev1));
    if (dispatch_prepare(pl.get(), record)) {
        (void)hipEventDestroy(pl->ev0);
        (void)hipEventDestroy(pl->ev1);
        return 1;
    }
    ctx->plans.insert(pl.get());
    *out = pl.release();
    return 0;
}

template <typename R> void sigma_of(double re, double im, double* s, double* db) {
    R r = R(re), i = R(im);
    R sg = R(4.0) * R(M_PI) * (r * r + i * i);
    *s = double(sg);
    *db = double(R(10.0) * std::log10(std::fmax(sg, R(1e-10))));
}

// Job sums (2 doubles per job) -> caller outputs; any may be NULL.
void write_results(int precision, const double* sums, int n, double* reim, double* sigma, double* dbsm) {
    for (int i = 0; i < n; ++i) {
        const double re = sums[2 * i], im = sums[2 * i + 1];
        if (reim) {
            reim[2 * i] = re;
            reim[2 * i + 1] = im;
        }
        double s = 0, d = 0;
        if (precision == SBR_FP32) sigma_of<float>(re, im, &s, &d);
        else sigma_of<double>(re, im, &s, &d);
        if (sigma) sigma[i] = s;
        if (dbsm) dbsm[i] = d;
    }
}

// Batched sweep with every device buffer carved out of the fixed scratch
// array at base (no device allocation). The host plan is built first, so a
// sweep too large for the array is refused before anything is launched.
template <typename R> int sweep_in_scratch(sbr_plan* pl, void* base, std::vector<double>& sums) {
    sbr_ctx* ctx = pl->ctx;
    const bool mixed = pl->precision == SBR_MIXED;
    HostPlan<R> hp;
    if (plan_host<R>(pl, false, true, hp)) return 1;
    if ((sizeof(R) == 4 ? ctx->tableDirty32 : ctx->tableDirty64) ||
        !(sizeof(R) == 4 ? ctx->meshTable32.p : ctx->meshTable64.p) ||
        (mixed && (ctx->tableDirty64 || !ctx->meshTable64.p)))
        return fail(&ctx->err, "mesh tables are not on the device");
    const int n = pl->nJobs;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += (std::max<size_t>(bytes, 1) + 255) / 256 * 256;
        return o;
    };
    const size_t oJobs = take(sizeof(Job<R>) * n);
    const size_t oJobs64 = take(mixed ? sizeof(Job<double>) * n : 0);
    const size_t oChunk = take(sizeof(int) * hp.chunkJob.size());
    const size_t oPart = take(partial_bytes(sizeof(R) == 8, mixed ? 1 : 0) * size_t(hp.passes));
    const size_t oOut = take(sizeof(double) * 2 * n);
    const size_t oTile = take(sizeof(unsigned long long));
    const size_t oCnt = take(sizeof(unsigned long long) * 16);
    if (off > kScratchBytes)
        return fail(&ctx->err, "sweep needs " + std::to_string(off) + " bytes of device scratch, more than the " +
                                   std::to_string(kScratchBytes) +
                                   "-byte array of the allocation-free sweep: split it, or use sbr_sweep");
    unsigned char* b = static_cast<unsigned char*>(base);
    DevPlan d;
    d.jobs = b + oJobs;
    d.jobs64 = mixed ? b + oJobs64 : nullptr;
    d.chunkJob = reinterpret_cast<const int*>(b + oChunk);
    d.partials = b + oPart;
    d.jobOut = reinterpret_cast<double*>(b + oOut);
    d.tileCounter = reinterpret_cast<unsigned long long*>(b + oTile);
    d.counters = reinterpret_cast<unsigned long long*>(b + oCnt);
    hipStream_t s = ctx->stream;
    HIP_TRY(&ctx->err, hipMemcpyAsync(d.jobs, hp.jobs.data(), sizeof(Job<R>) * n, hipMemcpyHostToDevice, s));
    if (mixed)
        HIP_TRY(&ctx->err,
                hipMemcpyAsync(d.jobs64, hp.jobs64.data(), sizeof(Job<double>) * n, hipMemcpyHostToDevice, s));
    HIP_TRY(&ctx->err, hipMemcpyAsync(b + oChunk, hp.chunkJob.data(), sizeof(int) * hp.chunkJob.size(),
                                      hipMemcpyHostToDevice, s));
    // the tile counter and the work counters are cleared by fill kernels in
    // plan_enqueue; every pass partial is written by its tile
    if (plan_enqueue<R>(pl, d, false, false, nullptr, nullptr, nullptr, nullptr, false)) return 1;
    sums.assign(size_t(2) * n, 0.0);
    HIP_TRY(&ctx->err, hipMemcpyAsync(sums.data(), d.jobOut, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(&ctx->err, hipStreamSynchronize(s));
    return 0;
}

}  // namespace

extern "C" {

const char* sbr_last_error(const sbr_ctx* ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

int sbr_abi_version(void) { return SBR_ABI_VERSION; }

int sbr_ctx_create(int device, sbr_ctx** out) {
    if (!out) return fail(nullptr, "null output pointer");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return fail(nullptr, "no HIP device available");
    if (device < 0 || device >= n) return fail(nullptr, "device index out of range");
    auto ctx = std::make_unique<sbr_ctx>();
    ctx->device = device;
    HIP_TRY(nullptr, hipSetDevice(device));
    HIP_TRY(nullptr, hipDeviceGetAttribute(&ctx->cuCount, hipDeviceAttributeMultiprocessorCount, device));
    HIP_TRY(nullptr, hipStreamCreateWithFlags(&ctx->ownStream, hipStreamNonBlocking));
    ctx->stream = ctx->ownStream;
    // device address of the allocation-free sweep's scratch array, looked up
    // once per device with the device current
    if (hipError_t e = scratch_init(device); e != hipSuccess)
        ctx->scratchErr = std::string("scratch array address lookup failed: ") + hipGetErrorString(e);
    *out = ctx.release();
    return 0;
}

int sbr_ctx_destroy(sbr_ctx* ctx) {
    if (!ctx) return 0;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    // plans still alive keep only their host side; every call on them fails
    for (sbr_plan* pl : ctx->plans) release_plan_device(pl);
    ctx->plans.clear();
    ctx->meshes.clear();
    if (ctx->stream != ctx->ownStream) (void)hipStreamSynchronize(ctx->ownStream);
    (void)hipStreamDestroy(ctx->ownStream);
    delete ctx;
    return 0;
}

int sbr_ctx_set_stream(sbr_ctx* ctx, void* stream) {
    if (!ctx) return fail(nullptr, "null context");
    HIP_TRY(&ctx->err, hipSetDevice(ctx->device));
    hipStream_t next = stream ? static_cast<hipStream_t>(stream) : ctx->ownStream;
    if (next == ctx->stream) return 0;
    // work already issued finishes before later work starts on the new stream
    HIP_TRY(&ctx->err, hipStreamSynchronize(ctx->stream));
    ctx->stream = next;
    return 0;
}

int sbr_ctx_cu_count(const sbr_ctx* ctx, int* out) {
    if (!ctx || !out) return fail(nullptr, "null argument");
    *out = ctx->cuCount;
    return 0;
}

int sbr_mesh_create(sbr_ctx* ctx, const float* xyz, int n_verts, const uint32_t* idx, int n_tris, double scale,
                    const sbr_bvh_options* opts, int* mesh_id) {
    if (!ctx) return fail(nullptr, "null context");
    if (!xyz || !idx || !mesh_id || n_verts <= 0 || n_tris <= 0)
        return fail(&ctx->err, "mesh contains no valid triangles");
    // leaf references pack the first triangle in 27 bits
    if (n_tris >= (1 << 27)) return fail(&ctx->err, "meshes are limited to 2^27 - 1 triangles");
    for (int64_t i = 0; i < int64_t(n_tris) * 3; ++i)
        if (idx[i] >= uint32_t(n_verts)) return fail(&ctx->err, "mesh contains out-of-range indices");
    auto m = std::make_unique<MeshEntry>();
    m->xyz.assign(xyz, xyz + size_t(n_verts) * 3);
    m->idx.assign(idx, idx + size_t(n_tris) * 3);
    m->scale = scale;
    for (int64_t i = 0; i < int64_t(n_verts); ++i) {
        double x = double(xyz[3 * i]) * scale, y = double(xyz[3 * i + 1]) * scale, z = double(xyz[3 * i + 2]) * scale;
        m->radius = std::max(m->radius, std::sqrt(x * x + y * y + z * z));
    }
    if (parse_options(opts, m->opt)) return fail(&ctx->err, g_err);
    HIP_TRY(&ctx->err, hipSetDevice(ctx->device));
    ctx->meshes.push_back(std::move(m));
    *mesh_id = int(ctx->meshes.size()) - 1;
    // build the FP32 tree eagerly: it is the one the default paths use
    if (ensure_built<float>(ctx, *ctx->meshes.back())) {
        ctx->meshes.pop_back();
        return 1;
    }
    return 0;
}

int sbr_mesh_destroy(sbr_ctx* ctx, int mesh_id) {
    if (!ctx) return fail(nullptr, "null context");
    if (mesh_id < 0 || mesh_id >= int(ctx->meshes.size()) || !ctx->meshes[mesh_id])
        return fail(&ctx->err, "bad or destroyed mesh id");
    HIP_TRY(&ctx->err, hipSetDevice(ctx->device));
    HIP_TRY(&ctx->err, hipStreamSynchronize(ctx->stream));
    ctx->meshes[mesh_id].reset();
    ctx->tableDirty32 = ctx->tableDirty64 = true;
    // tables keep their size: the destroyed entry is cleared in place
    if (ctx->meshTable32.p && upload_table<float>(ctx)) return 1;
    if (ctx->meshTable64.p && upload_table<double>(ctx)) return 1;
    return 0;
}

struct sbr_bvh {
    int precision = 0;
    Bvh<float> f;
    Bvh<double> d;
};

int sbr_bvh_build(const float* xyz, int n_verts, const uint32_t* idx, int n_tris, double scale,
                  const sbr_bvh_options* opts, int precision, sbr_bvh** out) {
    if (!xyz || !idx || !out || n_verts <= 0 || n_tris <= 0) return fail(nullptr, "mesh contains no valid triangles");
    if (precision != SBR_FP32 && precision != SBR_FP64) return fail(nullptr, "precision must be SBR_FP32 or SBR_FP64");
    for (int64_t i = 0; i < int64_t(n_tris) * 3; ++i)
        if (idx[i] >= uint32_t(n_verts)) return fail(nullptr, "mesh contains out-of-range indices");
    BuildOptions opt;
    if (parse_options(opts, opt)) return 1;
    auto b = std::make_unique<sbr_bvh>();
    b->precision = precision;
    auto build = [&](auto zero, auto& dst) {
        using R = decltype(zero);
        std::vector<R> soup(size_t(n_tris) * 9);
        const R s = static_cast<R>(scale);
        for (size_t f = 0; f < size_t(n_tris); ++f)
            for (int c = 0; c < 3; ++c)
                for (int k = 0; k < 3; ++k)
                    soup[9 * f + 3 * c + k] = static_cast<R>(xyz[3 * size_t(idx[3 * f + c]) + k]) * s;
        dst = build_bvh<R>(soup, opt);
    };
    if (precision == SBR_FP64) build(0.0, b->d);
    else build(0.0f, b->f);
    *out = b.release();
    return 0;
}

int sbr_bvh_info(const sbr_bvh* b, int32_t* out4) {
    if (!b || !out4) return fail(nullptr, "null argument");
    auto fill = [&](const auto& t) {
        out4[0] = int32_t(t.perm.size());
        out4[1] = int32_t(t.nodes.size());
        out4[2] = t.root;
        out4[3] = t.depth;
    };
    if (b->precision == SBR_FP64) fill(b->d);
    else fill(b->f);
    return 0;
}

int sbr_bvh_export(const sbr_bvh* b, double* tris9, int32_t* perm, double* bounds6, int32_t* links4) {
    if (!b) return fail(nullptr, "null argument");
    auto ex = [&](const auto& t) {
        if (tris9)
            for (size_t i = 0; i < t.tri9.size(); ++i) tris9[i] = double(t.tri9[i]);
        if (perm) std::copy(t.perm.begin(), t.perm.end(), perm);
        for (size_t i = 0; i < t.nodes.size(); ++i) {
            const auto& n = t.nodes[i];
            if (bounds6)
                for (int k = 0; k < 3; ++k) {
                    bounds6[6 * i + k] = n.lo[k];
                    bounds6[6 * i + 3 + k] = n.hi[k];
                }
            if (links4) {
                links4[4 * i] = n.left;
                links4[4 * i + 1] = n.right;
                links4[4 * i + 2] = n.firstTri;
                links4[4 * i + 3] = n.triCount;
            }
        }
    };
    if (b->precision == SBR_FP64) ex(b->d);
    else ex(b->f);
    return 0;
}

int sbr_bvh_destroy(sbr_bvh* b) {
    delete b;
    return 0;
}

static MeshEntry* get_mesh(sbr_ctx* ctx, int id, int precision) {
    if (id < 0 || id >= int(ctx->meshes.size()) || !ctx->meshes[id]) return nullptr;
    MeshEntry* m = ctx->meshes[id].get();
    if (precision == SBR_FP64 ? ensure_built<double>(ctx, *m) : ensure_built<float>(ctx, *m)) return nullptr;
    return m;
}

int sbr_mesh_info_get(sbr_ctx* ctx, int mesh_id, int precision, sbr_mesh_info* out) {
    if (!ctx || !out) return fail(nullptr, "null argument");
    MeshEntry* m = get_mesh(ctx, mesh_id, precision);
    if (!m) return fail(&ctx->err, ctx->err.empty() ? "bad mesh id" : ctx->err);
    auto fill = [&](auto& pm) {
        out->n_triangles = int(pm.bvh.perm.size());
        out->n_nodes = int(pm.bvh.nodes.size());
        out->n_gpu_nodes = pm.gpu.nInternal;
        out->depth = pm.bvh.depth;
        for (int k = 0; k < 3; ++k) {
            out->bounds[k] = pm.gpu.rootLo[k];
            out->bounds[3 + k] = pm.gpu.rootHi[k];
        }
        out->build_ms = pm.buildMs;
    };
    if (precision == SBR_FP64) fill(m->f64);
    else fill(m->f32);
    return 0;
}

int sbr_mesh_permutation(sbr_ctx* ctx, int mesh_id, int precision, int32_t* perm) {
    if (!ctx || !perm) return fail(nullptr, "null argument");
    MeshEntry* m = get_mesh(ctx, mesh_id, precision);
    if (!m) return fail(&ctx->err, "bad mesh id");
    const std::vector<int>& p = precision == SBR_FP64 ? m->f64.bvh.perm : m->f32.bvh.perm;
    std::copy(p.begin(), p.end(), perm);
    return 0;
}

int sbr_mesh_export_bvh(sbr_ctx* ctx, int mesh_id, int precision, double* bounds6, int32_t* links4, int* root) {
    if (!ctx || !bounds6 || !links4 || !root) return fail(nullptr, "null argument");
    MeshEntry* m = get_mesh(ctx, mesh_id, precision);
    if (!m) return fail(&ctx->err, "bad mesh id");
    auto ex = [&](auto& pm) {
        for (size_t i = 0; i < pm.bvh.nodes.size(); ++i) {
            const auto& n = pm.bvh.nodes[i];
            for (int k = 0; k < 3; ++k) {
                bounds6[6 * i + k] = n.lo[k];
                bounds6[6 * i + 3 + k] = n.hi[k];
            }
            links4[4 * i] = n.left;
            links4[4 * i + 1] = n.right;
            links4[4 * i + 2] = n.firstTri;
            links4[4 * i + 3] = n.triCount;
        }
        *root = pm.bvh.root;
    };
    if (precision == SBR_FP64) ex(m->f64);
    else ex(m->f32);
    return 0;
}

struct sbr_model {
    MeshData mesh;
};

int sbr_model_load(const char* path, sbr_model** out) {
    if (!path || !out) return fail(nullptr, "null argument");
    auto m = std::make_unique<sbr_model>();
    std::string err;
    if (!load_model(path, m->mesh, err)) return fail(nullptr, err.empty() ? "failed to load model" : err);
    *out = m.release();
    return 0;
}

int sbr_model_info(const sbr_model* m, int64_t* out2) {
    if (!m || !out2) return fail(nullptr, "null argument");
    out2[0] = int64_t(m->mesh.positions.size() / 3);
    out2[1] = int64_t(m->mesh.indices.size() / 3);
    return 0;
}

int sbr_model_copy(const sbr_model* m, float* xyz, uint32_t* idx) {
    if (!m) return fail(nullptr, "null argument");
    if (xyz) std::copy(m->mesh.positions.begin(), m->mesh.positions.end(), xyz);
    if (idx) std::copy(m->mesh.indices.begin(), m->mesh.indices.begin() + (m->mesh.indices.size() / 3) * 3, idx);
    return 0;
}

int sbr_model_destroy(sbr_model* m) {
    delete m;
    return 0;
}

int sbr_obj_parse_number(const char* token, double* out) {
    if (!token || !out) return fail(nullptr, "null argument");
    return obj_parse_number(token, token + std::strlen(token), out) ? 0 : fail(nullptr, "not a number");
}

int sbr_angle_frame(const sbr_sweep_params* p, int angle, int precision, double* out12, double* out_deg2) {
    if (!p || !out12) return fail(nullptr, "null argument");
    if (p->theta_samples <= 0 || p->phi_samples <= 0 || angle < 0 || angle >= p->theta_samples * p->phi_samples)
        return fail(nullptr, "angle index out of range");
    auto run = [&](auto zero) {
        using R = decltype(zero);
        R th, ph, llc[3], uv[3], vv[3], dir[3];
        sweep_degrees<R>(p, angle, th, ph);
        make_frame<R>(th, ph, R(p->grid_size), llc, uv, vv, dir);
        for (int k = 0; k < 3; ++k) {
            out12[k] = llc[k];
            out12[3 + k] = uv[k];
            out12[6 + k] = vv[k];
            out12[9 + k] = dir[k];
        }
        if (out_deg2) {
            out_deg2[0] = th;
            out_deg2[1] = ph;
        }
    };
    if (precision == SBR_FP64) run(0.0);
    else run(0.0f);
    return 0;
}

int sbr_sweep_jobs(const sbr_sweep_params* p, int mesh_id, int precision, sbr_job* jobs, int n_jobs) {
    if (!p || !jobs) return fail(nullptr, "null argument");
    if (p->theta_samples <= 0 || p->phi_samples <= 0) return fail(nullptr, "sample counts must be positive");
    if (n_jobs != p->theta_samples * p->phi_samples) return fail(nullptr, "n_jobs must equal theta*phi samples");
    for (int i = 0; i < n_jobs; ++i) {
        sbr_job& j = jobs[i];
        j.mesh = mesh_id;
        j.max_bounces = p->max_bounces;
        j.n_x = p->n_x;
        j.n_y = p->n_y;
        j.freq = p->freq;
        j.grid_size = p->grid_size;
        j.reflection_const = p->reflection_const;
        if (precision == SBR_FP64) {
            double th, ph;
            sweep_degrees<double>(p, i, th, ph);
            j.theta_deg = th;
            j.phi_deg = ph;
        } else {
            float th, ph;
            sweep_degrees<float>(p, i, th, ph);
            j.theta_deg = th;
            j.phi_deg = ph;
        }
    }
    return 0;
}

int sbr_plan_create(sbr_ctx* ctx, const sbr_job* jobs, int n_jobs, int precision, int method, sbr_plan** out) {
    return create_plan(ctx, jobs, n_jobs, precision, method, false, out);
}

int sbr_plan_set_tuning(sbr_plan* pl, const sbr_tuning* t) {
    if (!pl || !t) return fail(nullptr, "null argument");
    if (!pl->ctx) return fail(nullptr, "the plan's context was destroyed");
    if (t->lds_stack < 0 || t->wg_per_cu < 0) return fail(&pl->ctx->err, "negative tuning value");
    if (t->ray_pool < 0 || t->ray_pool > 2) return fail(&pl->ctx->err, "ray_pool must be 0, 1 or 2");
    if (t->traversal < 0 || t->traversal > 2) return fail(&pl->ctx->err, "traversal must be 0, 1 or 2");
    if (t->top_levels < 0 || t->top_levels > kTopLevels)
        return fail(&pl->ctx->err, "top_levels must be 0 .. " + std::to_string(kTopLevels));
    if (t->node_width != 0 && t->node_width != 2 && t->node_width != 4 && t->node_width != 8)
        return fail(&pl->ctx->err, "node_width must be 0, 2, 4 or 8");
    if (t->tile_rays != 0 && t->tile_rays != 16 && t->tile_rays != 32 && t->tile_rays != 64)
        return fail(&pl->ctx->err, "tile_rays must be 16, 32 or 64");
    if (t->cull < 0 || t->cull > 2) return fail(&pl->ctx->err, "cull must be 0, 1 or 2");
    drop_graph(pl);
    const sbr_tuning old = pl->tune;
    const int oldSide = pl->tileSide;
    pl->tune = *t;
    int ts = t->tile_rays ? t->tile_rays : default_tile(pl->method);
    // a new tile side changes the job tiling; a new node width (mixed mode)
    // may build that width's tree and must fit the traversal stack
    const bool widthChanged =
        pl->precision == SBR_MIXED && (old.node_width ? old.node_width : kWide) != (t->node_width ? t->node_width : kWide);
    if (ts != pl->tileSide || widthChanged) {
        pl->tileSide = ts;
        if (dispatch_prepare(pl, false)) {
            const std::string msg = pl->ctx->err;
            pl->tune = old;
            pl->tileSide = oldSide;
            (void)dispatch_prepare(pl, false);
            return fail(&pl->ctx->err, msg);
        }
    }
    return 0;
}

int sbr_plan_run(sbr_plan* pl) {
    if (!pl) return fail(nullptr, "null plan");
    if (!pl->ctx) return fail(nullptr, "the plan's context was destroyed");
    HIP_TRY(&pl->ctx->err, hipSetDevice(pl->ctx->device));
    return dispatch_launch(pl, false, false);
}

int sbr_plan_wait(sbr_plan* pl) {
    if (!pl) return fail(nullptr, "null plan");
    if (!pl->ctx) return fail(nullptr, "the plan's context was destroyed");
    HIP_TRY(&pl->ctx->err, hipStreamSynchronize(pl->ctx->stream));
    return 0;
}

int sbr_plan_last_ms(sbr_plan* pl, double* ms) {
    if (!pl || !ms) return fail(nullptr, "null argument");
    if (!pl->ctx) return fail(nullptr, "the plan's context was destroyed");
    if (!pl->ran) return fail(&pl->ctx->err, "plan has not run");
    float f = 0;
    HIP_TRY(&pl->ctx->err, hipEventSynchronize(pl->ev1));
    HIP_TRY(&pl->ctx->err, hipEventElapsedTime(&f, pl->ev0, pl->ev1));
    *ms = f;
    return 0;
}

int sbr_plan_results(sbr_plan* pl, double* reim, double* sigma, double* dbsm) {
    if (!pl) return fail(nullptr, "null plan");
    if (!pl->ctx) return fail(nullptr, "the plan's context was destroyed");
    if (!pl->ran) return fail(&pl->ctx->err, "plan has not run");
    sbr_ctx* ctx = pl->ctx;
    pl->out.assign(size_t(2) * pl->nJobs, 0.0);
    HIP_TRY(&ctx->err, hipMemcpyAsync(pl->out.data(), pl->dJobOut.p, sizeof(double) * 2 * pl->nJobs,
                                      hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(&ctx->err, hipStreamSynchronize(ctx->stream));
    write_results(pl->precision, pl->out.data(), pl->nJobs, reim, sigma, dbsm);
    return 0;
}

static int plan_count(sbr_plan* pl, sbr_counters* out, bool traced) {
    if (!pl || !out) return fail(nullptr, "null argument");
    if (!pl->ctx) return fail(nullptr, "the plan's context was destroyed");
    sbr_ctx* ctx = pl->ctx;
    HIP_TRY(&ctx->err, hipSetDevice(ctx->device));
    pl->countTraced = traced;
    const int rc = dispatch_launch(pl, false, true);
    pl->countTraced = false;
    if (rc) return 1;
    unsigned long long c[16] = {0};
    HIP_TRY(&ctx->err, hipMemcpyAsync(c, pl->dCounters.p, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(&ctx->err, hipStreamSynchronize(ctx->stream));
    // rays, visits, tests, bounces, queries (each query counted once, FP64
    // answers of undecided queries included), rays with an undecided query,
    // undecided queries
    out->rays = c[0];
    out->node_visits = c[1];
    out->tri_tests = c[2];
    out->ray_bounces = c[3];
    out->queries = c[4];
    out->retraced = c[5];
    out->fp64_queries = c[6];
    return 0;
}
int sbr_plan_count(sbr_plan* pl, sbr_counters* out) { return plan_count(pl, out, false); }
int sbr_plan_count_traced(sbr_plan* pl, sbr_counters* out) { return plan_count(pl, out, true); }

int sbr_plan_autotune(sbr_plan* pl, int sample_jobs, int repeats, sbr_tuning* best, double* best_ms, int* n_tried) {
    if (!pl) return fail(nullptr, "null plan");
    if (!pl->ctx) return fail(nullptr, "the plan's context was destroyed");
    sbr_ctx* ctx = pl->ctx;
    if (repeats < 1) return fail(&ctx->err, "repeats must be >= 1");
    HIP_TRY(&ctx->err, hipSetDevice(ctx->device));
    // probe plan: the plan itself, or an evenly strided job subset
    sbr_plan* probe = pl;
    sbr_plan* owned = nullptr;
    if (sample_jobs > 0 && sample_jobs < pl->nJobs) {
        std::vector<sbr_job> sub;
        sub.reserve(sample_jobs);
        for (int i = 0; i < sample_jobs; ++i) sub.push_back(pl->jobs[int64_t(i) * pl->nJobs / sample_jobs]);
        if (create_plan(ctx, sub.data(), int(sub.size()), pl->precision, pl->method, false, &owned)) return 1;
        probe = owned;
    }
    // stage 1: launch shape at the default traversal loop shape
    std::vector<sbr_tuning> cands;
    for (int pool : {1, 2})
        for (int sd : {8, 16, 32}) {
            if (pl->method == SBR_BATCHED) {
                for (int ts : {16, 32, 64})
                    for (int wg : {0, 2}) cands.push_back(sbr_tuning{wg, sd, ts, 0, pool, 0, 0, 0, 0});
            } else {
                for (int ts : {16, 32})
                    for (int g : {0, 1}) cands.push_back(sbr_tuning{0, sd, ts, g, pool, 0, 0, 0, 0});
            }
        }
    double bestT = 1e300;
    sbr_tuning bestC = cands[0];
    int rc = 0, tried = 0;
    auto trial = [&](const sbr_tuning& c) -> int {
        if (int e = sbr_plan_set_tuning(probe, &c)) return e;
        // one untimed run (graph capture, caches), then the timed ones
        if (int e = sbr_plan_run(probe)) return e;
        if (int e = sbr_plan_wait(probe)) return e;
        double t = 1e300;
        for (int r = 0; r < repeats; ++r) {
            auto t0 = std::chrono::steady_clock::now();
            if (int e = sbr_plan_run(probe)) return e;
            if (int e = sbr_plan_wait(probe)) return e;
            t = std::min(t, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
        }
        ++tried;
        if (t < bestT) {
            bestT = t;
            bestC = c;
        }
        return 0;
    };
    for (const sbr_tuning& c : cands)
        if ((rc = trial(c))) break;
    // stage 2: the other traversal loop shape at the best launch shape
    if (!rc) {
        sbr_tuning c = bestC;
        c.traversal = default_traversal() == 2 ? 1 : 2;
        rc = trial(c);
    }
    // stage 3 (mixed, one ray per lane): the tree's top levels in LDS, at
    // the best shape and with the smallest LDS stack (the staged records
    // and a deeper stack together leave room for fewer workgroups per CU)
    if (!rc && pl->precision == SBR_MIXED && bestC.ray_pool != 1) {
        const sbr_tuning base = bestC;
        std::vector<int> sds{base.lds_stack > 0 ? base.lds_stack : 8};
        if (sds[0] != 8) sds.push_back(8);
        for (int sd : sds)
            for (int k : {3, 4}) {
                sbr_tuning c = base;
                c.lds_stack = sd;
                c.top_levels = k;
                if (!rc) rc = trial(c);
            }
    }
    // stage 4 (mixed, one ray per lane): the binary and the eight-wide tree
    if (!rc && pl->precision == SBR_MIXED && bestC.ray_pool != 1) {
        const sbr_tuning base = bestC;
        for (int w : {2, 8}) {
            sbr_tuning c = base;
            c.node_width = w;
            if (!rc) rc = trial(c);
        }
    }
    // stage 5 (batched): the per-tile box tests instead of the coverage masks
    if (!rc && pl->method == SBR_BATCHED) {
        sbr_tuning c = bestC;
        c.cull = 2;
        rc = trial(c);
    }
    if (owned) sbr_plan_destroy(owned);
    if (rc) return rc;
    if (sbr_plan_set_tuning(pl, &bestC)) return 1;
    if (best) *best = bestC;
    if (best_ms) *best_ms = bestT;
    if (n_tried) *n_tried = tried;
    return 0;
}

int sbr_plan_destroy(sbr_plan* pl) {
    if (!pl) return 0;
    if (pl->ctx) {
        (void)hipSetDevice(pl->ctx->device);
        (void)hipStreamSynchronize(pl->ctx->stream);
        pl->ctx->plans.erase(pl);
        release_plan_device(pl);
    }
    delete pl;
    return 0;
}

int sbr_sweep(sbr_ctx* ctx, int mesh_id, const sbr_sweep_params* p, int precision, int method, double* reim,
              double* sigma, double* dbsm) {
    if (!ctx || !p) return fail(nullptr, "null argument");
    const int n = p->theta_samples * p->phi_samples;
    if (p->theta_samples <= 0 || p->phi_samples <= 0) return fail(&ctx->err, "sample counts must be positive");
    std::vector<sbr_job> jobs(n);
    if (sbr_sweep_jobs(p, mesh_id, precision, jobs.data(), n)) return fail(&ctx->err, g_err);
    sbr_plan* pl = nullptr;
    if (sbr_plan_create(ctx, jobs.data(), n, precision, method, &pl)) return 1;
    int rc = sbr_plan_run(pl);
    if (!rc) rc = sbr_plan_wait(pl);
    if (!rc) rc = sbr_plan_results(pl, reim, sigma, dbsm);
    sbr_plan_destroy(pl);
    return rc;
}

int sbr_sweep_noalloc(sbr_ctx* ctx, int mesh_id, const sbr_sweep_params* p, int precision, double* reim,
                      double* sigma, double* dbsm) {
    if (!ctx || !p) return fail(nullptr, "null argument");
    if (!valid_prec(precision)) return fail(&ctx->err, "unknown precision");
    if (p->theta_samples <= 0 || p->phi_samples <= 0) return fail(&ctx->err, "sample counts must be positive");
    if (mesh_id < 0 || mesh_id >= int(ctx->meshes.size()) || !ctx->meshes[mesh_id])
        return fail(&ctx->err, "bad or destroyed mesh id");
    HIP_TRY(&ctx->err, hipSetDevice(ctx->device));
    void* base = scratch_ptr(ctx->device);
    if (!base)
        return fail(&ctx->err, ctx->scratchErr.empty() ? "no scratch array address for this device" : ctx->scratchErr);
    const int n = p->theta_samples * p->phi_samples;
    // a host-only plan: its device buffers stay empty, the run uses the array
    sbr_plan pl;
    pl.ctx = ctx;
    pl.precision = precision;
    pl.method = SBR_BATCHED;
    pl.nJobs = n;
    pl.tileSide = default_tile(SBR_BATCHED);
    pl.jobs.resize(n);
    if (sbr_sweep_jobs(p, mesh_id, precision, pl.jobs.data(), n)) return fail(&ctx->err, g_err);
    pl.meshIds.push_back(mesh_id);
    std::vector<double> sums;
    const int rc = precision == SBR_FP64 ? sweep_in_scratch<double>(&pl, base, sums)
                                         : sweep_in_scratch<float>(&pl, base, sums);
    if (rc) return rc;
    write_results(precision, sums.data(), n, reim, sigma, dbsm);
    return 0;
}

int sbr_trace_rays(sbr_ctx* ctx, int mesh_id, const sbr_sweep_params* p, int angle, int precision, int32_t* hit_ids,
                   int32_t* bounces, double* path_len, double* reim_ray, int8_t* retraced) {
    if (!ctx || !p) return fail(nullptr, "null argument");
    const int n = p->theta_samples * p->phi_samples;
    if (angle < 0 || angle >= n) return fail(&ctx->err, "angle index out of range");
    std::vector<sbr_job> all(n);
    if (sbr_sweep_jobs(p, mesh_id, precision, all.data(), n)) return fail(&ctx->err, g_err);
    sbr_plan* pl = nullptr;
    if (create_plan(ctx, &all[angle], 1, precision, SBR_PER_ANGLE, true, &pl)) return 1;
    const size_t nr = size_t(p->n_x) * p->n_y;
    const size_t mb = std::max(p->max_bounces, 1);
    DevBuf dIds, dCnt, dL, dRe;
    int rc = 0;
    auto run = [&]() -> int {
        HIP_TRY(&ctx->err, dIds.ensure(nr * mb * sizeof(int32_t)));
        HIP_TRY(&ctx->err, dCnt.ensure(nr * sizeof(int32_t)));
        HIP_TRY(&ctx->err, dL.ensure(nr * sizeof(double)));
        HIP_TRY(&ctx->err, dRe.ensure(nr * 2 * sizeof(double)));
        HIP_TRY(&ctx->err, hipMemsetAsync(dIds.p, 0xff, nr * mb * sizeof(int32_t), ctx->stream));
        HIP_TRY(&ctx->err, hipMemsetAsync(dCnt.p, 0, nr * sizeof(int32_t), ctx->stream));
        HIP_TRY(&ctx->err, hipMemsetAsync(dL.p, 0, nr * sizeof(double), ctx->stream));
        HIP_TRY(&ctx->err, hipMemsetAsync(dRe.p, 0, nr * 2 * sizeof(double), ctx->stream));
        HIP_TRY(&ctx->err, pl->dRecFlag.ensure(nr));
        HIP_TRY(&ctx->err, hipMemsetAsync(pl->dRecFlag.p, 0, nr, ctx->stream));
        if (dispatch_launch(pl, true, false, static_cast<int32_t*>(dIds.p), static_cast<int32_t*>(dCnt.p),
                            static_cast<double*>(dL.p), static_cast<double*>(dRe.p)))
            return 1;
        HIP_TRY(&ctx->err, hipStreamSynchronize(ctx->stream));
        if (hit_ids && p->max_bounces > 0)
            HIP_TRY(&ctx->err, hipMemcpy(hit_ids, dIds.p, nr * p->max_bounces * sizeof(int32_t), hipMemcpyDeviceToHost));
        if (bounces) HIP_TRY(&ctx->err, hipMemcpy(bounces, dCnt.p, nr * sizeof(int32_t), hipMemcpyDeviceToHost));
        if (path_len) HIP_TRY(&ctx->err, hipMemcpy(path_len, dL.p, nr * sizeof(double), hipMemcpyDeviceToHost));
        if (reim_ray) HIP_TRY(&ctx->err, hipMemcpy(reim_ray, dRe.p, nr * 2 * sizeof(double), hipMemcpyDeviceToHost));
        if (retraced) HIP_TRY(&ctx->err, hipMemcpy(retraced, pl->dRecFlag.p, nr, hipMemcpyDeviceToHost));
        return 0;
    };
    rc = run();
    sbr_plan_destroy(pl);
    return rc;
}

int sbr_adjoint(sbr_ctx* ctx, int mesh_id, const sbr_job* jobs, int n_jobs, const double* dloss_dsigma,
                double* sigma, double* dloss_dxyz, double* dsigma_dfreq) {
    if (!ctx) return fail(nullptr, "null context");
    if (mesh_id < 0 || mesh_id >= int(ctx->meshes.size()) || !ctx->meshes[mesh_id])
        return fail(&ctx->err, "bad or destroyed mesh id");
    if (n_jobs < 0 || (n_jobs > 0 && !jobs)) return fail(&ctx->err, "bad job list");
    for (int i = 0; i < n_jobs; ++i)
        if (jobs[i].mesh != mesh_id) return fail(&ctx->err, "job " + std::to_string(i) + " is on another mesh");
    HIP_TRY(&ctx->err, hipSetDevice(ctx->device));
    MeshEntry& m = *ctx->meshes[mesh_id];
    const size_t nv = m.xyz.size() / 3;
    if (!m.adjVerts.p) {
        std::vector<double> v64(m.xyz.size());
        for (size_t k = 0; k < v64.size(); ++k) v64[k] = double(m.xyz[k]) * m.scale;
        HIP_TRY(&ctx->err, m.adjVerts.ensure(v64.size() * sizeof(double)));
        HIP_TRY(&ctx->err, hipMemcpy(m.adjVerts.p, v64.data(), v64.size() * sizeof(double), hipMemcpyHostToDevice));
        HIP_TRY(&ctx->err, m.adjIdx.ensure(m.idx.size() * sizeof(uint32_t)));
        HIP_TRY(&ctx->err,
                hipMemcpy(m.adjIdx.p, m.idx.data(), m.idx.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    DevBuf dV, dF, dIds, dCnt, dL, dRe;
    HIP_TRY(&ctx->err, dV.ensure(std::max<size_t>(nv * 3, 1) * sizeof(double)));
    HIP_TRY(&ctx->err, dF.ensure(std::max(n_jobs, 1) * sizeof(double)));
    HIP_TRY(&ctx->err, hipMemsetAsync(dV.p, 0, nv * 3 * sizeof(double), ctx->stream));
    HIP_TRY(&ctx->err, hipMemsetAsync(dF.p, 0, std::max(n_jobs, 1) * sizeof(double), ctx->stream));
    if (n_jobs > 0) {
        // forward of every direction in one record launch: mixed precision
        // (FP64 outcome per ray), hit faces recorded, job i's records after
        // job i-1's; then one adjoint launch over every job's rays, which
        // reads the jobs' sums from the forward's output (no host wait)
        sbr_plan* pl = nullptr;
        if (create_plan(ctx, jobs, n_jobs, SBR_MIXED, SBR_PER_ANGLE, true, &pl)) return 1;
        std::unique_ptr<sbr_plan, int (*)(sbr_plan*)> guard(pl, sbr_plan_destroy);
        std::vector<AdjJob> aj(n_jobs);
        size_t nrTot = 0, idTot = 0;
        int blocks = 0;
        for (int i = 0; i < n_jobs; ++i) {
            const sbr_job& jb = jobs[i];
            const size_t nr = size_t(jb.n_x) * jb.n_y;
            if (nr > size_t(INT32_MAX) - 255) return fail(&ctx->err, "job " + std::to_string(i) + " has too many rays");
            AdjJob& g = aj[i];
            g = AdjJob{};
            make_frame<double>(jb.theta_deg, jb.phi_deg, jb.grid_size, g.D.llc, g.D.uv, g.D.vv, g.D.dir);
            make_phys<double>(jb, g.D);
            g.D.nx = jb.n_x;
            g.D.ny = jb.n_y;
            g.D.maxB = jb.max_bounces;
            g.rayOff = int64_t(nrTot);
            g.idOff = int64_t(idTot);
            g.w = dloss_dsigma ? dloss_dsigma[i] : 0.0;
            g.freq = jb.freq;
            g.nRays = int(nr);
            g.blockBase = blocks;
            const int64_t nb = int64_t(nr + 255) / 256;
            if (int64_t(blocks) + nb > INT32_MAX) return fail(&ctx->err, "adjoint launch too large");
            blocks += int(nb);
            nrTot += nr;
            idTot += nr * size_t(jb.max_bounces);
        }
        DevBuf dAdj;
        HIP_TRY(&ctx->err, dAdj.ensure(sizeof(AdjJob) * aj.size()));
        HIP_TRY(&ctx->err, hipMemcpy(dAdj.p, aj.data(), sizeof(AdjJob) * aj.size(), hipMemcpyHostToDevice));
        HIP_TRY(&ctx->err, dIds.ensure(std::max<size_t>(idTot, 1) * sizeof(int32_t)));
        HIP_TRY(&ctx->err, dCnt.ensure(std::max<size_t>(nrTot, 1) * sizeof(int32_t)));
        HIP_TRY(&ctx->err, dL.ensure(std::max<size_t>(nrTot, 1) * sizeof(double)));
        HIP_TRY(&ctx->err, dRe.ensure(std::max<size_t>(nrTot, 1) * 2 * sizeof(double)));
        HIP_TRY(&ctx->err, pl->dRecFlag.ensure(std::max<size_t>(nrTot, 1)));
        HIP_TRY(&ctx->err, hipMemsetAsync(dIds.p, 0xff, idTot * sizeof(int32_t), ctx->stream));
        HIP_TRY(&ctx->err, hipMemsetAsync(dCnt.p, 0, nrTot * sizeof(int32_t), ctx->stream));
        if (dispatch_launch(pl, true, false, static_cast<int32_t*>(dIds.p), static_cast<int32_t*>(dCnt.p),
                            static_cast<double*>(dL.p), static_cast<double*>(dRe.p)))
            return 1;
        AdjArgs a{};
        a.jobs = static_cast<const AdjJob*>(dAdj.p);
        a.nJobs = n_jobs;
        a.verts = static_cast<const double*>(m.adjVerts.p);
        a.idx = static_cast<const unsigned*>(m.adjIdx.p);
        a.ids = static_cast<const int32_t*>(dIds.p);
        a.cnt = static_cast<const int32_t*>(dCnt.p);
        a.sums = static_cast<const double*>(pl->dJobOut.p);
        a.scale = m.scale;
        a.dvert = static_cast<double*>(dV.p);
        a.dfreq = static_cast<double*>(dF.p);
        HIP_TRY(&ctx->err, launch_adjoint(a, blocks, ctx->stream));
        std::vector<double> sg(n_jobs);
        if (sbr_plan_results(pl, nullptr, sg.data(), nullptr)) return 1;  // waits for the stream
        if (sigma)
            for (int i = 0; i < n_jobs; ++i) sigma[i] = sg[i];
    }
    if (dloss_dxyz && nv)
        HIP_TRY(&ctx->err, hipMemcpy(dloss_dxyz, dV.p, nv * 3 * sizeof(double), hipMemcpyDeviceToHost));
    if (dsigma_dfreq && n_jobs)
        HIP_TRY(&ctx->err, hipMemcpy(dsigma_dfreq, dF.p, n_jobs * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

}  // extern "C"
