// SPDX-License-Identifier: MIT
//
// TEST INFRASTRUCTURE — launcher for the reference's own HIP ray kernels.
//
// oracle/Makefile compiles the reference's src/cuda/rtKernels.cu,
// src/cuda/poKernels.cu and src/scene/bvhBuilder.cpp from the reference tree
// (with the reference's HIP flags, CMakeLists.txt:124-133) and links them
// with this file into oracle/_ref/libref_gpu_f{32,64}.so. Nothing of the
// reference is copied: this file only restates the host steps around the
// kernels so that tests can run them on the same inputs as libsbr:
//   * triangle soup Real(pos) * Real(scale)          src/app/rtRcsEntry.cu:118-148
//   * buildBvh + device upload                       src/app/rtRcsEntry.cu:158-196
//   * ray / hit buffers                              src/app/deviceBuffers.cu:12-29
//   * physics constants                              src/app/sweep.cu:48-53
//   * per-angle frame, memset, RT kernel, PO kernel,
//     accumulator copy, launch shapes                src/app/sweep.cu:94-95,336-398
// Each angle is run synchronously (the reference's two-stream overlap only
// hides the 16-byte accumulator copy); kernel times come from hipEvents
// around each launch, as the reference's "GPU Time (kernels only)" does.
// Only tests load this library.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "cuda/gpuConfig.cuh"
#include "cuda/poKernels.cuh"
#include "cuda/rtKernels.cuh"
#include "scene/bvhBuilder.hpp"

namespace {

struct RefGpu {
    bvhGpuData bvh;
    int cap = 0;
    vec3* hitNormal = nullptr;
    vec3* lastDir = nullptr;
    Real* hitDist = nullptr;
    int* hitCount = nullptr;
    cuRealComplex* accum = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
};

bool ok(hipError_t e) { return e == hipSuccess; }

bool ensure_rays(RefGpu* g, int n) {
    if (n <= g->cap) return true;
    (void)hipFree(g->hitNormal);
    (void)hipFree(g->lastDir);
    (void)hipFree(g->hitDist);
    (void)hipFree(g->hitCount);
    g->cap = 0;
    if (!ok(hipMalloc(&g->hitNormal, size_t(n) * sizeof(vec3))) ||
        !ok(hipMalloc(&g->lastDir, size_t(n) * sizeof(vec3))) ||
        !ok(hipMalloc(&g->hitDist, size_t(n) * sizeof(Real))) || !ok(hipMalloc(&g->hitCount, size_t(n) * sizeof(int))))
        return false;
    g->cap = n;
    return true;
}

struct Frame {
    vec3 llc, u, v, rayDir;
};

// sweep.cu:336-364 for sweep index idx; pd = thS thE phS phE grid, pi = nTh nPh
Frame frame(const double* pd, const int* pi, int idx) {
    const Real thS = Real(pd[0]), thE = Real(pd[1]), phS = Real(pd[2]), phE = Real(pd[3]), grid = Real(pd[4]);
    const int nTh = pi[0], nPh = pi[1];
    const int t = idx / nPh, p = idx % nPh;
    Real thetaDeg = thS + (nTh > 1 ? t * (thE - thS) / (nTh - 1) : 0);
    Real thetaRad = thetaDeg * Real(M_PI) / Real(180.0);
    Real phiDeg = phS + (nPh > 1 ? p * (phE - phS) / (nPh - 1) : 0);
    Real phiRad = phiDeg * Real(M_PI) / Real(180.0);
    vec3 dir(std::sin(thetaRad) * std::cos(phiRad), std::sin(thetaRad) * std::sin(phiRad), std::cos(thetaRad));
    vec3 up(0, 1, 0);
    if (std::fabs(dir.y()) > Real(0.99)) up = vec3(1, 0, 0);
    Frame f;
    f.u = unitVector(cross(up, dir)) * grid;
    f.v = unitVector(cross(dir, f.u)) * grid;
    f.llc = (dir * (Real(10.0) * grid)) - Real(0.5) * f.u - Real(0.5) * f.v;
    f.rayDir = -dir;
    return f;
}

// RT + PO kernels of one angle on the stream; event pairs time each launch
bool run_angle(RefGpu* g, const double* pd, const int* pi, int idx) {
    const int nx = pi[2], ny = pi[3], maxB = pi[4];
    const Real grid = Real(pd[4]);
    const Real c0 = Real(299792458.0);
    const Real lambda = c0 / Real(pd[5]);
    const Real k = Real(2.0) * Real(M_PI) / lambda;
    const Real rayArea = (grid * grid) / (nx * ny);
    const int nRays = nx * ny;
    Frame f = frame(pd, pi, idx);
    dim3 rtThreads(GPU_RT_BLOCK_X, GPU_RT_BLOCK_Y);
    dim3 rtBlocks(GPU_RT_GRID_X(nx), GPU_RT_GRID_Y(ny));
    if (!ok(hipMemsetAsync(g->accum, 0, sizeof(cuRealComplex), g->stream))) return false;
    if (!ok(hipEventRecord(g->ev[0], g->stream))) return false;
    hipLaunchKernelGGL(launchRaysMultiBounce, rtBlocks, rtThreads, 0, g->stream, g->hitNormal, g->lastDir,
                       g->hitDist, g->hitCount, nx, ny, f.llc, f.u, f.v, f.rayDir, g->bvh.triangles, g->bvh.nodes,
                       g->bvh.rootIndex, maxB);
    if (!ok(hipGetLastError()) || !ok(hipEventRecord(g->ev[1], g->stream))) return false;
    if (!ok(hipEventRecord(g->ev[2], g->stream))) return false;
    hipLaunchKernelGGL(integratePoMultiBounce, dim3(GPU_GRID_SIZE(nRays, GPU_PO_BLOCK_SIZE)),
                       dim3(GPU_PO_BLOCK_SIZE), 0, g->stream, g->hitNormal, g->lastDir, g->hitDist, g->hitCount,
                       nRays, k, f.rayDir, rayArea, g->accum, Real(pd[6]));
    if (!ok(hipGetLastError()) || !ok(hipEventRecord(g->ev[3], g->stream))) return false;
    return true;
}

}  // namespace

extern "C" {

// Precision of this build: 4 (float) or 8 (double).
int refgpu_real_size(void) { return int(sizeof(Real)); }

RefGpu* refgpu_create(const float* xyz, int nVerts, const uint32_t* idx, int nTris, double scale) {
    if (!xyz || !idx || nVerts <= 0 || nTris <= 0) return nullptr;
    std::vector<Triangle> tris;
    tris.reserve(nTris);
    const Real s = Real(scale);
    for (int i = 0; i < nTris; ++i) {
        Triangle t;
        vec3* v[3] = {&t.v0, &t.v1, &t.v2};
        for (int c = 0; c < 3; ++c) {
            const uint32_t vi = idx[3 * i + c];
            if (vi >= uint32_t(nVerts)) return nullptr;
            *v[c] = vec3(Real(xyz[3 * vi]) * s, Real(xyz[3 * vi + 1]) * s, Real(xyz[3 * vi + 2]) * s);
        }
        tris.push_back(t);
    }
    bvhBuildResult b = buildBvh(std::move(tris), bvhBuildOptions{});
    auto* g = new RefGpu();
    g->bvh.triangleCount = int(b.triangles.size());
    g->bvh.nodeCount = int(b.nodes.size());
    g->bvh.rootIndex = b.rootIndex;
    bool good = ok(hipStreamCreate(&g->stream)) &&
                ok(hipMalloc(&g->bvh.triangles, b.triangles.size() * sizeof(Triangle))) &&
                ok(hipMemcpy(g->bvh.triangles, b.triangles.data(), b.triangles.size() * sizeof(Triangle),
                             hipMemcpyHostToDevice)) &&
                ok(hipMalloc(&g->bvh.nodes, b.nodes.size() * sizeof(BvhNode))) &&
                ok(hipMemcpy(g->bvh.nodes, b.nodes.data(), b.nodes.size() * sizeof(BvhNode), hipMemcpyHostToDevice)) &&
                ok(hipMalloc(&g->accum, sizeof(cuRealComplex)));
    for (int e = 0; e < 4 && good; ++e) good = ok(hipEventCreate(&g->ev[e]));
    if (!good) {
        delete g;
        return nullptr;
    }
    return g;
}

void refgpu_destroy(RefGpu* g) {
    if (!g) return;
    (void)hipStreamSynchronize(g->stream);
    (void)hipFree(g->bvh.triangles);
    (void)hipFree(g->bvh.nodes);
    (void)hipFree(g->hitNormal);
    (void)hipFree(g->lastDir);
    (void)hipFree(g->hitDist);
    (void)hipFree(g->hitCount);
    (void)hipFree(g->accum);
    for (auto& e : g->ev)
        if (e) (void)hipEventDestroy(e);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
}

// The PO term of poKernels.cu:75-110 for every ray of the last RT launch,
// evaluated on the host in Real and summed in double in ray order: the
// reference's sum without its block reduction (whose second blockReduceSum
// call may overwrite the shared partials the first call is still reading).
bool host_po(RefGpu* g, const double* pd, const int* pi, double* re, double* im) {
    const int nx = pi[2], ny = pi[3], nr = nx * ny;
    std::vector<vec3> hn(nr), ld(nr);
    std::vector<Real> hd(nr);
    std::vector<int> hc(nr);
    if (!ok(hipMemcpy(hn.data(), g->hitNormal, nr * sizeof(vec3), hipMemcpyDeviceToHost)) ||
        !ok(hipMemcpy(ld.data(), g->lastDir, nr * sizeof(vec3), hipMemcpyDeviceToHost)) ||
        !ok(hipMemcpy(hd.data(), g->hitDist, nr * sizeof(Real), hipMemcpyDeviceToHost)) ||
        !ok(hipMemcpy(hc.data(), g->hitCount, nr * sizeof(int), hipMemcpyDeviceToHost)))
        return false;
    const Real grid = Real(pd[4]);
    const Real lambda = Real(299792458.0) / Real(pd[5]);
    const Real k = Real(2.0) * Real(M_PI) / lambda;
    const Real rayArea = (grid * grid) / (nx * ny);
    double sr = 0, si = 0;
    for (int r = 0; r < nr; ++r) {
        if (hc[r] <= 0) continue;
        Real refl = Real(1.0);
        for (int b = 0; b < hc[r]; ++b) refl *= Real(pd[6]);
        Real cosT = dot(hn[r], -ld[r]);
        if (!(cosT > Real(0.0))) continue;
        Real ph = std::fmod(Real(2.0) * k * hd[r], Real(2.0) * Real(M_PI));
        Real mag = (k * rayArea / (Real(4.0) * Real(M_PI))) * Real(2.0) * cosT * refl;
        sr += double(-mag * std::sin(ph));
        si += double(mag * std::cos(ph));
    }
    *re = sr;
    *im = si;
    return true;
}

// Sweep angles[0..n) of the sweep (pd = thS thE phS phE grid freq refl,
// pi = nTh nPh nx ny maxB): reim[2n] accumulator per angle, times[2] = total
// RT and PO kernel milliseconds. host_sum != 0: reim is the host PO sum of
// the RT kernel's outputs instead of the PO kernel's accumulator.
int refgpu_sweep(RefGpu* g, const double* pd, const int* pi, const int* angles, int n, int host_sum, double* reim,
                 double* times) {
    if (!g || !pd || !pi || (n > 0 && !angles)) return 1;
    if (!ensure_rays(g, pi[2] * pi[3])) return 2;
    double rt = 0, po = 0;
    for (int a = 0; a < n; ++a) {
        if (!run_angle(g, pd, pi, angles[a])) return 3;
        cuRealComplex h;
        if (!ok(hipMemcpyAsync(&h, g->accum, sizeof(h), hipMemcpyDeviceToHost, g->stream)) ||
            !ok(hipStreamSynchronize(g->stream)))
            return 4;
        float m0 = 0, m1 = 0;
        (void)hipEventElapsedTime(&m0, g->ev[0], g->ev[1]);
        (void)hipEventElapsedTime(&m1, g->ev[2], g->ev[3]);
        rt += m0;
        po += m1;
        if (reim) {
            if (host_sum) {
                if (!host_po(g, pd, pi, &reim[2 * a], &reim[2 * a + 1])) return 5;
            } else {
                reim[2 * a] = double(h.x);
                reim[2 * a + 1] = double(h.y);
            }
        }
    }
    if (times) {
        times[0] = rt;
        times[1] = po;
    }
    return 0;
}

// Per-ray outputs of one angle: hit counts and path lengths (row-major rays).
int refgpu_rays(RefGpu* g, const double* pd, const int* pi, int angle, int32_t* counts, double* dist) {
    if (!g || !pd || !pi) return 1;
    const int nr = pi[2] * pi[3];
    if (!ensure_rays(g, nr)) return 2;
    if (!run_angle(g, pd, pi, angle)) return 3;
    if (!ok(hipStreamSynchronize(g->stream))) return 4;
    std::vector<int> c(nr);
    std::vector<Real> d(nr);
    if (!ok(hipMemcpy(c.data(), g->hitCount, nr * sizeof(int), hipMemcpyDeviceToHost)) ||
        !ok(hipMemcpy(d.data(), g->hitDist, nr * sizeof(Real), hipMemcpyDeviceToHost)))
        return 4;
    for (int r = 0; r < nr; ++r) {
        if (counts) counts[r] = c[r];
        if (dist) dist[r] = double(d[r]);
    }
    return 0;
}

}  // extern "C"
