// Plain data shared by the host driver and the HIP kernels.
#pragma once

#include <cstddef>
#include <cstdint>

namespace sbr {

constexpr int kWgSize = 256;   // 4 wavefronts of 64
constexpr int kMaxStack = 64;  // deferred-left capacity; equals the reference's
                               // int stack[64] (src/cuda/rtKernels.cu:65-66)
constexpr int kMaxDepth = 62;  // deepest tree for which the reference never
                               // drops a push (stackPtr + 2 <= 64)
// Mixed mode's wide tree: W children per record (2, 4 or 8; kWide by
// default), up to W - 1 pushes per record visited.
constexpr int kWide = 4;
template <int W> constexpr int kRecBytes = (W * 28 + 63) / 64 * 64;  // 6 box floats + 1 ref per child, padded
constexpr int kWideRecBytes = kRecBytes<kWide>;
constexpr int kMaxStack4 = 96;  // deferred children of one mixed query (any width)
constexpr int kMaxDepth4 = kMaxStack4 / (kWide - 1);
// deepest wide tree of width w whose deferred children fit kMaxStack4
constexpr int max_depth_wide(int w) { return kMaxStack4 / (w - 1); }
// index of width 2, 4, 8 in per-width arrays
constexpr int width_slot(int w) { return w == 2 ? 0 : (w == 8 ? 2 : 1); }
constexpr int kEmptySlot = 0x7fffffff;  // unused child slot of a wide record
constexpr int kCullBoxes = 256;         // boxes of the tile and ray-patch emptiness tests
// Boxes of the batched runs' ray-patch coverage masks: a cut of the BVH
// (every triangle lies in one), at most this many
constexpr int kCovBoxes = 65536;
// Mask words a workgroup of the coverage kernel builds in LDS (64 KB)
constexpr int kCovLdsWords = 8192;
// LDS chunks of a job's coverage mask of maskX x maskY words: rows of at
// most kCovLdsWords words, as many rows as fit
inline int cov_chunks(int maskX, int maskY) {
    if (maskX <= 0 || maskY <= 0) return 0;
    const int cols = maskX < kCovLdsWords ? maskX : kCovLdsWords;
    int rows = kCovLdsWords / cols;
    if (rows < 1) rows = 1;
    if (rows > maskY) rows = maskY;
    return ((maskX + cols - 1) / cols) * ((maskY + rows - 1) / rows);
}
// Top levels of the wide tree that a workgroup can stage in LDS: at most
// kTopBytes of records (1 + 4 + 16 + 64 = 85 four-wide records)
constexpr int kTopLevels = 4;
constexpr int kTopRecs = 85;
constexpr int kTopBytes = kTopRecs * kWideRecBytes;
constexpr int kChunkTiles = 64;         // tiles per entry of the tile -> job table
// Fixed device scratch of the allocation-free sweep (sbr_sweep_noalloc)
constexpr size_t kScratchBytes = size_t(64) << 20;

// One sigma value (mesh, f, theta, phi) ready for the device. Frame and
// physics constants are in the working precision R; kd / twoPid carry the
// FP64 wavenumber and 2 pi for the adjoint.
template <typename R> struct alignas(16) Job {
    R llc[3], uv[3], vv[3], dir[3];  // launch frame, sweep.cu:349-364
    R k, pre, twoPi, refl;           // k, (k*dA/4pi)*2, 2pi, reflection const
    R invTwoPi;                      // 1 / twoPi rounded (phase reduction's quotient estimate)
    double kd, twoPid;
    double L0;        // mixed mode: distance the launch plane was moved towards the body
    float ex0;        // mixed mode: initial ray position error bound [m]
    int nx, ny, maxB, mesh;
    int i0, j0, tilesX, tilesY;      // culled ray rectangle in tiles
    // launch-plane coordinates of a point X: a = X . cullA[0..2] - cullA[3]
    // (ray (i, j) sits at a = (i + 0.5) / nx), b likewise with cullB
    float cullA[4], cullB[4];
    // bounds on the FP32 rounding of a and b over the mesh's boxes (the
    // products and sums grow with the coordinates' size against the grid)
    float cullErrA, cullErrB;
    int64_t tileBase;                // first global tile of this job
    // batched runs: this job's ray-patch coverage mask, one 64-bit word per
    // 64 x 64 ray block of its tile rectangle (maskX x maskY blocks from
    // (i0, j0), row-major from maskBase), bit (y * 8 + x) for the block's
    // 8 x 8 ray patch (x, y)
    int64_t maskBase;
    int maskX, maskY;
    // record runs: this job's first ray in the per-ray arrays and first
    // entry of the hit-face array (jobs one after another)
    int64_t recBase, recIdBase;
    int nTiles;
    // Partial sums are kept per 16x16 ray pass on a grid fixed by the culled
    // rectangle alone (passX x passY), so the summation order -- and sigma to
    // the last bit -- does not depend on the tile size or on the method.
    int passX, passY;
    int64_t passBase;                // first partial of this job
};

template <typename R> struct alignas(16) Mesh {
    const void* nodes;
    const void* tris;
    // FP32 tree only: the reference FP64 vertices of each triangle, in the
    // FP32 tree's order (v0 v1 v2 as 9 doubles + face id), for the mixed
    // mode's FP64 triangle tests; null when vertsExact
    const void* tris64;
    // FP32 tree only: every reference FP64 vertex equals its FP32 vertex
    // (model scale 1, or any scale that rounds no vertex), so the mixed
    // mode reads the FP32 record and widens it exactly
    int vertsExact;
    // FP32 tree only: the two-, four- and eight-wide collapses of the tree
    // (mixed mode; width_slot order; a width not built yet is null), their
    // root records, and the records of their top k levels, which come first
    // (topRecsW[slot][k], k = 0 .. kTopLevels)
    const void* nodesW[3];
    int rootRefW[3];
    int topRecsW[3][kTopLevels + 1];
    // up to kCullBoxes boxes (lo.xyz hi.xyz, FP32, rounded outwards) that
    // together hold every triangle, for the per-tile emptiness test
    const float* cull;
    int nCull;
    // up to kCovBoxes boxes of the same kind (a finer cut of the tree) for
    // the batched runs' coverage masks
    const float* cov;
    int nCov;
    R rootLo[3], rootHi[3];
    // FP32 tree only: the root box for the mixed mode, bounding the FP64
    // triangles (the FP32 root box widened when vertsExact is 0)
    float rootLo4[3], rootHi4[3];
    // 1 + the largest coordinate magnitude of the root box per axis, summed
    // (mixed mode's tie margin scale)
    double tieScale;
    int rootRef;
    int nTris;
};

// Kernel arguments (passed by value).
template <typename R> struct TraceArgs {
    const Job<R>* jobs;
    const Mesh<R>* meshes;
    int nJobs;
    int tileSide;            // rays per tile side (16, 32 or 64)
    int64_t totalTiles;
    int64_t tileOffset;      // first tile of a non-persistent launch
    void* partials;          // one accumulator per 16x16 ray pass
    double* jobOut;          // 2 doubles per job
    unsigned long long* tileCounter;  // persistent launch: next tile to take
    // job of the first tile of each chunk of kChunkTiles tiles (the job of
    // a tile lies between its chunk's entry and the next chunk's)
    const int* chunkJob;
    int64_t nChunks;
    // per-ray records (single-job debug launches)
    int32_t* recIds;
    int32_t* recCount;
    double* recL;
    double* recReIm;
    // work counters: rays, visits, tests, bounces, queries, rays with an
    // undecided query, undecided queries (answered by the full FP64 query)
    unsigned long long* counters;
    // mixed mode: reference FP64 jobs and meshes (FP64 ray state and the
    // queries the mixed query cannot vouch for), error-bound safety factor,
    // per-ray flags of rays with such a query (records only)
    const Job<double>* jobs64;
    const Mesh<double>* meshes64;
    float errK;
    int8_t* recFlag;
    // levels of the wide tree staged in LDS (mixed mode, 1 .. kTopLevels)
    int topLevels;
    // ray-patch coverage masks of the jobs (batched runs; null: tiles and
    // patches are tested against the kCullBoxes boxes instead); countCulled:
    // a counting run skips what the masks rule out, as a timed run does
    const unsigned long long* mask;
    int countCulled;
};

// Arguments of the frozen-path adjoint kernel (one launch per job).
// One job of an adjoint launch; its rays take blocks [blockBase, next
// job's blockBase), 256 rays each.
struct AdjJob {
    Job<double> D;           // frame, physics, n_x, n_y, max bounces
    int64_t rayOff, idOff;   // the job's first entry in the record arrays
    double w;                // upstream d loss / d sigma of the job
    double freq;
    int nRays;
    int blockBase;
};
struct AdjArgs {
    const AdjJob* jobs;
    int nJobs;
    const double* verts;  // n_verts * 3, reference FP64 vertices (scaled)
    const unsigned* idx;  // n_tris * 3
    const int32_t* ids;   // hit faces of every job's rays (max bounces per ray)
    const int32_t* cnt;   // bounce counts of every job's rays
    const double* sums;   // 2 per job: the forward complex sums
    double scale;
    double* dvert;        // n_verts * 3, d loss / d (input coordinate), accumulated
    double* dfreq;        // per job: d sigma / d f (unweighted), accumulated
};

}  // namespace sbr
