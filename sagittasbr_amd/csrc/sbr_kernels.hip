// HIP kernels of the SBR ray path for gfx950 (MI355X).
//
// One fused kernel replaces the reference's two per-angle launches
// (launchRaysMultiBounce, src/cuda/rtKernels.cu:105-158, and
// integratePoMultiBounce, src/cuda/poKernels.cu:66-122):
//   ray generation in registers -> closest-hit BVH traversal with
//   Moller-Trumbore -> multi-bounce mirror loop -> PO term in registers ->
//   wavefront shuffle reduction -> one partial per 16x16 ray pass -> a
//   finalize launch sums each job's partials in a fixed order
//   (deterministic, no float atomics, no per-ray buffers in HBM).
//
// Work unit: a tile of tileSide x tileSide parallel rays of one job (one
// sigma value). Each 256-thread workgroup is 4 wavefronts; a wavefront owns
// an 8x8 ray patch so its 64 rays are spatially coherent.
//
// Traversal decisions are those of the reference's stack DFS
// (rtKernels.cu:63-100): the BVH is stored with both children's boxes in
// the parent record, the right child is entered first, the left child is
// deferred on an LDS stack together with its slab entry distance, and on
// pop it is kept iff entry <= closest -- exactly the reference's
// "tmax >= tmin" test at the time the reference pops that node. Triangles
// are accepted with t <= closest, so the last-tested triangle wins a tie
// as in rtKernels.cu:36,85-88.
//
// Arithmetic: FP32 code uses fmaf at fixed places (dot, cross, ray origin,
// hit point, reflection, offset); FP64 code uses no fused operations. The
// file is compiled with -ffp-contract=off so nothing else is fused. Division
// and square root are IEEE correctly rounded (the mixed mode's FP32 slab
// reciprocals excepted: their error is part of its box bound).
//
// Mixed mode (MODE 1): FP32 box traversal of a wide tree (2, 4 or 8
// children per record; four by default) against an
// FP32 copy of the FP64 reference ray, every box grown by that copy's
// position error bound (so pruning is conservative), and the reference FP64
// Moller-Trumbore test of every triangle reached. A query whose answer could
// differ from the reference FP64 query's (a near-tie, or a hit on the edge
// of its triangle's box) is answered by that query itself on the FP64 tree,
// from the same FP64 ray state. Every query therefore has the FP64 outcome;
// the ray state, path length and phase are FP64 and the complex sum is FP32
// two-sum compensated.
#include <hip/hip_runtime.h>

#include <climits>
#include <mutex>
#include <type_traits>

#include "sbr_device.hpp"

namespace sbr {
namespace {

template <typename R> struct V3 {
    R x, y, z;
};

template <typename R> __device__ __forceinline__ V3<R> sub(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }

__device__ __forceinline__ float mn(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ double mn(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ float mx(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double mx(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ float rsqrt_ieee(float a) { return sqrtf(a); }
__device__ __forceinline__ double rsqrt_ieee(double a) { return sqrt(a); }
__device__ __forceinline__ float absr(float a) { return fabsf(a); }
__device__ __forceinline__ double absr(double a) { return fabs(a); }
__device__ __forceinline__ float fmodr(float a, float b) { return fmodf(a, b); }
__device__ __forceinline__ double fmodr(double a, double b) { return fmod(a, b); }
__device__ __forceinline__ void sincosr(float a, float* s, float* c) { sincosf(a, s, c); }
__device__ __forceinline__ void sincosr(double a, double* s, double* c) { sincos(a, s, c); }

template <typename R> __device__ __forceinline__ R dot3(V3<R> a, V3<R> b) {
    if constexpr (std::is_same_v<R, float>)
        return fmaf(a.z, b.z, fmaf(a.x, b.x, a.y * b.y));
    else
        return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <typename R> __device__ __forceinline__ V3<R> cross3(V3<R> a, V3<R> b) {
    if constexpr (std::is_same_v<R, float>)
        return {fmaf(a.y, b.z, -(a.z * b.y)), -fmaf(a.x, b.z, -(a.z * b.x)), fmaf(a.x, b.y, -(a.y * b.x))};
    else
        return {a.y * b.z - a.z * b.y, -(a.x * b.z - a.z * b.x), a.x * b.y - a.y * b.x};
}
// s*a + b
template <typename R> __device__ __forceinline__ V3<R> axpy(R s, V3<R> a, V3<R> b) {
    if constexpr (std::is_same_v<R, float>)
        return {fmaf(s, a.x, b.x), fmaf(s, a.y, b.y), fmaf(s, a.z, b.z)};
    else
        return {s * a.x + b.x, s * a.y + b.y, s * a.z + b.z};
}
template <typename R> __device__ __forceinline__ V3<R> unit3(V3<R> v) {
    R len = rsqrt_ieee(dot3(v, v));
    return {v.x / len, v.y / len, v.z / len};
}

// Slab test split in two: the closest-independent part here, the
// "entry <= closest" part at the caller (rtKernels.cu:46-61).
template <typename R>
__device__ __forceinline__ bool slab(R lx, R ly, R lz, R hx, R hy, R hz, V3<R> o, V3<R> inv, R& entry) {
    R t0x = (lx - o.x) * inv.x, t0y = (ly - o.y) * inv.y, t0z = (lz - o.z) * inv.z;
    R t1x = (hx - o.x) * inv.x, t1y = (hy - o.y) * inv.y, t1z = (hz - o.z) * inv.z;
    R tn = mx(mx(mn(t0x, t1x), mn(t0y, t1y)), mn(t0z, t1z));
    R tf = mn(mn(mx(t0x, t1x), mx(t0y, t1y)), mx(t0z, t1z));
    entry = mx(tn, R(0.001));
    return entry <= tf;
}

// Mixed-mode slab test of the box grown by dx on every side, one FMA per
// plane: (lo - dx - o) * inv = lo * inv - oLo * inv with oLo = o + dx, and
// (hi + dx - o) * inv = hi * inv - oHi * inv with oHi = o - dx (oLo, oHi
// rounded outwards and pLo = oLo * inv, pHi = oHi * inv formed once per
// query by the caller; inv is finite). Its rounding, eps (|o| + |t|) in
// position units, is covered by the position error bound dx.
// The same test with the slab interval's far end capped at bound: true iff
// entry <= far and entry <= bound (one compare; no value here is NaN).
__device__ __forceinline__ bool slab_grown(float lx, float ly, float lz, float hx, float hy, float hz, V3<float> pLo,
                                           V3<float> pHi, V3<float> inv, float bound, float& entry) {
    float t0x = fmaf(lx, inv.x, -pLo.x), t0y = fmaf(ly, inv.y, -pLo.y), t0z = fmaf(lz, inv.z, -pLo.z);
    float t1x = fmaf(hx, inv.x, -pHi.x), t1y = fmaf(hy, inv.y, -pHi.y), t1z = fmaf(hz, inv.z, -pHi.z);
    float tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fminf(t0z, t1z));
    float tf = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fminf(fmaxf(t0z, t1z), bound));
    entry = fmaxf(tn, 0.001f);
    return entry <= tf;
}
__device__ __forceinline__ bool slab_grown(float lx, float ly, float lz, float hx, float hy, float hz, V3<float> pLo,
                                           V3<float> pHi, V3<float> inv, float& entry) {
    return slab_grown(lx, ly, lz, hx, hy, hz, pLo, pHi, inv, 3.0e38f, entry);
}

// BVH and triangle records are read through global-address-space pointers:
// the mesh table holds them as plain pointers, and without the cast every
// fetch is a FLAT access (which also waits on the LDS counter).
template <typename T> using gptr = const __attribute__((address_space(1))) T*;
template <typename T> __device__ __forceinline__ gptr<T> as_global(const void* p) {
    return (gptr<T>)(p);
}
// Byte offsets into a mesh's node and triangle records are 32-bit (the host
// keeps every such array a plan reads below 4 GiB), so a load addresses the
// array's scalar base plus one 32-bit vector register.
template <typename T> __device__ __forceinline__ gptr<T> at_byte(const void* base, uint32_t off) {
    return reinterpret_cast<gptr<T>>(as_global<char>(base) + off);
}
// The same value in every lane, read from the first active lane (held in
// scalar registers).
__device__ __forceinline__ int64_t uniform64(int64_t v) {
    unsigned lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(static_cast<uint64_t>(v)));
    unsigned hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(static_cast<uint64_t>(v) >> 32));
    return static_cast<int64_t>((static_cast<uint64_t>(hi) << 32) | lo);
}
// A pointer every lane of the wavefront holds (a mesh's arrays: a wave
// traces rays of one job), kept in scalar registers; the compiler cannot
// tell that a pointer loaded from the mesh table is the same in every lane.
template <typename T> __device__ __forceinline__ T* uniform_ptr(T* p) {
    return reinterpret_cast<T*>(uniform64(reinterpret_cast<int64_t>(p)));
}

typedef float vf4 __attribute__((ext_vector_type(4)));
typedef double vd2 __attribute__((ext_vector_type(2)));
typedef int vi4 __attribute__((ext_vector_type(4)));

template <typename R> struct TriRec {
    V3<R> v0, e1, e2;
    int face;
};

// FP32 record: v0, v1, v2, face, 8 bytes of padding; the edges are formed
// here exactly as rtKernels.cu:13-14 forms them.
__device__ __forceinline__ TriRec<float> load_tri(const float* tris, int i) {
    const uint32_t o = static_cast<uint32_t>(i) * 48u;
    const vf4 a = *at_byte<vf4>(tris, o), b = *at_byte<vf4>(tris, o + 16u), c = *at_byte<vf4>(tris, o + 32u);
    V3<float> v0{a.x, a.y, a.z}, v1{a.w, b.x, b.y}, v2{b.z, b.w, c.x};
    return {v0, sub(v1, v0), sub(v2, v0), __float_as_int(c.y)};
}
// The same record's three vertices as stored, and the face id.
__device__ __forceinline__ void load_tri_raw64(const void* tris64, int i, V3<double>& v0, V3<double>& v1,
                                               V3<double>& v2, int& face) {
    const uint32_t o = static_cast<uint32_t>(i) * 80u;
    const vd2 a = *at_byte<vd2>(tris64, o), b = *at_byte<vd2>(tris64, o + 16u), c = *at_byte<vd2>(tris64, o + 32u),
              d = *at_byte<vd2>(tris64, o + 48u), e = *at_byte<vd2>(tris64, o + 64u);
    v0 = {a.x, a.y, b.x};
    v1 = {b.y, c.x, c.y};
    v2 = {d.x, d.y, e.x};
    face = static_cast<int>(__double_as_longlong(e.y));
}
// Reference FP64 vertices and face id of FP32-tree triangle i for the mixed
// mode: the FP32 record widened when it holds them exactly, else the FP64
// record.
__device__ __forceinline__ void load_tri_mixed(const Mesh<float>& m, int i, V3<double>& v0, V3<double>& v1,
                                               V3<double>& v2, int& face) {
    if (m.vertsExact) {
        const uint32_t o = static_cast<uint32_t>(i) * 48u;
        const vf4 a = *at_byte<vf4>(m.tris, o), b = *at_byte<vf4>(m.tris, o + 16u), c = *at_byte<vf4>(m.tris, o + 32u);
        v0 = {double(a.x), double(a.y), double(a.z)};
        v1 = {double(a.w), double(b.x), double(b.y)};
        v2 = {double(b.z), double(b.w), double(c.x)};
        face = __float_as_int(c.y);
    } else {
        load_tri_raw64(m.tris64, i, v0, v1, v2, face);
    }
}
__device__ __forceinline__ TriRec<double> load_tri(const double* tris, int i) {
    const uint32_t o = static_cast<uint32_t>(i) * 80u;
    const vd2 a = *at_byte<vd2>(tris, o), b = *at_byte<vd2>(tris, o + 16u), c = *at_byte<vd2>(tris, o + 32u),
              d = *at_byte<vd2>(tris, o + 48u), e = *at_byte<vd2>(tris, o + 64u);
    int face = static_cast<int>(__double_as_longlong(e.y) & 0xffffffffll);
    return {{a.x, a.y, b.x}, {b.y, c.x, c.y}, {d.x, d.y, e.x}, face};
}

template <typename R> struct NodeRec {
    R b[12];
    int l, r;
};
__device__ __forceinline__ NodeRec<float> load_node(const float* nodes, int i) {
    const uint32_t o = static_cast<uint32_t>(i) * 64u;
    const vf4 a = *at_byte<vf4>(nodes, o), b = *at_byte<vf4>(nodes, o + 16u), c = *at_byte<vf4>(nodes, o + 32u);
    const vi4 d = *at_byte<vi4>(nodes, o + 48u);
    return {{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w}, d.x, d.y};
}
__device__ __forceinline__ NodeRec<double> load_node(const double* nodes, int i) {
    const uint32_t o = static_cast<uint32_t>(i) * 128u;
    NodeRec<double> n;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const vd2 v = *at_byte<vd2>(nodes, o + 16u * k);
        n.b[2 * k] = v.x;
        n.b[2 * k + 1] = v.y;
    }
    const vi4 d = *at_byte<vi4>(nodes, o + 96u);
    n.l = d.x;
    n.r = d.y;
    return n;
}

// Wide record of the mixed mode's tree: lo.x[W] lo.y[W] lo.z[W] hi.x[W]
// hi.y[W] hi.z[W] ref[W] as 16-byte loads (W = 2: the last 8 bytes are
// the record's padding).
template <int W> struct NodeW {
    static constexpr int kLoads = (7 * W + 3) / 4;
    static_assert(16 * kLoads <= kRecBytes<W>, "loads stay in the record");
    float f[4 * kLoads];
    __device__ __forceinline__ int ref(int q) const { return __float_as_int(f[6 * W + q]); }
};
template <int W> __device__ __forceinline__ NodeW<W> load_nodew(const void* nodes, int i) {
    gptr<vf4> p = at_byte<vf4>(nodes, static_cast<uint32_t>(i) * static_cast<uint32_t>(kRecBytes<W>));
    NodeW<W> n;
#pragma unroll
    for (int k = 0; k < NodeW<W>::kLoads; ++k) {
        const vf4 v = p[k];
        n.f[4 * k] = v.x;
        n.f[4 * k + 1] = v.y;
        n.f[4 * k + 2] = v.z;
        n.f[4 * k + 3] = v.w;
    }
    return n;
}

// The same record from the workgroup's LDS copy of the tree's top levels
// (records of kRecBytes<W>, 16-byte aligned).
template <int W> __device__ __forceinline__ NodeW<W> load_nodew_lds(const vf4* top, int i) {
    const vf4* p = top + (kRecBytes<W> / 16) * i;
    NodeW<W> n;
#pragma unroll
    for (int k = 0; k < NodeW<W>::kLoads; ++k) {
        const vf4 v = p[k];
        n.f[4 * k] = v.x;
        n.f[4 * k + 1] = v.y;
        n.f[4 * k + 2] = v.z;
        n.f[4 * k + 3] = v.w;
    }
    return n;
}

// Sorting network on (key, value) pairs, ascending keys.
template <int W> __device__ __forceinline__ void sort_net(float* e, int* r) {
    // compare-exchange as selects (a branch costs exec-mask updates and moves)
    auto cs = [&](int a, int b) {
        const bool sw = e[b] < e[a];
        const float ea = sw ? e[b] : e[a], eb = sw ? e[a] : e[b];
        const int ra = sw ? r[b] : r[a], rb = sw ? r[a] : r[b];
        e[a] = ea;
        e[b] = eb;
        r[a] = ra;
        r[b] = rb;
    };
    if constexpr (W == 2) {
        cs(0, 1);
    } else if constexpr (W == 4) {
        cs(0, 1); cs(2, 3); cs(0, 2); cs(1, 3); cs(1, 2);
    } else {
        static_assert(W == 8, "wide records of 2, 4 or 8 children");
        cs(0, 1); cs(2, 3); cs(4, 5); cs(6, 7);
        cs(0, 2); cs(1, 3); cs(4, 6); cs(5, 7);
        cs(1, 2); cs(5, 6);
        cs(0, 4); cs(1, 5); cs(2, 6); cs(3, 7);
        cs(2, 4); cs(3, 5);
        cs(1, 2); cs(3, 4); cs(5, 6);
    }
}

// Moller-Trumbore, rtKernels.cu:10-44. Accepts t in [0.001, closest].
template <typename R>
__device__ __forceinline__ bool mt_test(const TriRec<R>& tr, V3<R> o, V3<R> d, R closest, R& tOut) {
    V3<R> p = cross3(d, tr.e2);
    R det = dot3(tr.e1, p);
    if (absr(det) < R(1.0e-6)) return false;
    R inv = R(1.0) / det;
    V3<R> tv = sub(o, tr.v0);
    R u = dot3(tv, p) * inv;
    if (u < R(0) || u > R(1)) return false;
    V3<R> q = cross3(tv, tr.e1);
    R v = dot3(d, q) * inv;
    if (v < R(0) || u + v > R(1)) return false;
    R t = dot3(tr.e2, q) * inv;
    if (t < R(0.001) || t > closest) return false;
    tOut = t;
    return true;
}

// Private overflow entry of the traversal stack: (slab entry, node ref)
// packed in one 64-bit (FP32) or 128-bit (FP64) word, so an overflow pop is
// one scratch load whose shape differs from the LDS pop (the compiler would
// otherwise merge the two into a FLAT access).
template <typename R> struct Ov;
template <> struct Ov<float> {
    using T = unsigned long long;
    __device__ static T pack(float e, int r) {
        return (static_cast<T>(static_cast<unsigned>(r)) << 32) | static_cast<T>(__float_as_uint(e));
    }
    __device__ static void unpack(T v, float& e, int& r) {
        e = __uint_as_float(static_cast<unsigned>(v));
        r = static_cast<int>(static_cast<unsigned>(v >> 32));
    }
};
template <> struct Ov<double> {
    using T = vd2;
    __device__ static T pack(double e, int r) { return T{e, __longlong_as_double(static_cast<long long>(r))}; }
    __device__ static void unpack(T v, double& e, int& r) {
        e = v.x;
        r = static_cast<int>(__double_as_longlong(v.y));
    }
};

struct Work {
    unsigned visits, tests;
};

// Reference value meaning "stack exhausted": below every leaf code
// ~((first << 4) | (n - 1)) because the host keeps first < 2^27.
constexpr int kNoRef = INT_MIN;

// Deferred-left stack of one lane: entries [0, SD) in LDS at
// [k * kWgSize + tid], deeper ones in the private overflow array. The lane
// keeps at = sp * kWgSize + tid, the LDS index of its next entry: since
// 0 <= tid < kWgSize, sp < SD exactly when at < SD * kWgSize, and sp is
// at / kWgSize, so one register serves as stack pointer and LDS address.
template <typename R, int SD> struct LaneStack {
    static_assert(kWgSize == 256, "sp = at >> 8");
    int* sRef;
    R* sE;
    typename Ov<R>::T* ov;
    int at;
    __device__ __forceinline__ void push(R e, int r) {
        if (SD > 0 && at < SD * kWgSize) {
            sRef[at] = r;
            sE[at] = e;
        } else {
            ov[(at >> 8) - SD] = Ov<R>::pack(e, r);
        }
        at += kWgSize;
    }
    // the next deferred child whose slab entry is <= lim, or kNoRef
    __device__ __forceinline__ int pop(R lim) {
        while (at >= kWgSize) {
            at -= kWgSize;
            R ee;
            int rr;
            if (SD > 0 && at < SD * kWgSize) {
                ee = sE[at];
                rr = sRef[at];
            } else {
                Ov<R>::unpack(ov[(at >> 8) - SD], ee, rr);
            }
            if (ee <= lim) return rr;
        }
        return kNoRef;
    }
};

// Loop shapes of the traversal (WW template flag). If-if (false): every
// iteration each lane takes its next step, a node step or a leaf's triangle
// tests, so a wavefront runs both kinds side by side and their memory
// latencies overlap. While-while (true): lanes walk internal nodes (pops
// included) until each has reached a leaf or emptied its stack, then the
// leaves are tested together, so the triangle code runs once for all lanes
// instead of beside every node step. Each lane's own sequence of node
// tests, triangle tests and pops -- and so every decision -- is that of the
// reference loop in both shapes.
// Closest hit (rtKernels.cu:63-100). Returns the BVH position of the hit
// triangle or -1. The host guarantees tree depth <= kMaxDepth so the stack
// stays below kMaxStack entries.
template <typename R, int SD, bool CNT, bool WW, bool PF = true>
__device__ __forceinline__ int closest_hit(const Mesh<R>& m, V3<R> o, V3<R> d, V3<R> inv, R& tHit, int* sRef, R* sE,
                                           typename Ov<R>::T* ov, int tid, Work& w) {
    // (the mesh's arrays in scalar registers)
    const R* nodes = uniform_ptr(static_cast<const R*>(m.nodes));
    const R* tris = uniform_ptr(static_cast<const R*>(m.tris));
    R closest = R(1e20);
    int hit = -1;
    if (CNT) w.visits += 1;
    R e;
    if (!slab(m.rootLo[0], m.rootLo[1], m.rootLo[2], m.rootHi[0], m.rootHi[1], m.rootHi[2], o, inv, e)) return -1;
    LaneStack<R, SD> st{sRef, sE, ov, tid};
    int cur = m.rootRef;
    // one node step: descend into a hit child (true) or not (false)
    auto node_step = [&]() -> bool {
        NodeRec<R> n = load_node(nodes, cur);
        if (CNT) w.visits += 2;
        R eL, eR;
        bool hL = slab(n.b[0], n.b[1], n.b[2], n.b[3], n.b[4], n.b[5], o, inv, eL) && eL <= closest;
        bool hR = slab(n.b[6], n.b[7], n.b[8], n.b[9], n.b[10], n.b[11], o, inv, eR) && eR <= closest;
        if (hR) {
            if (hL) st.push(eL, n.l);
            cur = n.r;
            return true;
        }
        if (hL) {
            cur = n.l;
            return true;
        }
        return false;
    };
    auto leaf_tests = [&]() {
        int code = ~cur;
        int first = code >> 4;
        int cnt = (code & 15) + 1;
        if (CNT) w.tests += cnt;
        if constexpr (std::is_same_v<R, float> && PF) {
            // the next triangle's nine vertex floats are loaded while this
            // one is tested (two 16-byte loads and one 4-byte load); the
            // edges are formed as load_tri forms them. PF false: the kernel
            // holds too many registers for the nine (the ray pool's)
            const uint32_t ob = static_cast<uint32_t>(first) * 48u;
            vf4 ra = *at_byte<vf4>(tris, ob), rb = *at_byte<vf4>(tris, ob + 16u);
            float rc = *at_byte<float>(tris, ob + 32u);
            for (int k = 0; k < cnt; ++k) {
                vf4 na = ra, nb = rb;
                float nc = rc;
                if (k + 1 < cnt) {
                    const uint32_t on = ob + 48u * static_cast<uint32_t>(k + 1);
                    na = *at_byte<vf4>(tris, on);
                    nb = *at_byte<vf4>(tris, on + 16u);
                    nc = *at_byte<float>(tris, on + 32u);
                }
                const V3<float> v0{ra.x, ra.y, ra.z}, v1{ra.w, rb.x, rb.y}, v2{rb.z, rb.w, rc};
                const TriRec<float> tr{v0, sub(v1, v0), sub(v2, v0), 0};
                R t;
                if (mt_test(tr, o, d, closest, t)) {
                    closest = t;
                    hit = first + k;
                }
                ra = na;
                rb = nb;
                rc = nc;
            }
        } else {
            for (int k = 0; k < cnt; ++k) {
                TriRec<R> tr = load_tri(tris, first + k);
                R t;
                if (mt_test(tr, o, d, closest, t)) {
                    closest = t;
                    hit = first + k;
                }
            }
        }
    };
    if constexpr (WW) {
        for (;;) {
            while (cur >= 0)
                if (!node_step()) cur = st.pop(closest);
            if (cur == kNoRef) break;
            leaf_tests();
            cur = st.pop(closest);
            if (cur == kNoRef) break;
        }
    } else {
        // one pop for the lanes that left a leaf and those at a dead end
        for (;;) {
            if (cur >= 0) {
                if (node_step()) continue;
            } else {
                leaf_tests();
            }
            cur = st.pop(closest);
            if (cur == kNoRef) break;
        }
    }
    tHit = closest;
    return hit;
}

// ------------------------------------------------- mixed-mode traversal
constexpr float kEps32 = 5.9604645e-8f;  // 2^-24

// Mixed-mode query: FP32 traversal of the wide tree, every box grown by the
// position error bound of the FP32 copy of the ray (so no box the FP64 ray
// reaches is missed), and the reference FP64 Moller-Trumbore test
// (rtKernels.cu:10-44) of every triangle reached, on the FP64 ray and the
// reference FP64 vertices. Returns the FP32-tree position of the nearest FP64
// hit (or -1) and its FP64 distance. The reference query returns that
// triangle too unless (a) another hit lies within the tie margin (the
// reference breaks exact ties by its own tree's test order) or (b) the hit
// point lies within the margin of the edge of the triangle's box on an axis
// the triangle spans (the reference's FP64 box tests could reject it by
// rounding; on an axis the triangle does not span, both slab planes of a
// leaf give the same value); amb is set then. Rounding of the reference's
// FP64 slab and distance arithmetic stays below 2^-51 (R + |o| + |P|) in
// position units (R: the mesh's largest coordinate magnitude per axis,
// summed), so the margin is 1e-12 of that. The FP32 origin may be moved
// along the ray by shift (first query): FP32 box distances are then shift
// shorter than the FP64 ones. Unless amb is set, a hit's FP64 vertices and
// face id are returned in hv0, hv1, hv2 and hFace.
// OCT (four-wide tree only): every lane running the query has the same
// direction signs. A slab's near and far planes are then known per axis for
// the whole wavefront: the near plane's distance is the smaller of the two
// (both are formed from the same rounded terms, and rounding is monotonic),
// so the box test reads the near and far planes from their own offsets in the
// record and needs no min / max per plane pair. Entries and decisions are
// bit-identical to the general test.
template <int SD, bool CNT, bool WW, bool TOP, int W, bool OCT = false>
__device__ __forceinline__ int closest_hit_mixed(const Mesh<float>& m, V3<float> o, V3<float> inv, V3<double> o64,
                                                 V3<double> d64, double shift, double& tHit, int* sRef, float* sE,
                                                 Ov<float>::T* ov, const vf4* top, int nTop, int tid, Work& w,
                                                 float ex, float ed, float K, bool& amb, V3<double>& hv0,
                                                 V3<double>& hv1, V3<double>& hv2, int& hFace) {
    const float o1 = fabsf(o.x) + fabsf(o.y) + fabsf(o.z);
    // far distance of anything in the root box from the origin
    float cx = 0.5f * (m.rootLo4[0] + m.rootHi4[0]), cy = 0.5f * (m.rootLo4[1] + m.rootHi4[1]),
          cz = 0.5f * (m.rootLo4[2] + m.rootHi4[2]);
    float hx = m.rootHi4[0] - cx, hy = m.rootHi4[1] - cy, hz = m.rootHi4[2] - cz;
    // (L1 norms bound the Euclidean ones)
    float tFar = (fabsf(o.x - cx) + fabsf(o.y - cy) + fabsf(o.z - cz) + hx + hy + hz) * (1.0f + 4.0f * kEps32);
    // position and distance error bounds of the FP32 box tests over the
    // query (the second tFar is the reciprocal's error beyond rounding)
    float dx = ex + ed * tFar + K * kEps32 * (o1 + 2.0f * tFar);
    float tErr = dx + 4.0f * K * kEps32 * tFar;
    double closest = 1e20, second = 1e20;
    int hit = -1;
    amb = false;
    float e;
    if (CNT) w.visits += 1;
    // o + dx and o - dx rounded outwards: the margin covers the rounding of
    // the sums and of the margin itself
    const float dxs = dx * (1.0f + 8.0f * kEps32) + 4.0f * kEps32 * o1;
    const V3<float> pLo{(o.x + dxs) * inv.x, (o.y + dxs) * inv.y, (o.z + dxs) * inv.z};
    const V3<float> pHi{(o.x - dxs) * inv.x, (o.y - dxs) * inv.y, (o.z - dxs) * inv.z};
    if (!slab_grown(m.rootLo4[0], m.rootLo4[1], m.rootLo4[2], m.rootHi4[0], m.rootHi4[1], m.rootHi4[2], pLo, pHi, inv,
                    e)) {
        tHit = closest;
        return -1;
    }
    // pruning bound in FP32 box distances: a box is dropped only when it
    // starts beyond the best FP64 hit by more than the box test's error (and
    // the tie margin, far below it)
    float lim = 3.0e38f;
    auto update_lim = [&]() { lim = float(closest - shift) * (1.0f + 4.0f * kEps32) + 2.0f * tErr; };
    LaneStack<float, SD> st{sRef, sE, ov, tid};
    // (the mesh's arrays in scalar registers)
    const void* nodes = uniform_ptr(m.nodesW[width_slot(W)]);
    const void* tris = uniform_ptr(m.tris);
    const void* tris64 = uniform_ptr(m.tris64);
    const bool vertsExact = __builtin_amdgcn_readfirstlane(m.vertsExact) != 0;
    // the key of a child not entered (3e38, above every finite bound), made
    // in a scalar register: as a literal the compiler holds it in a vector
    // register through the traversal
    float noKey;
    asm("s_mov_b32 %0, 0x7f61b1e6" : "=s"(noKey));
    int cur = m.rootRefW[width_slot(W)];
    // OCT: byte offsets of the near and far planes of each axis in a record
    // (planes lo x, lo y, lo z, hi x, hi y, hi z of the four children, 16
    // bytes each; a negative direction component enters at the hi plane), the
    // wavefront's sign bits read from its first lane, and the grown origin
    // terms of the near and far planes
    static_assert(!OCT || (W == 4 && kRecBytes<W> == 128 && NodeW<W>::kLoads == 7),
                  "the per-octant box test reads 128-byte four-wide records: six planes, then the references");
    const unsigned sgn = __builtin_amdgcn_readfirstlane((inv.x < 0.0f ? 1u : 0u) | (inv.y < 0.0f ? 2u : 0u) |
                                                        (inv.z < 0.0f ? 4u : 0u));
    const uint32_t nOx = (sgn & 1u) ? 48u : 0u, nOy = (sgn & 2u) ? 64u : 16u, nOz = (sgn & 4u) ? 80u : 32u;
    const uint32_t fOx = 48u - nOx, fOy = 80u - nOy, fOz = 112u - nOz;
    const V3<float> pN{(sgn & 1u) ? pHi.x : pLo.x, (sgn & 2u) ? pHi.y : pLo.y, (sgn & 4u) ? pHi.z : pLo.z};
    const V3<float> pF{(sgn & 1u) ? pLo.x : pHi.x, (sgn & 2u) ? pLo.y : pHi.y, (sgn & 4u) ? pLo.z : pHi.z};
    // wide tree: the hit children are visited nearest first (the outcome
    // does not depend on the order)
    auto node_step = [&]() -> bool {
        if (CNT) w.visits += W;
        if constexpr (OCT) {
            vf4 nx, ny, nz, fx, fy, fz, rf;
            if (TOP && cur < nTop) {
                const char* b = reinterpret_cast<const char*>(top) + static_cast<uint32_t>(cur) * 128u;
                nx = *reinterpret_cast<const vf4*>(b + nOx);
                ny = *reinterpret_cast<const vf4*>(b + nOy);
                nz = *reinterpret_cast<const vf4*>(b + nOz);
                fx = *reinterpret_cast<const vf4*>(b + fOx);
                fy = *reinterpret_cast<const vf4*>(b + fOy);
                fz = *reinterpret_cast<const vf4*>(b + fOz);
                rf = *reinterpret_cast<const vf4*>(b + 96);
            } else {
                // (each plane's scalar base plus the record's offset)
                const uint32_t ob = static_cast<uint32_t>(cur) * 128u;
                const char* nb = static_cast<const char*>(nodes);
                nx = *at_byte<vf4>(nb + nOx, ob);
                ny = *at_byte<vf4>(nb + nOy, ob);
                nz = *at_byte<vf4>(nb + nOz, ob);
                fx = *at_byte<vf4>(nb + fOx, ob);
                fy = *at_byte<vf4>(nb + fOy, ob);
                fz = *at_byte<vf4>(nb + fOz, ob);
                rf = *at_byte<vf4>(nb + 96, ob);
            }
            auto key = [&](int q) {
                const float tn = fmaxf(fmaxf(fmaxf(fmaf(nx[q], inv.x, -pN.x), fmaf(ny[q], inv.y, -pN.y)),
                                             fmaf(nz[q], inv.z, -pN.z)),
                                       0.001f);
                const float tf = fminf(fminf(fmaf(fx[q], inv.x, -pF.x), fmaf(fy[q], inv.y, -pF.y)),
                                       fmaf(fz[q], inv.z, -pF.z));
                return (tn <= tf && tn <= lim) ? tn : noKey;
            };
            // the four (key, child) pairs as plain scalars: held in arrays,
            // the compiler turns the sort's selects of children into selects
            // of vector indices
            float e0 = key(0), e1 = key(1), e2 = key(2), e3 = key(3);
            int r0 = __float_as_int(rf.x), r1 = __float_as_int(rf.y), r2 = __float_as_int(rf.z),
                r3 = __float_as_int(rf.w);
            auto cx = [](float& ea, float& eb, int& ra, int& rb) {
                const bool sw = eb < ea;
                const float lo = sw ? eb : ea, hi = sw ? ea : eb;
                const int rl = sw ? rb : ra, rh = sw ? ra : rb;
                ea = lo;
                eb = hi;
                ra = rl;
                rb = rh;
            };
            cx(e0, e1, r0, r1);
            cx(e2, e3, r2, r3);
            cx(e0, e2, r0, r2);
            cx(e1, e3, r1, r3);
            cx(e1, e2, r1, r2);
            if (!(e0 < 1.0e38f)) return false;
            if (e3 < 1.0e38f) st.push(e3, r3);
            if (e2 < 1.0e38f) st.push(e2, r2);
            if (e1 < 1.0e38f) st.push(e1, r1);
            cur = r0;
            return true;
        } else {
            float e[W];
            int r[W];
            // records below nTop are the top levels the workgroup staged in LDS
            const NodeW<W> n = (TOP && cur < nTop) ? load_nodew_lds<W>(top, cur) : load_nodew<W>(nodes, cur);
#pragma unroll
            for (int q = 0; q < W; ++q) {
                float ent;
                // an unused slot's box lies at +infinity: its entry is
                // +infinity (above the finite lim) or its slab interval is
                // empty
                const bool h = slab_grown(n.f[q], n.f[W + q], n.f[2 * W + q], n.f[3 * W + q], n.f[4 * W + q],
                                          n.f[5 * W + q], pLo, pHi, inv, lim, ent);
                e[q] = h ? ent : noKey;
                r[q] = n.ref(q);
            }
            sort_net<W>(e, r);
            if (!(e[0] < 1.0e38f)) return false;
#pragma unroll
            for (int q = W - 1; q >= 1; --q)
                if (e[q] < 1.0e38f) st.push(e[q], r[q]);
            cur = r[0];
            return true;
        }
    };
    auto leaf_tests = [&]() {
        int code = ~cur;
        int first = code >> 4;
        int cnt = (code & 15) + 1;
        if (CNT) w.tests += cnt;
        auto test = [&](const V3<double>& a, const V3<double>& b, const V3<double>& c, int k) {
            const TriRec<double> tr{a, sub(b, a), sub(c, a), 0};
            double t;
            if (mt_test(tr, o64, d64, 1e20, t)) {
                if (t < closest) {
                    second = closest;
                    closest = t;
                    hit = first + k;
                    update_lim();
                } else if (t < second) {
                    second = t;
                }
            }
        };
        if (vertsExact) {
            // FP32 records widened: the next triangle's record is loaded
            // while this one is tested
            // (the nine vertex floats of the 48-byte record: two 16-byte loads
            // and one 4-byte load)
            const uint32_t ob = static_cast<uint32_t>(first) * 48u;
            vf4 ra = *at_byte<vf4>(tris, ob), rb = *at_byte<vf4>(tris, ob + 16u);
            float rc = *at_byte<float>(tris, ob + 32u);
            for (int k = 0; k < cnt; ++k) {
                vf4 na = ra, nb = rb;
                float nc = rc;
                if (k + 1 < cnt) {
                    const uint32_t on = ob + 48u * static_cast<uint32_t>(k + 1);
                    na = *at_byte<vf4>(tris, on);
                    nb = *at_byte<vf4>(tris, on + 16u);
                    nc = *at_byte<float>(tris, on + 32u);
                }
                test(V3<double>{double(ra.x), double(ra.y), double(ra.z)},
                     V3<double>{double(ra.w), double(rb.x), double(rb.y)},
                     V3<double>{double(rb.z), double(rb.w), double(rc)}, k);
                ra = na;
                rb = nb;
                rc = nc;
            }
        } else {
            for (int k = 0; k < cnt; ++k) {
                V3<double> a, b, c;
                int f;
                load_tri_raw64(tris64, first + k, a, b, c, f);
                test(a, b, c, k);
            }
        }
    };
    if constexpr (WW) {
        for (;;) {
            while (cur >= 0)
                if (!node_step()) cur = st.pop(lim);
            if (cur == kNoRef) break;
            leaf_tests();
            cur = st.pop(lim);
            if (cur == kNoRef) break;
        }
    } else {
        // one pop for the lanes that left a leaf and those at a dead end
        for (;;) {
            if (cur >= 0) {
                if (node_step()) continue;
            } else {
                leaf_tests();
            }
            cur = st.pop(lim);
            if (cur == kNoRef) break;
        }
    }
    if (hit >= 0) {
        const double R = m.tieScale;
        const V3<double> P = axpy(closest, d64, o64);
        const double margin =
            1e-12 * (R + fabs(o64.x) + fabs(o64.y) + fabs(o64.z) + fabs(P.x) + fabs(P.y) + fabs(P.z));
        if (second - closest <= margin) {
            amb = true;
        } else {
            load_tri_mixed(m, hit, hv0, hv1, hv2, hFace);
            auto inside_axis = [&](double pa, double x0, double x1, double x2) {
                const double lo = fmin(x0, fmin(x1, x2)), hi = fmax(x0, fmax(x1, x2));
                return !(hi > lo) || (pa - lo > margin && hi - pa > margin);
            };
            amb = !inside_axis(P.x, hv0.x, hv1.x, hv2.x) || !inside_axis(P.y, hv0.y, hv1.y, hv2.y) ||
                  !inside_axis(P.z, hv0.z, hv1.z, hv2.z);
        }
    }
    tHit = closest;
    return hit;
}

// ---------------------------------------------------------------- sums
template <typename T> __device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Two-sum of (s1,c1) + (s2,c2): s = s1+s2 exactly split into s + err
__device__ __forceinline__ void two_sum_add(float& s, float& c, float s2, float c2) {
    float t = s + s2;
    float bp = t - s;
    float err = (s - (t - bp)) + (s2 - bp);
    s = t;
    c = c + c2 + err;
}
__device__ __forceinline__ void wave_two_sum(float& s, float& c) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        float s2 = __shfl_xor(s, off);
        float c2 = __shfl_xor(c, off);
        two_sum_add(s, c, s2, c2);
    }
}

// Accumulator per precision mode. MODE 0: plain sums in R; MODE 1 (float
// only): two-sum compensated FP32 (sum, error) pairs.
template <typename R, int MODE> struct Acc;
template <typename R> struct Acc<R, 0> {
    R re, im;
    static __device__ Acc zero() { return Acc{R(0), R(0)}; }
    __device__ void add_ray(R r, R i) { re += r; im += i; }
    __device__ void wave() { re = wave_sum(re); im = wave_sum(im); }
    __device__ void add(const Acc& o) { re += o.re; im += o.im; }
    __device__ double dre() const { return double(re); }
    __device__ double dim() const { return double(im); }
};
template <> struct Acc<float, 1> {
    float re, cre, im, cim;
    static __device__ Acc zero() { return Acc{0.f, 0.f, 0.f, 0.f}; }
    __device__ void add_ray(float r, float i) { two_sum_add(re, cre, r, 0.f); two_sum_add(im, cim, i, 0.f); }
    // an FP64 term enters as its FP32 (head, tail) pair
    __device__ void add_ray64(double r, double i) {
        float rh = float(r), ih = float(i);
        two_sum_add(re, cre, rh, float(r - double(rh)));
        two_sum_add(im, cim, ih, float(i - double(ih)));
    }
    __device__ void wave() { wave_two_sum(re, cre); wave_two_sum(im, cim); }
    __device__ void add(const Acc& o) { two_sum_add(re, cre, o.re, o.cre); two_sum_add(im, cim, o.im, o.cim); }
    __device__ double dre() const { return double(re) + double(cre); }
    __device__ double dim() const { return double(im) + double(cim); }
};

// Workgroup sum of one Acc per thread; result valid in thread 0.
template <typename A> __device__ __forceinline__ A wg_sum(A a, A* red, int tid) {
    a.wave();
    if ((tid & 63) == 0) red[tid >> 6] = a;
    __syncthreads();
    A r = A::zero();
    if (tid == 0) {
        r = red[0];
        for (int k = 1; k < kWgSize / 64; ++k) r.add(red[k]);
    }
    __syncthreads();
    return r;
}

// Values every lane has just read from LDS after a barrier, handed to the
// compiler as wave-uniform scalars so that the loop exit and the branches
// around barriers that depend on them are uniform branches.


// ------------------------------------------------------------ ray paths
template <typename R> struct PathOut {
    int count;
    R L;
    double Ld;
    R cosT;  // n_last . (-d_last), poKernels.cu:89-94, formed at the last hit
};

// Reference ray (rtKernels.cu:113-158) as a state machine: generation,
// then one closest-hit query plus bounce update per step, so a lane can be
// refilled with a new ray as soon as its current one ends.
template <typename R> struct RayState {
    V3<R> o, d;
    PathOut<R> P;
    int k;
};

template <typename R> __device__ __forceinline__ V3<R> ray_origin(const Job<R>& J, int i, int j) {
    R u = (R(i) + R(0.5)) / R(J.nx);
    R v = (R(j) + R(0.5)) / R(J.ny);
    V3<R> llc{J.llc[0], J.llc[1], J.llc[2]};
    V3<R> U{J.uv[0], J.uv[1], J.uv[2]}, Vv{J.vv[0], J.vv[1], J.vv[2]};
    return axpy(v, Vv, axpy(u, U, llc));
}

template <typename R> __device__ __forceinline__ void ray_init(const Job<R>& J, int i, int j, RayState<R>& st) {
    st.o = ray_origin(J, i, j);
    st.d = V3<R>{J.dir[0], J.dir[1], J.dir[2]};
    st.P = PathOut<R>{0, R(0), 0.0, R(0)};
    st.k = 0;
}

// One query of bounce st.k; returns true when the ray has ended (miss or
// max_bounces reached). Requires st.k < J.maxB.
template <typename R, int SD, bool CNT, bool REC, bool WW, bool PF = true>
__device__ __forceinline__ bool ray_step(const Job<R>& J, const Mesh<R>& M, RayState<R>& st, int* sRef, R* sE,
                                         typename Ov<R>::T* ov, int tid, int32_t* ids, Work& w,
                                         unsigned long long& queries) {
    V3<R> inv{R(1.0) / st.d.x, R(1.0) / st.d.y, R(1.0) / st.d.z};
    R t;
    int h = closest_hit<R, SD, CNT, WW, PF>(M, st.o, st.d, inv, t, sRef, sE, ov, tid, w);
    if (CNT) queries++;
    if (h < 0) return true;
    TriRec<R> tr = load_tri(static_cast<const R*>(M.tris), h);
    V3<R> n = unit3(cross3(tr.e1, tr.e2));
    V3<R> p = axpy(t, st.d, st.o);
    st.P.cosT = dot3(n, V3<R>{-st.d.x, -st.d.y, -st.d.z});
    st.P.count++;
    st.P.L += t;
    if (REC) ids[st.k] = tr.face;
    R s = R(2.0) * dot3(st.d, n);
    st.d = unit3(axpy(-s, n, st.d));
    st.o = axpy(R(1e-4), n, p);
    st.k++;
    return st.k >= J.maxB;
}

// Whole path of one ray in one call. SD == 0 keeps the whole traversal
// stack in private memory.
template <typename R, int SD, bool CNT, bool REC, bool WW>
__device__ __forceinline__ PathOut<R> trace_path(const Job<R>& J, const Mesh<R>& M, int i, int j, int* sRef, R* sE,
                                                 typename Ov<R>::T* ov, int tid, int32_t* ids, Work& w,
                                                 unsigned long long& queries) {
    RayState<R> st;
    ray_init(J, i, j, st);
    if (J.maxB > 0)
        while (!ray_step<R, SD, CNT, REC, WW>(J, M, st, sRef, sE, ov, tid, ids, w, queries)) {
        }
    st.P.Ld = double(st.P.L);
    return st.P;
}

// fmod(x, y) for x >= 0 and y > 0 given invY ~ 1 / y: the quotient from one
// multiply by invY is off by at most one below the size limit and is
// corrected, and the remainder x - n y is formed by one fused multiply-add,
// exact for the correct n (the remainder of a floating-point division is
// representable), so the value is fmod's. Larger quotients go to fmod.
__device__ __forceinline__ float fma_r(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double fma_r(double a, double b, double c) { return fma(a, b, c); }
template <typename R> __device__ __forceinline__ R fmod_pos(R x, R y, R invY) {
    constexpr R kMaxQ = std::is_same_v<R, double> ? R(1ull << 50) : R(1 << 20);
    R n = floor(x * invY);
    if (!(n < kMaxQ)) return fmodr(x, y);
    R r = fma_r(-n, y, x);
    if (r < R(0)) {
        n -= R(1);
        r = fma_r(-n, y, x);
    } else if (r >= y) {
        n += R(1);
        r = fma_r(-n, y, x);
    }
    return r;
}

// sin and cos of x in [0, 2 pi] in FP64 (the mixed mode's reduced phase).
// Cody-Waite reduction by the quadrant q = rint(2 x / pi), with pi / 2 split
// into two doubles, leaves |r| <= pi / 4 (+ 1e-16); there the Taylor series
// of sin to r^17 and of cos to r^16 are truncated below 1e-19, so the error
// is the FP64 evaluation's, about one unit in the last place of 1. Shorter
// than the general sincos, which also carries the large-argument reduction.
// An FP64 constant with a nonzero low word takes two registers. The compiler
// moves such constants out of the ray loops, where they stay live across the
// traversal and are spilled to scratch and reloaded every bounce; passed
// through an empty volatile asm they are formed at their use, two scalar
// moves each, on the scalar unit.
__device__ __forceinline__ double at_use(double c) {
    asm volatile("" : "+s"(c));
    return c;
}

__device__ __forceinline__ void sincos_2pi(double x, double* s, double* c) {
    constexpr double kTwoOverPi = 0.63661977236758134308;
    constexpr double kPio2Hi = 1.5707963267948966;     // pi / 2 rounded
    constexpr double kPio2Lo = 6.123233995736766e-17;  // pi / 2 - kPio2Hi
    const double q = rint(x * at_use(kTwoOverPi));
    double r = fma(-q, at_use(kPio2Hi), x);
    r = fma(-q, at_use(kPio2Lo), r);
    const double z = r * r;
    // sin r = r + r^3 (-1/3! + z (1/5! - z (1/7! - ...))), Horner in z
    double ps = at_use(1.0 / 355687428096000.0);  // 1/17!
    ps = fma(ps, z, at_use(-1.0 / 1307674368000.0));
    ps = fma(ps, z, at_use(1.0 / 6227020800.0));
    ps = fma(ps, z, at_use(-1.0 / 39916800.0));
    ps = fma(ps, z, at_use(1.0 / 362880.0));
    ps = fma(ps, z, at_use(-1.0 / 5040.0));
    ps = fma(ps, z, at_use(1.0 / 120.0));
    ps = fma(ps, z, at_use(-1.0 / 6.0));
    const double sr = fma(r * z, ps, r);
    // cos r = 1 + z (-1/2! + z (1/4! - ...))
    double pc = at_use(1.0 / 20922789888000.0);  // 1/16!
    pc = fma(pc, z, at_use(-1.0 / 87178291200.0));
    pc = fma(pc, z, at_use(1.0 / 479001600.0));
    pc = fma(pc, z, at_use(-1.0 / 3628800.0));
    pc = fma(pc, z, at_use(1.0 / 40320.0));
    pc = fma(pc, z, at_use(-1.0 / 720.0));
    pc = fma(pc, z, at_use(1.0 / 24.0));
    pc = fma(pc, z, at_use(-0.5));
    const double cr = fma(z, pc, 1.0);
    // x = q pi/2 + r: rotate by the quadrant
    const int n = static_cast<int>(q) & 3;
    const double a = (n & 1) ? cr : sr, b = (n & 1) ? sr : cr;
    *s = (n & 2) ? -a : a;
    *c = ((n + 1) & 2) ? -b : b;
}

// PO term (poKernels.cu:75-110) from a finished path. RED (FP64 only): sin
// and cos of the reduced phase by sincos_2pi.
template <typename R, bool RED = false>
__device__ __forceinline__ void po_term(const Job<R>& J, const PathOut<R>& P, R& re, R& im) {
    static_assert(!RED || std::is_same_v<R, double>, "reduced-phase sincos is FP64");
    re = R(0);
    im = R(0);
    if (P.count <= 0) return;
    R refl = R(1.0);
    for (int b = 0; b < P.count; ++b) refl *= J.refl;
    const R cosT = P.cosT;
    if (!(cosT > R(0))) return;
    const R ph = fmod_pos((R(2.0) * J.k) * P.L, J.twoPi, J.invTwoPi);
    R mag = (J.pre * cosT) * refl;
    R sn, cs;
    if constexpr (RED) sincos_2pi(ph, &sn, &cs);
    else sincosr(ph, &sn, &cs);
    re = -mag * sn;
    im = mag * cs;
}

// Mixed-mode ray. The ray state (origin, direction, path length, last
// normal) is carried in FP64 with the reference's FP64 operations, so it is
// the FP64 reference ray itself; each closest-hit query runs on an FP32 copy
// of it (the first one from a launch plane moved next to the body) with
// error-bounded box decisions and FP64 triangle tests.
__device__ __forceinline__ void ray_init_mixed(const Job<double>& D, int i, int j, RayState<double>& st) {
    st.o = ray_origin(D, i, j);
    st.d = V3<double>{D.dir[0], D.dir[1], D.dir[2]};
    st.P = PathOut<double>{0, 0.0, 0.0, 0.0};
    st.k = 0;
}

#ifndef SBR_OCTANT_TEST
#define SBR_OCTANT_TEST 1
#endif
constexpr bool kOctantTest = SBR_OCTANT_TEST != 0;

// One query of bounce st.k; returns true when the ray has ended (miss or
// max_bounces reached). A query whose answer the mixed query cannot vouch for
// (a near-tie or a hit on the edge of its box) is answered by the reference
// FP64 query on the FP64 tree from the same FP64 ray state -- the reference
// query itself -- and sets fell.
template <int SD, bool CNT, bool REC, bool WW, bool TOP, int W>
__device__ __forceinline__ bool ray_step_mixed(const Job<float>& J, const Mesh<float>& M, const Mesh<double>& M64,
                                               RayState<double>& st, int* sRef, float* sE, Ov<float>::T* ov,
                                               Ov<double>::T* ov64, const vf4* top, int nTop, int tid, int32_t* ids,
                                               Work& w,
                                               unsigned long long& queries, unsigned long long& fullQ, float K,
                                               bool& fell, double* sCos) {
    const V3<double> o = st.o, d = st.d;
    // FP32 copy of the ray for the box tests; the launch-plane move only
    // shifts the origin along the ray, which changes no decision
    const double shift = st.k == 0 ? J.L0 : 0.0;
    V3<float> of{float(o.x + shift * d.x), float(o.y + shift * d.y), float(o.z + shift * d.z)};
    V3<float> df{float(d.x), float(d.y), float(d.z)};
    // finite slab reciprocals: components below 1e-30 are taken as +-1e-30,
    // a direction change far inside the direction error bound. The hardware
    // reciprocal (relative error up to 2^-23, one unit more than a rounded
    // division) moves a slab distance t by up to 2^-23 t: the query's
    // position bound carries the extra 2^-24 t of it.
    auto away = [](float c) { return copysignf(fmaxf(fabsf(c), 1e-30f), c); };
    V3<float> inv{__builtin_amdgcn_rcpf(away(df.x)), __builtin_amdgcn_rcpf(away(df.y)),
                  __builtin_amdgcn_rcpf(away(df.z))};
    const float ex = K * kEps32 * (fabsf(of.x) + fabsf(of.y) + fabsf(of.z) + 1e-3f);
    const float ed = 2.0f * K * kEps32;
    double t;
    bool amb = false;
    V3<double> v0, v1, v2, e1, e2;
    int face;
    int h;
    // the same direction signs in every lane running this query (always so
    // in a first query: the job's rays are parallel): the per-octant box test
    const unsigned sg = (inv.x < 0.0f ? 1u : 0u) | (inv.y < 0.0f ? 2u : 0u) | (inv.z < 0.0f ? 4u : 0u);
    if (W == 4 && kOctantTest && __builtin_amdgcn_ballot_w64(sg != __builtin_amdgcn_readfirstlane(sg)) == 0)
        h = closest_hit_mixed<SD, CNT, WW, TOP, W, W == 4>(M, of, inv, o, d, shift, t, sRef, sE, ov, top, nTop, tid,
                                                            w, ex, ed, K, amb, v0, v1, v2, face);
    else
        h = closest_hit_mixed<SD, CNT, WW, TOP, W>(M, of, inv, o, d, shift, t, sRef, sE, ov, top, nTop, tid, w, ex,
                                                    ed, K, amb, v0, v1, v2, face);
    if (CNT) queries++;
    // FP64 edges of the chosen triangle (rtKernels.cu:13-14)
    if (!amb) {
        if (h < 0) return true;
        e1 = sub(v1, v0);
        e2 = sub(v2, v0);
    } else {
        fell = true;
        if (CNT) fullQ++;
        const V3<double> inv64{1.0 / d.x, 1.0 / d.y, 1.0 / d.z};
        const int h64 = closest_hit<double, 0, CNT, WW>(M64, o, d, inv64, t, nullptr, nullptr, ov64, tid, w);
        if (h64 < 0) return true;
        const TriRec<double> tr = load_tri(static_cast<const double*>(M64.tris), h64);
        e1 = tr.e1;
        e2 = tr.e2;
        face = tr.face;
    }
    // bounce update as ray_step forms it in FP64. t is the FP64
    // Moller-Trumbore distance of that triangle (both queries form it with
    // mt_test<double>; the FP64 tree's edges equal v1 - v0, v2 - v0 of the
    // same vertices)
    V3<double> n = unit3(cross3(e1, e2));
    V3<double> hp = axpy(t, d, o);
    // (the lane's LDS slot: only the last bounce's value is used, after the
    // path ends, and as a register it was spilled to scratch every bounce)
    sCos[tid] = dot3(n, V3<double>{-d.x, -d.y, -d.z});
    st.P.count++;
    st.P.L += t;
    if (REC) ids[st.k] = face;
    double s = 2.0 * dot3(d, n);
    st.d = unit3(axpy(-s, n, d));
    st.o = axpy(1e-4, n, hp);
    st.k++;
    return st.k >= J.maxB;
}

// Whole mixed-mode path in one call; fell is set when a query was answered
// on the FP64 tree.
template <int SD, bool CNT, bool REC, bool WW, bool TOP, int W>
__device__ __forceinline__ PathOut<double> trace_path_mixed(const Job<float>& J, const Job<double>& D,
                                                            const Mesh<float>& M, const Mesh<double>& M64, int i,
                                                            int j, int* sRef, float* sE, Ov<float>::T* ov,
                                                            Ov<double>::T* ov64, const vf4* top, int nTop, int tid,
                                                            int32_t* ids, Work& w,
                                                            unsigned long long& queries, unsigned long long& fullQ,
                                                            float K, bool& fell, double* sCos) {
    RayState<double> st;
    ray_init_mixed(D, i, j, st);
    if (J.maxB > 0)
        while (!ray_step_mixed<SD, CNT, REC, WW, TOP, W>(J, M, M64, st, sRef, sE, ov, ov64, top, nTop, tid, ids, w,
                                                      queries, fullQ, K, fell, sCos)) {
        }
    if (st.P.count > 0) st.P.cosT = sCos[tid];
    st.P.Ld = st.P.L;
    return st.P;
}

// Rectangle of launch-plane coordinates (a, b) that the projection of one of
// the mesh's cull boxes along the ray direction covers. The boxes together
// hold every triangle, so a ray whose (a, b) lies outside every box's
// rectangle hits nothing. Boxes past the mesh's count give an empty
// rectangle.
struct CullRect {
    float amin, amax, bmin, bmax;
};
template <typename R> __device__ __forceinline__ CullRect box_rect(const Job<R>& J, gptr<float> c) {
    float amin = -J.cullA[3], amax = -J.cullA[3], bmin = -J.cullB[3], bmax = -J.cullB[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float lo = c[k], hi = c[3 + k];
        const float pa = lo * J.cullA[k], qa = hi * J.cullA[k];
        const float pb = lo * J.cullB[k], qb = hi * J.cullB[k];
        amin += fminf(pa, qa);
        amax += fmaxf(pa, qa);
        bmin += fminf(pb, qb);
        bmax += fmaxf(pb, qb);
    }
    return CullRect{amin, amax, bmin, bmax};
}
template <typename R> __device__ __forceinline__ CullRect cull_rect_of(const Job<R>& J, const Mesh<R>& M, int box) {
    if (box >= M.nCull) return CullRect{3.0e38f, -3.0e38f, 3.0e38f, -3.0e38f};
    return box_rect(J, as_global<float>(M.cull) + 6 * box);
}

// Whether some ray of the ray rectangle [ia, ib] x [ja, jb] of a job can
// reach one of the kCullBoxes rectangles in rects (lane q tests rectangles
// q, q + 64, ...; the result is wave-uniform). The rectangles are widened
// by half a ray spacing, 1e-5 of the grid and the host's bound on their FP32
// rounding (which grows with the mesh's distance from the origin in grid
// sizes) and compared with the rectangle of ray centres. Rays of a rectangle
// that reaches none contribute exactly zero and are not traced.
template <typename R>
__device__ __forceinline__ bool rays_reach_mesh(const Job<R>& J, const CullRect* rects, int ia, int ib, int ja, int jb,
                                                int lane) {
    const float ma = 1e-5f + 0.5f / float(J.nx) + J.cullErrA, mb = 1e-5f + 0.5f / float(J.ny) + J.cullErrB;
    const float aLo = (float(ia) + 0.5f) / float(J.nx) - ma, aHi = (float(ib) + 0.5f) / float(J.nx) + ma;
    const float bLo = (float(ja) + 0.5f) / float(J.ny) - mb, bHi = (float(jb) + 0.5f) / float(J.ny) + mb;
    bool reach = false;
#pragma unroll
    for (int q = lane; q < kCullBoxes; q += 64) {
        const CullRect r = rects[q];
        reach = reach || (r.amax >= aLo && r.amin <= aHi && r.bmax >= bLo && r.bmin <= bHi);
    }
    return __ballot(reach) != 0ull;
}

// Coverage masks of a batched run. A job's mask words are built in LDS by
// one workgroup per chunk of at most kCovLdsWords words (a rectangle of its
// 64 x 64-ray blocks; blockIdx.x is the chunk, blockIdx.y strides over the
// jobs), whose threads take the mesh's coverage boxes in turn: a box's
// rectangle, widened as rays_reach_mesh widens the cull rectangles, becomes
// the range of rays whose centres it can hold, one more ray on each side
// for the FP32 rounding of that conversion (far below one ray spacing at
// any grid size an int ray index allows), clipped to the job's mask region;
// the 8 x 8 ray patches that range touches in the chunk get their bits set.
// The chunk's words are then stored, so every word of every job is written
// once per run and needs no clearing. The coverage boxes are a cut of the
// BVH, so a patch whose bit stays clear holds no ray that can reach a
// triangle: its rays contribute exactly zero.
template <typename R>
__global__ __launch_bounds__(256) void coverage_kernel(const Job<R>* jobs, const Mesh<R>* meshes, int nJobs,
                                                       unsigned long long* mask) {
    __shared__ unsigned long long sMask[kCovLdsWords];
    const int tid = static_cast<int>(threadIdx.x);
    for (int jid = blockIdx.y; jid < nJobs; jid += gridDim.y) {
        const Job<R>& J = jobs[jid];
        if (J.maskX == 0) continue;  // (uniform: no tiles)
        // the chunk: cw x ch blocks from block (cx0, cy0)
        const int cols = min(J.maskX, kCovLdsWords), rows = min(max(kCovLdsWords / cols, 1), J.maskY);
        const int nCx = (J.maskX + cols - 1) / cols, nCy = (J.maskY + rows - 1) / rows;
        const int c = static_cast<int>(blockIdx.x);
        if (c >= nCx * nCy) continue;  // (uniform)
        const int cx0 = (c % nCx) * cols, cy0 = (c / nCx) * rows;
        const int cw = min(cols, J.maskX - cx0), ch = min(rows, J.maskY - cy0);
        for (int k = tid; k < cw * ch; k += 256) sMask[k] = 0ull;
        __syncthreads();
        const Mesh<R>& M = meshes[J.mesh];
        const float ma = 1e-5f + 0.5f / float(J.nx) + J.cullErrA, mb = 1e-5f + 0.5f / float(J.ny) + J.cullErrB;
        // the chunk's first and last ray column and row (inside the grid)
        const int iLo = J.i0 + 64 * cx0, iHi = min(J.i0 + 64 * (cx0 + cw), J.nx) - 1;
        const int jLo = J.j0 + 64 * cy0, jHi = min(J.j0 + 64 * (cy0 + ch), J.ny) - 1;
        for (int box = tid; box < M.nCov; box += 256) {
            const CullRect r = box_rect(J, as_global<float>(M.cov) + 6 * box);
            // (NaN-free: fmaxf/fminf return the non-NaN operand, and an
            // empty rectangle keeps lo > hi)
            const float fa0 = fmaxf(floorf((r.amin - ma) * float(J.nx) - 0.5f) - 1.0f, float(iLo));
            const float fa1 = fminf(ceilf((r.amax + ma) * float(J.nx) - 0.5f) + 1.0f, float(iHi));
            const float fb0 = fmaxf(floorf((r.bmin - mb) * float(J.ny) - 0.5f) - 1.0f, float(jLo));
            const float fb1 = fminf(ceilf((r.bmax + mb) * float(J.ny) - 0.5f) + 1.0f, float(jHi));
            if (!(fa0 <= fa1) || !(fb0 <= fb1)) continue;
            // patch columns and rows relative to the job's mask region
            const int gx0 = (int(fa0) - J.i0) >> 3, gx1 = (int(fa1) - J.i0) >> 3;
            const int gy0 = (int(fb0) - J.j0) >> 3, gy1 = (int(fb1) - J.j0) >> 3;
            for (int by = gy0 >> 3; by <= gy1 >> 3; ++by) {
                const int ly0 = max(gy0 - 8 * by, 0), ly1 = min(gy1 - 8 * by, 7);
                for (int bx = gx0 >> 3; bx <= gx1 >> 3; ++bx) {
                    const int lx0 = max(gx0 - 8 * bx, 0), lx1 = min(gx1 - 8 * bx, 7);
                    const unsigned long long row = (0xffull >> (7 - lx1)) & (0xffull << lx0);
                    unsigned long long bits = 0;
                    for (int ly = ly0; ly <= ly1; ++ly) bits |= row << (8 * ly);
                    atomicOr(&sMask[(by - cy0) * cw + (bx - cx0)], bits);
                }
            }
        }
        __syncthreads();
        for (int k = tid; k < cw * ch; k += 256)
            mask[J.maskBase + int64_t(cy0 + k / cw) * J.maskX + cx0 + k % cw] = sMask[k];
        // (the next job's chunk reuses the LDS words)
        __syncthreads();
    }
}

// Bits of one tile's 8 x 8 ray patches in its job's coverage mask, bit
// (y * P + x) for the tile's patch (x, y), P = ts / 8 patches per side (ts
// = 16, 32 or 64, so the tile lies inside one 64 x 64 block).
template <typename R>
__device__ __forceinline__ unsigned long long tile_patch_bits(const Job<R>& J, const unsigned long long* mask, int tx,
                                                              int ty, int ts) {
    const int P = ts >> 3;
    const int gx = tx * P, gy = ty * P;  // the tile's first patch column and row
    const unsigned long long w = mask[J.maskBase + int64_t(gy >> 3) * J.maskX + (gx >> 3)];
    if (P == 8) return w;
    const unsigned long long row = (1ull << P) - 1ull;
    unsigned long long bits = 0;
    for (int y = 0; y < P; ++y) bits |= ((w >> (8 * ((gy & 7) + y) + (gx & 7))) & row) << (P * y);
    return bits;
}

// Per-ray result slot of the ray-pool variant: the ray's (re, im) term in
// the precision the accumulator adds it in.
template <typename R, int MODE> struct SlotOf {
    using T = double;
};
template <> struct SlotOf<float, 0> {
    using T = float;
};
constexpr int kPoolRays = 1024;  // rays per pool batch (4 passes of 16x16)

// POOL: lanes are refilled from the tile's ray pool as soon as their ray
// ends (one closest-hit query per loop iteration, so lanes of a wave never
// wait for a longer path); each ray's term goes to its own LDS slot and the
// slots are reduced per 16x16 pass exactly as without the pool.
// Minimum waves per SIMD the register allocator must allow, per precision
// mode (1 = no constraint); overridable at build time. Measured on the
// trihedral+duct sweep and a 10 GHz airframe sweep (LDS stack 8): the
// unconstrained mixed kernel takes ~150 VGPRs (3 waves) and runs 12-16%
// slower; held to 4 waves (128 VGPRs) it beats 5 waves (9% on the airframe,
// 26% on the trihedral+duct sweep, with its per-query spills) and 6 waves
// (80 VGPRs, spills in the traversal loop: -8%). FP64 gains ~10% at 4 waves.
// The FP32 kernels fit 6 waves (at most 80 VGPRs) without spills; asked for
// 8 the allocator misses that target and the persistent kernel can land at
// 85 VGPRs, 5 waves (batched FP32 on an airframe sample +8.7%). Deeper LDS
// stacks are bound by LDS instead.
#ifndef SBR_WPE_MIXED
#define SBR_WPE_MIXED 4
#endif
#ifndef SBR_WPE_FP64
#define SBR_WPE_FP64 4
#endif
#ifndef SBR_WPE_FP32
#define SBR_WPE_FP32 6
#endif
template <typename R, int MODE> constexpr int kWavesPerEu =
    MODE == 1 ? SBR_WPE_MIXED : (std::is_same_v<R, double> ? SBR_WPE_FP64 : SBR_WPE_FP32);

template <typename R, int MODE, int SD, bool REC, bool CNT, bool PERSIST, bool POOL, bool WW, bool TOP = false,
          int W = kWide>
__global__ __launch_bounds__(kWgSize) __attribute__((amdgpu_waves_per_eu(kWavesPerEu<R, MODE>))) void trace_kernel(
    TraceArgs<R> a) {
    static_assert(!TOP || (MODE == 1 && !REC && !CNT && !POOL), "top-level staging: mixed mode, timed runs");
    static_assert(W == kWide || (MODE == 1 && !POOL), "node widths other than the default: mixed mode, no pool");
    using ST = typename SlotOf<R, MODE>::T;
    using RS = typename std::conditional<MODE == 1, double, R>::type;  // ray state precision
    // (256-byte aligned: a stack entry's LDS address is then 4 x its index,
    // the arrays' places going into the instructions' offset fields)
    __shared__ __attribute__((aligned(256))) int sRef[SD * kWgSize];
    __shared__ __attribute__((aligned(256))) R sE[SD * kWgSize];
    __shared__ Acc<R, MODE> red[kWgSize / 64];
    // wavefront sums of each 16x16 pass of a tile (at most 4 x 4 passes)
    __shared__ Acc<R, MODE> sPass[16 * (kWgSize / 64)];
    __shared__ ST sRes[POOL ? 2 * kPoolRays : 1];
    __shared__ int64_t sNext;
    __shared__ unsigned sPool;
    // next 8x8 ray patch of the current tile (timed runs, one ray per lane)
    __shared__ int sPatch;
    // projected cull boxes of the current tile's job (one per thread)
    __shared__ CullRect sCull[REC ? 1 : kCullBoxes];
    // the top levels of the current mesh's W-wide tree (TOP)
    __shared__ vf4 sTop[TOP ? kTopBytes / 16 : 1];
    // each lane's last-bounce incidence cosine (mixed mode)
    __shared__ double sCos[MODE == 1 ? kWgSize : 1];
    int stagedMesh = -1, nTop = 0;
    static_assert(kCullBoxes == kWgSize, "one cull box per thread");

    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    unsigned long long cRays = 0, cVisits = 0, cTests = 0, cBounces = 0, cQueries = 0, cUndecided = 0, cFull = 0;
    typename Ov<R>::T ov[(MODE == 1 ? kMaxStack4 : kMaxStack) - SD];
    // mixed mode: private stack of the FP64 queries that answer ambiguous ones
    Ov<double>::T ov64[MODE == 1 ? kMaxStack : 1];

    // Persistent mode pops tiles from a global counter. The pop for the next
    // tile is made by thread 0 at the loop tail only, so the loop head holds
    // no thread-0-only code: a loop-head `if (tid == 0)` next to a loop-tail
    // `if (tid == 0)` store lets the compiler thread the two into a nested
    // loop whose inner back edge skips the pop for lanes 1..63 of wave 0
    // (seen in the generated code as a loop that never ends).
    int64_t tile = a.tileOffset + blockIdx.x;
    // The pool counter is never reset: every lane ends each batch with
    // exactly one failed fetch, so a batch of n rays advances it by
    // n + kWgSize and every thread can track the base itself.
    unsigned poolBase = 0;
    if (POOL && tid == 0) sPool = 0;
    if (tid == 0) sPatch = 0;
    if (PERSIST) {
        if (tid == 0) sNext = static_cast<int64_t>(atomicAdd(a.tileCounter, 1ull));
    }
    if (PERSIST || POOL) __syncthreads();
    if (PERSIST) tile = uniform64(sNext);
    while (tile < a.totalTiles) {

        // the last job whose first tile is <= tile, searched between the
        // jobs of this tile's chunk and of the next chunk
        int jid;
        {
            const int64_t c = tile / kChunkTiles;
            int lo = a.chunkJob[c], hi = c + 1 < a.nChunks ? a.chunkJob[c + 1] : a.nJobs - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (a.jobs[mid].tileBase <= tile) lo = mid;
                else hi = mid - 1;
            }
            jid = lo;
        }
        const Job<R>& J = a.jobs[jid];
        const Mesh<R>& M = a.meshes[J.mesh];
        const int tl = static_cast<int>(tile - J.tileBase);
        const int ts = a.tileSide;
        const int tx = tl % J.tilesX, ty = tl / J.tilesX;
        // tiles and 8x8 ray patches none of whose rays can reach the mesh
        // are not traced (not in record runs, which account for every ray,
        // nor in counting runs unless they count what a timed run traces).
        // Runs with coverage masks read the job's mask; the others test the
        // kCullBoxes rectangles, into which every wave may write once it has
        // passed the previous tile's last barrier.
        bool skip = false;
        const bool culled = !REC && (!CNT || a.countCulled != 0);
        const bool useMask = culled && a.mask != nullptr;
        // coverage bits of the tile's patches (bit y * ts/8 + x)
        unsigned long long pbits = ~0ull;
        if constexpr (TOP) {
            // copy the mesh's top levels once per change of mesh (made visible
            // by the barrier below)
            if (J.mesh != stagedMesh) {
                stagedMesh = J.mesh;
                // (a prefix of the breadth-first top records when the levels
                // do not fit)
                nTop = __builtin_amdgcn_readfirstlane(
                    min(M.topRecsW[width_slot(W)][a.topLevels], kTopBytes / kRecBytes<W>));
                const gptr<vf4> src = as_global<vf4>(M.nodesW[width_slot(W)]);
                for (int k = tid; k < nTop * (kRecBytes<W> / 16); k += kWgSize) sTop[k] = src[k];
            }
        }
        if (useMask) {
            // (the same for every lane: kept in scalar registers)
            pbits = static_cast<unsigned long long>(uniform64(static_cast<int64_t>(tile_patch_bits(J, a.mask, tx, ty, ts))));
            skip = pbits == 0ull;
            // (makes the staged top levels visible)
            if constexpr (TOP) __syncthreads();
        } else if constexpr (!REC) {
            if (culled) {
                sCull[tid] = cull_rect_of(J, M, tid);
                __syncthreads();
                const int ia = J.i0 + tx * ts, ja = J.j0 + ty * ts;
                skip = !rays_reach_mesh(J, sCull, ia, min(ia + ts - 1, J.nx - 1), ja, min(ja + ts - 1, J.ny - 1),
                                        lane);
            }
        }

        using AccT = Acc<R, MODE>;
        AccT* parts = static_cast<AccT*>(a.partials);
        if (skip) {
            // every pass of the tile sums to zero: store the zero partials
            // directly (skip is the same for the whole workgroup)
            const int perSide = ts >> 4;
            if (tid < perSide * perSide) {
                const int sx = (tx * ts + (tid % perSide) * 16) >> 4, sy = (ty * ts + (tid / perSide) * 16) >> 4;
                if (sx < J.passX && sy < J.passY) parts[J.passBase + int64_t(sy) * J.passX + sx] = AccT::zero();
            }
            // every wave has read sNext of the previous tile before thread 0
            // overwrites it below (a traced tile has the barrier before its
            // pass sums)
            __syncthreads();
        } else if constexpr (POOL) {
            const int perSide = ts >> 4;
            const int nPass = perSide * perSide;
            for (int b0 = 0; b0 < nPass; b0 += kPoolRays / 256) {
                const int nb = min(kPoolRays / 256, nPass - b0);
                const unsigned nRays = static_cast<unsigned>(nb) << 8;
                // slot r -> pass b0 + r/256, and within the pass the same
                // 8x8-per-wave layout as the one-ray-per-lane path
                unsigned r = 0;
                int ri = 0, rj = 0;
                auto fetch = [&]() -> bool {
                    for (;;) {
                        r = atomicAdd(&sPool, 1u) - poolBase;
                        if (r >= nRays) return false;
                        const int pIdx = b0 + static_cast<int>(r >> 8);
                        const int q = static_cast<int>(r & 255u);
                        const int w2 = q >> 6, l2 = q & 63;
                        ri = J.i0 + tx * ts + (pIdx % perSide) * 16 + (w2 & 1) * 8 + (l2 & 7);
                        rj = J.j0 + ty * ts + (pIdx / perSide) * 16 + (w2 >> 1) * 8 + (l2 >> 3);
                        // (rays of patches the coverage mask rules out are not traced)
                        const int pb = ((pIdx / perSide) * 2 + (w2 >> 1)) * (ts >> 3) + (pIdx % perSide) * 2 + (w2 & 1);
                        if (ri < J.nx && rj < J.ny && ((pbits >> pb) & 1ull)) return true;
                        sRes[2 * r] = ST(0);
                        sRes[2 * r + 1] = ST(0);
                    }
                };
                RayState<RS> st;
                Work w{0, 0};
                bool fell = false;
                bool alive = fetch();
                auto start = [&]() {
                    if (CNT) cRays++;
                    if constexpr (MODE == 1) ray_init_mixed(a.jobs64[jid], ri, rj, st);
                    else ray_init(J, ri, rj, st);
                };
                if (alive) start();
                while (__ballot(alive) != 0ull) {
                    if (alive) {
                        bool done;
                        if constexpr (MODE == 1) {
                            done = ray_step_mixed<SD, CNT, false, WW, false, kWide>(
                                J, M, a.meshes64[J.mesh], st, sRef, sE, ov, ov64, nullptr, 0, tid, nullptr, w,
                                cQueries, cFull, a.errK, fell, sCos);
                        } else {
                            done = ray_step<R, SD, CNT, false, WW, false>(J, M, st, sRef, sE, ov, tid, nullptr, w,
                                                                   cQueries);
                        }
                        if (done) {
                            ST re = ST(0), im = ST(0);
                            if constexpr (MODE == 1) {
                                if (st.P.count > 0) st.P.cosT = sCos[tid];
                                st.P.Ld = st.P.L;
                                po_term<double, true>(a.jobs64[jid], st.P, re, im);
                                if (CNT && fell) cUndecided++;
                                fell = false;
                            } else {
                                po_term<R>(J, st.P, re, im);
                            }
                            if (CNT) cBounces += st.P.count;
                            sRes[2 * r] = re;
                            sRes[2 * r + 1] = im;
                            alive = fetch();
                            if (alive) start();
                        }
                    }
                }
                if (CNT) {
                    cVisits += w.visits;
                    cTests += w.tests;
                }
                poolBase += nRays + kWgSize;
                __syncthreads();
                for (int pp = 0; pp < nb; ++pp) {
                    AccT acc = AccT::zero();
                    const ST sr = sRes[2 * ((pp << 8) + tid)], si = sRes[2 * ((pp << 8) + tid) + 1];
                    if constexpr (MODE == 1) acc.add_ray64(sr, si);
                    else acc.add_ray(sr, si);
                    const AccT psum = wg_sum(acc, red, tid);
                    const int pIdx = b0 + pp;
                    const int sx = (tx * ts + (pIdx % perSide) * 16) >> 4, sy = (ty * ts + (pIdx / perSide) * 16) >> 4;
                    if (tid == 0 && sx < J.passX && sy < J.passY)
                        parts[J.passBase + int64_t(sy) * J.passX + sx] = psum;
                }
            }
        } else
        // a tile of side ts is covered by (ts/16)^2 passes of 16x16 rays,
        // each of four 8x8 patches (its quadrants). A wavefront sums one patch
        // at a time (a butterfly over its 64 rays); after the tile the four
        // patch sums of each pass are added in quadrant order and stored at
        // the pass's place in the job's pass grid. In timed runs the
        // wavefronts take the tile's patches one by one from an LDS counter,
        // so they reach the end of the tile together; record and counting
        // runs give wavefront w quadrant w of every pass. The sums do not
        // depend on which wavefront traced a patch.
        {
        const int perSide = ts >> 4;
        constexpr bool kDyn = !REC && !CNT;
        const int nPatch = perSide * perSide * 4;
        for (int p = kDyn ? 0 : wave;; p += kDyn ? 0 : 4) {
            if constexpr (kDyn) {
                // all 64 lanes add one (the compiler makes that a single LDS
                // add of 64 whose result lane k sees plus k), so the counter
                // advances by 64 per pop and lane 0 holds 64 x the patch. A
                // pop guarded by `lane == 0` here is threaded by the compiler
                // with the `lane == 0` store at the loop tail into a nested
                // loop in which lanes 1..63 go round again without a pop and
                // never leave it.
                const int got = atomicAdd(&sPatch, 1);
                p = __builtin_amdgcn_readfirstlane(got) >> 6;
            }
            if (p >= nPatch) break;
            const int pass = p >> 2, quad = p & 3;
            const int px = (pass % perSide) * 16, py = (pass / perSide) * 16;
            {
                AccT acc = AccT::zero();
                const int i = J.i0 + tx * ts + px + (quad & 1) * 8 + (lane & 7);
                const int j = J.j0 + ty * ts + py + (quad >> 1) * 8 + (lane >> 3);
                bool reach = true;
                if (useMask) {
                    reach = ((pbits >> (((py >> 3) + (quad >> 1)) * (ts >> 3) + (px >> 3) + (quad & 1))) & 1ull) != 0;
                } else if constexpr (!REC) {
                    if (culled) {
                        const int pi = J.i0 + tx * ts + px + (quad & 1) * 8,
                                  pj = J.j0 + ty * ts + py + (quad >> 1) * 8;
                        reach = rays_reach_mesh(J, sCull, pi, min(pi + 7, J.nx - 1), pj, min(pj + 7, J.ny - 1), lane);
                    }
                }
                do {
                    if (!reach || i >= J.nx || j >= J.ny) break;
                    if (CNT) cRays++;
                    const int ray = j * J.nx + i;
                    int32_t* ids = REC ? a.recIds + J.recIdBase + size_t(ray) * J.maxB : nullptr;
                    Work w{0, 0};
                    R re = 0, im = 0;
                    double dre = 0.0, dimg = 0.0;
                    PathOut<R> P;
                    bool fell = false;
                    if constexpr (MODE == 1) {
                        const Job<double>& D = a.jobs64[jid];
                        const PathOut<double> PD = trace_path_mixed<SD, CNT, REC, WW, TOP, W>(
                            J, D, M, a.meshes64[J.mesh], i, j, sRef, sE, ov, ov64, sTop, nTop, tid, ids, w, cQueries,
                            cFull, a.errK, fell, sCos);
                        po_term<double, true>(D, PD, dre, dimg);
                        P.count = PD.count;
                        P.Ld = PD.Ld;
                    } else {
                        P = trace_path<R, SD, CNT, REC, WW>(J, M, i, j, sRef, sE, ov, tid, ids, w, cQueries);
                        po_term<R>(J, P, re, im);
                    }
                    if (CNT) {
                        cVisits += w.visits;
                        cTests += w.tests;
                    }
                    if constexpr (MODE == 1) {
                        if (CNT && fell) cUndecided++;
                        if (REC) a.recFlag[J.recBase + ray] = fell ? 1 : 0;
                    }
                    if (CNT) cBounces += P.count;
                    if (REC) {
                        for (int k = P.count; k < J.maxB; ++k) ids[k] = -1;
                        const int64_t rr = J.recBase + ray;
                        a.recCount[rr] = P.count;
                        a.recL[rr] = MODE == 1 ? P.Ld : double(P.L);
                        a.recReIm[2 * rr] = MODE == 1 ? dre : double(re);
                        a.recReIm[2 * rr + 1] = MODE == 1 ? dimg : double(im);
                    }
                    if constexpr (MODE == 1) acc.add_ray64(dre, dimg);
                    else acc.add_ray(re, im);
                } while (false);
                acc.wave();
                if (lane == 0) sPass[pass * (kWgSize / 64) + quad] = acc;
            }
        }
        __syncthreads();
        // every wavefront has taken its last patch of the tile: the counter
        // restarts for the next one (made visible by the next tile's barriers)
        if (kDyn && tid == 0) sPatch = 0;
        if (tid < perSide * perSide) {
            AccT r = sPass[tid * (kWgSize / 64)];
#pragma unroll
            for (int k = 1; k < kWgSize / 64; ++k) r.add(sPass[tid * (kWgSize / 64) + k]);
            const int sx = (tx * ts + (tid % perSide) * 16) >> 4, sy = (ty * ts + (tid / perSide) * 16) >> 4;
            if (sx < J.passX && sy < J.passY) parts[J.passBase + int64_t(sy) * J.passX + sx] = r;
        }
        }

        // the next tile (persistent), broadcast through LDS; the job sums
        // are formed after the launch (finalize_kernel), so a tile ends
        // without fences: a device-scope release per tile writes the XCD's
        // dirty L2 lines (the traversal stacks' scratch) back every time
        int64_t next = a.totalTiles;
        if constexpr (PERSIST) {
            if (tid == 0) sNext = static_cast<int64_t>(atomicAdd(a.tileCounter, 1ull));
            __syncthreads();
            next = uniform64(sNext);
        }
        tile = next;
    }

    if (CNT) {
        unsigned long long c[7] = {cRays, cVisits, cTests, cBounces, cQueries, cUndecided, cFull};
#pragma unroll
        for (int q = 0; q < 7; ++q) {
            unsigned long long v = wave_sum(c[q]);
            if (lane == 0) atomicAdd(&a.counters[q], v);
        }
    }
}

// Per-job sums of the pass partials in grid order, one workgroup per job
// (grid-strided), after the trace launches of a run: the order is fixed by
// the job's pass grid alone, so sigma does not depend on the launch shape.
template <typename R, int MODE>
__global__ __launch_bounds__(kWgSize) void finalize_kernel(const Job<R>* jobs, int nJobs, const void* partials,
                                                          double* jobOut) {
    using AccT = Acc<R, MODE>;
    __shared__ AccT red[kWgSize / 64];
    const AccT* parts = static_cast<const AccT*>(partials);
    for (int jid = blockIdx.x; jid < nJobs; jid += gridDim.x) {
        const Job<R>& J = jobs[jid];
        AccT acc = AccT::zero();
        const int nPass = J.passX * J.passY;
        for (int q = threadIdx.x; q < nPass; q += kWgSize) acc.add(parts[J.passBase + q]);
        const AccT tot = wg_sum(acc, red, static_cast<int>(threadIdx.x));
        if (threadIdx.x == 0) {
            jobOut[2 * jid] = tot.dre();
            jobOut[2 * jid + 1] = tot.dim();
        }
    }
}

__global__ void zero_words(unsigned* p, size_t n) {
    size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ------------------------------------------------------------- adjoint
// Frozen-path adjoint of sigma (no counterpart in the reference). For each
// ray the recorded hit faces fix the path; the bounce geometry is recomputed
// in FP64 exactly as the mixed/FP64 ray computes it (rtKernels.cu:13-41,
// 144-152) and differentiated in reverse:
//   sigma = 4 pi |S|^2, S = sum_r c_r, c_r = mag_r (-sin ph_r, cos ph_r),
//   ph_r = 2 k L_r (mod 2 pi), mag_r = pre cos(theta_r) R^n,
//   L_r = sum_b t_b, t_b = ((v0 - o_b).c_b) / (d_b.c_b) (the Moller-Trumbore
//   distance in plane form), c_b = e1 x e2, n_b = c_b / |c_b|,
//   d_{b+1} = w/|w|, w = d_b - 2 (d_b.n_b) n_b, o_{b+1} = p_b + 1e-4 n_b.
// d sigma / d f follows from k = 2 pi f / c0 and pre ~ k (poKernels.cu:98-101).

struct AdjBounce {
    V3<double> v0, e1, e2, c, n, p, dn;
    double t, cl, s, wl;
};

__device__ __forceinline__ V3<double> vtx64(const AdjArgs& a, int f, int k) {
    const unsigned vi = a.idx[3 * f + k];
    return {a.verts[3 * size_t(vi)], a.verts[3 * size_t(vi) + 1], a.verts[3 * size_t(vi) + 2]};
}

// Geometry of one bounce on face f from (o, d), with the forward's operations.
__device__ __forceinline__ AdjBounce adj_bounce(const AdjArgs& a, int f, V3<double> o, V3<double> d) {
    AdjBounce b;
    b.v0 = vtx64(a, f, 0);
    b.e1 = sub(vtx64(a, f, 1), b.v0);
    b.e2 = sub(vtx64(a, f, 2), b.v0);
    V3<double> pv = cross3(d, b.e2);
    double det = dot3(b.e1, pv);
    double invd = 1.0 / det;
    V3<double> q = cross3(sub(o, b.v0), b.e1);
    b.t = dot3(b.e2, q) * invd;
    b.c = cross3(b.e1, b.e2);
    b.cl = sqrt(dot3(b.c, b.c));
    b.n = unit3(b.c);
    b.p = axpy(b.t, d, o);
    b.s = dot3(d, b.n);
    V3<double> wv = axpy(-(2.0 * b.s), b.n, d);
    b.wl = sqrt(dot3(wv, wv));
    b.dn = unit3(wv);
    return b;
}

__device__ __forceinline__ void adj_add(double* p, V3<double> g, double s) {
    atomicAdd(p, g.x * s);
    atomicAdd(p + 1, g.y * s);
    atomicAdd(p + 2, g.z * s);
}

__global__ __launch_bounds__(256) void adjoint_kernel(AdjArgs a) {
    __shared__ double red[4];
    // the job of this block: the last one whose first block is <= blockIdx.x
    int lo = 0, hi = a.nJobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.jobs[mid].blockBase <= static_cast<int>(blockIdx.x)) lo = mid;
        else hi = mid - 1;
    }
    const AdjJob& G = a.jobs[lo];
    const int ray = (static_cast<int>(blockIdx.x) - G.blockBase) * 256 + static_cast<int>(threadIdx.x);
    double dfr = 0.0;
    const int B = ray < G.nRays ? a.cnt[G.rayOff + ray] : 0;
    if (B > 0) {
        const Job<double>& D = G.D;
        const int32_t* fids = a.ids + G.idOff + size_t(ray) * D.maxB;
        const int i = ray % D.nx, j = ray / D.nx;
        const V3<double> o0 = ray_origin(D, i, j);
        const V3<double> d0{D.dir[0], D.dir[1], D.dir[2]};
        // forward: path length and the last hit's cos(theta)
        V3<double> o = o0, d = d0;
        double L = 0.0, cosT = 0.0;
        for (int b = 0; b < B; ++b) {
            AdjBounce g = adj_bounce(a, fids[b], o, d);
            cosT = dot3(g.n, V3<double>{-d.x, -d.y, -d.z});
            L += g.t;
            d = g.dn;
            o = axpy(1e-4, g.n, g.p);
        }
        if (cosT > 0.0) {
            double refl = 1.0;
            for (int b = 0; b < B; ++b) refl *= D.refl;
            const double ph = fmod((2.0 * D.kd) * L, D.twoPid);
            const double mag = (D.pre * cosT) * refl;
            double sn, cs;
            sincos(ph, &sn, &cs);
            const double cre = -mag * sn, cim = mag * cs;
            // d sigma / d (Re c, Im c) = 8 pi (Re S, Im S)
            const double gre = 8.0 * M_PI * a.sums[2 * lo], gim = 8.0 * M_PI * a.sums[2 * lo + 1];
            const double rot = gre * (-cim) + gim * cre;  // g . (i c)
            const double par = gre * cre + gim * cim;     // g . c
            dfr = (par + 2.0 * L * D.kd * rot) / G.freq;
            const double gL = G.w * 2.0 * D.kd * rot;
            const double gC = G.w * par / cosT;
            if (G.w != 0.0) {
                V3<double> ob{0, 0, 0}, db{0, 0, 0};  // adjoints of o_{b+1}, d_{b+1}
                for (int b = B - 1; b >= 0; --b) {
                    // o_b, d_b: forward again from the ray start
                    V3<double> ob_ = o0, db_ = d0;
                    for (int k = 0; k < b; ++k) {
                        AdjBounce g = adj_bounce(a, fids[k], ob_, db_);
                        db_ = g.dn;
                        ob_ = axpy(1e-4, g.n, g.p);
                    }
                    const AdjBounce g = adj_bounce(a, fids[b], ob_, db_);
                    // o_{b+1} = p + 1e-4 n
                    V3<double> pbar = ob;
                    V3<double> nbar = axpy(1e-4, ob, V3<double>{0, 0, 0});
                    // d_{b+1} = w / |w|
                    const double pr = dot3(db, g.dn);
                    V3<double> wbar{(db.x - pr * g.dn.x) / g.wl, (db.y - pr * g.dn.y) / g.wl,
                                    (db.z - pr * g.dn.z) / g.wl};
                    // w = d - 2 s n
                    V3<double> dbar = wbar;
                    const double sbar = -2.0 * dot3(g.n, wbar);
                    nbar = axpy(-2.0 * g.s, wbar, nbar);
                    // s = d . n
                    dbar = axpy(sbar, g.n, dbar);
                    nbar = axpy(sbar, db_, nbar);
                    // cos(theta) = -n . d at the last hit
                    if (b == B - 1) {
                        nbar = axpy(-gC, db_, nbar);
                        dbar = axpy(-gC, g.n, dbar);
                    }
                    // p = o + t d
                    V3<double> obar = pbar;
                    const double tbar = gL + dot3(pbar, db_);
                    dbar = axpy(g.t, pbar, dbar);
                    // t = ((v0 - o).c) / (d.c)
                    const double den = dot3(db_, g.c);
                    const double tr = tbar / den;
                    obar = axpy(-tr, g.c, obar);
                    dbar = axpy(-tr * g.t, g.c, dbar);
                    V3<double> v0bar = axpy(tr, g.c, V3<double>{0, 0, 0});
                    V3<double> cbar = axpy(tr, sub(g.v0, g.p), V3<double>{0, 0, 0});
                    // n = c / |c|
                    const double nn = dot3(nbar, g.n);
                    cbar = axpy(1.0 / g.cl, V3<double>{nbar.x - nn * g.n.x, nbar.y - nn * g.n.y, nbar.z - nn * g.n.z},
                                cbar);
                    // c = e1 x e2
                    const V3<double> e1bar = cross3(g.e2, cbar), e2bar = cross3(cbar, g.e1);
                    v0bar = sub(v0bar, V3<double>{e1bar.x + e2bar.x, e1bar.y + e2bar.y, e1bar.z + e2bar.z});
                    const int f = fids[b];
                    adj_add(a.dvert + 3 * size_t(a.idx[3 * f]), v0bar, a.scale);
                    adj_add(a.dvert + 3 * size_t(a.idx[3 * f + 1]), e1bar, a.scale);
                    adj_add(a.dvert + 3 * size_t(a.idx[3 * f + 2]), e2bar, a.scale);
                    ob = obar;
                    db = dbar;
                }
            }
        }
    }
    // d sigma / d f: wave sums, then one atomic per workgroup
    dfr = wave_sum(dfr);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = dfr;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(a.dfreq + lo, red[0] + red[1] + red[2] + red[3]);
}

template <typename R, int MODE, int SD, bool REC, bool CNT, bool PERSIST, bool POOL, bool WW, bool TOP = false,
          int W = kWide>
hipError_t launch_one(const TraceArgs<R>& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL((trace_kernel<R, MODE, SD, REC, CNT, PERSIST, POOL, WW, TOP, W>), dim3(grid), dim3(kWgSize), 0,
                       s, a);
    return hipGetLastError();
}

// Timed mixed-mode launches (one ray per lane) on the W-wide tree; top:
// stage the tree's top levels in LDS.
template <int SD, int W>
hipError_t launch_mixed_w(const TraceArgs<float>& a, bool persist, bool ww, bool top, int grid, hipStream_t s) {
    if (top) {
        if (ww)
            return persist ? launch_one<float, 1, SD, false, false, true, false, true, true, W>(a, grid, s)
                           : launch_one<float, 1, SD, false, false, false, false, true, true, W>(a, grid, s);
        return persist ? launch_one<float, 1, SD, false, false, true, false, false, true, W>(a, grid, s)
                       : launch_one<float, 1, SD, false, false, false, false, false, true, W>(a, grid, s);
    }
    if (ww)
        return persist ? launch_one<float, 1, SD, false, false, true, false, true, false, W>(a, grid, s)
                       : launch_one<float, 1, SD, false, false, false, false, true, false, W>(a, grid, s);
    return persist ? launch_one<float, 1, SD, false, false, true, false, false, false, W>(a, grid, s)
                   : launch_one<float, 1, SD, false, false, false, false, false, false, W>(a, grid, s);
}

// Record and counting runs use the while-while shape and the default node
// width only (their outputs do not depend on either); width: node width of
// the mixed mode's tree in timed runs without the ray pool.
template <typename R, int MODE, int SD, bool POOL>
hipError_t launch_pool(const TraceArgs<R>& a, bool cnt, bool persist, bool ww, bool top, int width, int grid,
                       hipStream_t s) {
    if (cnt) return persist ? launch_one<R, MODE, SD, false, true, true, POOL, true>(a, grid, s)
                            : launch_one<R, MODE, SD, false, true, false, POOL, true>(a, grid, s);
    if constexpr (MODE == 1 && !POOL) {
        if (width == 2) return launch_mixed_w<SD, 2>(a, persist, ww, top, grid, s);
        if (width == 8) return launch_mixed_w<SD, 8>(a, persist, ww, top, grid, s);
        return launch_mixed_w<SD, kWide>(a, persist, ww, top, grid, s);
    }
    if (ww)
        return persist ? launch_one<R, MODE, SD, false, false, true, POOL, true>(a, grid, s)
                       : launch_one<R, MODE, SD, false, false, false, POOL, true>(a, grid, s);
    return persist ? launch_one<R, MODE, SD, false, false, true, POOL, false>(a, grid, s)
                   : launch_one<R, MODE, SD, false, false, false, POOL, false>(a, grid, s);
}

template <typename R, int MODE, int SD>
hipError_t launch_sd(const TraceArgs<R>& a, bool rec, bool cnt, bool persist, bool pool, bool ww, bool top, int width,
                     int grid, hipStream_t s) {
    if (rec) return launch_one<R, MODE, SD, true, false, false, false, true>(a, grid, s);
    return pool ? launch_pool<R, MODE, SD, true>(a, cnt, persist, ww, false, kWide, grid, s)
                : launch_pool<R, MODE, SD, false>(a, cnt, persist, ww, top, width, grid, s);
}

}  // namespace

// Host entry used by the driver. sd: traversal stack entries kept in LDS
// per ray (8, 16 or 32); deeper entries spill to scratch. ww: while-while
// traversal loop (else if-if). width: node width (2, 4 or 8) of the mixed
// mode's tree in timed runs without the ray pool.
template <typename R>
hipError_t launch_trace(const TraceArgs<R>& a, int mode, int sd, bool rec, bool cnt, bool persist, bool pool, bool ww,
                        bool top, int width, int grid, hipStream_t s) {
    if constexpr (std::is_same_v<R, float>) {
        if (mode == 1) {
            if (sd <= 8) return launch_sd<float, 1, 8>(a, rec, cnt, persist, pool, ww, top, width, grid, s);
            if (sd <= 16) return launch_sd<float, 1, 16>(a, rec, cnt, persist, pool, ww, top, width, grid, s);
            return launch_sd<float, 1, 32>(a, rec, cnt, persist, pool, ww, top, width, grid, s);
        }
    }
    if (sd <= 8) return launch_sd<R, 0, 8>(a, rec, cnt, persist, pool, ww, false, kWide, grid, s);
    if (sd <= 16) return launch_sd<R, 0, 16>(a, rec, cnt, persist, pool, ww, false, kWide, grid, s);
    return launch_sd<R, 0, 32>(a, rec, cnt, persist, pool, ww, false, kWide, grid, s);
}

template hipError_t launch_trace<float>(const TraceArgs<float>&, int, int, bool, bool, bool, bool, bool, bool, int,
                                        int, hipStream_t);
template hipError_t launch_trace<double>(const TraceArgs<double>&, int, int, bool, bool, bool, bool, bool, bool, int,
                                         int, hipStream_t);

template <typename R>
hipError_t launch_finalize(const Job<R>* jobs, int nJobs, int mode, const void* partials, double* jobOut,
                           hipStream_t s) {
    if (nJobs <= 0) return hipSuccess;
    const dim3 grid(static_cast<unsigned>(nJobs < 65535 ? nJobs : 65535));
    if constexpr (std::is_same_v<R, float>) {
        if (mode == 1) {
            hipLaunchKernelGGL((finalize_kernel<float, 1>), grid, dim3(kWgSize), 0, s, jobs, nJobs, partials, jobOut);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((finalize_kernel<R, 0>), grid, dim3(kWgSize), 0, s, jobs, nJobs, partials, jobOut);
    return hipGetLastError();
}
template hipError_t launch_finalize<float>(const Job<float>*, int, int, const void*, double*, hipStream_t);
template hipError_t launch_finalize<double>(const Job<double>*, int, int, const void*, double*, hipStream_t);

hipError_t launch_adjoint(const AdjArgs& a, int blocks, hipStream_t s) {
    if (blocks <= 0 || a.nJobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(adjoint_kernel, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_zero(unsigned* p, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    unsigned grid = static_cast<unsigned>((n + 255) / 256);
    hipLaunchKernelGGL(zero_words, dim3(grid), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

// Coverage masks of a batched run's jobs; maxChunks: the most LDS chunks
// of kCovLdsWords mask words any job of the run has.
template <typename R>
hipError_t launch_coverage(const Job<R>* jobs, const Mesh<R>* meshes, int nJobs, int maxChunks,
                           unsigned long long* mask, hipStream_t s) {
    if (nJobs <= 0 || maxChunks <= 0) return hipSuccess;
    const dim3 grid(static_cast<unsigned>(maxChunks), static_cast<unsigned>(nJobs < 65535 ? nJobs : 65535));
    hipLaunchKernelGGL((coverage_kernel<R>), grid, dim3(256), 0, s, jobs, meshes, nJobs, mask);
    return hipGetLastError();
}
template hipError_t launch_coverage<float>(const Job<float>*, const Mesh<float>*, int, int, unsigned long long*,
                                           hipStream_t);
template hipError_t launch_coverage<double>(const Job<double>*, const Mesh<double>*, int, int, unsigned long long*,
                                            hipStream_t);

template <int SD, int W> const void* mixed_fn(bool top) {
    return top ? reinterpret_cast<const void*>(&trace_kernel<float, 1, SD, false, false, true, false, true, true, W>)
               : reinterpret_cast<const void*>(&trace_kernel<float, 1, SD, false, false, true, false, true, false, W>);
}

template <typename R, int MODE, int SD> int occ_one(bool pool, bool top, int width) {
    int n = 0;
    const void* f = pool ? reinterpret_cast<const void*>(&trace_kernel<R, MODE, SD, false, false, true, true, true>)
                         : reinterpret_cast<const void*>(&trace_kernel<R, MODE, SD, false, false, true, false, true>);
    if constexpr (MODE == 1) {
        if (!pool) f = width == 2 ? mixed_fn<SD, 2>(top) : (width == 8 ? mixed_fn<SD, 8>(top) : mixed_fn<SD, kWide>(top));
    }
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, kWgSize, 0);
    return e == hipSuccess ? n : 1;
}

// Resident workgroups per CU of the persistent kernel variant.
template <typename R> int trace_occupancy(int mode, int sd, bool pool, bool top, int width) {
    if constexpr (std::is_same_v<R, float>) {
        if (mode == 1)
            return sd <= 8 ? occ_one<float, 1, 8>(pool, top, width)
                           : (sd <= 16 ? occ_one<float, 1, 16>(pool, top, width)
                                       : occ_one<float, 1, 32>(pool, top, width));
    }
    return sd <= 8 ? occ_one<R, 0, 8>(pool, false, kWide)
                   : (sd <= 16 ? occ_one<R, 0, 16>(pool, false, kWide) : occ_one<R, 0, 32>(pool, false, kWide));
}
template int trace_occupancy<float>(int, int, bool, bool, int);
template int trace_occupancy<double>(int, int, bool, bool, int);

// ------------------------------------------------ fixed scratch memory
// Device storage of the allocation-free sweep (sbr_sweep_noalloc): its job
// records, tile table, per-pass partials, job sums and counters live in this
// array instead of buffers the library allocates. Each device that loads the
// library holds its own copy until the process ends. In host code the name
// `scratch` is only a stand-in for the symbol, not a device address: the
// address is looked up once per device (scratch_init, with that device
// current) and only that address is handed to kernels and copies.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratchMutex;
void* g_scratchPtr[kMaxScratchDevices] = {};
}  // namespace

hipError_t scratch_init(int device) {
    if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (g_scratchPtr[device]) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratchPtr[device] = p;
    return hipSuccess;
}

void* scratch_ptr(int device) {
    if (device < 0 || device >= kMaxScratchDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    return g_scratchPtr[device];
}

// Size in bytes of one per-pass partial for (precision mode).
size_t partial_bytes(bool fp64, int mode) {
    if (fp64) return sizeof(Acc<double, 0>);
    return mode == 1 ? sizeof(Acc<float, 1>) : sizeof(Acc<float, 0>);
}

}  // namespace sbr
