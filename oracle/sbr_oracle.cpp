// SPDX-License-Identifier: MIT
//
// TEST INFRASTRUCTURE — the CPU oracle for the per-angle SBR ray path.
//
// This file is a plain C++ restatement of the reference algorithm. Only the
// test suite, the smoke check and the CPU-baseline leg of the benchmark load
// it; the product library never links it.
//
// What it restates (citations are into the reference repository):
//   * triangle soup assembly        src/app/rtRcsEntry.cu:118-148
//   * SAH/median BVH build          src/scene/bvhBuilder.cpp:22-360
//   * per-angle launch frame        src/app/sweep.cu:336-364
//   * physics constants             src/app/sweep.cu:50-55
//   * ray grid                      src/cuda/rtKernels.cu:113-124
//   * Moller-Trumbore               src/cuda/rtKernels.cu:10-44
//   * slab test                     src/cuda/rtKernels.cu:46-61
//   * stack DFS closest hit         src/cuda/rtKernels.cu:63-100
//   * multi-bounce loop             src/cuda/rtKernels.cu:131-158
//   * PO term                       src/cuda/poKernels.cu:75-110
//   * sigma / dBsm                  src/app/sweep.cu:228-231
//   * vec3 arithmetic               include/RT/vec3.h:48,99-119,171-174
//
// Arithmetic policy. Every routine is templated on the real type R and on a
// flag F. With F=false every a*b+c is a separate multiply and add (the file
// is compiled with -ffp-contract=off), i.e. the reference source read
// literally; <double,false> is the FP64 oracle. With F=true the same formulas
// use fused multiply-adds at fixed places (dot: fma(z,z',fma(x,x',y*y')),
// cross: fma(a,b,-(c*d)), ray origin / hit point / reflection / offset:
// one fma per component). <float,true> is the FP32 twin that the GPU FP32
// kernel must match bit for bit.
//
// Pins (see tests/test_oracle_*.py): the BVH against the reference builder
// compiled from its own sources (oracle/_ref), closest hits against brute
// force over all triangles, Moller-Trumbore known answers, the flat-plate
// closed form sigma = 4*pi*A^2/lambda^2, trihedral retro-reflection, the
// sphere optical limit / Mie series, and the reference's sphere result file.
// sin/cos come from the host libm: parity of the PO phase below the ulp level
// is unpinned (tolerance only).

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <type_traits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

template <typename R> struct V3 {
    R x, y, z;
};

template <typename R> inline V3<R> mk(R x, R y, R z) { return V3<R>{x, y, z}; }
template <typename R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename R> inline V3<R> neg(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <typename R> inline R comp(V3<R> a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// vec3.h:110-112 : x*x' + y*y' + z*z'
template <typename R, bool F> inline R dot(V3<R> a, V3<R> b) {
    if (F) return std::fma(a.z, b.z, std::fma(a.x, b.x, a.y * b.y));
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
// vec3.h:115-119
template <typename R, bool F> inline V3<R> cross(V3<R> a, V3<R> b) {
    if (F)
        return {std::fma(a.y, b.z, -(a.z * b.y)), -std::fma(a.x, b.z, -(a.z * b.x)),
                std::fma(a.x, b.y, -(a.y * b.x))};
    return {a.y * b.z - a.z * b.y, -(a.x * b.z - a.z * b.x), a.x * b.y - a.y * b.x};
}
// vec3.h:48 length, :171-174 unitVector = v / length (three divisions)
template <typename R, bool F> inline V3<R> unit(V3<R> v) {
    R len = std::sqrt(dot<R, F>(v, v));
    return {v.x / len, v.y / len, v.z / len};
}
// s*a + b per component
template <typename R, bool F> inline V3<R> axpy(R s, V3<R> a, V3<R> b) {
    if (F) return {std::fma(s, a.x, b.x), std::fma(s, a.y, b.y), std::fma(s, a.z, b.z)};
    return {s * a.x + b.x, s * a.y + b.y, s * a.z + b.z};
}

template <typename R> struct Tri {
    V3<R> v0, v1, v2;
};

// include/scene/bvh.h:13-21
template <typename R> struct Node {
    V3<R> bmin, bmax;
    int left, right, firstTri, triCount;
};

template <typename R> struct Scene {
    std::vector<Tri<R>> tris;  // BVH order
    std::vector<int> perm;     // BVH position -> original face index
    std::vector<Node<R>> nodes;
    int root = 0;
    int depth = 0;
};

template <typename R> inline R flt_max() { return std::numeric_limits<R>::max(); }

// ---------------------------------------------------------------- BVH build
// bvhBuilder.cpp:22-62
template <typename R> inline V3<R> vmin(V3<R> a, V3<R> b) {
    return {std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)};
}
template <typename R> inline V3<R> vmax(V3<R> a, V3<R> b) {
    return {std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)};
}
template <typename R> inline V3<R> tmin3(const Tri<R>& t) { return vmin(vmin(t.v0, t.v1), t.v2); }
template <typename R> inline V3<R> tmax3(const Tri<R>& t) { return vmax(vmax(t.v0, t.v1), t.v2); }
template <typename R> inline V3<R> centroid(const Tri<R>& t) {
    V3<R> s = (t.v0 + t.v1) + t.v2;
    return {s.x / R(3.0), s.y / R(3.0), s.z / R(3.0)};
}
template <typename R> inline R area(V3<R> lo, V3<R> hi) {
    V3<R> d = hi - lo;
    R dx = std::fmax(R(0), d.x), dy = std::fmax(R(0), d.y), dz = std::fmax(R(0), d.z);
    return R(2.0) * (dx * dy + dy * dz + dz * dx);
}
template <typename R> inline int max_axis(V3<R> v) {
    if (v.x >= v.y && v.x >= v.z) return 0;
    if (v.y >= v.z) return 1;
    return 2;
}

template <typename R> struct Builder {
    std::vector<Tri<R>>& tris;
    std::vector<int>& perm;
    std::vector<Node<R>>& nodes;
    int leafSize, binCount;
    bool sah;
    int maxDepth = 0;

    void bounds(int s, int e, V3<R>& lo, V3<R>& hi) {
        lo = mk(flt_max<R>(), flt_max<R>(), flt_max<R>());
        hi = mk(-flt_max<R>(), -flt_max<R>(), -flt_max<R>());
        for (int i = s; i < e; ++i) {
            lo = vmin(lo, tmin3(tris[i]));
            hi = vmax(hi, tmax3(tris[i]));
        }
    }
    void cbounds(int s, int e, V3<R>& lo, V3<R>& hi) {
        lo = mk(flt_max<R>(), flt_max<R>(), flt_max<R>());
        hi = mk(-flt_max<R>(), -flt_max<R>(), -flt_max<R>());
        for (int i = s; i < e; ++i) {
            V3<R> c = centroid(tris[i]);
            lo = vmin(lo, c);
            hi = vmax(hi, c);
        }
    }
    // The reference reorders Triangle structs; the oracle reorders an index
    // array with the same std:: algorithm and predicate, then gathers. The
    // libstdc++ algorithms used only look at predicate / comparison results,
    // so the resulting permutation is the same.
    template <typename Fn> int part(int s, int e, Fn pred) {
        std::vector<int> idx(e - s);
        for (int i = s; i < e; ++i) idx[i - s] = i;
        auto mid = std::partition(idx.begin(), idx.end(), [&](int i) { return pred(tris[i]); });
        apply(s, idx);
        return s + int(mid - idx.begin());
    }
    void apply(int s, const std::vector<int>& idx) {
        std::vector<Tri<R>> t(idx.size());
        std::vector<int> p(idx.size());
        for (size_t k = 0; k < idx.size(); ++k) {
            t[k] = tris[idx[k]];
            p[k] = perm[idx[k]];
        }
        for (size_t k = 0; k < idx.size(); ++k) {
            tris[s + k] = t[k];
            perm[s + k] = p[k];
        }
    }
    // bvhBuilder.cpp:93-113
    bool split_median(int s, int e, int& mid) {
        V3<R> lo, hi;
        cbounds(s, e, lo, hi);
        int axis = max_axis(hi - lo);
        mid = s + (e - s) / 2;
        std::vector<int> idx(e - s);
        for (int i = s; i < e; ++i) idx[i - s] = i;
        std::nth_element(idx.begin(), idx.begin() + (mid - s), idx.end(), [&](int a, int b) {
            return comp(centroid(tris[a]), axis) < comp(centroid(tris[b]), axis);
        });
        apply(s, idx);
        return mid > s && mid < e;
    }
    // bvhBuilder.cpp:115-243
    bool split_sah(int s, int e, V3<R> blo, V3<R> bhi, int& mid) {
        int n = e - s;
        V3<R> clo, chi;
        cbounds(s, e, clo, chi);
        V3<R> ext = chi - clo;
        R parentArea = area(blo, bhi);
        if (parentArea <= R(0)) return false;
        int bestAxis = -1, bestSplit = -1;
        R bestCost = flt_max<R>();
        for (int axis = 0; axis < 3; ++axis) {
            R extent = comp(ext, axis);
            if (extent <= R(1.0e-6)) continue;
            std::vector<V3<R>> bl(binCount, mk(flt_max<R>(), flt_max<R>(), flt_max<R>()));
            std::vector<V3<R>> bh(binCount, mk(-flt_max<R>(), -flt_max<R>(), -flt_max<R>()));
            std::vector<int> bc(binCount, 0);
            for (int i = s; i < e; ++i) {
                R off = (comp(centroid(tris[i]), axis) - comp(clo, axis)) / extent;
                int b = static_cast<int>(off * binCount);
                if (b < 0) b = 0;
                else if (b >= binCount) b = binCount - 1;
                bl[b] = vmin(bl[b], tmin3(tris[i]));
                bh[b] = vmax(bh[b], tmax3(tris[i]));
                bc[b] += 1;
            }
            std::vector<V3<R>> ll(binCount), lh(binCount), rl(binCount), rh(binCount);
            std::vector<int> lc(binCount), rc(binCount);
            V3<R> lo = mk(flt_max<R>(), flt_max<R>(), flt_max<R>());
            V3<R> hi = mk(-flt_max<R>(), -flt_max<R>(), -flt_max<R>());
            int cnt = 0;
            for (int i = 0; i < binCount; ++i) {
                if (bc[i] > 0) {
                    lo = vmin(lo, bl[i]);
                    hi = vmax(hi, bh[i]);
                    cnt += bc[i];
                }
                ll[i] = lo; lh[i] = hi; lc[i] = cnt;
            }
            lo = mk(flt_max<R>(), flt_max<R>(), flt_max<R>());
            hi = mk(-flt_max<R>(), -flt_max<R>(), -flt_max<R>());
            cnt = 0;
            for (int i = binCount - 1; i >= 0; --i) {
                if (bc[i] > 0) {
                    lo = vmin(lo, bl[i]);
                    hi = vmax(hi, bh[i]);
                    cnt += bc[i];
                }
                rl[i] = lo; rh[i] = hi; rc[i] = cnt;
            }
            for (int i = 0; i < binCount - 1; ++i) {
                int nl = lc[i], nr = rc[i + 1];
                if (nl == 0 || nr == 0) continue;
                R cost = (area(ll[i], lh[i]) * nl + area(rl[i + 1], rh[i + 1]) * nr) / parentArea;
                if (cost < bestCost) {
                    bestCost = cost;
                    bestAxis = axis;
                    bestSplit = i;
                }
            }
        }
        if (bestAxis < 0) return false;
        if (bestCost >= static_cast<R>(n)) return false;
        R extent = comp(ext, bestAxis);
        if (extent <= R(1.0e-6)) return false;
        R frac = static_cast<R>(bestSplit + 1) / static_cast<R>(binCount);
        R splitPos = comp(clo, bestAxis) + extent * frac;
        mid = part(s, e, [&](const Tri<R>& t) { return comp(centroid(t), bestAxis) < splitPos; });
        return mid > s && mid < e;
    }
    // bvhBuilder.cpp:250-312, serial: node indices in pre-order
    int build(int s, int e, int depth) {
        if (depth > maxDepth) maxDepth = depth;
        V3<R> lo, hi;
        bounds(s, e, lo, hi);
        int id = int(nodes.size());
        nodes.push_back(Node<R>{lo, hi, -1, -1, s, e - s});
        if (e - s <= leafSize) return id;
        int mid = -1;
        bool ok = false;
        if (sah) ok = split_sah(s, e, lo, hi, mid);
        if (!ok) ok = split_median(s, e, mid);
        if (!ok) return id;
        int l = build(s, mid, depth + 1);
        int r = build(mid, e, depth + 1);
        nodes[id].left = l;
        nodes[id].right = r;
        nodes[id].firstTri = 0;
        nodes[id].triCount = 0;
        return id;
    }
};

template <typename R>
Scene<R>* make_scene(const float* xyz, int nVerts, const uint32_t* idx, int nTris, double scale,
                     int leafSize, int binCount, int sah) {
    auto* sc = new Scene<R>();
    R s = static_cast<R>(scale);
    sc->tris.resize(nTris);
    sc->perm.resize(nTris);
    for (int f = 0; f < nTris; ++f) {
        uint32_t i0 = idx[3 * f], i1 = idx[3 * f + 1], i2 = idx[3 * f + 2];
        if (i0 >= uint32_t(nVerts) || i1 >= uint32_t(nVerts) || i2 >= uint32_t(nVerts)) {
            delete sc;
            return nullptr;
        }
        auto vtx = [&](uint32_t i) {
            return mk(static_cast<R>(xyz[3 * i]) * s, static_cast<R>(xyz[3 * i + 1]) * s,
                      static_cast<R>(xyz[3 * i + 2]) * s);
        };
        sc->tris[f] = Tri<R>{vtx(i0), vtx(i1), vtx(i2)};
        sc->perm[f] = f;
    }
    if (nTris > 0) {
        if (leafSize < 1) leafSize = 1;
        if (binCount < 2) binCount = 2;
        Builder<R> b{sc->tris, sc->perm, sc->nodes, leafSize, binCount, sah != 0};
        sc->nodes.reserve(size_t(2) * nTris);
        sc->root = b.build(0, nTris, 0);
        sc->depth = b.maxDepth;
    }
    return sc;
}

// ----------------------------------------------------------------- tracing
template <typename R> struct Hit {
    R t, u, v;
    int tri;  // BVH position
};

// rtKernels.cu:10-44. Returns true and fills t,u,v on acceptance.
template <typename R, bool F>
inline bool hit_tri(const Tri<R>& tri, V3<R> o, V3<R> d, R tMin, R tMax, R& t, R& uu, R& vv) {
    const R eps = R(1.0e-6);
    V3<R> e1 = tri.v1 - tri.v0, e2 = tri.v2 - tri.v0;
    V3<R> p = cross<R, F>(d, e2);
    R det = dot<R, F>(e1, p);
    if (std::fabs(det) < eps) return false;
    R inv = R(1.0) / det;
    V3<R> tv = o - tri.v0;
    R u = dot<R, F>(tv, p) * inv;
    if (u < R(0) || u > R(1)) return false;
    V3<R> q = cross<R, F>(tv, e1);
    R v = dot<R, F>(d, q) * inv;
    if (v < R(0) || u + v > R(1)) return false;
    R tt = dot<R, F>(e2, q) * inv;
    if (tt < tMin || tt > tMax) return false;
    t = tt; uu = u; vv = v;
    return true;
}

// rtKernels.cu:46-61 (1/d recomputed per node in the reference; same values)
template <typename R>
inline bool hit_box(const Node<R>& n, V3<R> o, V3<R> inv, R tMin, R tMax) {
    V3<R> t0 = {(n.bmin.x - o.x) * inv.x, (n.bmin.y - o.y) * inv.y, (n.bmin.z - o.z) * inv.z};
    V3<R> t1 = {(n.bmax.x - o.x) * inv.x, (n.bmax.y - o.y) * inv.y, (n.bmax.z - o.z) * inv.z};
    R lo = std::fmax(std::fmax(std::fmin(t0.x, t1.x), std::fmin(t0.y, t1.y)), std::fmin(t0.z, t1.z));
    R hi = std::fmin(std::fmin(std::fmax(t0.x, t1.x), std::fmax(t0.y, t1.y)), std::fmax(t0.z, t1.z));
    lo = std::fmax(lo, tMin);
    hi = std::fmin(hi, tMax);
    return hi >= lo;
}

struct Counters {
    long long visits = 0, tests = 0, drops = 0;
};

// rtKernels.cu:63-100 (stack of 64, push left then right, drop when full)
template <typename R, bool F>
inline bool closest(const Scene<R>& sc, V3<R> o, V3<R> d, R tMin, R tMax, Hit<R>& h, Counters& c) {
    int stack[64];
    int sp = 0;
    stack[sp++] = sc.root;
    bool any = false;
    R best = tMax;
    V3<R> inv = {R(1.0) / d.x, R(1.0) / d.y, R(1.0) / d.z};
    while (sp > 0) {
        const Node<R>& n = sc.nodes[stack[--sp]];
        c.visits++;
        if (!hit_box(n, o, inv, tMin, best)) continue;
        if (n.triCount > 0) {
            for (int i = 0; i < n.triCount; ++i) {
                int ti = n.firstTri + i;
                R t, u, v;
                c.tests++;
                if (hit_tri<R, F>(sc.tris[ti], o, d, tMin, best, t, u, v)) {
                    any = true;
                    best = t;
                    h = Hit<R>{t, u, v, ti};
                }
            }
        } else if (sp + 2 <= 64) {
            stack[sp++] = n.left;
            stack[sp++] = n.right;
        } else {
            c.drops++;
        }
    }
    return any;
}

// brute force closest hit over every triangle, in BVH order (test pin)
template <typename R, bool F>
inline bool closest_brute(const Scene<R>& sc, V3<R> o, V3<R> d, R tMin, R tMax, Hit<R>& h) {
    bool any = false;
    R best = tMax;
    for (int ti = 0; ti < int(sc.tris.size()); ++ti) {
        R t, u, v;
        if (hit_tri<R, F>(sc.tris[ti], o, d, tMin, best, t, u, v)) {
            any = true;
            best = t;
            h = Hit<R>{t, u, v, ti};
        }
    }
    return any;
}

template <typename R> struct Frame {
    V3<R> llc, uvec, vvec, dir;  // dir = ray direction (= -radar direction)
};

// sweep.cu:349-364 for given angles in degrees
template <typename R> Frame<R> frame_deg(R thDeg, R phDeg, R g) {
    const R pi = R(M_PI);
    R th = thDeg * pi / R(180.0);
    R ph = phDeg * pi / R(180.0);
    V3<R> dir = {std::sin(th) * std::cos(ph), std::sin(th) * std::sin(ph), std::cos(th)};
    V3<R> up = {R(0), R(1), R(0)};
    if (std::fabs(dir.y) > R(0.99)) up = {R(1), R(0), R(0)};
    V3<R> u = unit<R, false>(cross<R, false>(up, dir));
    u = {u.x * g, u.y * g, u.z * g};
    V3<R> v = unit<R, false>(cross<R, false>(dir, u));
    v = {v.x * g, v.y * g, v.z * g};
    R off = R(10.0) * g;
    V3<R> a = {dir.x * off, dir.y * off, dir.z * off};
    V3<R> hu = R(0.5) * u, hv = R(0.5) * v;
    Frame<R> f;
    f.llc = (a - hu) - hv;
    f.uvec = u;
    f.vvec = v;
    f.dir = neg(dir);
    return f;
}

// sweep.cu:336-364
template <typename R>
Frame<R> frame(double thS, double thE, int nTh, double phS, double phE, int nPh, double grid, int idx) {
    int t = idx / nPh, p = idx % nPh;
    R ts = R(thS), te = R(thE), ps = R(phS), pe = R(phE), g = R(grid);
    R thDeg = ts + (nTh > 1 ? t * (te - ts) / (nTh - 1) : R(0));
    R phDeg = ps + (nPh > 1 ? p * (pe - ps) / (nPh - 1) : R(0));
    return frame_deg<R>(thDeg, phDeg, g);
}

struct Params {
    double thS, thE, phS, phE, grid, freq, refl;
    int nTh, nPh, nx, ny, maxB;
};

template <typename R> struct Phys {
    R k, rayArea, twoPi, pre;
};

// sweep.cu:50-55 and the constant part of poKernels.cu:98-101
template <typename R> Phys<R> phys(const Params& p) {
    Phys<R> ph;
    const R c0 = R(299792458.0);
    R lambda = c0 / R(p.freq);
    ph.k = R(2.0) * R(M_PI) / lambda;
    R g = R(p.grid);
    ph.rayArea = (g * g) / R(p.nx * p.ny);
    ph.twoPi = R(2.0) * R(M_PI);
    ph.pre = (ph.k * ph.rayArea / (R(4.0) * R(M_PI))) * R(2.0);
    return ph;
}

template <typename R> struct RayResult {
    int count = 0;
    R L = 0;
    V3<R> n{0, 0, 0}, d{0, 0, 0};
    R re = 0, im = 0;
};

// rtKernels.cu:113-158 and poKernels.cu:75-110 for one ray; ids (original
// face indices, -1 past the last hit) are written when ids != nullptr.
template <typename R, bool F>
RayResult<R> trace_ray(const Scene<R>& sc, const Frame<R>& fr, const Params& p, const Phys<R>& ph,
                       int i, int j, int* ids, Counters& c) {
    R u = (R(i) + R(0.5)) / R(p.nx);
    R v = (R(j) + R(0.5)) / R(p.ny);
    V3<R> o = axpy<R, F>(v, fr.vvec, axpy<R, F>(u, fr.uvec, fr.llc));
    V3<R> d = fr.dir;
    RayResult<R> rr;
    const R EPS = R(1e-4);
    if (ids)
        for (int k = 0; k < p.maxB; ++k) ids[k] = -1;
    for (int k = 0; k < p.maxB; ++k) {
        Hit<R> h;
        if (sc.nodes.empty() || !closest<R, F>(sc, o, d, R(0.001), R(1e20), h, c)) break;
        const Tri<R>& tr = sc.tris[h.tri];
        V3<R> e1 = tr.v1 - tr.v0, e2 = tr.v2 - tr.v0;
        V3<R> n = unit<R, F>(cross<R, F>(e1, e2));
        V3<R> hp = axpy<R, F>(h.t, d, o);
        rr.n = n;
        rr.d = d;
        rr.count++;
        rr.L += h.t;
        if (ids) ids[k] = sc.perm[h.tri];
        R s = R(2.0) * dot<R, F>(d, n);
        d = unit<R, F>(axpy<R, F>(-s, n, d));
        o = axpy<R, F>(EPS, n, hp);
    }
    if (rr.count > 0) {
        R refl = R(1.0);
        for (int b = 0; b < rr.count; ++b) refl *= R(p.refl);
        R cosT = dot<R, F>(rr.n, neg(rr.d));
        if (cosT > R(0)) {
            R phase = std::fmod(R(2.0) * ph.k * rr.L, ph.twoPi);
            R mag = ph.pre * cosT * refl;
            rr.re = -mag * std::sin(phase);
            rr.im = mag * std::cos(phase);
        }
    }
    return rr;
}

// ------------------------------------------------------- frozen paths (FP64)
// The ray of rtKernels.cu:113-158 in FP64 with its hit faces given (no
// traversal): each bounce's distance is the Moller-Trumbore distance on the
// recorded face, computed with hit_tri's operations; the term is that of
// poKernels.cu:75-110. Used as the finite-difference reference of the
// adjoint (derivatives of sigma along frozen paths).
struct FrozenMesh {
    const double* verts;  // n_verts*3, scaled
    const uint32_t* idx;  // n_tris*3
};

inline V3<double> fvtx(const FrozenMesh& m, int f, int k) {
    const uint32_t vi = m.idx[3 * f + k];
    return mk(m.verts[3 * size_t(vi)], m.verts[3 * size_t(vi) + 1], m.verts[3 * size_t(vi) + 2]);
}

struct FrozenJob {
    Frame<double> fr;
    Phys<double> ph;
    int nx, ny, maxB;
    double refl;
};

inline FrozenJob frozen_job(const double* pj, const int* ij) {
    // pj: theta_deg, phi_deg, grid, freq, refl; ij: nx, ny, maxB
    FrozenJob J;
    J.fr = frame_deg<double>(pj[0], pj[1], pj[2]);
    Params p{};
    p.grid = pj[2];
    p.freq = pj[3];
    p.refl = pj[4];
    p.nx = ij[0];
    p.ny = ij[1];
    p.maxB = ij[2];
    J.ph = phys<double>(p);
    J.nx = ij[0];
    J.ny = ij[1];
    J.maxB = ij[2];
    J.refl = pj[4];
    return J;
}

inline void frozen_ray(const FrozenMesh& m, const FrozenJob& J, int ray, const int32_t* fids, int cnt, double& re,
                       double& im) {
    using R = double;
    const int i = ray % J.nx, j = ray / J.nx;
    R u = (R(i) + R(0.5)) / R(J.nx);
    R v = (R(j) + R(0.5)) / R(J.ny);
    V3<R> o = axpy<R, false>(v, J.fr.vvec, axpy<R, false>(u, J.fr.uvec, J.fr.llc));
    V3<R> d = J.fr.dir;
    V3<R> nl{0, 0, 0}, dl{0, 0, 0};
    R L = 0;
    for (int b = 0; b < cnt; ++b) {
        V3<R> v0 = fvtx(m, fids[b], 0);
        V3<R> e1 = fvtx(m, fids[b], 1) - v0, e2 = fvtx(m, fids[b], 2) - v0;
        V3<R> pv = cross<R, false>(d, e2);
        R det = dot<R, false>(e1, pv);
        R inv = R(1.0) / det;
        V3<R> tv = o - v0;
        V3<R> q = cross<R, false>(tv, e1);
        R t = dot<R, false>(e2, q) * inv;
        V3<R> n = unit<R, false>(cross<R, false>(e1, e2));
        V3<R> hp = axpy<R, false>(t, d, o);
        nl = n;
        dl = d;
        L += t;
        R s = R(2.0) * dot<R, false>(d, n);
        d = unit<R, false>(axpy<R, false>(-s, n, d));
        o = axpy<R, false>(R(1e-4), n, hp);
    }
    re = 0;
    im = 0;
    if (cnt <= 0) return;
    R refl = R(1.0);
    for (int b = 0; b < cnt; ++b) refl *= R(J.refl);
    R cosT = dot<R, false>(nl, neg(dl));
    if (!(cosT > R(0))) return;
    R phase = std::fmod(R(2.0) * J.ph.k * L, J.ph.twoPi);
    R mag = J.ph.pre * cosT * refl;
    re = -mag * std::sin(phase);
    im = mag * std::cos(phase);
}

}  // namespace orc

using namespace orc;

// ---------------------------------------------------------------- C entry
// All entry points return 0 on success, nonzero on a bad argument.
extern "C" {

struct orc_scene {
    int prec;  // 0 float, 1 double
    void* s;
};

orc_scene* orc_scene_create(const float* xyz, int nVerts, const uint32_t* idx, int nTris, double scale,
                            int precision, int leafSize, int binCount, int sah) {
    auto* o = new orc_scene{precision, nullptr};
    if (precision == 1)
        o->s = make_scene<double>(xyz, nVerts, idx, nTris, scale, leafSize, binCount, sah);
    else
        o->s = make_scene<float>(xyz, nVerts, idx, nTris, scale, leafSize, binCount, sah);
    if (!o->s) {
        delete o;
        return nullptr;
    }
    return o;
}

void orc_scene_destroy(orc_scene* o) {
    if (!o) return;
    if (o->prec == 1) delete static_cast<Scene<double>*>(o->s);
    else delete static_cast<Scene<float>*>(o->s);
    delete o;
}

// counts: [nTris, nNodes, root, depth]
void orc_scene_info(const orc_scene* o, int* out4) {
    if (o->prec == 1) {
        auto* s = static_cast<Scene<double>*>(o->s);
        out4[0] = int(s->tris.size()); out4[1] = int(s->nodes.size()); out4[2] = s->root; out4[3] = s->depth;
    } else {
        auto* s = static_cast<Scene<float>*>(o->s);
        out4[0] = int(s->tris.size()); out4[1] = int(s->nodes.size()); out4[2] = s->root; out4[3] = s->depth;
    }
}

}  // extern "C"
template <typename R> static void export_scene(const Scene<R>* s, double* tris9, int* perm, double* bounds6, int* links4) {
    for (size_t i = 0; i < s->tris.size(); ++i) {
        const Tri<R>& t = s->tris[i];
        double* o = tris9 + 9 * i;
        o[0] = t.v0.x; o[1] = t.v0.y; o[2] = t.v0.z;
        o[3] = t.v1.x; o[4] = t.v1.y; o[5] = t.v1.z;
        o[6] = t.v2.x; o[7] = t.v2.y; o[8] = t.v2.z;
        perm[i] = s->perm[i];
    }
    for (size_t i = 0; i < s->nodes.size(); ++i) {
        const Node<R>& n = s->nodes[i];
        double* b = bounds6 + 6 * i;
        b[0] = n.bmin.x; b[1] = n.bmin.y; b[2] = n.bmin.z;
        b[3] = n.bmax.x; b[4] = n.bmax.y; b[5] = n.bmax.z;
        int* l = links4 + 4 * i;
        l[0] = n.left; l[1] = n.right; l[2] = n.firstTri; l[3] = n.triCount;
    }
}

extern "C" {
void orc_scene_export(const orc_scene* o, double* tris9, int* perm, double* bounds6, int* links4) {
    if (o->prec == 1) export_scene(static_cast<Scene<double>*>(o->s), tris9, perm, bounds6, links4);
    else export_scene(static_cast<Scene<float>*>(o->s), tris9, perm, bounds6, links4);
}

static Params mkparams(const double* pd, const int* pi) {
    // pd: thS thE phS phE grid freq refl ; pi: nTh nPh nx ny maxB
    Params p;
    p.thS = pd[0]; p.thE = pd[1]; p.phS = pd[2]; p.phE = pd[3]; p.grid = pd[4]; p.freq = pd[5]; p.refl = pd[6];
    p.nTh = pi[0]; p.nPh = pi[1]; p.nx = pi[2]; p.ny = pi[3]; p.maxB = pi[4];
    return p;
}

// frame of one angle: out12 = llc, u, v, rayDir (as doubles of the R values);
// out_const4 = k, rayArea, 2pi, prefactor
void orc_frame(int precision, const double* pd, const int* pi, int angle, double* out12, double* out4) {
    Params p = mkparams(pd, pi);
    auto put = [&](auto f, auto ph) {
        double* o = out12;
        o[0] = f.llc.x; o[1] = f.llc.y; o[2] = f.llc.z;
        o[3] = f.uvec.x; o[4] = f.uvec.y; o[5] = f.uvec.z;
        o[6] = f.vvec.x; o[7] = f.vvec.y; o[8] = f.vvec.z;
        o[9] = f.dir.x; o[10] = f.dir.y; o[11] = f.dir.z;
        out4[0] = ph.k; out4[1] = ph.rayArea; out4[2] = ph.twoPi; out4[3] = ph.pre;
    };
    if (precision == 1)
        put(frame<double>(p.thS, p.thE, p.nTh, p.phS, p.phE, p.nPh, p.grid, angle), phys<double>(p));
    else
        put(frame<float>(p.thS, p.thE, p.nTh, p.phS, p.phE, p.nPh, p.grid, angle), phys<float>(p));
}

}  // extern "C"
template <typename R, bool F>
static void trace_angle_t(const Scene<R>* s, const Params& p, int angle, int* ids, int* counts, double* L,
                          double* reim_ray, double* sum2, long long* ctr3) {
    Frame<R> fr = frame<R>(p.thS, p.thE, p.nTh, p.phS, p.phE, p.nPh, p.grid, angle);
    Phys<R> ph = phys<R>(p);
    const int nRays = p.nx * p.ny;
    int nthreads = 1;
#ifdef _OPENMP
    nthreads = omp_get_max_threads();
#endif
    std::vector<double> pre(nthreads, 0.0), pim(nthreads, 0.0);
    std::vector<Counters> pc(nthreads);
#pragma omp parallel for schedule(static)
    for (int r = 0; r < nRays; ++r) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        int i = r % p.nx, j = r / p.nx;
        RayResult<R> rr = trace_ray<R, F>(*s, fr, p, ph, i, j, ids ? ids + size_t(r) * p.maxB : nullptr, pc[tid]);
        if (counts) counts[r] = rr.count;
        if (L) L[r] = double(rr.L);
        if (reim_ray) {
            reim_ray[2 * r] = double(rr.re);
            reim_ray[2 * r + 1] = double(rr.im);
        }
        pre[tid] += double(rr.re);
        pim[tid] += double(rr.im);
    }
    double re = 0, im = 0;
    Counters c;
    for (int t = 0; t < nthreads; ++t) {
        re += pre[t];
        im += pim[t];
        c.visits += pc[t].visits; c.tests += pc[t].tests; c.drops += pc[t].drops;
    }
    if (sum2) { sum2[0] = re; sum2[1] = im; }
    if (ctr3) { ctr3[0] = c.visits; ctr3[1] = c.tests; ctr3[2] = c.drops; }
}

// One angle, every ray. Outputs (any may be NULL): ids[nRays*maxB] original
// face index per bounce, counts[nRays], L[nRays], reim_ray[2*nRays],
// sum2 = (sum Re, sum Im) over rays in double, ctr3 = node visits, triangle
// tests, stack drops. fused selects the F variant (see file header).
extern "C" {
int orc_trace_angle(const orc_scene* o, const double* pd, const int* pi, int angle, int fused, int* ids,
                    int* counts, double* L, double* reim_ray, double* sum2, long long* ctr3) {
    Params p = mkparams(pd, pi);
    if (p.nx <= 0 || p.ny <= 0 || p.maxB < 0) return 1;
    if (angle < 0 || angle >= p.nTh * p.nPh) return 1;
    if (o->prec == 1) {
        auto* s = static_cast<Scene<double>*>(o->s);
        if (fused) trace_angle_t<double, true>(s, p, angle, ids, counts, L, reim_ray, sum2, ctr3);
        else trace_angle_t<double, false>(s, p, angle, ids, counts, L, reim_ray, sum2, ctr3);
    } else {
        auto* s = static_cast<Scene<float>*>(o->s);
        if (fused) trace_angle_t<float, true>(s, p, angle, ids, counts, L, reim_ray, sum2, ctr3);
        else trace_angle_t<float, false>(s, p, angle, ids, counts, L, reim_ray, sum2, ctr3);
    }
    return 0;
}

// sigma and dBsm from a complex sum (sweep.cu:228-231), evaluated in R.
void orc_sigma(int precision, double re, double im, double* sigma, double* dbsm) {
    if (precision == 1) {
        double s = 4.0 * M_PI * (re * re + im * im);
        *sigma = s;
        *dbsm = 10.0 * std::log10(std::fmax(s, 1e-10));
    } else {
        float r = float(re), i = float(im);
        float s = 4.0f * float(M_PI) * (r * r + i * i);
        *sigma = s;
        *dbsm = 10.0f * std::log10(std::fmax(s, 1e-10f));
    }
}

// Single ray with an explicit origin and direction (edge-tie probes).
// ids[maxB]; returns the bounce count. out_L receives the path length.
}  // extern "C"
template <typename R, bool F>
static int ray_t(const Scene<R>* s, const double* o3, const double* d3, int maxB, int* ids, double* outL) {
    Params p{};
    p.maxB = maxB; p.nx = 1; p.ny = 1; p.refl = 1.0; p.freq = 1e9; p.grid = 1.0;
    Frame<R> fr;
    fr.llc = mk(R(o3[0]), R(o3[1]), R(o3[2]));
    fr.uvec = mk(R(0), R(0), R(0));
    fr.vvec = mk(R(0), R(0), R(0));
    fr.dir = mk(R(d3[0]), R(d3[1]), R(d3[2]));
    Phys<R> ph = phys<R>(p);
    Counters c;
    // i=j=0 with zero u/v vectors gives origin = llc exactly
    RayResult<R> rr = trace_ray<R, F>(*s, fr, p, ph, 0, 0, ids, c);
    if (outL) *outL = double(rr.L);
    return rr.count;
}

extern "C" {
int orc_trace_one(const orc_scene* o, const double* o3, const double* d3, int maxB, int fused, int* ids, double* outL) {
    if (o->prec == 1) {
        auto* s = static_cast<Scene<double>*>(o->s);
        return fused ? ray_t<double, true>(s, o3, d3, maxB, ids, outL) : ray_t<double, false>(s, o3, d3, maxB, ids, outL);
    }
    auto* s = static_cast<Scene<float>*>(o->s);
    return fused ? ray_t<float, true>(s, o3, d3, maxB, ids, outL) : ray_t<float, false>(s, o3, d3, maxB, ids, outL);
}

// Closest hit of one ray by BVH and by brute force. out: [bvhTri, bruteTri]
// as original face indices (-1 = miss), tOut: [tBvh, tBrute].
}  // extern "C"
template <typename R>
static void closest_pair_t(const Scene<R>* s, const double* o3, const double* d3, int* out2, double* t2) {
    V3<R> oo = mk(R(o3[0]), R(o3[1]), R(o3[2]));
    V3<R> dd = mk(R(d3[0]), R(d3[1]), R(d3[2]));
    Hit<R> h1{}, h2{};
    Counters c;
    bool a = !s->nodes.empty() && closest<R, false>(*s, oo, dd, R(0.001), R(1e20), h1, c);
    bool b = closest_brute<R, false>(*s, oo, dd, R(0.001), R(1e20), h2);
    out2[0] = a ? s->perm[h1.tri] : -1;
    out2[1] = b ? s->perm[h2.tri] : -1;
    t2[0] = a ? double(h1.t) : 0.0;
    t2[1] = b ? double(h2.t) : 0.0;
}

extern "C" {
int orc_closest_pair(const orc_scene* o, const double* o3, const double* d3, int* out2, double* t2) {
    if (o->prec == 1) closest_pair_t(static_cast<Scene<double>*>(o->s), o3, d3, out2, t2);
    else closest_pair_t(static_cast<Scene<float>*>(o->s), o3, d3, out2, t2);
    return 0;
}

// Moller-Trumbore on one triangle (known-answer pin). Returns 1 on a hit.
int orc_hit_triangle(int precision, const double* v9, const double* o3, const double* d3, double tMin, double tMax,
                     double* tuv) {
    auto run = [&](auto zero) {
        using R = decltype(zero);
        Tri<R> tr{mk(R(v9[0]), R(v9[1]), R(v9[2])), mk(R(v9[3]), R(v9[4]), R(v9[5])), mk(R(v9[6]), R(v9[7]), R(v9[8]))};
        R t = 0, u = 0, v = 0;
        bool h = hit_tri<R, false>(tr, mk(R(o3[0]), R(o3[1]), R(o3[2])), mk(R(d3[0]), R(d3[1]), R(d3[2])), R(tMin),
                                   R(tMax), t, u, v);
        tuv[0] = t; tuv[1] = u; tuv[2] = v;
        return h ? 1 : 0;
    };
    return precision == 1 ? run(0.0) : run(0.0f);
}

// Sweep of many angles (CPU baseline). Writes 2 doubles per angle.
int orc_sweep(const orc_scene* o, const double* pd, const int* pi, int fused, const int* angles, int nAngles,
              double* reim, long long* ctr3) {
    long long tot[3] = {0, 0, 0};
    for (int a = 0; a < nAngles; ++a) {
        long long c[3];
        int rc = orc_trace_angle(o, pd, pi, angles[a], fused, nullptr, nullptr, nullptr, nullptr, reim + 2 * a, c);
        if (rc) return rc;
        tot[0] += c[0]; tot[1] += c[1]; tot[2] += c[2];
    }
    if (ctr3) { ctr3[0] = tot[0]; ctr3[1] = tot[1]; ctr3[2] = tot[2]; }
    return 0;
}

// Complex sum of one job along frozen paths (ids: rays*maxB original face
// ids, cnt: rays). pj: theta_deg, phi_deg, grid, freq, refl; ij: nx, ny, maxB.
int orc_frozen_sum(const double* verts, const uint32_t* idx, const double* pj, const int* ij, const int32_t* ids,
                   const int32_t* cnt, double* reim) {
    FrozenMesh m{verts, idx};
    FrozenJob J = frozen_job(pj, ij);
    const int nr = J.nx * J.ny;
    double sr = 0, si = 0;
    for (int r = 0; r < nr; ++r) {
        double a, b;
        frozen_ray(m, J, r, ids + size_t(r) * J.maxB, cnt[r], a, b);
        sr += a;
        si += b;
    }
    reim[0] = sr;
    reim[1] = si;
    return 0;
}

// Central differences of sum_j w_j sigma_j along frozen paths with respect
// to vertex coordinates coords[c] = 3*vertex + axis (step h in the scaled
// coordinates; with richardson != 0 the O(h^4) combination
// (4 D(h/2) - D(h)) / 3). Only rays whose recorded path touches the vertex
// are re-evaluated; the others cancel exactly. pj: 5 doubles per job, ij: 3 ints
// per job; ids/cnt: per job, rays*maxB / rays. Outputs grad[nCoords] and
// sigma[nJobs] (may be NULL).
int orc_frozen_fd(const double* verts, int nVerts, const uint32_t* idx, int nTris, int nJobs, const double* pj,
                  const int* ij, const int32_t* ids, const int32_t* cnt, const double* w, int nCoords,
                  const int* coords, double h, int richardson, double* grad, double* sigma) {
    FrozenMesh m{verts, idx};
    std::vector<FrozenJob> J(nJobs);
    std::vector<size_t> rayBase(nJobs + 1, 0), idBase(nJobs + 1, 0);
    for (int k = 0; k < nJobs; ++k) {
        J[k] = frozen_job(pj + 5 * k, ij + 3 * k);
        rayBase[k + 1] = rayBase[k] + size_t(J[k].nx) * J[k].ny;
        idBase[k + 1] = idBase[k] + size_t(J[k].nx) * J[k].ny * J[k].maxB;
    }
    // base sums
    std::vector<double> S(2 * nJobs, 0.0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < nJobs; ++k) {
        int ijk[3] = {J[k].nx, J[k].ny, J[k].maxB};
        orc_frozen_sum(verts, idx, pj + 5 * k, ijk, ids + idBase[k], cnt + rayBase[k], &S[2 * k]);
    }
    if (sigma)
        for (int k = 0; k < nJobs; ++k) sigma[k] = 4.0 * M_PI * (S[2 * k] * S[2 * k] + S[2 * k + 1] * S[2 * k + 1]);
    // rays touching each target vertex
    std::vector<int> target(nVerts, -1);
    std::vector<int> tv;
    for (int c = 0; c < nCoords; ++c) {
        int v = coords[c] / 3;
        if (v < 0 || v >= nVerts) return 1;
        if (target[v] < 0) {
            target[v] = int(tv.size());
            tv.push_back(v);
        }
    }
    std::vector<std::vector<std::pair<int, int>>> touch(tv.size());
    for (int k = 0; k < nJobs; ++k) {
        const size_t nr = rayBase[k + 1] - rayBase[k];
        for (size_t r = 0; r < nr; ++r) {
            const int c = cnt[rayBase[k] + r];
            const int32_t* f = ids + idBase[k] + r * J[k].maxB;
            int last = -1;
            for (int b = 0; b < c; ++b) {
                if (f[b] < 0 || f[b] >= nTris) return 2;
                for (int q = 0; q < 3; ++q) {
                    int t = target[idx[3 * f[b] + q]];
                    if (t >= 0 && t != last) {
                        auto& lst = touch[t];
                        if (lst.empty() || lst.back() != std::make_pair(k, int(r))) lst.emplace_back(k, int(r));
                    }
                }
            }
        }
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int c = 0; c < nCoords; ++c) {
        const int v = coords[c] / 3, ax = coords[c] % 3;
        const auto& lst = touch[target[v]];
        std::vector<double> vv(verts, verts + size_t(nVerts) * 3);
        FrozenMesh mp{vv.data(), idx};
        // central difference with step hs; sigma(+) - sigma(-) is formed as
        // 4 pi [2 Re(conj(S) (dS+ - dS-)) + |dS+|^2 - |dS-|^2], which has
        // no cancellation of |S|^2
        auto central = [&](double hs) {
            std::vector<double> dP(2 * nJobs, 0.0), dM(2 * nJobs, 0.0);
            for (int sgn = -1; sgn <= 1; sgn += 2) {
                vv[3 * size_t(v) + ax] = verts[3 * size_t(v) + ax] + sgn * hs;
                std::vector<double>& dS = sgn > 0 ? dP : dM;
                for (const auto& kr : lst) {
                    const int k = kr.first, r = kr.second;
                    const int32_t* f = ids + idBase[k] + size_t(r) * J[k].maxB;
                    const int cc = cnt[rayBase[k] + r];
                    double a0, b0, a1, b1;
                    frozen_ray(m, J[k], r, f, cc, a0, b0);
                    frozen_ray(mp, J[k], r, f, cc, a1, b1);
                    dS[2 * k] += a1 - a0;
                    dS[2 * k + 1] += b1 - b0;
                }
            }
            vv[3 * size_t(v) + ax] = verts[3 * size_t(v) + ax];
            double sum = 0;
            for (int k = 0; k < nJobs; ++k) {
                if (w[k] == 0.0) continue;
                const double xr = dP[2 * k] - dM[2 * k], xi = dP[2 * k + 1] - dM[2 * k + 1];
                const double ds = 4.0 * M_PI *
                                  (2.0 * (S[2 * k] * xr + S[2 * k + 1] * xi) +
                                   (dP[2 * k] * dP[2 * k] + dP[2 * k + 1] * dP[2 * k + 1]) -
                                   (dM[2 * k] * dM[2 * k] + dM[2 * k + 1] * dM[2 * k + 1]));
                sum += w[k] * ds;
            }
            return sum / (2.0 * hs);
        };
        if (richardson) grad[c] = (4.0 * central(0.5 * h) - central(h)) / 3.0;
        else grad[c] = central(h);
    }
    return 0;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
