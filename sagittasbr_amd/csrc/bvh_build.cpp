// Host BVH build (see bvh_build.hpp). Compiled with -ffp-contract=off so the
// SAH cost arithmetic is the literal reference expression.
#include "bvh_build.hpp"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <queue>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace sbr {
namespace {

template <typename R> struct V {
    R x, y, z;
    R operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <typename R> inline V<R> vmin(V<R> a, V<R> b) {
    return {std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)};
}
template <typename R> inline V<R> vmax(V<R> a, V<R> b) {
    return {std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)};
}
template <typename R> inline V<R> big() {
    R m = std::numeric_limits<R>::max();
    return {m, m, m};
}
template <typename R> inline V<R> nbig() {
    R m = -std::numeric_limits<R>::max();
    return {m, m, m};
}
// surface area of a box, bvhBuilder.cpp:46-52
template <typename R> inline R box_area(V<R> lo, V<R> hi) {
    R dx = std::fmax(R(0), hi.x - lo.x);
    R dy = std::fmax(R(0), hi.y - lo.y);
    R dz = std::fmax(R(0), hi.z - lo.z);
    return R(2.0) * (dx * dy + dy * dz + dz * dx);
}

template <typename R> struct Builder {
    // per original triangle (indexed by face id)
    std::vector<V<R>> tlo, thi, cen;
    std::vector<int> order;  // the permutation being built
    std::vector<RefNode<R>> nodes;
    std::atomic<int> counter{0};
    std::atomic<int> maxDepth{0};
    BuildOptions opt;

    void range_bounds(int s, int e, V<R>& lo, V<R>& hi) const {
        lo = big<R>();
        hi = nbig<R>();
        for (int i = s; i < e; ++i) {
            int t = order[i];
            lo = vmin(lo, tlo[t]);
            hi = vmax(hi, thi[t]);
        }
    }
    void centroid_bounds(int s, int e, V<R>& lo, V<R>& hi) const {
        lo = big<R>();
        hi = nbig<R>();
        for (int i = s; i < e; ++i) {
            V<R> c = cen[order[i]];
            lo = vmin(lo, c);
            hi = vmax(hi, c);
        }
    }

    // median split on the widest centroid axis (bvhBuilder.cpp:93-113)
    bool split_median(int s, int e, int& mid) {
        V<R> clo, chi;
        centroid_bounds(s, e, clo, chi);
        V<R> ext{chi.x - clo.x, chi.y - clo.y, chi.z - clo.z};
        int axis = (ext.x >= ext.y && ext.x >= ext.z) ? 0 : (ext.y >= ext.z ? 1 : 2);
        mid = s + (e - s) / 2;
        const std::vector<V<R>>& c = cen;
        std::nth_element(order.begin() + s, order.begin() + mid, order.begin() + e,
                         [&c, axis](int a, int b) { return c[a][axis] < c[b][axis]; });
        return mid > s && mid < e;
    }

    // binned SAH (bvhBuilder.cpp:115-243)
    bool split_sah(int s, int e, V<R> plo, V<R> phi, int& mid) {
        const int nb = opt.binCount;
        const int n = e - s;
        V<R> clo, chi;
        centroid_bounds(s, e, clo, chi);
        V<R> ext{chi.x - clo.x, chi.y - clo.y, chi.z - clo.z};
        R parent = box_area(plo, phi);
        if (parent <= R(0)) return false;
        int bestAxis = -1, bestSplit = -1;
        R bestCost = std::numeric_limits<R>::max();
        std::vector<V<R>> blo(nb), bhi(nb), llo(nb), lhi(nb), rlo(nb), rhi(nb);
        std::vector<int> bcnt(nb), lcnt(nb), rcnt(nb);
        for (int axis = 0; axis < 3; ++axis) {
            R extent = ext[axis];
            if (extent <= R(1.0e-6)) continue;
            R base = clo[axis];
            for (int b = 0; b < nb; ++b) {
                blo[b] = big<R>();
                bhi[b] = nbig<R>();
                bcnt[b] = 0;
            }
            for (int i = s; i < e; ++i) {
                int t = order[i];
                R off = (cen[t][axis] - base) / extent;
                int b = static_cast<int>(off * nb);
                b = b < 0 ? 0 : (b >= nb ? nb - 1 : b);
                blo[b] = vmin(blo[b], tlo[t]);
                bhi[b] = vmax(bhi[b], thi[t]);
                bcnt[b] += 1;
            }
            V<R> lo = big<R>(), hi = nbig<R>();
            int c = 0;
            for (int b = 0; b < nb; ++b) {
                if (bcnt[b] > 0) {
                    lo = vmin(lo, blo[b]);
                    hi = vmax(hi, bhi[b]);
                    c += bcnt[b];
                }
                llo[b] = lo;
                lhi[b] = hi;
                lcnt[b] = c;
            }
            lo = big<R>();
            hi = nbig<R>();
            c = 0;
            for (int b = nb - 1; b >= 0; --b) {
                if (bcnt[b] > 0) {
                    lo = vmin(lo, blo[b]);
                    hi = vmax(hi, bhi[b]);
                    c += bcnt[b];
                }
                rlo[b] = lo;
                rhi[b] = hi;
                rcnt[b] = c;
            }
            for (int b = 0; b + 1 < nb; ++b) {
                int nl = lcnt[b], nr = rcnt[b + 1];
                if (nl == 0 || nr == 0) continue;
                R cost = (box_area(llo[b], lhi[b]) * nl + box_area(rlo[b + 1], rhi[b + 1]) * nr) / parent;
                if (cost < bestCost) {
                    bestCost = cost;
                    bestAxis = axis;
                    bestSplit = b;
                }
            }
        }
        if (bestAxis < 0 || bestCost >= static_cast<R>(n)) return false;
        R extent = ext[bestAxis];
        if (extent <= R(1.0e-6)) return false;
        R pos = clo[bestAxis] + extent * (static_cast<R>(bestSplit + 1) / static_cast<R>(nb));
        const std::vector<V<R>>& c = cen;
        auto it = std::partition(order.begin() + s, order.begin() + e,
                                 [&c, bestAxis, pos](int t) { return c[t][bestAxis] < pos; });
        mid = static_cast<int>(it - order.begin());
        return mid > s && mid < e;
    }

    // bvhBuilder.cpp:250-312
    int node(int s, int e, int depth) {
        int d = maxDepth.load(std::memory_order_relaxed);
        while (depth > d && !maxDepth.compare_exchange_weak(d, depth)) {
        }
        V<R> lo, hi;
        range_bounds(s, e, lo, hi);
        int id = counter.fetch_add(1);
        RefNode<R>& nd = nodes[id];
        nd.lo[0] = lo.x; nd.lo[1] = lo.y; nd.lo[2] = lo.z;
        nd.hi[0] = hi.x; nd.hi[1] = hi.y; nd.hi[2] = hi.z;
        nd.left = -1;
        nd.right = -1;
        nd.firstTri = s;
        nd.triCount = e - s;
        if (e - s <= opt.leafSize) return id;
        int mid = -1;
        bool ok = opt.sah && split_sah(s, e, lo, hi, mid);
        if (!ok) ok = split_median(s, e, mid);
        if (!ok) return id;
        int l = -1, r = -1;
        bool par = opt.enableParallel && (e - s) >= opt.minTrianglesForParallel && depth < opt.maxParallelDepth;
        if (par) {
#pragma omp task shared(l)
            l = node(s, mid, depth + 1);
#pragma omp task shared(r)
            r = node(mid, e, depth + 1);
#pragma omp taskwait
        } else {
            l = node(s, mid, depth + 1);
            r = node(mid, e, depth + 1);
        }
        RefNode<R>& nd2 = nodes[id];
        nd2.left = l;
        nd2.right = r;
        nd2.firstTri = 0;
        nd2.triCount = 0;
        return id;
    }
};

}  // namespace

template <typename R> Bvh<R> build_bvh(const std::vector<R>& soup9, const BuildOptions& optIn) {
    Bvh<R> out;
    const int n = static_cast<int>(soup9.size() / 9);
    if (n == 0) return out;
    Builder<R> b;
    b.opt = optIn;
    if (b.opt.leafSize < 1) b.opt.leafSize = 1;
    if (b.opt.binCount < 2) b.opt.binCount = 2;
    b.tlo.resize(n);
    b.thi.resize(n);
    b.cen.resize(n);
    b.order.resize(n);
#pragma omp parallel for schedule(static) if (n > 65536)
    for (int t = 0; t < n; ++t) {
        const R* p = &soup9[size_t(9) * t];
        V<R> v0{p[0], p[1], p[2]}, v1{p[3], p[4], p[5]}, v2{p[6], p[7], p[8]};
        b.tlo[t] = vmin(vmin(v0, v1), v2);
        b.thi[t] = vmax(vmax(v0, v1), v2);
        // centroid (v0 + v1 + v2) / 3, bvhBuilder.cpp:42-44
        b.cen[t] = V<R>{((v0.x + v1.x) + v2.x) / R(3.0), ((v0.y + v1.y) + v2.y) / R(3.0),
                        ((v0.z + v1.z) + v2.z) / R(3.0)};
        b.order[t] = t;
    }
    b.nodes.resize(size_t(2) * n - 1);
    int root = 0;
#ifdef _OPENMP
    if (b.opt.enableParallel) {
#pragma omp parallel
        {
#pragma omp single nowait
            root = b.node(0, n, 0);
        }
    } else {
        root = b.node(0, n, 0);
    }
#else
    root = b.node(0, n, 0);
#endif
    const int nn = b.counter.load();
    // pre-order renumbering (left subtree before right)
    std::vector<int> pre;
    pre.reserve(nn);
    std::vector<int> stack{root};
    while (!stack.empty()) {
        int id = stack.back();
        stack.pop_back();
        pre.push_back(id);
        const RefNode<R>& nd = b.nodes[id];
        if (nd.triCount == 0) {
            stack.push_back(nd.right);
            stack.push_back(nd.left);
        }
    }
    std::vector<int> pos(nn, -1);
    for (int i = 0; i < static_cast<int>(pre.size()); ++i) pos[pre[i]] = i;
    out.nodes.resize(pre.size());
    for (size_t i = 0; i < pre.size(); ++i) {
        RefNode<R> nd = b.nodes[pre[i]];
        if (nd.triCount == 0) {
            nd.left = pos[nd.left];
            nd.right = pos[nd.right];
        }
        out.nodes[i] = nd;
    }
    out.root = 0;
    out.depth = b.maxDepth.load();
    out.perm = b.order;
    out.tri9.resize(soup9.size());
    for (int i = 0; i < n; ++i)
        std::memcpy(&out.tri9[size_t(9) * i], &soup9[size_t(9) * b.order[i]], 9 * sizeof(R));
    return out;
}

template <typename R> bool to_gpu_layout(const Bvh<R>& b, GpuBvh<R>& out, std::string& err) {
    const int nn = static_cast<int>(b.nodes.size());
    const int nt = static_cast<int>(b.perm.size());
    if (nn == 0 || nt == 0) {
        err = "empty mesh";
        return false;
    }
    if (nt >= (1 << 27)) {
        err = "too many triangles for the 27-bit leaf encoding";
        return false;
    }
    // internal nodes get consecutive record indices in pre-order
    std::vector<int> rec(nn, -1);
    int ni = 0;
    for (int i = 0; i < nn; ++i)
        if (b.nodes[i].triCount == 0) rec[i] = ni++;
    for (int i = 0; i < nn; ++i)
        if (b.nodes[i].triCount > 16) {
            err = "BVH leaf holds more than 16 triangles (leaf_size must be <= 16)";
            return false;
        }
    auto ref_of = [&](int i) {
        const RefNode<R>& c = b.nodes[i];
        return c.triCount == 0 ? rec[i] : encode_leaf(c.firstTri, c.triCount);
    };
    const size_t recBytes = sizeof(R) == 4 ? 64 : 128;
    out.nodes.assign(recBytes * std::max(ni, 1), 0);
    for (int i = 0; i < nn; ++i) {
        const RefNode<R>& nd = b.nodes[i];
        if (nd.triCount != 0) continue;
        uint8_t* p = out.nodes.data() + recBytes * rec[i];
        R box[12];
        const RefNode<R>& L = b.nodes[nd.left];
        const RefNode<R>& Rn = b.nodes[nd.right];
        for (int k = 0; k < 3; ++k) {
            box[k] = L.lo[k];
            box[3 + k] = L.hi[k];
            box[6 + k] = Rn.lo[k];
            box[9 + k] = Rn.hi[k];
        }
        std::memcpy(p, box, sizeof(box));
        int refs[2] = {ref_of(nd.left), ref_of(nd.right)};
        std::memcpy(p + sizeof(box), refs, sizeof(refs));
    }
    const size_t triBytes = sizeof(R) == 4 ? 48 : 80;
    out.tris.assign(triBytes * nt, 0);
    for (int t = 0; t < nt; ++t) {
        const R* v = &b.tri9[size_t(9) * t];
        uint8_t* p = out.tris.data() + triBytes * t;
        int face = b.perm[t];
        if (sizeof(R) == 4) {
            // FP32: the three vertices (edges are formed per test, as the
            // reference does) and the face id; the last 8 bytes stay zero
            std::memcpy(p, v, 9 * sizeof(R));
            std::memcpy(p + 9 * sizeof(R), &face, sizeof(int));
        } else {
            // FP64: v0, e1 = v1 - v0, e2 = v2 - v0 (exact for float inputs)
            R rec9[9] = {v[0], v[1], v[2], v[3] - v[0], v[4] - v[1], v[5] - v[2], v[6] - v[0], v[7] - v[1], v[8] - v[2]};
            std::memcpy(p, rec9, sizeof(rec9));
            std::memcpy(p + sizeof(rec9), &face, sizeof(int));
        }
    }
    const RefNode<R>& root = b.nodes[b.root];
    for (int k = 0; k < 3; ++k) {
        out.rootLo[k] = root.lo[k];
        out.rootHi[k] = root.hi[k];
    }
    out.rootRef = ref_of(b.root);
    out.nInternal = ni;
    return true;
}

namespace {
template <int W> bool to_gpu_layout_w(const Bvh<float>& b, GpuBvhW& out, std::string& err) {
    constexpr int F = kRecBytes<W> / 4;  // floats per record
    out.width = W;
    const int nn = static_cast<int>(b.nodes.size());
    if (nn == 0) {
        err = "empty mesh";
        return false;
    }
    auto area = [&](int i) {
        const RefNode<float>& n = b.nodes[i];
        const double x = double(n.hi[0]) - n.lo[0], y = double(n.hi[1]) - n.lo[1], z = double(n.hi[2]) - n.lo[2];
        return x * y + y * z + z * x;
    };
    auto leaf_ref = [&](int i) { return encode_leaf(b.nodes[i].firstTri, b.nodes[i].triCount); };
    if (b.nodes[b.root].triCount != 0) {
        out.nodes.assign(kRecBytes<W>, 0);
        out.rootRef = leaf_ref(b.root);
        out.nInternal = 0;
        out.depth = 0;
        return true;
    }
    std::vector<std::array<float, F>> recs;
    int maxDepth = 0;
    // explicit work list: (binary node, record slot to fill, depth)
    struct Item {
        int node, rec, depth;
    };
    std::vector<Item> work;
    recs.emplace_back();
    work.push_back({b.root, 0, 1});
    while (!work.empty()) {
        const Item it = work.back();
        work.pop_back();
        maxDepth = std::max(maxDepth, it.depth);
        // open internal members, largest surface area first, up to W
        int kids[W];
        kids[0] = b.nodes[it.node].left;
        kids[1] = b.nodes[it.node].right;
        int nk = 2;
        while (nk < W) {
            int best = -1;
            double bestA = -1.0;
            for (int q = 0; q < nk; ++q)
                if (b.nodes[kids[q]].triCount == 0 && area(kids[q]) > bestA) {
                    bestA = area(kids[q]);
                    best = q;
                }
            if (best < 0) break;
            const int open = kids[best];
            for (int q = nk; q > best + 1; --q) kids[q] = kids[q - 1];
            kids[best] = b.nodes[open].left;
            kids[best + 1] = b.nodes[open].right;
            ++nk;
        }
        std::array<float, F> r{};
        int refs[W];
        for (int q = 0; q < W; ++q) {
            refs[q] = kEmptySlot;
            if (q >= nk) {
                // a box at +infinity on every axis: its slab distances are
                // infinite, so no ray enters it
                for (int k = 0; k < 6; ++k) r[W * k + q] = std::numeric_limits<float>::infinity();
                continue;
            }
            const RefNode<float>& c = b.nodes[kids[q]];
            for (int k = 0; k < 3; ++k) {
                r[W * k + q] = c.lo[k];
                r[W * (3 + k) + q] = c.hi[k];
            }
            if (c.triCount != 0) {
                refs[q] = leaf_ref(kids[q]);
            } else {
                refs[q] = static_cast<int>(recs.size());
                recs.emplace_back();
                work.push_back({kids[q], refs[q], it.depth + 1});
            }
        }
        std::memcpy(&r[6 * W], refs, sizeof(refs));
        recs[it.rec] = r;
    }
    // the top kTopLevels levels first, breadth-first (so a prefix of the
    // records is the top k levels, for staging in LDS); every deeper record
    // keeps its place relative to the others (subtrees stay together)
    const int nr = static_cast<int>(recs.size());
    std::vector<int> order, level{0};
    std::vector<int> newIdx(nr, -1);
    out.topRecs[0] = 0;
    for (int d = 0; d < kTopLevels; ++d) {
        std::vector<int> next;
        for (int r : level) {
            newIdx[r] = static_cast<int>(order.size());
            order.push_back(r);
            for (int q = 0; q < W; ++q) {
                int ref;
                std::memcpy(&ref, &recs[r][6 * W + q], sizeof(int));
                if (ref >= 0 && ref != kEmptySlot) next.push_back(ref);
            }
        }
        out.topRecs[d + 1] = static_cast<int>(order.size());
        level.swap(next);
    }
    for (int r = 0; r < nr; ++r)
        if (newIdx[r] < 0) {
            newIdx[r] = static_cast<int>(order.size());
            order.push_back(r);
        }
    out.nodes.resize(size_t(nr) * kRecBytes<W>);
    for (int k = 0; k < nr; ++k) {
        std::array<float, F> r = recs[order[k]];
        for (int q = 0; q < W; ++q) {
            int ref;
            std::memcpy(&ref, &r[6 * W + q], sizeof(int));
            if (ref >= 0 && ref != kEmptySlot) ref = newIdx[ref];
            std::memcpy(&r[6 * W + q], &ref, sizeof(int));
        }
        std::memcpy(out.nodes.data() + size_t(k) * kRecBytes<W>, r.data(), kRecBytes<W>);
    }
    out.rootRef = newIdx[0];
    out.nInternal = nr;
    out.depth = maxDepth;
    return true;
}
}  // namespace

bool to_gpu_layout_wide(const Bvh<float>& b, int width, GpuBvhW& out, std::string& err) {
    switch (width) {
        case 2: return to_gpu_layout_w<2>(b, out, err);
        case 4: return to_gpu_layout_w<4>(b, out, err);
        case 8: return to_gpu_layout_w<8>(b, out, err);
        default: err = "node width must be 2, 4 or 8"; return false;
    }
}

template <typename R> std::vector<float> cull_boxes(const Bvh<R>& b, int maxBoxes) {
    std::vector<int> open{b.root};
    auto area = [&](int i) {
        const RefNode<R>& n = b.nodes[i];
        const double x = double(n.hi[0]) - n.lo[0], y = double(n.hi[1]) - n.lo[1], z = double(n.hi[2]) - n.lo[2];
        return x * y + y * z + z * x;
    };
    for (;;) {
        int best = -1;
        double bestA = -1.0;
        for (size_t q = 0; q < open.size(); ++q)
            if (b.nodes[open[q]].triCount == 0 && area(open[q]) > bestA) {
                bestA = area(open[q]);
                best = int(q);
            }
        if (best < 0 || int(open.size()) + 1 > maxBoxes) break;
        const int n = open[best];
        open[best] = b.nodes[n].left;
        open.push_back(b.nodes[n].right);
    }
    std::vector<float> out;
    for (int n : open) {
        const RefNode<R>& nd = b.nodes[n];
        for (int k = 0; k < 3; ++k) {
            float lo = static_cast<float>(nd.lo[k]);
            if (double(lo) > double(nd.lo[k])) lo = std::nextafter(lo, -INFINITY);
            out.push_back(lo);
        }
        for (int k = 0; k < 3; ++k) {
            float hi = static_cast<float>(nd.hi[k]);
            if (double(hi) < double(nd.hi[k])) hi = std::nextafter(hi, INFINITY);
            out.push_back(hi);
        }
    }
    return out;
}

template <typename R> std::vector<float> cover_boxes(const Bvh<R>& b, int maxBoxes) {
    auto area = [&](int i) {
        const RefNode<R>& n = b.nodes[i];
        const double x = double(n.hi[0]) - n.lo[0], y = double(n.hi[1]) - n.lo[1], z = double(n.hi[2]) - n.lo[2];
        return x * y + y * z + z * x;
    };
    // open internal nodes by area (largest first); leaves are final
    std::priority_queue<std::pair<double, int>> open;
    std::vector<int> done;
    int count = 1;
    if (b.nodes[b.root].triCount == 0) open.push({area(b.root), b.root});
    else done.push_back(b.root);
    while (!open.empty() && count + 1 <= maxBoxes) {
        const int n = open.top().second;
        open.pop();
        ++count;
        for (int c : {b.nodes[n].left, b.nodes[n].right}) {
            if (b.nodes[c].triCount == 0) open.push({area(c), c});
            else done.push_back(c);
        }
    }
    while (!open.empty()) {
        done.push_back(open.top().second);
        open.pop();
    }
    std::vector<float> out;
    out.reserve(done.size() * 6);
    for (int n : done) {
        const RefNode<R>& nd = b.nodes[n];
        for (int k = 0; k < 3; ++k) {
            float lo = static_cast<float>(nd.lo[k]);
            if (double(lo) > double(nd.lo[k])) lo = std::nextafter(lo, -INFINITY);
            out.push_back(lo);
        }
        for (int k = 0; k < 3; ++k) {
            float hi = static_cast<float>(nd.hi[k]);
            if (double(hi) < double(nd.hi[k])) hi = std::nextafter(hi, INFINITY);
            out.push_back(hi);
        }
    }
    return out;
}

template std::vector<float> cover_boxes<float>(const Bvh<float>&, int);
template std::vector<float> cover_boxes<double>(const Bvh<double>&, int);
template std::vector<float> cull_boxes<float>(const Bvh<float>&, int);
template std::vector<float> cull_boxes<double>(const Bvh<double>&, int);
template Bvh<float> build_bvh<float>(const std::vector<float>&, const BuildOptions&);
template Bvh<double> build_bvh<double>(const std::vector<double>&, const BuildOptions&);
template bool to_gpu_layout<float>(const Bvh<float>&, GpuBvh<float>&, std::string&);
template bool to_gpu_layout<double>(const Bvh<double>&, GpuBvh<double>&, std::string&);

}  // namespace sbr
