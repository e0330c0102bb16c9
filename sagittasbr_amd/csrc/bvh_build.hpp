// Host BVH construction and conversion to the GPU child-box layout.
//
// The build restates the reference's binned-SAH builder
// (src/scene/bvhBuilder.cpp:22-360, options include/scene/bvhBuilder.hpp:13-20)
// so that the tree, and therefore the traversal order and edge-tie winners,
// are those of the reference. Node numbering is made deterministic by a
// pre-order renumbering after the (OpenMP task parallel) build.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "sbr_device.hpp"

namespace sbr {

struct BuildOptions {
    int leafSize = 4;
    int binCount = 16;
    int maxParallelDepth = 6;
    int minTrianglesForParallel = 4096;
    bool enableParallel = true;
    bool sah = true;
};

// Reference-layout node (include/scene/bvh.h:13-21).
template <typename R> struct RefNode {
    R lo[3], hi[3];
    int left, right, firstTri, triCount;
};

template <typename R> struct Bvh {
    std::vector<R> tri9;     // BVH order, v0 v1 v2
    std::vector<int> perm;   // BVH position -> original face index
    std::vector<RefNode<R>> nodes;  // pre-order
    int root = 0;
    int depth = 0;
};

// soup9: nTris*9 values in original face order (already scaled, in R).
template <typename R> Bvh<R> build_bvh(const std::vector<R>& soup9, const BuildOptions& opt);

// GPU child-box layout. One record per internal node holding both children's
// boxes and references; leaves live inline in their parent's child slot.
//   child ref >= 0 : index of an internal-node record
//   child ref <  0 : leaf, ~ref = (firstTri << 4) | (triCount - 1)
// FP32 record: 64 B  = lmin[3] lmax[3] rmin[3] rmax[3] lref rref pad pad
// FP64 record: 128 B = 12 doubles, lref rref, pad to 128
// FP32 triangle: 48 B = v0.xyz v1.xyz v2.xyz faceId |e1|+|e2| |e1||e2|
// FP64 triangle: 80 B = v0.xyz e1.xyz e2.xyz faceId (int in the last slot)
template <typename R> struct GpuBvh {
    std::vector<uint8_t> nodes;
    std::vector<uint8_t> tris;
    R rootLo[3], rootHi[3];
    int rootRef = 0;
    int nInternal = 0;
};

template <typename R> bool to_gpu_layout(const Bvh<R>& b, GpuBvh<R>& out, std::string& err);

// Wide FP32 layout for the mixed mode's certified queries, whose outcome
// does not depend on the tree or its visiting order (only on which triangles
// are tested). Made by collapsing the binary tree: each record takes up to
// W descendants (W = 2, 4 or 8), opening the child with the largest surface
// area first; leaves are the binary tree's leaves (same references). Record
// of kRecBytes<W>: lo.x[W] lo.y[W] lo.z[W] hi.x[W] hi.y[W] hi.z[W] ref[W],
// pad; an unused slot holds ref kEmptySlot and a box at +infinity on every
// axis.
struct GpuBvhW {
    int width = kWide;
    std::vector<uint8_t> nodes;
    int rootRef = 0;
    int nInternal = 0;
    int depth = 0;  // records on the longest root-to-leaf path
    // the records of the top kTopLevels levels come first, in breadth-first
    // order: topRecs[k] records hold levels 0 .. k-1
    int topRecs[kTopLevels + 1] = {};
};
// width: 2, 4 or 8
bool to_gpu_layout_wide(const Bvh<float>& b, int width, GpuBvhW& out, std::string& err);

// Up to maxBoxes boxes that together hold every triangle: the binary tree's
// frontier reached by opening the largest-area internal node first. Written
// as FP32 lo.xyz hi.xyz, rounded outwards.
template <typename R> std::vector<float> cull_boxes(const Bvh<R>& b, int maxBoxes);
// The same cut for large maxBoxes (a heap instead of a scan per step; ties
// between equal areas may open in another order than cull_boxes opens them).
template <typename R> std::vector<float> cover_boxes(const Bvh<R>& b, int maxBoxes);

inline int encode_leaf(int firstTri, int count) { return ~((firstTri << 4) | (count - 1)); }

}  // namespace sbr
