// TEST INFRASTRUCTURE — a C entry point around the reference's own BVH
// builder (src/scene/bvhBuilder.cpp:315-360), compiled from the reference
// sources by oracle/Makefile into oracle/_ref/. It only copies data in and
// out; the build itself is the reference code.
#include <cstdint>
#include <vector>

#include "scene/bvhBuilder.hpp"

extern "C" {

// tris9: nTris*9 floats (v0,v1,v2) in original order, already scaled.
// Outputs are sized by the caller for 2*nTris-1 nodes.
// Returns the node count; *root receives the root index.
int ref_bvh_build(const float* tris9, int nTris, int parallel, double* outTris9, double* bounds6, int* links4,
                  int* root) {
    std::vector<Triangle> tris(nTris);
    for (int i = 0; i < nTris; ++i) {
        const float* t = tris9 + 9 * i;
        tris[i].v0 = vec3(Real(t[0]), Real(t[1]), Real(t[2]));
        tris[i].v1 = vec3(Real(t[3]), Real(t[4]), Real(t[5]));
        tris[i].v2 = vec3(Real(t[6]), Real(t[7]), Real(t[8]));
    }
    bvhBuildOptions opt;
    opt.enableParallel = parallel != 0;
    bvhBuildResult r = buildBvh(std::move(tris), opt);
    for (size_t i = 0; i < r.triangles.size(); ++i) {
        const Triangle& t = r.triangles[i];
        double* o = outTris9 + 9 * i;
        o[0] = t.v0.x(); o[1] = t.v0.y(); o[2] = t.v0.z();
        o[3] = t.v1.x(); o[4] = t.v1.y(); o[5] = t.v1.z();
        o[6] = t.v2.x(); o[7] = t.v2.y(); o[8] = t.v2.z();
    }
    for (size_t i = 0; i < r.nodes.size(); ++i) {
        const BvhNode& n = r.nodes[i];
        double* b = bounds6 + 6 * i;
        b[0] = n.boundsMin.x(); b[1] = n.boundsMin.y(); b[2] = n.boundsMin.z();
        b[3] = n.boundsMax.x(); b[4] = n.boundsMax.y(); b[5] = n.boundsMax.z();
        int* l = links4 + 4 * i;
        l[0] = n.left; l[1] = n.right; l[2] = n.firstTri; l[3] = n.triCount;
    }
    *root = r.rootIndex;
    return int(r.nodes.size());
}

int ref_real_size(void) { return int(sizeof(Real)); }
}
