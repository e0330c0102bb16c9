// Mesh file loaders for the drop-in surface: Wavefront OBJ and glTF 2.0
// (.gltf with external or data-URI buffers, and binary .glb).
//
// Behaviour follows the reference's modelLoader (src/modelLoader.cpp):
//   OBJ  (loadObjInternal, :59-178, through tinyobjloader): faces are
//        triangulated as tinyobj triangulates them (quads split on the
//        shorter diagonal, larger polygons by its built-in ear clipping),
//        vertices are deduplicated by (position, normal, texcoord) index
//        triple in first-use order. Numbers are parsed with tinyobj's
//        own accumulation (tinyOBJLoader.hpp:897-1028), which is not
//        correctly rounded, then narrowed to float -- so positions are
//        bit-identical to the reference's.
//   glTF (loadGltfInternal, :286-413): every TRIANGLES primitive of every
//        mesh, FLOAT POSITION accessors only (others yield no vertices),
//        u8/u16/u32 indices or implicit 0..n-1, node transforms ignored.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace sbr {

struct MeshData {
    std::vector<float> positions;  // x0 y0 z0 x1 ...
    std::vector<uint32_t> indices;  // 3 per triangle
};

// Extension-based dispatch (modelLoader.cpp:419-448). Returns false and
// fills err on failure.
bool load_model(const std::string& path, MeshData& out, std::string& err);
bool load_obj(const std::string& path, MeshData& out, std::string& err);
bool load_gltf(const std::string& path, MeshData& out, std::string& err);

// tinyobj's float parser semantics (see header comment); exposed for tests.
bool obj_parse_number(const char* s, const char* end, double* out);

}  // namespace sbr
