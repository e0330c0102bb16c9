// TEST INFRASTRUCTURE — a C entry point around the reference's own model
// loader (src/modelLoader.cpp, with the tinyobjloader and tinygltf it
// vendors under include/model/), compiled from the reference sources by
// oracle/Makefile into oracle/_ref/. It only copies the loaded positions and
// indices out; the loading itself is the reference code.
#include <cstdint>
#include <cstring>
#include <string>

#include "modelLoader.hpp"

extern "C" {

// Loads path with modelLoader::loadModel (FileType::Auto). Returns an opaque
// handle (NULL on failure, with the loader's message in err[0..errLen)).
void* refload_open(const char* path, char* err, int errLen) {
    auto* m = new modelLoader::MeshData();
    std::string e;
    if (!modelLoader::loadModel(path, *m, e)) {
        if (err && errLen > 0) {
            std::strncpy(err, e.c_str(), size_t(errLen - 1));
            err[errLen - 1] = '\0';
        }
        delete m;
        return nullptr;
    }
    return m;
}

// out2 = number of vertices, number of triangles
void refload_info(void* h, int64_t* out2) {
    const auto* m = static_cast<const modelLoader::MeshData*>(h);
    out2[0] = int64_t(m->positions.size() / 3);
    out2[1] = int64_t(m->indices.size() / 3);
}

void refload_copy(void* h, float* xyz, uint32_t* idx) {
    const auto* m = static_cast<const modelLoader::MeshData*>(h);
    std::memcpy(xyz, m->positions.data(), m->positions.size() * sizeof(float));
    std::memcpy(idx, m->indices.data(), (m->indices.size() / 3) * 3 * sizeof(uint32_t));
}

void refload_close(void* h) { delete static_cast<modelLoader::MeshData*>(h); }

}  // extern "C"
