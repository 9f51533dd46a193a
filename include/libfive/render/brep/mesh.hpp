// The C++ meshing entry: libfive::Mesh
// (libfive/include/libfive/render/brep/mesh.hpp:23-65 with the BRep<3>
// storage of brep.hpp:22-107): verts[0] is a reserved zero vertex, branes
// are index triples into verts.  Vertices and triangles are small value
// types with x() / y() / z(), operator[] and operator() where the
// reference uses Eigen vectors.
#pragma once

#include <cstdint>
#include <list>
#include <memory>
#include <string>
#include <vector>

#include "libfive/tree/tree.hpp"
#include "libfive/render/brep/region.hpp"
#include "libfive/render/brep/settings.hpp"

namespace libfive {

class Evaluator;

struct MeshVertex {
    float v[3];
    float x() const { return v[0]; }
    float y() const { return v[1]; }
    float z() const { return v[2]; }
    float& operator[](unsigned i) { return v[i]; }
    float operator[](unsigned i) const { return v[i]; }
    float operator()(unsigned i) const { return v[i]; }
};

struct MeshTriangle {
    uint32_t v[3];
    uint32_t& operator[](unsigned i) { return v[i]; }
    uint32_t operator[](unsigned i) const { return v[i]; }
    uint32_t operator()(unsigned i) const { return v[i]; }
};

class Mesh {
public:
    Mesh() { verts.push_back(MeshVertex{{0.0f, 0.0f, 0.0f}}); }

    std::vector<MeshVertex> verts;
    std::vector<MeshTriangle> branes;

    // Mesh::render (mesh.cpp:35-125): dual contouring of t over r.
    // Returns nullptr if settings.cancel is set during the render, if
    // min_feature is not positive, or if the GPU reports an error (with the
    // message in libfive_last_error()).
    static std::unique_ptr<Mesh> render(const Tree& t, const Region<3>& r,
                                        const BRepSettings& settings);
    // The same for the evaluator's tree and free-variable values.
    static std::unique_ptr<Mesh> render(Evaluator* es, const Region<3>& r,
                                        const BRepSettings& settings);

    // Binary STL (mesh.cpp:139-198).
    bool saveSTL(const std::string& filename) const;
    static bool saveSTL(const std::string& filename,
                        const std::list<const Mesh*>& meshes);
};

}   // namespace libfive
