// Binary STL writer shared by the C API and the C++ Mesh class
// (Mesh::saveSTL, libfive/src/render/brep/mesh.cpp:139-198).
#pragma once

#include <exception>
#include <vector>

#include "../../include/libfive.h"

namespace libfive {
namespace amd {

// Writes every triangle of the meshes in order, with zero normals; false
// if the file cannot be written.
bool save_stl(const char* filename, const std::vector<const libfive_mesh*>& meshes);

// Records a failure for libfive_last_error() and prints it on stderr.
void set_last_error(const char* where, const std::exception& e);

}   // namespace amd
}   // namespace libfive
