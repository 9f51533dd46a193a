#include "mesher.hpp"

#include "runtime.hpp"

namespace libfive {
namespace amd {

static thread_local MeshStats g_stats;

const MeshStats& last_mesh_stats() { return g_stats; }

bool render_mesh(const Tree&, const double*, const double*,
                 const MeshSettings&, HostMesh&) {
    throw Error("render_mesh: the GPU mesher is not built yet");
}

}   // namespace amd
}   // namespace libfive
