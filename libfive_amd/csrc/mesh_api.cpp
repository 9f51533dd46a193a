// The C++ meshing entry, libfive::Mesh::render (libfive/include/libfive/
// render/brep/mesh.hpp:23-65, libfive/src/render/brep/mesh.cpp:35-125),
// over the same GPU mesher as libfive_tree_render_mesh.
#include "../../include/libfive/render/brep/mesh.hpp"
#include "../../include/libfive/eval/evaluator.hpp"

#include <stdexcept>

#include "mesher.hpp"
#include "stl.hpp"

namespace libfive {

namespace {

std::unique_ptr<Mesh> render_impl(const Tree& t,
                                  const std::map<Tree::Id, float>* vars,
                                  const Region<3>& r, const BRepSettings& s) {
    // The reference returns nullptr for an invalid min_feature (mesh.hpp:27).
    if (!(s.min_feature > 0.0)) return nullptr;
    try {
        if (s.alg != DUAL_CONTOURING) {
            throw std::runtime_error("only the DUAL_CONTOURING algorithm is available");
        }
        amd::MeshSettings ms;
        ms.min_feature = s.min_feature;
        ms.max_err = s.max_err;
        ms.cancel_flag = &s.cancel;
        ms.vars = vars;
        const double lo[3] = {r.lower[0], r.lower[1], r.lower[2]};
        const double hi[3] = {r.upper[0], r.upper[1], r.upper[2]};
        std::unique_ptr<Mesh> out(new Mesh);
        amd::HostMesh hm;
        hm.want_keys = false;
        size_t ntris = 0;
        hm.alloc = [&](size_t nv, size_t nt, float** v, uint32_t** tr) {
            out->verts.resize(nv);
            out->branes.resize(nt ? nt : 1);
            ntris = nt;
            *v = out->verts[0].v;
            *tr = out->branes[0].v;
            return false;
        };
        if (!amd::render_mesh(t, lo, hi, ms, hm)) return nullptr;
        out->branes.resize(ntris);
        return out;
    } catch (const std::exception& e) {
        amd::set_last_error("Mesh::render", e);
        return nullptr;
    }
}

libfive_mesh view(const Mesh& m) {
    static_assert(sizeof(MeshVertex) == sizeof(libfive_vec3), "vertex layout");
    static_assert(sizeof(MeshTriangle) == sizeof(libfive_tri), "triangle layout");
    libfive_mesh v;
    v.verts = reinterpret_cast<libfive_vec3*>(const_cast<MeshVertex*>(m.verts.data()));
    v.tris = reinterpret_cast<libfive_tri*>(const_cast<MeshTriangle*>(m.branes.data()));
    v.vert_count = uint32_t(m.verts.size());
    v.tri_count = uint32_t(m.branes.size());
    return v;
}

}   // namespace

std::unique_ptr<Mesh> Mesh::render(const Tree& t, const Region<3>& r,
                                   const BRepSettings& settings) {
    return render_impl(t, nullptr, r, settings);
}

std::unique_ptr<Mesh> Mesh::render(Evaluator* es, const Region<3>& r,
                                   const BRepSettings& settings) {
    return render_impl(es->tree(), &es->varValues(), r, settings);
}

bool Mesh::saveSTL(const std::string& filename) const {
    const libfive_mesh v = view(*this);
    return amd::save_stl(filename.c_str(), {&v});
}

bool Mesh::saveSTL(const std::string& filename,
                   const std::list<const Mesh*>& meshes) {
    std::vector<libfive_mesh> views;
    views.reserve(meshes.size());
    for (const Mesh* m : meshes) views.push_back(view(*m));
    std::vector<const libfive_mesh*> ptrs;
    for (const auto& v : views) ptrs.push_back(&v);
    return amd::save_stl(filename.c_str(), ptrs);
}

}   // namespace libfive
