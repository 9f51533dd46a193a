// The C++ meshing entry (include/libfive/render/brep/mesh.hpp) on the GPU:
// the reference's known answers for a sphere (libfive/test/api.cpp:107-146),
// the same mesh as the C API, the Evaluator overload with a free variable,
// binary STL output, and cancellation (libfive/test/xtree.cpp:240-272: a
// cancelled render returns nullptr within 150 ms).  Prints one line per
// check and exits non-zero on the first failure.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <future>
#include <string>
#include <thread>

#include "libfive.h"
#include "libfive/eval/evaluator.hpp"
#include "libfive/render/brep/mesh.hpp"

using namespace libfive;

#define CHECK(cond, ...)                                                     \
    do {                                                                     \
        if (!(cond)) {                                                       \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__);                 \
            std::printf(__VA_ARGS__);                                        \
            std::printf("\n");                                               \
            return 1;                                                        \
        }                                                                    \
    } while (0)

static Tree sphere(double r) {
    const Tree x = Tree::X(), y = Tree::Y(), z = Tree::Z();
    const Tree s = Tree::binary(Opcode::OP_ADD,
                                Tree::binary(Opcode::OP_ADD, Tree::unary(Opcode::OP_SQUARE, x),
                                             Tree::unary(Opcode::OP_SQUARE, y)),
                                Tree::unary(Opcode::OP_SQUARE, z));
    return Tree::binary(Opcode::OP_SUB, Tree::unary(Opcode::OP_SQRT, s), Tree(r));
}

static Tree gyroid(double k) {
    const Tree x = Tree::binary(Opcode::OP_MUL, Tree::X(), Tree(k));
    const Tree y = Tree::binary(Opcode::OP_MUL, Tree::Y(), Tree(k));
    const Tree z = Tree::binary(Opcode::OP_MUL, Tree::Z(), Tree(k));
    auto cs = [](const Tree& a, const Tree& b) {
        return Tree::binary(Opcode::OP_MUL, Tree::unary(Opcode::OP_COS, a),
                            Tree::unary(Opcode::OP_SIN, b));
    };
    const Tree g = Tree::binary(Opcode::OP_ADD, Tree::binary(Opcode::OP_ADD, cs(x, y), cs(y, z)),
                                cs(z, x));
    return Tree::binary(Opcode::OP_SUB, Tree::unary(Opcode::OP_ABS, g), Tree(0.1));
}

int main() {
    if (!libfive_gpu_available()) {
        std::printf("SKIP no GPU\n");
        return 0;
    }
    // Sphere: every vertex at radius 1 (api.cpp:107-146 checks r^2 in
    // (0.99, 1.01) at resolution 10), verts[0] the zero dummy.
    BRepSettings settings;
    settings.min_feature = 0.1;
    // Bounds exact in float, so that the C API's float region is the same.
    const Region<3> r({-1.5, -1.5, -1.5}, {1.5, 1.5, 1.5});
    auto m = Mesh::render(sphere(1.0), r, settings);
    CHECK(m != nullptr, "render returned nullptr: %s", libfive_last_error());
    CHECK(m->verts.size() > 100 && m->branes.size() > 100, "too few elements");
    CHECK(m->verts[0].x() == 0 && m->verts[0].y() == 0 && m->verts[0].z() == 0, "dummy vertex");
    for (size_t i = 1; i < m->verts.size(); ++i) {
        const auto& v = m->verts[i];
        const double r2 = double(v.x()) * v.x() + double(v.y()) * v.y() + double(v.z()) * v.z();
        CHECK(r2 > 0.99 && r2 < 1.01, "vertex %zu at r^2 = %f", i, r2);
    }
    for (const auto& t : m->branes) {
        for (int j = 0; j < 3; ++j) CHECK(t[j] >= 1 && t[j] < m->verts.size(), "index");
    }
    std::printf("sphere: %zu verts, %zu triangles on the unit sphere\n",
                m->verts.size(), m->branes.size());

    // The same mesh as the C API (same renderer, same numbering).
    const Tree s1 = sphere(1.0);
    libfive_tree ct = reinterpret_cast<libfive_tree>(const_cast<Tree::Data*>(s1.get()));
    const libfive_region3 R{{-1.5f, 1.5f}, {-1.5f, 1.5f}, {-1.5f, 1.5f}};
    libfive_mesh* cm = libfive_tree_render_mesh(ct, R, 10.0f);
    CHECK(cm != nullptr, "C API render failed");
    CHECK(cm->vert_count == m->verts.size() && cm->tri_count == m->branes.size(),
          "C API mesh size %u/%u vs %zu/%zu", cm->vert_count, cm->tri_count,
          m->verts.size(), m->branes.size());
    CHECK(std::memcmp(cm->verts, m->verts.data(), 12 * size_t(cm->vert_count)) == 0, "verts differ");
    CHECK(std::memcmp(cm->tris, m->branes.data(), 12 * size_t(cm->tri_count)) == 0, "tris differ");
    libfive_mesh_delete(cm);
    std::printf("C API: identical vertices and triangles\n");

    // Evaluator overload: the radius is a free variable.
    const Tree rv = Tree::var();
    const Tree x = Tree::X(), y = Tree::Y(), z = Tree::Z();
    const Tree ss = Tree::binary(Opcode::OP_ADD,
                                 Tree::binary(Opcode::OP_ADD, Tree::unary(Opcode::OP_SQUARE, x),
                                              Tree::unary(Opcode::OP_SQUARE, y)),
                                 Tree::unary(Opcode::OP_SQUARE, z));
    const Tree sv = Tree::binary(Opcode::OP_SUB, Tree::unary(Opcode::OP_SQRT, ss), rv);
    Evaluator ev(sv, {{rv.id(), 0.5f}});
    auto mv = Mesh::render(&ev, r, settings);
    CHECK(mv != nullptr, "evaluator render failed: %s", libfive_last_error());
    for (size_t i = 1; i < mv->verts.size(); ++i) {
        const auto& v = mv->verts[i];
        const double rr = std::sqrt(double(v.x()) * v.x() + double(v.y()) * v.y() +
                                    double(v.z()) * v.z());
        CHECK(std::fabs(rr - 0.5) < 0.01, "radius %f, want 0.5", rr);
    }
    CHECK(ev.updateVars({{rv.id(), 1.0f}}), "updateVars reports a change");
    auto mv2 = Mesh::render(&ev, r, settings);
    CHECK(mv2 && mv2->verts.size() == m->verts.size(), "variable update not meshed");
    std::printf("Evaluator: radius 0.5 then 1.0\n");

    // Binary STL: 80-byte header, count, 50 bytes per triangle.
    const std::string path = "mesh_api_test.stl";
    CHECK(m->saveSTL(path), "saveSTL failed");
    FILE* f = std::fopen(path.c_str(), "rb");
    CHECK(f != nullptr, "no STL file");
    std::fseek(f, 0, SEEK_END);
    const long size = std::ftell(f);
    std::fclose(f);
    std::remove(path.c_str());
    CHECK(size == long(84 + 50 * m->branes.size()), "STL size %ld", size);
    std::printf("saveSTL: %ld bytes\n", size);

    // Cancellation (xtree.cpp:240-272): a long render cancelled right after
    // it starts returns nullptr within 150 ms.
    BRepSettings big;
    big.min_feature = 2.4 / 1024;
    const Region<3> rg({-1.2, -1.2, -1.2}, {1.2, 1.2, 1.2});
    const Tree g = gyroid(10.0);
    auto t0 = std::chrono::steady_clock::now();
    auto full = Mesh::render(g, rg, big);
    const double full_ms = std::chrono::duration<double, std::milli>(
        std::chrono::steady_clock::now() - t0).count();
    CHECK(full != nullptr, "uncancelled render failed: %s", libfive_last_error());
    auto fut = std::async(std::launch::async, [&]() { return Mesh::render(g, rg, big); });
    const auto start = std::chrono::steady_clock::now();
    big.cancel.store(true);
    auto res = fut.get();
    const double ms = std::chrono::duration<double, std::milli>(
        std::chrono::steady_clock::now() - start).count();
    CHECK(res == nullptr, "cancelled render returned a mesh");
    CHECK(ms < 150.0, "cancel took %.1f ms", ms);
    std::printf("cancel: nullptr after %.1f ms (the full render takes %.1f ms)\n", ms, full_ms);

    // A cancel requested part way through the octree build.
    big.cancel.store(false);
    auto fut2 = std::async(std::launch::async, [&]() { return Mesh::render(g, rg, big); });
    std::this_thread::sleep_for(std::chrono::milliseconds(30));
    const auto start2 = std::chrono::steady_clock::now();
    big.cancel.store(true);
    auto res2 = fut2.get();
    const double ms2 = std::chrono::duration<double, std::milli>(
        std::chrono::steady_clock::now() - start2).count();
    CHECK(res2 == nullptr, "cancelled render returned a mesh");
    CHECK(ms2 < 150.0, "mid-render cancel took %.1f ms", ms2);
    std::printf("cancel mid-render: nullptr after %.1f ms\n", ms2);
    std::printf("OK\n");
    return 0;
}
