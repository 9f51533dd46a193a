// extern "C" entry points (include/libfive.h).  Tree handling follows
// libfive/src/libfive.cpp:51-231; evaluation and meshing go to the GPU.
#include <algorithm>
#include <atomic>
#include <thread>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <sstream>
#include <string>

#include "../../include/libfive.h"
#include "autotune.hpp"
#include "eval_kernels.hpp"
#include "contours.hpp"
#include "heightmap.hpp"
#include "interval_kernels.hpp"
#include "mesher.hpp"
#include "tape_jit.hpp"

#include <map>
#include <mutex>
#include <vector>
#include "runtime.hpp"
#include "stl.hpp"
#include "tree.hpp"

using namespace libfive;
using namespace libfive::amd;

static thread_local std::string g_last_error;

static const TreeData* D(libfive_tree t) {
    return reinterpret_cast<const TreeData*>(t);
}
static libfive_tree H(const TreeData* d) {
    return reinterpret_cast<libfive_tree>(const_cast<TreeData*>(d));
}

static void set_error(const char* where, const std::exception& e) {
    g_last_error = std::string(where) + ": " + e.what();
    std::fprintf(stderr, "%s\n", g_last_error.c_str());
}

namespace libfive {
namespace amd {
void set_last_error(const char* where, const std::exception& e) { set_error(where, e); }
}   // namespace amd
}   // namespace libfive

// Vertex keys of the meshes handed out, released by libfive_mesh_delete.
static std::mutex g_keys_mu;
static std::map<const libfive_mesh*, std::vector<uint32_t>> g_keys;

struct libfive_evaluator_ {
    Tree tree;
    std::map<const void*, float> vars;
};

struct libfive_tape_ {
    std::unique_ptr<GpuTape> t;
    std::unique_ptr<SpecializedTape> spec;
    bool use_spec = false;
};

extern "C" {

const char* libfive_last_error(void) { return g_last_error.c_str(); }

int libfive_gpu_available(void) { return gpu_available() ? 1 : 0; }

// Mesh arrays.  A large mesh lives in one pooled page-locked block (the
// vertices, then the triangles at the next 256-byte boundary) that the
// device copies into directly and libfive_mesh_delete hands back to the
// pool; a small one uses new[], as the reference does.
static std::mutex g_blocks_mu;
static std::map<const libfive_mesh*, std::pair<void*, size_t>> g_blocks;
constexpr size_t kPinnedMeshMin = size_t(4) << 20;

static bool alloc_mesh_arrays(libfive_mesh* m, size_t nv, size_t nt, float** v,
                              uint32_t** t) {
    const size_t vb = nv * sizeof(libfive_vec3);
    const size_t tb = std::max<size_t>(nt, 1) * sizeof(libfive_tri);
    const size_t toff = (vb + 255) & ~size_t(255);
    bool pinned = false;
    if (toff + tb >= kPinnedMeshMin) {
        size_t granted = 0;
        char* p = static_cast<char*>(pinned_alloc(toff + tb, &granted));
        {
            std::lock_guard<std::mutex> lock(g_blocks_mu);
            g_blocks[m] = {p, granted};
        }
        m->verts = reinterpret_cast<libfive_vec3*>(p);
        m->tris = reinterpret_cast<libfive_tri*>(p + toff);
        pinned = true;
    } else {
        m->verts = new libfive_vec3[nv];
        m->tris = new libfive_tri[nt ? nt : 1];
    }
    m->vert_count = uint32_t(nv);
    m->tri_count = uint32_t(nt);
    *v = &m->verts[0].x;
    *t = &m->tris[0].a;
    return pinned;
}

static void free_mesh_arrays(libfive_mesh* m) {
    {
        std::lock_guard<std::mutex> lock(g_blocks_mu);
        auto it = g_blocks.find(m);
        if (it != g_blocks.end()) {
            pinned_free(it->second.first, it->second.second);
            g_blocks.erase(it);
            m->verts = nullptr;
            m->tris = nullptr;
            return;
        }
    }
    delete[] m->verts;
    delete[] m->tris;
    m->verts = nullptr;
    m->tris = nullptr;
}

void libfive_mesh_delete(libfive_mesh* m) {
    if (!m) return;
    {
        std::lock_guard<std::mutex> lock(g_keys_mu);
        g_keys.erase(m);
    }
    free_mesh_arrays(m);
    delete m;
}

const uint32_t* libfive_mesh_vertex_keys(const libfive_mesh* m) {
    std::lock_guard<std::mutex> lock(g_keys_mu);
    auto it = g_keys.find(m);
    return it == g_keys.end() ? nullptr : it->second.data();
}

int libfive_opcode_enum(const char* op) {
    auto o = Opcode::fromScmString(op);
    return (o == Opcode::INVALID || o == Opcode::LAST_OP) ? -1 : int(o);
}

int libfive_opcode_args(int op) {
    return (op >= 0 && op < Opcode::LAST_OP)
        ? Opcode::args(Opcode::Opcode(op)) : -1;
}

libfive_tree libfive_tree_x(void) { return H(Tree::X().release()); }
libfive_tree libfive_tree_y(void) { return H(Tree::Y().release()); }
libfive_tree libfive_tree_z(void) { return H(Tree::Z().release()); }
libfive_tree libfive_tree_const(float f) { return H(Tree(f).release()); }
libfive_tree libfive_tree_var(void) { return H(Tree::var().release()); }

bool libfive_tree_is_var(libfive_tree t) {
    return D(t)->op() == Opcode::VAR_FREE;
}

float libfive_tree_get_const(libfive_tree t, bool* success) {
    if (D(t)->op() == Opcode::CONSTANT) {
        if (success) *success = true;
        return D(t)->value();
    }
    if (success) *success = false;
    return 0;
}

static bool opcode_is_valid(int op, int expected_args) {
    return op >= 0 && op < Opcode::LAST_OP &&
           Opcode::args(Opcode::Opcode(op)) == expected_args;
}

libfive_tree libfive_tree_nullary(int op) {
    return opcode_is_valid(op, 0)
        ? H(Tree::nullary(Opcode::Opcode(op)).release()) : nullptr;
}

libfive_tree libfive_tree_unary(int op, libfive_tree a) {
    return (opcode_is_valid(op, 1) && a)
        ? H(Tree::unary(Opcode::Opcode(op), Tree(D(a))).release()) : nullptr;
}

libfive_tree libfive_tree_binary(int op, libfive_tree a, libfive_tree b) {
    return (opcode_is_valid(op, 2) && a && b)
        ? H(Tree::binary(Opcode::Opcode(op), Tree(D(a)), Tree(D(b))).release())
        : nullptr;
}

const void* libfive_tree_id(libfive_tree t) { return t; }

void libfive_tree_delete(libfive_tree ptr) {
    if (ptr) Tree::reclaim(D(ptr));
}

libfive_tree libfive_tree_remap(libfive_tree p, libfive_tree x,
                                libfive_tree y, libfive_tree z) {
    return (p && x && y && z)
        ? H(Tree(D(p)).remap(Tree(D(x)), Tree(D(y)), Tree(D(z))).release())
        : nullptr;
}

libfive_tree libfive_tree_optimized(libfive_tree t) {
    return H(Tree(D(t)).optimized().release());
}

char* libfive_tree_print(libfive_tree t) {
    std::stringstream ss;
    ss << Tree(D(t));
    const auto str = ss.str();
    auto out = static_cast<char*>(std::malloc(str.size() + 1));
    std::memcpy(out, str.c_str(), str.size() + 1);
    return out;
}

void libfive_free_str(char* ptr) { std::free(ptr); }

////////////////////////////////////////////////////////////////////////////////
// Point evaluation

static float eval_point(libfive_tree t, libfive_vec3 p, float grad[3]) {
    Context& c = context();
    auto gt = compile_tree(Tree(D(t)));
    std::lock_guard<std::mutex> lock(c.mu);
    c.h_small[0] = p.x;
    c.h_small[1] = p.y;
    c.h_small[2] = p.z;
    float* d = c.d_small;
    hip_check(hipMemcpyAsync(d, c.h_small, 3 * sizeof(float),
                             hipMemcpyHostToDevice, c.stream), "copy point");
    const auto& h = gt->host;
    if (grad) {
        hip_check(launch_eval_derivs(gt->d.p, gt->size(), h.root_reg,
                                     h.num_regs, tape_is_fast(h), d, d + 1, d + 2, d + 4, d + 5,
                                     d + 6, d + 7, 1, c.stream),
                  "eval_d");
    } else {
        hip_check(launch_eval_values(gt->d.p, gt->size(), h.root_reg,
                                     h.num_regs, tape_is_fast(h), d, d + 1, d + 2, d + 4, 1,
                                     c.stream),
                  "eval_f");
    }
    hip_check(hipMemcpyAsync(c.h_small + 4, d + 4, 4 * sizeof(float),
                             hipMemcpyDeviceToHost, c.stream), "copy result");
    hip_check(hipStreamSynchronize(c.stream), "eval sync");
    if (grad) {
        grad[0] = c.h_small[5];
        grad[1] = c.h_small[6];
        grad[2] = c.h_small[7];
    }
    return c.h_small[4];
}

float libfive_tree_eval_f(libfive_tree t, libfive_vec3 p) {
    try {
        return eval_point(t, p, nullptr);
    } catch (const std::exception& e) {
        set_error("libfive_tree_eval_f", e);
        return std::numeric_limits<float>::quiet_NaN();
    }
}

libfive_vec3 libfive_tree_eval_d(libfive_tree t, libfive_vec3 p) {
    try {
        float g[3];
        eval_point(t, p, g);
        return {g[0], g[1], g[2]};
    } catch (const std::exception& e) {
        set_error("libfive_tree_eval_d", e);
        const float n = std::numeric_limits<float>::quiet_NaN();
        return {n, n, n};
    }
}

libfive_interval libfive_tree_eval_r(libfive_tree t, libfive_region3 r) {
    float lo = 0, hi = 0;
    const int rc = libfive_tree_eval_interval_array(
        t, &r.X.lower, &r.X.upper, &r.Y.lower, &r.Y.upper, &r.Z.lower,
        &r.Z.upper, 1, &lo, &hi, nullptr);
    if (rc) {
        const float n = std::numeric_limits<float>::quiet_NaN();
        return {n, n};
    }
    return {lo, hi};
}

int libfive_tree_eval_array(libfive_tree t, const float* x, const float* y,
                            const float* z, size_t n, float* out) {
    try {
        Context& c = context();
        auto gt = compile_tree(Tree(D(t)));
        DevBuf<float> buf(4 * n);
        std::lock_guard<std::mutex> lock(c.mu);
        const size_t b = n * sizeof(float);
        hip_check(hipMemcpyAsync(buf.p, x, b, hipMemcpyHostToDevice, c.stream), "copy x");
        hip_check(hipMemcpyAsync(buf.p + n, y, b, hipMemcpyHostToDevice, c.stream), "copy y");
        hip_check(hipMemcpyAsync(buf.p + 2 * n, z, b, hipMemcpyHostToDevice, c.stream), "copy z");
        hip_check(launch_eval_values(gt->d.p, gt->size(), gt->host.root_reg,
                                     gt->host.num_regs, tape_is_fast(gt->host), buf.p, buf.p + n,
                                     buf.p + 2 * n, buf.p + 3 * n, n, c.stream),
                  "eval_values");
        hip_check(hipMemcpyAsync(out, buf.p + 3 * n, b, hipMemcpyDeviceToHost, c.stream), "copy out");
        hip_check(hipStreamSynchronize(c.stream), "eval_array sync");
        return 0;
    } catch (const std::exception& e) {
        set_error("libfive_tree_eval_array", e);
        return -1;
    }
}

int libfive_tree_eval_deriv_array(libfive_tree t, const float* x,
                                  const float* y, const float* z, size_t n,
                                  float* value, float* dx, float* dy,
                                  float* dz) {
    try {
        Context& c = context();
        auto gt = compile_tree(Tree(D(t)));
        DevBuf<float> buf(7 * n);
        std::lock_guard<std::mutex> lock(c.mu);
        const size_t b = n * sizeof(float);
        float* p = buf.p;
        hip_check(hipMemcpyAsync(p, x, b, hipMemcpyHostToDevice, c.stream), "copy x");
        hip_check(hipMemcpyAsync(p + n, y, b, hipMemcpyHostToDevice, c.stream), "copy y");
        hip_check(hipMemcpyAsync(p + 2 * n, z, b, hipMemcpyHostToDevice, c.stream), "copy z");
        hip_check(launch_eval_derivs(gt->d.p, gt->size(), gt->host.root_reg,
                                     gt->host.num_regs, tape_is_fast(gt->host), p, p + n, p + 2 * n,
                                     p + 3 * n, p + 4 * n, p + 5 * n, p + 6 * n,
                                     n, c.stream),
                  "eval_derivs");
        hip_check(hipMemcpyAsync(value, p + 3 * n, b, hipMemcpyDeviceToHost, c.stream), "copy v");
        hip_check(hipMemcpyAsync(dx, p + 4 * n, b, hipMemcpyDeviceToHost, c.stream), "copy dx");
        hip_check(hipMemcpyAsync(dy, p + 5 * n, b, hipMemcpyDeviceToHost, c.stream), "copy dy");
        hip_check(hipMemcpyAsync(dz, p + 6 * n, b, hipMemcpyDeviceToHost, c.stream), "copy dz");
        hip_check(hipStreamSynchronize(c.stream), "eval_deriv_array sync");
        return 0;
    } catch (const std::exception& e) {
        set_error("libfive_tree_eval_deriv_array", e);
        return -1;
    }
}

static int eval_ambiguous(const char* where, libfive_tree t, bool derivs,
                          const float* x, const float* y, const float* z,
                          size_t n, uint8_t* out) {
    try {
        Context& c = context();
        auto gt = compile_tree(Tree(D(t)));
        DevBuf<float> buf(3 * n);
        DevBuf<uint8_t> flags(n ? n : 1);
        std::lock_guard<std::mutex> lock(c.mu);
        const size_t b = n * sizeof(float);
        float* p = buf.p;
        hip_check(hipMemcpyAsync(p, x, b, hipMemcpyHostToDevice, c.stream), "copy x");
        hip_check(hipMemcpyAsync(p + n, y, b, hipMemcpyHostToDevice, c.stream), "copy y");
        hip_check(hipMemcpyAsync(p + 2 * n, z, b, hipMemcpyHostToDevice, c.stream), "copy z");
        hip_check(launch_eval_ambiguous(gt->d.p, gt->size(), gt->host.root_reg,
                                        gt->host.num_regs, derivs, p, p + n,
                                        p + 2 * n, flags.p, n, c.stream),
                  "eval_ambiguous");
        hip_check(hipMemcpyAsync(out, flags.p, n, hipMemcpyDeviceToHost, c.stream),
                  "copy flags");
        hip_check(hipStreamSynchronize(c.stream), "eval_ambiguous sync");
        return 0;
    } catch (const std::exception& e) {
        set_error(where, e);
        return -1;
    }
}

int libfive_tree_eval_ambiguous_array(libfive_tree t, const float* x,
                                      const float* y, const float* z,
                                      size_t n, uint8_t* out) {
    return eval_ambiguous("libfive_tree_eval_ambiguous_array", t, false, x, y,
                          z, n, out);
}

int libfive_tree_eval_deriv_ambiguous_array(libfive_tree t, const float* x,
                                            const float* y, const float* z,
                                            size_t n, uint8_t* out) {
    return eval_ambiguous("libfive_tree_eval_deriv_ambiguous_array", t, true, x,
                          y, z, n, out);
}

int libfive_tree_eval_interval_array(libfive_tree t, const float* xlo,
                                     const float* xhi, const float* ylo,
                                     const float* yhi, const float* zlo,
                                     const float* zhi, size_t n,
                                     float* lower, float* upper,
                                     uint8_t* maybe_nan) {
    try {
        Context& c = context();
        auto gt = compile_tree(Tree(D(t)));
        DevBuf<float> buf(8 * n);
        DevBuf<uint8_t> nanbuf(n ? n : 1);
        std::lock_guard<std::mutex> lock(c.mu);
        const size_t b = n * sizeof(float);
        float* p = buf.p;
        const float* src[6] = {xlo, xhi, ylo, yhi, zlo, zhi};
        for (int k = 0; k < 6; ++k) {
            hip_check(hipMemcpyAsync(p + k * n, src[k], b, hipMemcpyHostToDevice,
                                     c.stream), "copy box");
        }
        hip_check(launch_eval_intervals(gt->d.p, gt->size(), gt->host.root_reg,
                                        gt->host.num_regs, p, p + n, p + 2 * n,
                                        p + 3 * n, p + 4 * n, p + 5 * n,
                                        p + 6 * n, p + 7 * n, nanbuf.p, n,
                                        c.stream),
                  "eval_intervals");
        hip_check(hipMemcpyAsync(lower, p + 6 * n, b, hipMemcpyDeviceToHost, c.stream), "copy lo");
        hip_check(hipMemcpyAsync(upper, p + 7 * n, b, hipMemcpyDeviceToHost, c.stream), "copy hi");
        if (maybe_nan) {
            hip_check(hipMemcpyAsync(maybe_nan, nanbuf.p, n, hipMemcpyDeviceToHost,
                                     c.stream), "copy nan");
        }
        hip_check(hipStreamSynchronize(c.stream), "eval_interval sync");
        return 0;
    } catch (const std::exception& e) {
        set_error("libfive_tree_eval_interval_array", e);
        return -1;
    }
}

////////////////////////////////////////////////////////////////////////////////
// Compiled tapes

libfive_tape libfive_tape_new(libfive_tree t) {
    try {
        auto out = new libfive_tape_;
        out->t = compile_tree(Tree(D(t)));
        return out;
    } catch (const std::exception& e) {
        set_error("libfive_tape_new", e);
        return nullptr;
    }
}

void libfive_tape_delete(libfive_tape t) { delete t; }

uint32_t libfive_tape_clauses(libfive_tape t) { return t->t->host.ref_length; }
uint32_t libfive_tape_registers(libfive_tape t) { return t->t->host.num_regs; }

// Evaluation calls of at least kTuneMin points tune their launch schedule
// on their first kTuneSlice points (autotune.hpp).
static constexpr size_t kTuneMin = size_t(1) << 20;
static constexpr size_t kTuneSlice = size_t(1) << 22;
static const unsigned kBlocks[3] = {256, 128, 64};
static const std::vector<std::string> kBlockNames = {
    "256 lanes per workgroup", "128 lanes per workgroup", "64 lanes per workgroup"};

static std::string spec_key(const char* kernel, const SpecializedTape& k) {
    return std::string(kernel) + "/tape=" + std::to_string(k.serial);
}

int libfive_tape_eval_values_device(libfive_tape t, const float* x,
                                    const float* y, const float* z, size_t n,
                                    float* out, void* stream) {
    try {
        Context& c = context();
        hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c.stream;
        const auto& h = t->t->host;
        // Launch schedules are tuned on the first large call (a slice of
        // its own points; every schedule gives the same values).
        const size_t slice = std::min<size_t>(n, kTuneSlice);
        if (t->spec && t->use_spec) {
            const SpecializedTape& k = *t->spec;
            const int c = n < kTuneMin ? std::max(0, tune::known(spec_key("lf_values", k)))
                : tune::pick(spec_key("lf_values", k), kBlockNames, [&](int b) {
                      return launch_specialized_values(k, x, y, z, out, slice, s, kBlocks[b]);
                  }, s);
            hip_check(launch_specialized_values(k, x, y, z, out, n, s, kBlocks[c % 3]),
                      "eval_values (specialised)");
            if (k.redo) {
                hip_check(launch_eval_redo(t->t->d.p, t->t->size(), h.root_reg, h.num_regs,
                                           x, y, z, out, nullptr, nullptr, nullptr, n, s),
                          "eval_values (redo)");
            }
            return 0;
        }
        int ppl = default_value_points(h.num_regs);
        const std::vector<int> opts = value_points_options(h.num_regs);
        if (!value_points_forced() && opts.size() > 1) {
            std::vector<std::string> names;
            for (int p : opts) names.push_back(std::to_string(p) + " points per lane");
            const std::string key = "eval_values/banks=" + std::to_string((h.num_regs + 31) / 32);
            const int c = n < kTuneMin ? tune::known(key) : tune::pick(key, names, [&](int k) {
                return launch_eval_values_ppl(t->t->d.p, t->t->size(), h.root_reg, h.num_regs, tape_is_fast(h),
                                              x, y, z, out, slice, s, opts[k]);
            }, s);
            if (c >= 0 && c < int(opts.size())) ppl = opts[c];
        }
        hip_check(launch_eval_values_ppl(t->t->d.p, t->t->size(), h.root_reg,
                                         h.num_regs, tape_is_fast(h), x, y, z, out, n, s, ppl),
                  "eval_values");
        return 0;
    } catch (const std::exception& e) {
        set_error("libfive_tape_eval_values_device", e);
        return -1;
    }
}

int libfive_tape_eval_derivs_device(libfive_tape t, const float* x,
                                    const float* y, const float* z, size_t n,
                                    float* value, float* dx, float* dy,
                                    float* dz, void* stream) {
    try {
        Context& c = context();
        hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c.stream;
        const auto& h = t->t->host;
        if (t->spec && t->use_spec) {
            const SpecializedTape& k = *t->spec;
            const size_t slice = std::min<size_t>(n, kTuneSlice);
            const int c = n < kTuneMin ? std::max(0, tune::known(spec_key("lf_derivs", k)))
                : tune::pick(spec_key("lf_derivs", k), kBlockNames, [&](int b) {
                      return launch_specialized_derivs(k, x, y, z, value, dx, dy, dz, slice, s,
                                                       kBlocks[b]);
                  }, s);
            hip_check(launch_specialized_derivs(k, x, y, z, value, dx, dy, dz, n, s,
                                                kBlocks[c % 3]),
                      "eval_derivs (specialised)");
            if (k.redo) {
                hip_check(launch_eval_redo(t->t->d.p, t->t->size(), h.root_reg, h.num_regs,
                                           x, y, z, value, dx, dy, dz, n, s),
                          "eval_derivs (redo)");
            }
            return 0;
        }
        hip_check(launch_eval_derivs(t->t->d.p, t->t->size(), h.root_reg,
                                     h.num_regs, tape_is_fast(h), x, y, z, value, dx, dy, dz, n,
                                     s),
                  "eval_derivs");
        return 0;
    } catch (const std::exception& e) {
        set_error("libfive_tape_eval_derivs_device", e);
        return -1;
    }
}

int libfive_tape_eval_values(libfive_tape t, const float* x, const float* y,
                             const float* z, size_t n, float* out) {
    try {
        Context& c = context();
        DevBuf<float> buf(4 * n);
        std::lock_guard<std::mutex> lock(c.mu);
        const size_t b = n * sizeof(float);
        float* p = buf.p;
        hip_check(hipMemcpyAsync(p, x, b, hipMemcpyHostToDevice, c.stream), "copy x");
        hip_check(hipMemcpyAsync(p + n, y, b, hipMemcpyHostToDevice, c.stream), "copy y");
        hip_check(hipMemcpyAsync(p + 2 * n, z, b, hipMemcpyHostToDevice, c.stream), "copy z");
        if (libfive_tape_eval_values_device(t, p, p + n, p + 2 * n, n, p + 3 * n,
                                            c.stream)) {
            return -1;
        }
        hip_check(hipMemcpyAsync(out, p + 3 * n, b, hipMemcpyDeviceToHost, c.stream), "copy out");
        hip_check(hipStreamSynchronize(c.stream), "tape eval sync");
        return 0;
    } catch (const std::exception& e) {
        set_error("libfive_tape_eval_values", e);
        return -1;
    }
}

int libfive_tape_eval_values_noalloc(libfive_tape t, const float* x,
                                     const float* y, const float* z, size_t n,
                                     float* out) {
    try {
        Context& c = context();
        if (n > kNoAllocMaxPoints) {
            throw Error("at most " + std::to_string(kNoAllocMaxPoints) +
                        " points per call, got " + std::to_string(n));
        }
        if (!c.scratch) throw Error("no staging memory (" + c.scratch_error + ")");
        if (n == 0) return 0;
        // Calls serialise on the library lock: the staging array is shared
        // by every call on this GPU and keeps the last call's contents.
        std::lock_guard<std::mutex> lock(c.mu);
        const size_t b = n * sizeof(float);
        float* px = c.scratch;
        float* py = px + kNoAllocMaxPoints;
        float* pz = py + kNoAllocMaxPoints;
        float* po = pz + kNoAllocMaxPoints;
        hip_check(hipMemcpyAsync(px, x, b, hipMemcpyHostToDevice, c.stream), "copy x");
        hip_check(hipMemcpyAsync(py, y, b, hipMemcpyHostToDevice, c.stream), "copy y");
        hip_check(hipMemcpyAsync(pz, z, b, hipMemcpyHostToDevice, c.stream), "copy z");
        const auto& h = t->t->host;
        if (t->spec && t->use_spec) {
            hip_check(launch_specialized_values(*t->spec, px, py, pz, po, n, c.stream),
                      "eval_values (specialised)");
            if (t->spec->redo) {
                hip_check(launch_eval_redo(t->t->d.p, t->t->size(), h.root_reg, h.num_regs,
                                           px, py, pz, po, nullptr, nullptr, nullptr, n,
                                           c.stream),
                          "eval_values (redo)");
            }
        } else {
            hip_check(launch_eval_values(t->t->d.p, t->t->size(), h.root_reg,
                                         h.num_regs, tape_is_fast(h), px, py, pz, po, n, c.stream),
                      "eval_values");
        }
        hip_check(hipMemcpyAsync(out, po, b, hipMemcpyDeviceToHost, c.stream), "copy out");
        hip_check(hipStreamSynchronize(c.stream), "tape eval sync");
        return 0;
    } catch (const std::exception& e) {
        set_error("libfive_tape_eval_values_noalloc", e);
        return -1;
    }
}

int libfive_tape_eval_derivs(libfive_tape t, const float* x, const float* y,
                             const float* z, size_t n, float* value, float* dx,
                             float* dy, float* dz) {
    try {
        Context& c = context();
        DevBuf<float> buf(7 * n);
        std::lock_guard<std::mutex> lock(c.mu);
        const size_t b = n * sizeof(float);
        float* p = buf.p;
        hip_check(hipMemcpyAsync(p, x, b, hipMemcpyHostToDevice, c.stream), "copy x");
        hip_check(hipMemcpyAsync(p + n, y, b, hipMemcpyHostToDevice, c.stream), "copy y");
        hip_check(hipMemcpyAsync(p + 2 * n, z, b, hipMemcpyHostToDevice, c.stream), "copy z");
        if (libfive_tape_eval_derivs_device(t, p, p + n, p + 2 * n, n, p + 3 * n,
                                            p + 4 * n, p + 5 * n, p + 6 * n,
                                            c.stream)) {
            return -1;
        }
        float* outs[4] = {value, dx, dy, dz};
        for (int k = 0; k < 4; ++k) {
            hip_check(hipMemcpyAsync(outs[k], p + (3 + k) * n, b,
                                     hipMemcpyDeviceToHost, c.stream), "copy out");
        }
        hip_check(hipStreamSynchronize(c.stream), "tape eval sync");
        return 0;
    } catch (const std::exception& e) {
        set_error("libfive_tape_eval_derivs", e);
        return -1;
    }
}

int libfive_tape_specialize(libfive_tape t, double* compile_ms) {
    try {
        if (!t->spec) t->spec = specialize(t->t->host);
        t->use_spec = true;
        if (compile_ms) *compile_ms = t->spec->compile_ms;
        return 0;
    } catch (const std::exception& e) {
        set_error("libfive_tape_specialize", e);
        return -1;
    }
}

void libfive_tape_use_specialized(libfive_tape t, int on) {
    t->use_spec = (on != 0);
}

size_t libfive_tree_specialized_code_size(libfive_tree t) {
    try {
        Deck deck{Tree(D(t))};
        return compile_specialized(compile_deck(deck)).size();
    } catch (const std::exception& e) {
        set_error("libfive_tree_specialized_code_size", e);
        return 0;
    }
}

void libfive_set_points_per_lane(int p) { set_points_per_lane(p); }

int libfive_tree_tape_info(libfive_tree t, uint32_t* clauses,
                           uint32_t* registers) {
    try {
        Deck deck{Tree(D(t))};
        const DeviceTape dt = compile_deck(deck);
        if (clauses) *clauses = dt.ref_length;
        if (registers) *registers = dt.num_regs;
        return 0;
    } catch (const std::exception& e) {
        set_error("libfive_tree_tape_info", e);
        return -1;
    }
}

int libfive_tree_deck(libfive_tree t, uint32_t* clauses, size_t* n_clauses,
                      uint32_t* const_ids, float* const_values,
                      size_t* n_consts, uint32_t* xyz_ids,
                      uint32_t* num_slots) {
    try {
        Deck deck{Tree(D(t))};
        if (clauses) {
            for (size_t i = 0; i < deck.clauses.size(); ++i) {
                clauses[4 * i + 0] = deck.clauses[i].op;
                clauses[4 * i + 1] = deck.clauses[i].id;
                clauses[4 * i + 2] = deck.clauses[i].a;
                clauses[4 * i + 3] = deck.clauses[i].b;
            }
        }
        if (n_clauses) *n_clauses = deck.clauses.size();
        if (const_ids && const_values) {
            for (size_t i = 0; i < deck.constants.size(); ++i) {
                const_ids[i] = deck.constants[i].first;
                const_values[i] = deck.constants[i].second;
            }
        }
        if (n_consts) *n_consts = deck.constants.size();
        if (xyz_ids) {
            xyz_ids[0] = deck.X;
            xyz_ids[1] = deck.Y;
            xyz_ids[2] = deck.Z;
        }
        if (num_slots) *num_slots = uint32_t(deck.num_clauses);
        return 0;
    } catch (const std::exception& e) {
        set_error("libfive_tree_deck", e);
        return -1;
    }
}

int libfive_tree_deck_vars(libfive_tree t, uint32_t* slot_ids, const void** var_ids,
                           size_t* n) {
    try {
        Deck deck{Tree(D(t))};
        if (slot_ids && var_ids) {
            for (size_t i = 0; i < deck.vars.size(); ++i) {
                slot_ids[i] = deck.vars[i];
                var_ids[i] = deck.var_nodes[i];
            }
        }
        if (n) *n = deck.vars.size();
        return 0;
    } catch (const std::exception& e) {
        set_error("libfive_tree_deck_vars", e);
        return -1;
    }
}

////////////////////////////////////////////////////////////////////////////////
// Meshing

void libfive_mesh_settings_default(libfive_mesh_settings* s) {
    s->min_feature = 0.1;
    s->max_err = 1e-8;
    s->precision = 0;
    s->collect_stats = 0;
    s->cancel = nullptr;
}

size_t libfive_last_mesh_level_counts(uint64_t* out, size_t cap) {
    const auto& st = last_mesh_stats();
    const size_t n = st.level_counts.size() / 3;
    for (size_t i = 0; out && i < 3 * std::min(n, cap); ++i) out[i] = st.level_counts[i];
    return n;
}

size_t libfive_last_mesh_tape_lengths(uint32_t* out, size_t cap) {
    const auto& st = last_mesh_stats();
    for (size_t i = 0; out && i < std::min(st.tape_lengths.size(), cap); ++i) {
        out[i] = st.tape_lengths[i];
    }
    return st.tape_lengths.size();
}

size_t libfive_last_mesh_cells(uint32_t* out, size_t cap) {
    const auto& st = last_mesh_stats();
    const size_t n = st.cells.size() / 5;
    for (size_t i = 0; out && i < 5 * std::min(n, cap); ++i) out[i] = st.cells[i];
    return n;
}

void libfive_last_mesh_times(double* ms4) {
    const auto& st = last_mesh_stats();
    ms4[0] = st.ms_build;
    ms4[1] = st.ms_collapse;
    ms4[2] = st.ms_emit;
    ms4[3] = st.ms_total;
}

// Renders into a new libfive_mesh whose arrays (alloc_mesh_arrays, released
// by libfive_mesh_delete) receive the copy from the device.  Returns NULL
// when cancelled.
static libfive_mesh* render_to_c(const Tree& tree, const double lo[3],
                                 const double hi[3], const MeshSettings& ms,
                                 bool keys, hipStream_t* stream) {
    std::unique_ptr<libfive_mesh> out(new libfive_mesh{nullptr, nullptr, 0, 0});
    struct Arrays {
        libfive_mesh* m;
        ~Arrays() {
            if (m) free_mesh_arrays(m);
        }
    } arrays{out.get()};
    HostMesh hm;
    hm.want_keys = keys;
    hm.alloc = [&](size_t nv, size_t nt, float** v, uint32_t** t) {
        return alloc_mesh_arrays(out.get(), nv, nt, v, t);
    };
    const bool done = stream ? render_mesh_on(tree, lo, hi, ms, hm, *stream)
                             : render_mesh(tree, lo, hi, ms, hm);
    if (!done) return nullptr;
    arrays.m = nullptr;
    if (keys) {
        std::lock_guard<std::mutex> lock(g_keys_mu);
        g_keys[out.get()] = std::move(hm.keys);
    }
    return out.release();
}

static libfive_mesh* render(libfive_tree tree, libfive_region3 R,
                            const libfive_mesh_settings& s, const char* who) {
    try {
        MeshSettings ms;
        ms.min_feature = s.min_feature;
        ms.max_err = s.max_err;
        ms.precision = s.precision;
        if (s.precision > 1) {
            throw Error("precision " + std::to_string(s.precision) +
                        " is not available (0: FP64 QEF accumulation, 1: compensated FP32)");
        }
        ms.collect_stats = s.collect_stats != 0;
        ms.cancel = s.cancel;
        const double lo[3] = {R.X.lower, R.Y.lower, R.Z.lower};
        const double hi[3] = {R.X.upper, R.Y.upper, R.Z.upper};
        libfive_mesh* m = render_to_c(Tree(D(tree)), lo, hi, ms,
                                      s.collect_stats != 0, nullptr);
        if (!m) std::fprintf(stderr, "libfive_tree_render_mesh: got empty mesh\n");
        return m;
    } catch (const std::exception& e) {
        set_error(who, e);
        return nullptr;
    }
}

int libfive_trees_render_mesh_batch(const libfive_tree* trees,
                                    const libfive_region3* regions,
                                    const float* res, size_t n, int streams,
                                    libfive_mesh** out) {
    for (size_t i = 0; i < n; ++i) out[i] = nullptr;
    if (n == 0) return 0;
    std::vector<hipStream_t> ss;
    try {
        context();
        const int k = int(std::max<size_t>(1, std::min<size_t>(
            n, size_t(streams > 0 ? streams : 8))));
        for (int i = 0; i < k; ++i) {
            hipStream_t st = nullptr;
            hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking),
                      "batch stream");
            ss.push_back(st);
        }
        // Largest models first (cells grow with res^3) so that the last
        // models to start are the short ones.
        std::vector<size_t> order(n);
        for (size_t i = 0; i < n; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
            return res[a] > res[b];
        });
        std::atomic<size_t> next{0};
        std::atomic<int> failed{0};
        std::mutex err_mu;
        std::string first_error;
        auto worker = [&](int w) {
            for (;;) {
                const size_t j = next.fetch_add(1);
                if (j >= n) break;
                const size_t i = order[j];
                try {
                    MeshSettings ms;
                    ms.min_feature = 1 / res[i];
                    const libfive_region3& R = regions[i];
                    const double lo[3] = {R.X.lower, R.Y.lower, R.Z.lower};
                    const double hi[3] = {R.X.upper, R.Y.upper, R.Z.upper};
                    out[i] = render_to_c(Tree(D(trees[i])), lo, hi, ms, false, &ss[w]);
                    if (!out[i]) failed++;
                } catch (const std::exception& e) {
                    failed++;
                    std::lock_guard<std::mutex> lock(err_mu);
                    if (first_error.empty()) {
                        first_error = "model " + std::to_string(i) + ": " + e.what();
                    }
                }
            }
        };
        std::vector<std::thread> pool;
        for (int w = 0; w < k; ++w) pool.emplace_back(worker, w);
        for (auto& th : pool) th.join();
        for (auto st : ss) (void)hipStreamDestroy(st);
        ss.clear();
        if (failed) {
            set_error("libfive_trees_render_mesh_batch",
                      Error(first_error.empty() ? "cancelled" : first_error));
        }
        return failed;
    } catch (const std::exception& e) {
        for (auto st : ss) (void)hipStreamDestroy(st);
        set_error("libfive_trees_render_mesh_batch", e);
        return int(n);
    }
}

int libfive_trees_render_mesh_merged(const libfive_tree* trees,
                                     const libfive_region3* regions,
                                     const float* res, size_t n, libfive_mesh** out) {
    for (size_t i = 0; i < n; ++i) out[i] = nullptr;
    if (n == 0) return 0;
    std::vector<std::unique_ptr<libfive_mesh>> meshes(n);
    auto free_arrays = [&]() {
        for (auto& m : meshes) {
            if (m) {
                free_mesh_arrays(m.get());
                m.reset();
            }
        }
    };
    try {
        std::vector<Tree> tr;
        tr.reserve(n);
        for (size_t i = 0; i < n; ++i) tr.emplace_back(D(trees[i]));
        std::vector<BatchItem> items(n);
        std::vector<HostMesh> hms(n);
        for (size_t i = 0; i < n; ++i) {
            const libfive_region3& R = regions[i];
            items[i].tree = &tr[i];
            items[i].lo[0] = R.X.lower; items[i].lo[1] = R.Y.lower; items[i].lo[2] = R.Z.lower;
            items[i].hi[0] = R.X.upper; items[i].hi[1] = R.Y.upper; items[i].hi[2] = R.Z.upper;
            items[i].min_feature = 1 / res[i];      // float division, as libfive.cpp:343
            meshes[i].reset(new libfive_mesh{nullptr, nullptr, 0, 0});
            hms[i].want_keys = false;
            libfive_mesh* m = meshes[i].get();
            hms[i].alloc = [m](size_t nv, size_t nt, float** v, uint32_t** t) {
                return alloc_mesh_arrays(m, nv, nt, v, t);
            };
        }
        Context& ctx = context();
        std::lock_guard<std::mutex> lock(ctx.mu);
        render_meshes_merged(items.data(), n, hms.data(), ctx.stream);
        for (size_t i = 0; i < n; ++i) {
            if (!meshes[i]->verts) throw Error("model " + std::to_string(i) + " was not meshed");
        }
        for (size_t i = 0; i < n; ++i) out[i] = meshes[i].release();
        return 0;
    } catch (const std::exception& e) {
        free_arrays();
        set_error("libfive_trees_render_mesh_merged", e);
        return int(n);
    }
}

void libfive_autotune_enable(int on) { tune::set_enabled(on != 0); }

void libfive_autotune_reset() { tune::reset(); }

void libfive_autotune_force(const char* key, int choice) {
    if (key) tune::force(key, choice);
}

size_t libfive_autotune_report(char* buf, size_t cap) {
    std::string out = "[";
    const auto ds = tune::decisions();
    for (size_t i = 0; i < ds.size(); ++i) {
        const auto& d = ds[i];
        out += i ? ", " : "";
        out += "{\"key\": \"" + d.key + "\", \"choice\": \"" + d.names[size_t(d.choice)] +
               "\", \"candidates_ms\": {";
        for (size_t c = 0; c < d.names.size(); ++c) {
            char num[32];
            std::snprintf(num, sizeof(num), "%.4f", d.ms[c]);
            out += (c ? ", \"" : "\"") + d.names[c] + "\": " + num;
        }
        out += "}}";
    }
    out += "]";
    if (buf && cap) {
        const size_t k = std::min(cap - 1, out.size());
        std::memcpy(buf, out.data(), k);
        buf[k] = '\0';
    }
    return out.size() + 1;
}

libfive_mesh* libfive_tree_render_mesh(libfive_tree tree, libfive_region3 R,
                                       float res) {
    libfive_mesh_settings s;
    libfive_mesh_settings_default(&s);
    s.min_feature = 1 / res;     // float division, as libfive.cpp:343
    return render(tree, R, s, "libfive_tree_render_mesh");
}

libfive_mesh* libfive_tree_render_mesh_st(libfive_tree tree, libfive_region3 R,
                                          float res) {
    libfive_mesh_settings s;
    libfive_mesh_settings_default(&s);
    s.min_feature = 1 / res;
    return render(tree, R, s, "libfive_tree_render_mesh_st");
}

libfive_mesh_coords* libfive_tree_render_mesh_coords(libfive_tree tree,
                                                     libfive_region3 R,
                                                     float res) {
    libfive_mesh* m = libfive_tree_render_mesh(tree, R, res);
    if (!m) return nullptr;
    auto out = new libfive_mesh_coords;
    out->vert_count = m->vert_count;
    out->verts = m->verts;                    // new[]-allocated, moved over
    out->coord_index_count = 4 * m->tri_count;
    out->coord_indices = new int32_t[out->coord_index_count ? out->coord_index_count : 1];
    for (uint32_t i = 0; i < m->tri_count; ++i) {
        out->coord_indices[4 * i] = int32_t(m->tris[i].a);
        out->coord_indices[4 * i + 1] = int32_t(m->tris[i].b);
        out->coord_indices[4 * i + 2] = int32_t(m->tris[i].c);
        out->coord_indices[4 * i + 3] = -1;
    }
    m->verts = nullptr;
    libfive_mesh_delete(m);
    return out;
}

void libfive_mesh_coords_delete(libfive_mesh_coords* m) {
    if (!m) return;
    delete[] m->verts;
    delete[] m->coord_indices;
    delete m;
}

}   // extern "C"

// Binary STL of several meshes (Mesh::saveSTL, mesh.cpp:139-198).
bool libfive::amd::save_stl(const char* f, const std::vector<const libfive_mesh*>& meshes) {
    const std::string name(f ? f : "");
    if (name.size() < 4 || (name.compare(name.size() - 4, 4, ".stl") != 0 &&
                            name.compare(name.size() - 4, 4, ".STL") != 0)) {
        std::fprintf(stderr, "libfive_tree_save_mesh: filename \"%s\" does not end in .stl\n",
                     name.c_str());
    }
    FILE* fp = std::fopen(name.c_str(), "wb");
    if (!fp) {
        std::fprintf(stderr, "libfive_tree_save_mesh: could not open %s\n", name.c_str());
        return false;
    }
    char header[80];
    std::memset(header, ' ', sizeof(header));
    const char* text = "This is a binary STL exported from libfive.";
    std::memcpy(header, text, std::strlen(text));
    uint32_t count = 0;
    for (const auto* m : meshes) count += m->tri_count;
    bool ok = std::fwrite(header, 1, 80, fp) == 80 &&
              std::fwrite(&count, 4, 1, fp) == 1;
    std::vector<char> rec(50 * size_t(1024));
    for (const auto* m : meshes) {
        for (uint32_t i = 0; ok && i < m->tri_count;) {
            const uint32_t n = std::min<uint32_t>(1024, m->tri_count - i);
            for (uint32_t k = 0; k < n; ++k) {
                char* r = rec.data() + 50 * k;
                std::memset(r, 0, 50);             // zero normal and attribute
                const libfive_tri& t = m->tris[i + k];
                const uint32_t idx[3] = {t.a, t.b, t.c};
                for (int j = 0; j < 3; ++j) std::memcpy(r + 12 + 12 * j, &m->verts[idx[j]], 12);
            }
            ok = std::fwrite(rec.data(), 50, n, fp) == n;
            i += n;
        }
    }
    ok = (std::fclose(fp) == 0) && ok;
    return ok;
}

extern "C" {

bool libfive_tree_save_mesh(libfive_tree tree, libfive_region3 R, float res,
                            const char* f) {
    libfive_mesh* m = libfive_tree_render_mesh(tree, R, res);
    if (!m) return false;
    const bool ok = save_stl(f, {m});
    libfive_mesh_delete(m);
    return ok;
}

bool libfive_tree_save_meshes(libfive_tree trees[], libfive_region3 R,
                              float res, float quality, const char* f) {
    std::vector<libfive_mesh*> ms;
    bool ok = true;
    for (size_t i = 0; trees[i] != nullptr; ++i) {
        libfive_mesh_settings s;
        libfive_mesh_settings_default(&s);
        s.min_feature = 1 / res;
        s.max_err = std::pow(10.0, -double(quality));
        libfive_mesh* m = render(trees[i], R, s, "libfive_tree_save_meshes");
        if (!m) {
            ok = false;
            break;
        }
        ms.push_back(m);
    }
    if (ok) ok = save_stl(f, std::vector<const libfive_mesh*>(ms.begin(), ms.end()));
    for (auto* m : ms) libfive_mesh_delete(m);
    return ok;
}

static ContourResult slice(libfive_tree tree, libfive_region2 R, float z,
                           double min_feature, double max_err) {
    const double lo[2] = {R.X.lower, R.Y.lower}, hi[2] = {R.X.upper, R.Y.upper};
    return render_contours(Tree(D(tree)), lo, hi, z, min_feature, max_err);
}

libfive_contours* libfive_tree_render_slice(libfive_tree tree, libfive_region2 R,
                                            float z, float res) {
    try {
        const ContourResult r = slice(tree, R, z, 1 / res, 1e-8);
        auto out = new libfive_contours;
        out->count = uint32_t(r.contours.size());
        out->cs = new libfive_contour[out->count ? out->count : 1];
        for (uint32_t i = 0; i < out->count; ++i) {
            const auto& c = r.contours[i];
            out->cs[i].count = uint32_t(c.size());
            out->cs[i].pts = new libfive_vec2[c.size() ? c.size() : 1];
            for (size_t j = 0; j < c.size(); ++j) {
                out->cs[i].pts[j] = {r.verts[c[j]][0], r.verts[c[j]][1]};
            }
        }
        return out;
    } catch (const std::exception& e) {
        set_error("libfive_tree_render_slice", e);
        return nullptr;
    }
}

libfive_contours3* libfive_tree_render_slice3(libfive_tree tree, libfive_region2 R,
                                              float z, float res) {
    try {
        const ContourResult r = slice(tree, R, z, 1 / res, 1e-8);
        auto out = new libfive_contours3;
        out->count = uint32_t(r.contours.size());
        out->cs = new libfive_contour3[out->count ? out->count : 1];
        for (uint32_t i = 0; i < out->count; ++i) {
            const auto& c = r.contours[i];
            out->cs[i].count = uint32_t(c.size());
            out->cs[i].pts = new libfive_vec3[c.size() ? c.size() : 1];
            for (size_t j = 0; j < c.size(); ++j) {
                out->cs[i].pts[j] = {r.verts[c[j]][0], r.verts[c[j]][1], z};
            }
        }
        return out;
    } catch (const std::exception& e) {
        set_error("libfive_tree_render_slice3", e);
        return nullptr;
    }
}

void libfive_contours_delete(libfive_contours* cs) {
    if (!cs) return;
    for (uint32_t i = 0; i < cs->count; ++i) delete[] cs->cs[i].pts;
    delete[] cs->cs;
    delete cs;
}

void libfive_contours3_delete(libfive_contours3* cs) {
    if (!cs) return;
    for (uint32_t i = 0; i < cs->count; ++i) delete[] cs->cs[i].pts;
    delete[] cs->cs;
    delete cs;
}

libfive_slice_segments* libfive_tree_render_slice_segments(
        libfive_tree tree, libfive_region2 R, float z, double min_feature,
        double max_err) {
    try {
        const ContourResult r = slice(tree, R, z, min_feature, max_err);
        auto out = new libfive_slice_segments;
        out->vert_count = uint32_t(r.verts.size());
        out->seg_count = uint32_t(r.segs.size());
        out->verts = new float[2 * r.verts.size()];
        out->keys = new uint32_t[4 * r.verts.size()];
        out->segs = new uint32_t[2 * r.segs.size() + 1];
        for (size_t i = 0; i < r.verts.size(); ++i) {
            out->verts[2 * i] = r.verts[i][0];
            out->verts[2 * i + 1] = r.verts[i][1];
            for (int k = 0; k < 4; ++k) out->keys[4 * i + k] = r.keys[i][k];
        }
        for (size_t i = 0; i < r.segs.size(); ++i) {
            out->segs[2 * i] = r.segs[i][0];
            out->segs[2 * i + 1] = r.segs[i][1];
        }
        out->cell_count = uint32_t(r.cells.size() / 4);
        out->cells = new uint32_t[r.cells.size() + 1];
        std::copy(r.cells.begin(), r.cells.end(), out->cells);
        return out;
    } catch (const std::exception& e) {
        set_error("libfive_tree_render_slice_segments", e);
        return nullptr;
    }
}

void libfive_slice_segments_delete(libfive_slice_segments* s) {
    if (!s) return;
    delete[] s->verts;
    delete[] s->keys;
    delete[] s->segs;
    delete[] s->cells;
    delete s;
}

libfive_pixels* libfive_tree_render_pixels(libfive_tree tree, libfive_region2 R,
                                           float z, float res) {
    try {
        PixelImage img = render_pixels(Tree(D(tree)), R.X.lower, R.X.upper,
                                       R.Y.lower, R.Y.upper, z, res);
        auto out = new libfive_pixels;
        out->width = img.width;
        out->height = img.height;
        const size_t n = size_t(img.width) * img.height;
        out->pixels = new bool[n ? n : 1];
        for (size_t i = 0; i < n; ++i) out->pixels[i] = img.filled[i] != 0;
        return out;
    } catch (const std::exception& e) {
        set_error("libfive_tree_render_pixels", e);
        return nullptr;
    }
}

void libfive_pixels_delete(libfive_pixels* ps) {
    if (!ps) return;
    delete[] ps->pixels;
    delete ps;
}

void libfive_vars_delete(libfive_vars* j) {
    if (!j) return;
    delete[] j->vars;
    delete[] j->values;
    delete j;
}

libfive_evaluator libfive_tree_evaluator(libfive_tree tree, libfive_vars vars) {
    auto* e = new libfive_evaluator_{Tree(D(tree)), {}};
    for (uint32_t i = 0; i < vars.size; ++i) e->vars[vars.vars[i]] = vars.values[i];
    return e;
}

bool libfive_evaluator_update_vars(libfive_evaluator ev, libfive_vars vars) {
    // A variable counts only if the tree uses it (setVar's result).
    const Deck deck(ev->tree);
    bool changed = false;
    for (uint32_t i = 0; i < vars.size; ++i) {
        const void* id = vars.vars[i];
        const bool used = std::find(deck.var_nodes.begin(), deck.var_nodes.end(),
                                    static_cast<const TreeData*>(id)) != deck.var_nodes.end();
        auto it = ev->vars.find(id);
        const float old = (it == ev->vars.end()) ? 0.0f : it->second;
        if (used && !(old == vars.values[i])) changed = true;
        ev->vars[id] = vars.values[i];
    }
    return changed;
}

bool libfive_evaluator_save_mesh(libfive_evaluator ev, libfive_region3 R,
                                 const char* f) {
    try {
        MeshSettings ms;                  // BRepSettings defaults
        ms.vars = &ev->vars;
        const double lo[3] = {R.X.lower, R.Y.lower, R.Z.lower};
        const double hi[3] = {R.X.upper, R.Y.upper, R.Z.upper};
        libfive_mesh* m = render_to_c(ev->tree, lo, hi, ms, false, nullptr);
        if (!m) return false;
        const bool ok = save_stl(f, {m});
        libfive_mesh_delete(m);
        return ok;
    } catch (const std::exception& e) {
        set_error("libfive_evaluator_save_mesh", e);
        return false;
    }
}

void libfive_evaluator_delete(libfive_evaluator ptr) { delete ptr; }

libfive_mesh* libfive_tree_render_mesh_settings(libfive_tree tree,
                                                libfive_region3 R,
                                                const libfive_mesh_settings* s) {
    return render(tree, R, *s, "libfive_tree_render_mesh_settings");
}

}   // extern "C"
