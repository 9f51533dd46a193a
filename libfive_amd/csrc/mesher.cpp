// GPU dual-contouring mesher: host orchestration of the level-synchronous
// passes in mesh_kernels.hip.
//
//   build     root interval pass, then per level: link the ambiguous cells
//             (prefix sum), interval-evaluate and prune their children; the
//             finest level goes to the leaf kernel
//             (WorkerPool::build/run, worker_pool.inl:36-257)
//   collapse  levels 1 .. root, bottom-up (collectChildren,
//             dc_tree.inl:581-735)
//   emit      quads per level (count, prefix sums, write), vertex gather,
//             copy to the host (Dual<3>::walk + DCMesher,
//             dual.hpp:127-203, dc_mesher.cpp:19-169)
#include "mesher.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>

#include "deck.hpp"
#include "mesh_kernels.hpp"
#include "runtime.hpp"

namespace libfive {
namespace amd {

using namespace mesh;

static thread_local MeshStats g_stats;

const MeshStats& last_mesh_stats() { return g_stats; }

int resolution_level(const double lo[3], const double hi[3], double min_feature) {
    double md = hi[0] - lo[0];
    md = std::min(md, hi[1] - lo[1]);
    md = std::min(md, hi[2] - lo[2]);
    const float r = logf(float(md / min_feature));
    const float m = fmaxf(0.0f, r);
    return int(std::ceil(m / logf(2.0f)));
}

namespace {

struct LevelStore {
    PoolBuf<uint32_t> key;
    PoolBuf<uint8_t> type;
    PoolBuf<int32_t> child;
    PoolBuf<uint32_t> tape;
    PoolBuf<LeafRec> leaf;
    uint32_t n = 0;

    void alloc(uint32_t count) {
        n = count;
        key.alloc(3 * size_t(count));
        type.alloc(count);
        child.alloc(count);
        tape.alloc(count);
        leaf.alloc(count);
    }
    LevelView view(uint32_t base) {
        LevelView v;
        v.key = key.p;
        v.type = type.p;
        v.child = child.p;
        v.tape = tape.p;
        v.leaf = leaf.p;
        v.n = n;
        v.cell_base = base;
        return v;
    }
};

void wait(hipStream_t s, const char* what);

template <typename T>
void grow(PoolBuf<T>& b, size_t used, size_t need, hipStream_t s) {
    if (need <= b.n) return;
    PoolBuf<T> nb(std::max(need, 2 * b.n));
    if (used) {
        hip_check(hipMemcpyAsync(nb.p, b.p, used * sizeof(T),
                                 hipMemcpyDeviceToDevice, s), "grow copy");
    }
    wait(s, "buffer growth");
    b.swap(nb);
}

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(
        std::chrono::steady_clock::now() - t0).count();
}

// Waits for the stream with a time limit, so that a GPU step that does not
// finish is reported by name instead of blocking the caller forever.
void wait(hipStream_t s, const char* what) {
    static const bool trace = std::getenv("LIBFIVE_MESH_TRACE") != nullptr;
    if (trace) {
        std::fprintf(stderr, "render_mesh: waiting for %s\n", what);
        std::fflush(stderr);
    }
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        const hipError_t e = hipStreamQuery(s);
        if (e == hipSuccess) return;
        if (e != hipErrorNotReady) hip_check(e, what);
        if (ms_since(t0) > 60000.0) {
            throw Error(std::string("render_mesh: GPU step '") + what +
                        "' did not finish within 60 s");
        }
        std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
}

void check_errors(uint32_t err) {
    if (err & ERR_FEATURE) {
        throw Error("render_mesh: a point on a min/max feature has more "
                    "features than the GPU feature evaluation holds");
    }
    if (err & (ERR_POOL | ERR_TAPES)) {
        throw Error("render_mesh: tape storage exhausted");
    }
    if (err) throw Error("render_mesh: internal error " + std::to_string(err));
}

// Parents per launch of the leaf corner kernel (four times as many edge
// tasks per launch of the edge kernel): about 10 ms of GPU time each on the
// gyroid at 512^3.
constexpr uint32_t kLeafChunk = 1u << 18;

// Evaluates the leaf cells of `groups` (DCTree::evalLeaf): one pass for
// all cells, then the cells that met an exact zero on a tied MIN/MAX again
// with the feature evaluation, in batches that bound the workspace.
void run_leaves(const Params& P, const uint32_t* groups, uint32_t ngroups,
                uint32_t per_group, const LevelView& l1, const LevelView& l0,
                uint32_t max_len, bool fast, const MeshSettings& settings,
                PoolBuf<Counters>& counters, Counters& hc, hipStream_t s) {
    const uint32_t cells = ngroups * per_group;
    PoolBuf<uint32_t> deferred(cells);
    hip_check(hipMemsetAsync(&counters.p->deferred, 0, sizeof(uint32_t), s), "deferred");
    if (per_group == 8) {
        // corners per parent, then the parents' lattice edges, then one
        // vertex solve per ambiguous leaf
        PoolBuf<LeafGroupRec> grec(ngroups);
        // per-group counts and their scans (edges, tasks, leaves)
        PoolBuf<uint32_t> cnt(3 * size_t(ngroups)), base(3 * size_t(ngroups)),
            scan_tmp(scan_scratch_words(ngroups));
        PoolBuf<uint64_t> tasks(size_t(ngroups) * kLeafTasksPerGroup);
        PoolBuf<uint32_t> leaves(cells);
        // With a cancel flag the kernels are launched in chunks of at most
        // kLeafChunk parents (tasks below), and the host waits after each
        // chunk and checks the flag, so that a request is seen within one
        // chunk's run time.
        const bool poll = settings.cancel || settings.cancel_flag;
        const uint32_t chunk = poll ? kLeafChunk : ngroups;
        for (uint32_t g0 = 0; g0 < ngroups; g0 += chunk) {
            const uint32_t ng = std::min(chunk, ngroups - g0);
            hip_check(launch_leaf_corners(P, fast, groups, g0, ng, ngroups, l1, l0, grec.p,
                                          cnt.p, deferred.p, s),
                      "leaf corner kernel");
            if (poll) {
                wait(s, "leaf corner kernel");
                if (settings.cancelled()) return;     // the caller checks too
            }
        }
        for (int k = 0; k < 3; ++k) {
            hip_check(exclusive_scan(cnt.p + size_t(k) * ngroups, base.p + size_t(k) * ngroups,
                                     ngroups, scan_tmp.p, s), "leaf scan");
        }
        hip_check(launch_leaf_scatter(P, ngroups, grec.p, cnt.p, base.p, tasks.p, leaves.p, s),
                  "leaf scatter kernel");
        hip_check(hipMemcpyAsync(&hc, counters.p, sizeof(hc), hipMemcpyDeviceToHost, s),
                  "counters");
        wait(s, "leaf corner kernel");
        check_errors(hc.error);
        if (settings.cancelled()) return;
        if (hc.tasks > size_t(ngroups) * kLeafTasksPerGroup || hc.leaves > cells ||
            hc.edges > 54ull * ngroups || hc.deferred > cells) {
            throw Error("render_mesh: inconsistent leaf counts");
        }
        PoolBuf<EdgeRec> edges(std::max<uint32_t>(hc.edges, 1u));
        const uint32_t tchunk = poll ? 4 * kLeafChunk : std::max<uint32_t>(hc.tasks, 1u);
        for (uint32_t t0 = 0; t0 < hc.tasks; t0 += tchunk) {
            const uint32_t nt = std::min(tchunk, hc.tasks - t0);
            hip_check(launch_leaf_edges(P, fast, groups, l1, grec.p, tasks.p + t0, nt, edges.p,
                                        deferred.p, s), "leaf edge kernel");
            if (poll) {
                wait(s, "leaf edge kernel");
                if (settings.cancelled()) return;
            }
        }
        hip_check(launch_leaf_solve(P, groups, l1, l0, grec.p, leaves.p, hc.leaves,
                                    edges.p, s), "leaf solve kernel");
        hip_check(hipMemcpyAsync(&hc, counters.p, sizeof(hc), hipMemcpyDeviceToHost, s),
                  "counters");
        wait(s, "leaf edge and solve kernels");
    } else {
        hip_check(launch_leaf(P, groups, ngroups, per_group, l1, l0, deferred.p, s),
                  "leaf kernel");
        hip_check(hipMemcpyAsync(&hc, counters.p, sizeof(hc), hipMemcpyDeviceToHost, s),
                  "counters");
        wait(s, "leaf kernel");
    }
    check_errors(hc.error);
    const uint32_t nd = hc.deferred;
    if (nd == 0) return;
    if (nd > cells) throw Error("render_mesh: inconsistent deferred cell count");
    const size_t ws_bytes = leaf_feature_workspace(max_len, P.nregs);
    const size_t batch = std::max<size_t>(1, std::min<size_t>(nd, (size_t(1) << 30) / ws_bytes));
    PoolBuf<uint8_t> ws(batch * ws_bytes);
    for (size_t first = 0; first < nd; first += batch) {
        const uint32_t count = uint32_t(std::min<size_t>(batch, nd - first));
        hip_check(launch_leaf_features(P, groups, ngroups, per_group, l1, l0,
                                       deferred.p + first, count, ws.p, ws_bytes, s),
                  "leaf feature kernel");
        hip_check(hipMemcpyAsync(&hc, counters.p, sizeof(hc), hipMemcpyDeviceToHost, s),
                  "counters");
        wait(s, "leaf feature kernel");
        check_errors(hc.error);
    }
}

}   // namespace

bool render_mesh(const Tree& tree, const double lo[3], const double hi[3],
                 const MeshSettings& settings, HostMesh& out) {
    Context& ctx = context();
    std::lock_guard<std::mutex> lock(ctx.mu);
    return render_mesh_on(tree, lo, hi, settings, out, ctx.stream);
}

// The lookup tables live in __constant__ memory: uploaded once per process
// and completed before any render uses them.
static void ensure_tables() {
    static std::once_flag once;
    std::call_once(once, [] {
        hipStream_t s = context().stream;
        hip_check(upload_tables(s), "mesh tables");
        hip_check(hipStreamSynchronize(s), "mesh tables");
    });
}

static bool render_body(const Tree& tree, const double lo[3], const double hi[3],
                        const MeshSettings& settings, HostMesh& out, hipStream_t s);

bool render_mesh_on(const Tree& tree, const double lo[3], const double hi[3],
                    const MeshSettings& settings, HostMesh& out, hipStream_t s) {
    pool_defer_begin();
    try {
        const bool r = render_body(tree, lo, hi, settings, out, s);
        pool_defer_end(s);
        return r;
    } catch (...) {
        pool_defer_end(s);
        throw;
    }
}

static bool render_body(const Tree& tree, const double lo[3], const double hi[3],
                        const MeshSettings& settings, HostMesh& out, hipStream_t s) {
    const auto t_start = std::chrono::steady_clock::now();
    g_stats = MeshStats();
    ensure_tables();

    const Deck deck(tree);
    std::vector<float> var_values(deck.vars.size(), 0.0f);
    if (settings.vars) {
        for (size_t i = 0; i < deck.var_nodes.size(); ++i) {
            auto it = settings.vars->find(deck.var_nodes[i]);
            if (it != settings.vars->end()) var_values[i] = it->second;
        }
    }
    const DeviceTape base = compile_deck(deck, &var_values);
    const uint32_t nregs = base.num_regs;
    if (nregs > 256) {
        throw Error("render_mesh: the tape needs " + std::to_string(nregs) +
                    " registers; the mesher supports 256");
    }
    // The call-free leaf kernels serve tapes of the common opcodes only
    // (tape::fast_opcode; pe_value_op in point_eval.hpp handles the same set).
    bool fast = true;
    for (const tape::Clause& c : base.clauses) fast = fast && tape::fast_opcode(tape::op(c));
    const int top = resolution_level(lo, hi, settings.min_feature);
    if (top > 20) throw Error("render_mesh: resolution too fine (more than 20 levels)");
    g_stats.top = top;

    // Clause pool and tape records; the base tape is record 0.
    const uint32_t nbase = uint32_t(base.clauses.size());
    const uint32_t base_padded = tape::padded_length(nbase);
    PoolBuf<tape::Clause> pool(std::max<size_t>(size_t(base_padded) * 64, 1u << 20));
    {
        std::vector<tape::Clause> h(base_padded, tape::Clause{0, 0, 0, 0});
        std::copy(base.clauses.begin(), base.clauses.end(), h.begin());
        hip_check(hipMemcpyAsync(pool.p, h.data(), h.size() * sizeof(tape::Clause),
                                 hipMemcpyHostToDevice, s), "upload tape");
    }
    PoolBuf<TapeRec> tapes(1u << 16);
    {
        TapeRec r{};
        r.off = 0;
        r.len = nbase;
        r.parent = kNoTape;
        r.flags = 0;
        r.ref_len = base.ref_length;
        hip_check(hipMemcpyAsync(tapes.p, &r, sizeof(r), hipMemcpyHostToDevice, s),
                  "upload tape record");
    }
    PoolBuf<Counters> counters(1);
    Counters hc{};
    hc.pool_top = base_padded;
    hc.tape_count = 1;
    hip_check(hipMemcpyAsync(counters.p, &hc, sizeof(hc), hipMemcpyHostToDevice, s),
              "counters");

    Params P{};
    for (int a = 0; a < 3; ++a) {
        P.lo[a] = lo[a];
        P.hi[a] = hi[a];
    }
    P.top = top;
    P.max_err = settings.max_err;
    P.nregs = std::max<uint32_t>(nregs, tape::kFirstFreeReg);
    P.root_reg = base.root_reg;
    P.slot_x = deck.X;
    P.slot_y = deck.Y;
    P.slot_z = deck.Z;
    P.qef_f32c = settings.precision == 1 ? 1u : 0u;
    P.counters = counters.p;
    auto refresh = [&]() {
        P.pool = pool.p;
        P.pool_cap = uint32_t(std::min<size_t>(pool.n, 0xffffffffu));
        P.tapes = tapes.p;
        P.tape_cap = uint32_t(tapes.n);
    };
    refresh();

    std::vector<LevelStore> levels(top + 1);
    std::vector<PoolBuf<uint32_t>> groups(top + 2);
    // Per-level scan buffers stay allocated until the render ends, so that
    // no device memory is released while kernels are in flight.
    std::vector<PoolBuf<uint32_t>> flags(top + 1), basei(top + 1), scan_tmp(top + 1);
    std::vector<PoolBuf<uint8_t>> iws(top + 1);      // interval work arrays of long tapes

    // ---- build ------------------------------------------------------------
    levels[top].alloc(1);
    if (top == 0) {
        // The root is a leaf: evaluate it directly with the base tape.
        LevelStore virt;
        virt.alloc(1);
        const uint32_t zero_key[3] = {0, 0, 0};
        const int32_t child0 = 0;
        const uint32_t tape0 = 0;
        hip_check(hipMemcpyAsync(virt.key.p, zero_key, sizeof(zero_key), hipMemcpyHostToDevice, s), "root");
        hip_check(hipMemcpyAsync(virt.child.p, &child0, sizeof(child0), hipMemcpyHostToDevice, s), "root");
        hip_check(hipMemcpyAsync(virt.tape.p, &tape0, sizeof(tape0), hipMemcpyHostToDevice, s), "root");
        groups[1].alloc(1);
        hip_check(hipMemsetAsync(groups[1].p, 0, sizeof(uint32_t), s), "root");
        run_leaves(P, groups[1].p, 1, 1, virt.view(0), levels[0].view(0), nbase,
                   fast, settings, counters, hc, s);
    } else {
        hip_check(hipMemsetAsync(&counters.p->max_len, 0, sizeof(uint32_t), s), "max_len");
        iws[top].alloc(interval_workspace_bytes(P.nregs, nbase));
        hip_check(launch_interval(P, top, nullptr, 1, levels[top].view(0),
                                  levels[top].view(0), nbase, iws[top].p, iws[top].n, s),
                  "interval kernel");
        for (int L = top; L >= 1; --L) {
            LevelStore& cur = levels[L];
            flags[L].alloc(cur.n);
            basei[L].alloc(cur.n);
            scan_tmp[L].alloc(scan_scratch_words(cur.n));
            groups[L].alloc(cur.n);
            hip_check(launch_flags(cur.view(0), flags[L].p, s), "flags");
            hip_check(exclusive_scan(flags[L].p, basei[L].p, cur.n, scan_tmp[L].p, s), "scan");
            hip_check(launch_link(cur.view(0), basei[L].p, groups[L].p, s), "link");
            uint32_t last[2] = {0, 0};
            hip_check(hipMemcpyAsync(&last[0], basei[L].p + cur.n - 1, 4, hipMemcpyDeviceToHost, s), "count");
            hip_check(hipMemcpyAsync(&last[1], flags[L].p + cur.n - 1, 4, hipMemcpyDeviceToHost, s), "count");
            hip_check(hipMemcpyAsync(&hc, counters.p, sizeof(hc), hipMemcpyDeviceToHost, s), "counters");
            wait(s, L == top ? "root interval pass" : "interval pass");
            check_errors(hc.error);
            const uint64_t namb = uint64_t(last[0]) + last[1];
            if (namb > cur.n) throw Error("render_mesh: inconsistent cell count");
            if (namb == 0) break;
            if (namb * 8u > 0xffffffffull / 4) throw Error("render_mesh: too many cells");
            levels[L - 1].alloc(uint32_t(namb * 8u));
            if (L - 1 > 0) {
                const uint32_t max_len = hc.max_len;
                if (max_len == 0 || max_len > hc.pool_top) {
                    throw Error("render_mesh: inconsistent tape length");
                }
                grow(pool, hc.pool_top,
                     size_t(hc.pool_top) + size_t(namb) * 8u * (max_len + kPushExtra), s);
                grow(tapes, hc.tape_count, size_t(hc.tape_count) + size_t(namb) * 8u, s);
                refresh();
                hip_check(hipMemsetAsync(&counters.p->max_len, 0, sizeof(uint32_t), s), "max_len");
                iws[L - 1].alloc(interval_workspace_bytes(P.nregs, max_len));
                hip_check(launch_interval(P, L - 1, groups[L].p, uint32_t(namb), cur.view(0),
                                          levels[L - 1].view(0), max_len, iws[L - 1].p,
                                          iws[L - 1].n, s),
                          "interval kernel");
            } else {
                run_leaves(P, groups[1].p, uint32_t(namb), 8, cur.view(0),
                           levels[0].view(0), std::max<uint32_t>(hc.max_len, 1),
                           fast, settings, counters, hc, s);
            }
            if (settings.cancelled()) {
                wait(s, "cancel");
                return false;
            }
        }
    }

    if (settings.cancelled()) {
        wait(s, "cancel");
        return false;
    }

    // Global cell numbering (level 0 first).
    std::vector<uint32_t> cell_base(top + 1, 0);
    uint64_t total_cells = 0;
    for (int L = 0; L <= top; ++L) {
        cell_base[L] = uint32_t(total_cells);
        total_cells += levels[L].n;
    }
    if (total_cells * 4 > 0xffffffffull) throw Error("render_mesh: too many cells");

    if (settings.collect_stats) {
        // Per-level state counts and pruned tape lengths, before collapse.
        std::vector<TapeRec> trs(hc.tape_count);
        hip_check(hipMemcpyAsync(trs.data(), tapes.p, trs.size() * sizeof(TapeRec),
                                 hipMemcpyDeviceToHost, s), "stats");
        for (int L = top; L >= 0; --L) {
            std::vector<uint8_t> ty(levels[L].n);
            std::vector<uint32_t> tp(levels[L].n), ky(3 * size_t(levels[L].n));
            if (levels[L].n) {
                hip_check(hipMemcpyAsync(ty.data(), levels[L].type.p, ty.size(),
                                         hipMemcpyDeviceToHost, s), "stats");
                hip_check(hipMemcpyAsync(tp.data(), levels[L].tape.p, tp.size() * 4,
                                         hipMemcpyDeviceToHost, s), "stats");
                hip_check(hipMemcpyAsync(ky.data(), levels[L].key.p, ky.size() * 4,
                                         hipMemcpyDeviceToHost, s), "stats");
            }
            wait(s, "statistics copy");
            uint64_t cnt[3] = {0, 0, 0};
            for (size_t i = 0; i < ty.size(); ++i) {
                if (ty[i] > AMBIGUOUS || tp[i] >= trs.size()) {
                    throw Error("render_mesh: inconsistent cell record");
                }
                cnt[ty[i]]++;
                g_stats.cells.insert(g_stats.cells.end(),
                                     {ky[3 * i], ky[3 * i + 1], ky[3 * i + 2], uint32_t(L),
                                      uint32_t(ty[i])});
                if (L > 0 && ty[i] == AMBIGUOUS) {
                    g_stats.tape_lengths.push_back(trs[tp[i]].ref_len);
                }
            }
            for (int k = 0; k < 3; ++k) g_stats.level_counts.push_back(cnt[k]);
            if (L == 0) g_stats.leaves = cnt[AMBIGUOUS];
        }
    }
    g_stats.ms_build = ms_since(t_start);

    // ---- collapse -----------------------------------------------------------
    const auto t_collapse = std::chrono::steady_clock::now();
    std::vector<LevelView> views(top + 1);
    for (int L = 0; L <= top; ++L) views[L] = levels[L].view(cell_base[L]);
    for (int L = 1; L <= top; ++L) {
        if (levels[L].n && levels[L - 1].n) {
            hip_check(launch_collapse(P, L, views[L], views[L - 1], s), "collapse kernel");
        }
        // Levels are collapsed in order on one stream; the flag is checked
        // once the work queued so far has finished.
        if (settings.cancelled()) {
            wait(s, "cancel");
            return false;
        }
    }
    wait(s, "collapse");
    g_stats.ms_collapse = ms_since(t_collapse);

    // ---- emit -----------------------------------------------------------------
    const auto t_emit = std::chrono::steady_clock::now();
    PoolBuf<LevelView> d_views(top + 1);
    hip_check(hipMemcpyAsync(d_views.p, views.data(), views.size() * sizeof(LevelView),
                             hipMemcpyHostToDevice, s), "views");
    P.levels = d_views.p;
    PoolBuf<uint32_t> tri_count(total_cells), tri_off(total_cells);
    PoolBuf<uint32_t> vslot(4 * total_cells), vid(4 * total_cells);
    PoolBuf<uint32_t> scan_a(scan_scratch_words(total_cells));
    PoolBuf<uint32_t> scan_b(scan_scratch_words(4 * total_cells));
    hip_check(hipMemsetAsync(vslot.p, 0, vslot.n * sizeof(uint32_t), s), "vslot");
    for (int L = 0; L <= top; ++L) {
        hip_check(launch_quads(P, views[L], 0, tri_count.p, vslot.p, nullptr, nullptr,
                               nullptr, s), "quad count");
    }
    hip_check(exclusive_scan(tri_count.p, tri_off.p, total_cells, scan_a.p, s), "scan");
    hip_check(exclusive_scan(vslot.p, vid.p, vslot.n, scan_b.p, s), "scan");
    uint32_t tail[4] = {0, 0, 0, 0};
    hip_check(hipMemcpyAsync(&tail[0], tri_off.p + total_cells - 1, 4, hipMemcpyDeviceToHost, s), "tail");
    hip_check(hipMemcpyAsync(&tail[1], tri_count.p + total_cells - 1, 4, hipMemcpyDeviceToHost, s), "tail");
    hip_check(hipMemcpyAsync(&tail[2], vid.p + vslot.n - 1, 4, hipMemcpyDeviceToHost, s), "tail");
    hip_check(hipMemcpyAsync(&tail[3], vslot.p + vslot.n - 1, 4, hipMemcpyDeviceToHost, s), "tail");
    hip_check(hipMemcpyAsync(&hc, counters.p, sizeof(hc), hipMemcpyDeviceToHost, s), "counters");
    wait(s, "quad count");
    check_errors(hc.error);
    if (settings.cancelled()) return false;
    const uint64_t ntris64 = uint64_t(tail[0]) + tail[1];
    const uint64_t nverts64 = uint64_t(tail[2]) + tail[3] + 1;    // with the dummy vertex
    if (ntris64 > 24 * total_cells || nverts64 > 4 * total_cells + 1) {
        throw Error("render_mesh: inconsistent mesh size");
    }
    const uint32_t ntris = uint32_t(ntris64), nverts = uint32_t(nverts64);
    PoolBuf<uint32_t> d_tris(3 * size_t(std::max<uint32_t>(ntris, 1)));
    PoolBuf<float> d_verts(3 * size_t(nverts));
    PoolBuf<uint32_t> d_keys(out.want_keys ? 5 * size_t(nverts) : 0);
    hip_check(hipMemsetAsync(d_verts.p, 0, 3 * sizeof(float), s), "dummy");
    if (out.want_keys) {
        hip_check(hipMemsetAsync(d_keys.p, 0, 5 * sizeof(uint32_t), s), "dummy");
    }
    for (int L = 0; L <= top; ++L) {
        hip_check(launch_quads(P, views[L], 1, tri_count.p, vslot.p, tri_off.p, vid.p,
                               d_tris.p, s), "quad write");
        hip_check(launch_verts(views[L], L, vslot.p, vid.p, d_verts.p, d_keys.p, s),
                  "vertex gather");
    }
    if (settings.cancelled()) {
        wait(s, "cancel");
        return false;
    }
    float* vdst = nullptr;
    uint32_t* tdst = nullptr;
    if (out.alloc) {
        out.alloc(nverts, ntris, &vdst, &tdst);
    } else {
        out.verts.resize(3 * size_t(nverts));
        out.tris.resize(3 * size_t(ntris));
        vdst = out.verts.data();
        tdst = out.tris.data();
    }
    // Device -> pinned staging at full link speed, then host threads copy
    // into the caller's (pageable) arrays.
    const size_t vbytes = 3 * size_t(nverts) * sizeof(float);
    const size_t tbytes = 3 * size_t(ntris) * sizeof(uint32_t);
    size_t granted = 0;
    char* stage = static_cast<char*>(pinned_alloc(vbytes + tbytes, &granted));
    struct StageRelease {
        char* p;
        size_t n;
        ~StageRelease() { pinned_free(p, n); }
    } stage_release{stage, granted};
    hip_check(hipMemcpyAsync(stage, d_verts.p, vbytes, hipMemcpyDeviceToHost, s),
              "copy verts");
    if (ntris) {
        hip_check(hipMemcpyAsync(stage + vbytes, d_tris.p, tbytes,
                                 hipMemcpyDeviceToHost, s), "copy tris");
    }
    if (out.want_keys) {
        out.keys.resize(5 * size_t(nverts));
        hip_check(hipMemcpyAsync(out.keys.data(), d_keys.p, out.keys.size() * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, s), "copy keys");
    }
    wait(s, "mesh copy");
    parallel_memcpy(vdst, stage, vbytes);
    if (ntris) parallel_memcpy(tdst, stage + vbytes, tbytes);
    g_stats.ms_emit = ms_since(t_emit);
    g_stats.ms_total = ms_since(t_start);
    return true;
}

}   // namespace amd
}   // namespace libfive
