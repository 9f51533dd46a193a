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
#include <atomic>
#include <exception>
#include <mutex>
#include <thread>

#include "autotune.hpp"
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
    LevelView view(const uint32_t* cbase = nullptr) {
        LevelView v;
        v.key = key.p;
        v.type = type.p;
        v.child = child.p;
        v.tape = tape.p;
        v.leaf = leaf.p;
        v.n = n;
        v.pad = 0;
        v.cbase = cbase;
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
    // LIBFIVE_MESH_TRACE=wait names every wait; any other value prints the
    // phase times of each render.
    static const bool trace = [] {
        const char* e = std::getenv("LIBFIVE_MESH_TRACE");
        return e != nullptr && std::string(e) == "wait";
    }();
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

// Clears the per-level maxima and sums the interval pass accumulates
// (Counters::max_len, Counters::push_need).
void reset_level_counters(Counters* c, hipStream_t s) {
    hip_check(hipMemsetAsync(&c->max_len, 0, sizeof(uint32_t), s), "max_len");
    hip_check(hipMemsetAsync(&c->push_need, 0, sizeof(c->push_need), s), "push_need");
}

// The pushes of a level take at most need clauses (Counters::push_need);
// need is also bounded by namb x 8 x (max_len + kPushExtra).
size_t push_budget(const Counters& hc, uint64_t namb) {
    const unsigned long long bound =
        (unsigned long long)namb * 8ull * (hc.max_len + kPushExtra);
    if (hc.push_need > bound) throw Error("render_mesh: inconsistent push budget");
    return size_t(hc.push_need);
}

void check_errors(uint32_t err) {
    if (err & ERR_FEATURE) {
        throw Error("render_mesh: a point on a min/max feature has more "
                    "features than the GPU feature evaluation holds");
    }
    if (err & (ERR_POOL | ERR_TAPES)) {
        throw Error("render_mesh: tape storage exhausted");
    }
    if (err & ERR_WIDE) {
        throw Error("render_mesh: the interval pass of a long tape met an invalid tape "
                    "or made no progress");
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
        // The corner kernel keeps a group's deferred bit when it runs again
        // (the autotuner reruns a slice), so the bits start cleared.
        hip_check(hipMemsetAsync(grec.p, 0, size_t(ngroups) * sizeof(LeafGroupRec), s), "grec");
        const int nv = register_banks(P.nregs);
        const std::string tag = "/banks=" + std::to_string(nv) + "/fast=" + (fast ? "1" : "0");
        const std::vector<std::string> wave_names = {"1 wave per workgroup",
                                                     "4 waves per workgroup"};
        auto waves_of = [](int c) { return c == 1 ? 4 : 1; };
        const int cw = nv == 8 ? 0 : tune::pick(
            "leaf_corners" + tag, wave_names, [&](int c) {
                return launch_leaf_corners(P, fast, groups, 0, std::min(ngroups, 16384u), ngroups,
                                           l1, l0, grec.p, cnt.p, deferred.p, s, waves_of(c));
            }, s);
        for (uint32_t g0 = 0; g0 < ngroups; g0 += chunk) {
            const uint32_t ng = std::min(chunk, ngroups - g0);
            hip_check(launch_leaf_corners(P, fast, groups, g0, ng, ngroups, l1, l0, grec.p,
                                          cnt.p, deferred.p, s, waves_of(cw)),
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
        const int ew = (nv == 8 || hc.tasks == 0) ? 0 : tune::pick(
            "leaf_edges" + tag, wave_names, [&](int c) {
                return launch_leaf_edges(P, fast, groups, l1, grec.p, tasks.p,
                                         std::min(hc.tasks, 65536u), edges.p, deferred.p, s,
                                         waves_of(c));
            }, s);
        for (uint32_t t0 = 0; t0 < hc.tasks; t0 += tchunk) {
            const uint32_t nt = std::min(tchunk, hc.tasks - t0);
            hip_check(launch_leaf_edges(P, fast, groups, l1, grec.p, tasks.p + t0, nt, edges.p,
                                        deferred.p, s, waves_of(ew)), "leaf edge kernel");
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
static void ensure_tables_once() {
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
    ensure_tables_once();

    const Deck deck(tree);
    std::vector<float> var_values(deck.vars.size(), 0.0f);
    if (settings.vars) {
        for (size_t i = 0; i < deck.var_nodes.size(); ++i) {
            auto it = settings.vars->find(deck.var_nodes[i]);
            if (it != settings.vars->end()) var_values[i] = it->second;
        }
    }
    const DeviceTape base = compile_deck(deck, &var_values);
    const double ms_front = ms_since(t_start);
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
        run_leaves(P, groups[1].p, 1, 1, virt.view(), levels[0].view(), nbase,
                   fast, settings, counters, hc, s);
    } else {
        reset_level_counters(counters.p, s);
        iws[top].alloc(interval_workspace_bytes(P.nregs, nbase, 1));
        hip_check(launch_interval(P, top, nullptr, 1, levels[top].view(),
                                  levels[top].view(), nbase, iws[top].p, iws[top].n, s),
                  "interval kernel");
        for (int L = top; L >= 1; --L) {
            LevelStore& cur = levels[L];
            flags[L].alloc(cur.n);
            basei[L].alloc(cur.n);
            scan_tmp[L].alloc(scan_scratch_words(cur.n));
            groups[L].alloc(cur.n);
            hip_check(launch_flags(cur.view(), flags[L].p, s), "flags");
            hip_check(exclusive_scan(flags[L].p, basei[L].p, cur.n, scan_tmp[L].p, s), "scan");
            hip_check(launch_link(cur.view(), basei[L].p, groups[L].p, s), "link");
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
                grow(pool, hc.pool_top, size_t(hc.pool_top) + push_budget(hc, namb), s);
                grow(tapes, hc.tape_count, size_t(hc.tape_count) + size_t(namb) * 8u, s);
                refresh();
                reset_level_counters(counters.p, s);
                iws[L - 1].alloc(interval_workspace_bytes(P.nregs, max_len, uint32_t(namb)));
                hip_check(launch_interval(P, L - 1, groups[L].p, uint32_t(namb), cur.view(),
                                          levels[L - 1].view(), max_len, iws[L - 1].p,
                                          iws[L - 1].n, s),
                          "interval kernel");
            } else {
                run_leaves(P, groups[1].p, uint32_t(namb), 8, cur.view(),
                           levels[0].view(), std::max<uint32_t>(hc.max_len, 1),
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
    PoolBuf<uint32_t> d_cbase(top + 1);
    hip_check(hipMemcpyAsync(d_cbase.p, cell_base.data(), cell_base.size() * sizeof(uint32_t),
                             hipMemcpyHostToDevice, s), "cell numbering");

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
    for (int L = 0; L <= top; ++L) views[L] = levels[L].view(d_cbase.p + L);
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
    const double ms_counted = ms_since(t_emit);
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
        hip_check(launch_verts(P, views[L], L, vslot.p, vid.p, d_verts.p, d_keys.p, s),
                  "vertex gather");
    }
    if (settings.cancelled()) {
        wait(s, "cancel");
        return false;
    }
    float* vdst = nullptr;
    uint32_t* tdst = nullptr;
    bool dst_pinned = false;
    if (out.alloc) {
        dst_pinned = out.alloc(nverts, ntris, &vdst, &tdst);
    } else {
        out.verts.resize(3 * size_t(nverts));
        out.tris.resize(3 * size_t(ntris));
        vdst = out.verts.data();
        tdst = out.tris.data();
    }
    const double ms_alloced = ms_since(t_emit);
    // Device -> page-locked memory at full link speed: the caller's arrays
    // when they are page-locked, else a staging block that host threads
    // then copy into the caller's (pageable) arrays.
    const size_t vbytes = 3 * size_t(nverts) * sizeof(float);
    const size_t tbytes = 3 * size_t(ntris) * sizeof(uint32_t);
    size_t granted = 0;
    char* stage = dst_pinned ? nullptr
                             : static_cast<char*>(pinned_alloc(vbytes + tbytes, &granted));
    struct StageRelease {
        char* p;
        size_t n;
        ~StageRelease() {
            if (p) pinned_free(p, n);
        }
    } stage_release{stage, granted};
    char* vto = dst_pinned ? reinterpret_cast<char*>(vdst) : stage;
    char* tto = dst_pinned ? reinterpret_cast<char*>(tdst) : stage + vbytes;
    hip_check(hipMemcpyAsync(vto, d_verts.p, vbytes, hipMemcpyDeviceToHost, s), "copy verts");
    if (ntris) {
        hip_check(hipMemcpyAsync(tto, d_tris.p, tbytes, hipMemcpyDeviceToHost, s), "copy tris");
    }
    if (out.want_keys) {
        out.keys.resize(5 * size_t(nverts));
        hip_check(hipMemcpyAsync(out.keys.data(), d_keys.p, out.keys.size() * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, s), "copy keys");
    }
    wait(s, "mesh copy");
    const double ms_copied = ms_since(t_emit);
    if (!dst_pinned) {
        parallel_memcpy(vdst, stage, vbytes);
        if (ntris) parallel_memcpy(tdst, stage + vbytes, tbytes);
    }
    g_stats.ms_emit = ms_since(t_emit);
    g_stats.ms_total = ms_since(t_start);
    if (std::getenv("LIBFIVE_MESH_TRACE") != nullptr) {
        std::fprintf(stderr, "render: front end %.1f ms, build %.1f ms, collapse %.1f ms, "
                     "emit %.1f ms (counted %.1f, arrays %.1f, copied %.1f; %u verts, %u tris), "
                     "total %.1f ms\n", ms_front, g_stats.ms_build - ms_front,
                     g_stats.ms_collapse, g_stats.ms_emit, ms_counted, ms_alloced, ms_copied,
                     nverts, ntris, g_stats.ms_total);
    }
    return true;
}

// ---- many models through one set of passes -----------------------------------

namespace {

struct MergedModel {
    DeviceTape tape;
    uint32_t slot_x = 0, slot_y = 0, slot_z = 0;
    int top = 0;
    size_t index = 0;            // position in the caller's list
};

// Runs f(i) for i in [0, n) on up to 16 host threads.
template <typename F>
void parallel_for(size_t n, F f) {
    const size_t k = std::min<size_t>(n, std::max<unsigned>(1u, std::min(16u,
                                      std::thread::hardware_concurrency())));
    if (k <= 1) {
        for (size_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<size_t> next{0};
    std::exception_ptr err;
    std::mutex mu;
    std::vector<std::thread> th;
    for (size_t w = 0; w < k; ++w) {
        th.emplace_back([&]() {
            for (;;) {
                const size_t i = next.fetch_add(1);
                if (i >= n) break;
                try {
                    f(i);
                } catch (...) {
                    std::lock_guard<std::mutex> lock(mu);
                    if (!err) err = std::current_exception();
                }
            }
        });
    }
    for (auto& t : th) t.join();
    if (err) std::rethrow_exception(err);
}

void merged_body(const BatchItem* items, const std::vector<size_t>& which, HostMesh* outs,
                 hipStream_t s) {
    const auto t_start = std::chrono::steady_clock::now();
    g_stats = MeshStats();
    ensure_tables_once();
    const size_t nm = which.size();
    // Front end per model (Deck, register allocation), on host threads.
    std::vector<MergedModel> mm(nm);
    parallel_for(nm, [&](size_t k) {
        const BatchItem& it = items[which[k]];
        const Deck deck(*it.tree);
        std::vector<float> var_values(deck.vars.size(), 0.0f);
        mm[k].tape = compile_deck(deck, &var_values);
        mm[k].slot_x = deck.X;
        mm[k].slot_y = deck.Y;
        mm[k].slot_z = deck.Z;
        mm[k].top = resolution_level(it.lo, it.hi, it.min_feature);
        mm[k].index = which[k];
    });
    // Roots first by level, highest first: every level's cells are then in
    // model order (children of earlier models, then the roots that join).
    std::stable_sort(mm.begin(), mm.end(), [](const MergedModel& a, const MergedModel& b) {
        return a.top > b.top;
    });
    uint32_t nregs = tape::kFirstFreeReg;
    bool fast = true;
    size_t pool_clauses = 0;
    for (const MergedModel& m : mm) {
        if (m.tape.num_regs > 256) {
            throw Error("render_mesh: the tape needs " + std::to_string(m.tape.num_regs) +
                        " registers; the mesher supports 256");
        }
        if (m.top > 20) throw Error("render_mesh: resolution too fine (more than 20 levels)");
        nregs = std::max(nregs, m.tape.num_regs);
        for (const tape::Clause& c : m.tape.clauses) fast = fast && tape::fast_opcode(tape::op(c));
        pool_clauses += tape::padded_length(uint32_t(m.tape.clauses.size()));
    }
    const int top = mm.front().top;
    const double ms_front = ms_since(t_start);

    // Base tapes (record k = model k) and model records.
    PoolBuf<tape::Clause> pool(std::max<size_t>(pool_clauses * 64, 1u << 20));
    PoolBuf<TapeRec> tapes(std::max<size_t>(1u << 16, 4 * nm));
    std::vector<ModelRec> hmodels(nm);
    {
        std::vector<tape::Clause> h(pool_clauses, tape::Clause{0, 0, 0, 0});
        std::vector<TapeRec> recs(nm);
        uint32_t off = 0;
        for (size_t k = 0; k < nm; ++k) {
            const MergedModel& m = mm[k];
            std::copy(m.tape.clauses.begin(), m.tape.clauses.end(), h.begin() + off);
            TapeRec r{};
            r.off = off;
            r.len = uint32_t(m.tape.clauses.size());
            r.parent = kNoTape;
            r.ref_len = m.tape.ref_length;
            r.model = uint32_t(k);
            recs[k] = r;
            off += tape::padded_length(r.len);
            const BatchItem& it = items[m.index];
            ModelRec& M = hmodels[k];
            for (int a = 0; a < 3; ++a) {
                M.lo[a] = it.lo[a];
                M.hi[a] = it.hi[a];
            }
            M.top = m.top;
            M.root_reg = m.tape.root_reg;
            M.slot_x = m.slot_x;
            M.slot_y = m.slot_y;
            M.slot_z = m.slot_z;
            M.root_cell = 0;
            M.slot0 = 0;
        }
        hip_check(hipMemcpyAsync(pool.p, h.data(), h.size() * sizeof(tape::Clause),
                                 hipMemcpyHostToDevice, s), "upload tapes");
        hip_check(hipMemcpyAsync(tapes.p, recs.data(), recs.size() * sizeof(TapeRec),
                                 hipMemcpyHostToDevice, s), "upload tape records");
        wait(s, "upload tapes");
    }
    // Every record is on the device from the start (k_child_ranges reads
    // the top level of all models at every level); each level's pass then
    // re-uploads the records of the models rooted there, with their root
    // cells set.
    PoolBuf<ModelRec> d_models(nm);
    hip_check(hipMemcpyAsync(d_models.p, hmodels.data(), nm * sizeof(ModelRec),
                             hipMemcpyHostToDevice, s), "model records");
    PoolBuf<Counters> counters(1);
    Counters hc{};
    hc.pool_top = uint32_t(pool_clauses);
    hc.tape_count = uint32_t(nm);
    hip_check(hipMemcpyAsync(counters.p, &hc, sizeof(hc), hipMemcpyHostToDevice, s), "counters");

    Params P{};
    const ModelRec& M0 = hmodels[0];
    for (int a = 0; a < 3; ++a) {
        P.lo[a] = M0.lo[a];
        P.hi[a] = M0.hi[a];
    }
    P.top = M0.top;
    P.max_err = MeshSettings().max_err;
    P.nregs = nregs;
    P.root_reg = M0.root_reg;
    P.slot_x = M0.slot_x;
    P.slot_y = M0.slot_y;
    P.slot_z = M0.slot_z;
    P.qef_f32c = 0;
    P.counters = counters.p;
    P.models = d_models.p;
    auto refresh = [&]() {
        P.pool = pool.p;
        P.pool_cap = uint32_t(std::min<size_t>(pool.n, 0xffffffffu));
        P.tapes = tapes.p;
        P.tape_cap = uint32_t(tapes.n);
    };
    refresh();

    std::vector<LevelStore> levels(top + 1);
    std::vector<PoolBuf<uint32_t>> groups(top + 2), flags(top + 1), basei(top + 1),
        scan_tmp(top + 1), ranges(top + 1);
    std::vector<PoolBuf<uint8_t>> iws(top + 1);
    std::vector<uint32_t> namb(top + 2, 0);
    uint32_t max_len_above = 0;
    size_t push_above = 0;           // push budget of the level above's cells
    const MeshSettings settings;

    // ---- build ------------------------------------------------------------
    size_t next_root = 0;            // models are in decreasing order of top
    for (int L = top; L >= 1; --L) {
        const size_t first_root = next_root;
        uint32_t root_len = 0;
        while (next_root < nm && mm[next_root].top == L) {
            hmodels[next_root].root_cell = 8u * namb[L + 1] + uint32_t(next_root - first_root);
            root_len = std::max<uint32_t>(root_len, uint32_t(mm[next_root].tape.clauses.size()));
            ++next_root;
        }
        const uint32_t nroots = uint32_t(next_root - first_root);
        const uint64_t ncells = 8ull * namb[L + 1] + nroots;
        if (ncells > 0xffffffffull / 4) throw Error("render_mesh: too many cells");
        levels[L].alloc(uint32_t(ncells));
        ranges[L].alloc(2 * nm);
        if (nroots) {
            hip_check(hipMemcpyAsync(d_models.p + first_root, &hmodels[first_root],
                                     nroots * sizeof(ModelRec), hipMemcpyHostToDevice, s),
                      "model records");
        }
        if (ncells) {
            // room for the pushes of this level's evaluations
            grow(pool, hc.pool_top, size_t(hc.pool_top) + push_above +
                 size_t(nroots) * (root_len + kPushExtra), s);
            grow(tapes, hc.tape_count, size_t(hc.tape_count) + ncells, s);
            refresh();
            reset_level_counters(counters.p, s);
            iws[L].alloc(std::max(interval_workspace_bytes(P.nregs, max_len_above, namb[L + 1]),
                                  interval_workspace_bytes(P.nregs, root_len, nroots)));
            if (namb[L + 1]) {
                hip_check(launch_interval(P, L, groups[L + 1].p, namb[L + 1], levels[L + 1].view(),
                                          levels[L].view(), max_len_above, iws[L].p, iws[L].n, s),
                          "interval kernel");
            }
            if (nroots) {
                hip_check(launch_interval(P, L, nullptr, nroots, levels[L].view(), levels[L].view(),
                                          root_len, iws[L].p, iws[L].n, s, uint32_t(first_root)),
                          "root interval kernel");
            }
        }
        const bool has_upper = L < top && levels[L + 1].n > 0;
        hip_check(launch_child_ranges(d_models.p, uint32_t(nm), L,
                                      has_upper ? ranges[L + 1].p : nullptr,
                                      has_upper ? basei[L + 1].p : nullptr,
                                      has_upper ? flags[L + 1].p : nullptr,
                                      has_upper ? levels[L + 1].n : 0u, ranges[L].p, s),
                  "cell ranges");
        LevelStore& cur = levels[L];
        if (cur.n == 0) {
            namb[L] = 0;
            max_len_above = 0;
            push_above = 0;
            continue;
        }
        flags[L].alloc(cur.n);
        basei[L].alloc(cur.n);
        scan_tmp[L].alloc(scan_scratch_words(cur.n));
        groups[L].alloc(cur.n);
        hip_check(launch_flags(cur.view(), flags[L].p, s), "flags");
        hip_check(exclusive_scan(flags[L].p, basei[L].p, cur.n, scan_tmp[L].p, s), "scan");
        hip_check(launch_link(cur.view(), basei[L].p, groups[L].p, s), "link");
        uint32_t last[2] = {0, 0};
        hip_check(hipMemcpyAsync(&last[0], basei[L].p + cur.n - 1, 4, hipMemcpyDeviceToHost, s), "count");
        hip_check(hipMemcpyAsync(&last[1], flags[L].p + cur.n - 1, 4, hipMemcpyDeviceToHost, s), "count");
        hip_check(hipMemcpyAsync(&hc, counters.p, sizeof(hc), hipMemcpyDeviceToHost, s), "counters");
        wait(s, "interval pass");
        check_errors(hc.error);
        const uint64_t na = uint64_t(last[0]) + last[1];
        if (na > cur.n) throw Error("render_mesh: inconsistent cell count");
        if (na * 8u > 0xffffffffull / 4) throw Error("render_mesh: too many cells");
        namb[L] = uint32_t(na);
        max_len_above = hc.max_len;
        if (na && (max_len_above == 0 || max_len_above > hc.pool_top)) {
            throw Error("render_mesh: inconsistent tape length");
        }
        push_above = push_budget(hc, na);
    }
    if (next_root != nm) throw Error("render_mesh: a model's root level was not reached");
    // level 0: the children of the ambiguous level-1 cells
    levels[0].alloc(8u * namb[1]);
    ranges[0].alloc(2 * nm);
    if (namb[1]) {
        run_leaves(P, groups[1].p, namb[1], 8, levels[1].view(), levels[0].view(),
                   std::max<uint32_t>(max_len_above, 1), fast, settings, counters, hc, s);
    }
    {
        const bool has_upper = levels[1].n > 0;
        hip_check(launch_child_ranges(d_models.p, uint32_t(nm), 0,
                                      has_upper ? ranges[1].p : nullptr,
                                      has_upper ? basei[1].p : nullptr,
                                      has_upper ? flags[1].p : nullptr,
                                      has_upper ? levels[1].n : 0u, ranges[0].p, s),
                  "cell ranges");
    }

    // Global numbering: model by model, level 0 first within a model.
    std::vector<uint32_t> hr(size_t(top + 1) * 2 * nm);
    for (int L = 0; L <= top; ++L) {
        hip_check(hipMemcpyAsync(hr.data() + size_t(L) * 2 * nm, ranges[L].p,
                                 2 * nm * sizeof(uint32_t), hipMemcpyDeviceToHost, s), "ranges");
    }
    wait(s, "cell ranges");
    std::vector<uint32_t> cbase(size_t(top + 1) * nm);
    uint64_t total_cells = 0;
    for (size_t k = 0; k < nm; ++k) {
        hmodels[k].slot0 = uint32_t(4 * total_cells);
        for (int L = 0; L <= top; ++L) {
            const uint32_t b = hr[size_t(L) * 2 * nm + k], e = hr[size_t(L) * 2 * nm + nm + k];
            if (e < b || e > levels[L].n) {
                throw Error("render_mesh: inconsistent cell ranges (level " + std::to_string(L) +
                            ", model " + std::to_string(k) + ": " + std::to_string(b) + ".." +
                            std::to_string(e) + " of " + std::to_string(levels[L].n) + ")");
            }
            // index i of the level maps to total_cells + (i - b)
            cbase[size_t(L) * nm + k] = uint32_t(total_cells - b);
            total_cells += e - b;
        }
        if (total_cells * 4 > 0xffffffffull) throw Error("render_mesh: too many cells");
    }
    {
        uint64_t all = 0;
        for (int L = 0; L <= top; ++L) all += levels[L].n;
        if (all != total_cells) throw Error("render_mesh: cell ranges do not cover the levels");
    }
    PoolBuf<uint32_t> d_cbase(cbase.size());
    hip_check(hipMemcpyAsync(d_cbase.p, cbase.data(), cbase.size() * sizeof(uint32_t),
                             hipMemcpyHostToDevice, s), "cell numbering");
    hip_check(hipMemcpyAsync(d_models.p, hmodels.data(), nm * sizeof(ModelRec),
                             hipMemcpyHostToDevice, s), "model records");

    g_stats.ms_build = ms_since(t_start);

    // ---- collapse -----------------------------------------------------------
    const auto t_collapse = std::chrono::steady_clock::now();
    std::vector<LevelView> views(top + 1);
    for (int L = 0; L <= top; ++L) views[L] = levels[L].view(d_cbase.p + size_t(L) * nm);
    for (int L = 1; L <= top; ++L) {
        if (levels[L].n && levels[L - 1].n) {
            hip_check(launch_collapse(P, L, views[L], views[L - 1], s), "collapse kernel");
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
    PoolBuf<uint32_t> d_bases(2 * (nm + 1));
    hip_check(launch_model_bases(d_models.p, uint32_t(nm), tri_off.p, tri_count.p,
                                 uint32_t(total_cells), vid.p, vslot.p, uint32_t(vslot.n),
                                 d_bases.p, s), "model bases");
    std::vector<uint32_t> bases(2 * (nm + 1));
    hip_check(hipMemcpyAsync(bases.data(), d_bases.p, bases.size() * sizeof(uint32_t),
                             hipMemcpyDeviceToHost, s), "model bases");
    hip_check(hipMemcpyAsync(&hc, counters.p, sizeof(hc), hipMemcpyDeviceToHost, s), "counters");
    wait(s, "quad count");
    check_errors(hc.error);
    const uint32_t* tb = bases.data();
    const uint32_t* vb = bases.data() + nm + 1;
    const uint64_t ntris = tb[nm], nverts = vb[nm];
    if (ntris > 24 * total_cells || nverts > 4 * total_cells) {
        throw Error("render_mesh: inconsistent mesh size");
    }
    for (size_t k = 0; k < nm; ++k) {
        if (tb[k + 1] < tb[k] || vb[k + 1] < vb[k]) throw Error("render_mesh: inconsistent mesh size");
    }
    const size_t nv_all = size_t(nverts) + nm;            // a dummy vertex per model
    PoolBuf<uint32_t> d_tris(3 * std::max<size_t>(ntris, 1));
    PoolBuf<float> d_verts(3 * nv_all);
    bool want_keys = false;
    for (size_t k = 0; k < nm; ++k) want_keys = want_keys || outs[mm[k].index].want_keys;
    PoolBuf<uint32_t> d_keys(want_keys ? 5 * nv_all : 0);
    for (int L = 0; L <= top; ++L) {
        hip_check(launch_quads(P, views[L], 1, tri_count.p, vslot.p, tri_off.p, vid.p,
                               d_tris.p, s), "quad write");
        hip_check(launch_verts(P, views[L], L, vslot.p, vid.p, d_verts.p, d_keys.p, s),
                  "vertex gather");
    }
    const size_t vbytes = 3 * nv_all * sizeof(float);
    const size_t tbytes = 3 * size_t(ntris) * sizeof(uint32_t);
    const size_t kbytes = want_keys ? 5 * nv_all * sizeof(uint32_t) : 0;
    size_t granted = 0;
    char* stage = static_cast<char*>(pinned_alloc(vbytes + tbytes + kbytes, &granted));
    struct StageRelease {
        char* p;
        size_t n;
        ~StageRelease() { pinned_free(p, n); }
    } stage_release{stage, granted};
    hip_check(hipMemcpyAsync(stage, d_verts.p, vbytes, hipMemcpyDeviceToHost, s), "copy verts");
    if (ntris) {
        hip_check(hipMemcpyAsync(stage + vbytes, d_tris.p, tbytes, hipMemcpyDeviceToHost, s),
                  "copy tris");
    }
    if (kbytes) {
        hip_check(hipMemcpyAsync(stage + vbytes + tbytes, d_keys.p, kbytes,
                                 hipMemcpyDeviceToHost, s), "copy keys");
    }
    wait(s, "mesh copy");
    check_errors(hc.error);
    const double ms_device = ms_since(t_emit);
    // Each model's block: its dummy vertex and vertices at vb[k] + k, its
    // triangles (already numbered within the model) at tb[k].
    const float* sv = reinterpret_cast<const float*>(stage);
    const uint32_t* st = reinterpret_cast<const uint32_t*>(stage + vbytes);
    const uint32_t* sk = reinterpret_cast<const uint32_t*>(stage + vbytes + tbytes);
    parallel_for(nm, [&](size_t k) {
        HostMesh& out = outs[mm[k].index];
        const size_t nv = size_t(vb[k + 1] - vb[k]) + 1, nt = tb[k + 1] - tb[k];
        const size_t v0 = size_t(vb[k]) + k;
        float* vdst = nullptr;
        uint32_t* tdst = nullptr;
        if (out.alloc) {
            out.alloc(nv, nt, &vdst, &tdst);
        } else {
            out.verts.resize(3 * nv);
            out.tris.resize(3 * nt);
            vdst = out.verts.data();
            tdst = out.tris.data();
        }
        std::memcpy(vdst, sv + 3 * v0, 3 * nv * sizeof(float));
        if (nt) std::memcpy(tdst, st + 3 * size_t(tb[k]), 3 * nt * sizeof(uint32_t));
        if (out.want_keys) out.keys.assign(sk + 5 * v0, sk + 5 * (v0 + nv));
    });
    g_stats.ms_emit = ms_since(t_emit);
    g_stats.ms_total = ms_since(t_start);
    if (std::getenv("LIBFIVE_MESH_TRACE") != nullptr) {
        std::fprintf(stderr, "merged render: %zu models, front end %.1f ms, build %.1f ms, "
                     "collapse %.1f ms, emit %.1f ms (device part %.1f), total %.1f ms\n",
                     nm, ms_front, g_stats.ms_build - ms_front, g_stats.ms_collapse,
                     g_stats.ms_emit, ms_device, g_stats.ms_total);
    }
}

}   // namespace

void render_meshes_merged(const BatchItem* items, size_t n, HostMesh* outs, hipStream_t s) {
    std::vector<size_t> merged;
    for (size_t i = 0; i < n; ++i) {
        if (resolution_level(items[i].lo, items[i].hi, items[i].min_feature) >= 1) {
            merged.push_back(i);
        } else {
            MeshSettings ms;
            ms.min_feature = items[i].min_feature;
            render_mesh_on(*items[i].tree, items[i].lo, items[i].hi, ms, outs[i], s);
        }
    }
    if (merged.empty()) return;
    pool_defer_begin();
    try {
        merged_body(items, merged, outs, s);
        pool_defer_end(s);
    } catch (...) {
        pool_defer_end(s);
        throw;
    }
}

}   // namespace amd
}   // namespace libfive
