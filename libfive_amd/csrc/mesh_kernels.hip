// GPU dual-contouring mesher kernels.
//
// Follows the reference's octree mesher (libfive/src/render/brep/
// worker_pool.inl:36-257, dc/dc_tree.inl:94-813, dc/dc_tree3.cpp:17-154,
// dual.hpp:127-203, dc/dc_mesher.cpp:19-169) restated as level-synchronous
// passes:
//   k_interval  one wave per ambiguous parent: interval-evaluate its 8
//               children with the parent's tape (one child per lane), record
//               MIN/MAX keep decisions with wave ballots, then prune a tape
//               per child (Tape::push) by a backward liveness walk
//   k_leaf      one wave per finest cell: corner states,
//               vertex patches, the 4 x 16-sample edge search, derivatives
//               and the FP64 QEF vertex solve (DCTree::evalLeaf)
//   k_collapse  one wave per branch cell, bottom-up (collectChildren)
//   k_quads     one lane per ambiguous leaf: the minimal edges it owns,
//               located by descending the octree from the root, emitted with
//               the reference's winding and triangulation (DCMesher::load)
//
// Compile with -ffp-contract=off: every rounding must match the CPU
// restatement of the reference.
#include <hip/hip_runtime.h>

#include "mesh_kernels.hpp"
#include "interval_ops.hpp"
#include "mesh_tables.hpp"
#include "point_eval.hpp"
#include "feature_eval.hpp"
#include "qef.hpp"
#include "regfile.hpp"
#include "autotune.hpp"

#include <algorithm>

namespace libfive {
namespace amd {


namespace mesh {

using namespace tape;
namespace OC = Opcode;

__constant__ int8_t c_mv[256 * 4 * 12 * 2];
__constant__ int8_t c_me[8 * 8];
__constant__ int8_t c_mp[256 * 24];
__constant__ uint8_t c_manifold[256];

hipError_t upload_tables(hipStream_t s) {
    const meshing::Tables& t = meshing::tables();
    hipError_t e;
    if ((e = hipMemcpyToSymbolAsync(HIP_SYMBOL(c_mv), t.v, sizeof(t.v), 0,
                                    hipMemcpyHostToDevice, s)) != hipSuccess) return e;
    if ((e = hipMemcpyToSymbolAsync(HIP_SYMBOL(c_me), t.e, sizeof(t.e), 0,
                                    hipMemcpyHostToDevice, s)) != hipSuccess) return e;
    if ((e = hipMemcpyToSymbolAsync(HIP_SYMBOL(c_mp), t.p, sizeof(t.p), 0,
                                    hipMemcpyHostToDevice, s)) != hipSuccess) return e;
    return hipMemcpyToSymbolAsync(HIP_SYMBOL(c_manifold), t.manifold,
                                  sizeof(t.manifold), 0, hipMemcpyHostToDevice, s);
}

__device__ __forceinline__ int mv(uint32_t mask, int patch, int k, int end) {
    return c_mv[((mask * 4 + patch) * 12 + k) * 2 + end];
}

// Region of a cell: the chain of midpoint subdivisions from the root
// (Region::subdivide, region.hpp:89-114: mid = (lower + upper) / 2).
__device__ __forceinline__ void cell_bounds(const Params& P, const uint32_t key[3],
                                            int level, double lo[3], double hi[3]) {
    for (int a = 0; a < 3; ++a) {
        double l = P.lo[a], h = P.hi[a];
        for (int k = P.top; k > level; --k) {
            const double mid = (l + h) / 2;
            if ((key[a] >> (k - 1)) & 1u) l = mid;
            else h = mid;
        }
        lo[a] = l;
        hi[a] = h;
    }
}

__device__ __forceinline__ uint64_t ballot(bool b) { return __ballot(b); }

// The parameters of one model: P's own when P.models is null (one model),
// else those of model m.
__device__ __forceinline__ Params model_view(const Params& P, uint32_t m) {
    Params Q = P;
    if (P.models != nullptr) {
        const ModelRec& M = P.models[m];
        for (int a = 0; a < 3; ++a) {
            Q.lo[a] = M.lo[a];
            Q.hi[a] = M.hi[a];
        }
        Q.top = M.top;
        Q.root_reg = M.root_reg;
        Q.slot_x = M.slot_x;
        Q.slot_y = M.slot_y;
        Q.slot_z = M.slot_z;
        Q.root_cell = M.root_cell;
        Q.slot0 = M.slot0;
    }
    return Q;
}

////////////////////////////////////////////////////////////////////////////////
// Interval classification and pruning

// Synchronises the lanes of one wavefront for LDS exchanges (the interval
// kernel runs one wave per workgroup, so this is all the workgroup needs).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Tapes are walked four clauses per (scalar) load, the next batch loaded
// before the current one is used.  Every tape in the pool is followed by
// at least 7 clauses of padding or other tapes, and batch b covers clauses
// [4b, 4b + 4).
struct Batch4 {
    Clause c[4];
};

// Interval evaluation of up to 8 boxes with one tape, one box per lane
// (lanes 0-7; `active` marks the lanes in use), then Tape::push
// (tape.cpp:23-158) for each box whose result is neither empty nor filled,
// and not NaN unless push_nan.  Returns the tape that the box's children
// use: the pushed tape, or tape_id when the push would change nothing.
//
// Work arrays: choice (per clause, the lanes that keep a and those that
// keep b at a MIN/MAX; 2 bytes) and lv (per clause, the lanes for which it
// is live; 1 byte), in LDS or, for long tapes, in a global workspace; need
// (ceil(nregs / 32) words per lane, 8 lanes) and bytes (16 * nregs: per
// lane register renaming and alias counts) in LDS.
//
// Tapes and records are read through pool_in and recs_in, which alias
// P.pool and P.tapes: a pass only reads tapes made by earlier passes and
// only writes new ones, and saying so lets the compiler keep the clause
// reads on the scalar unit.
template <int NV>
__device__ __forceinline__ uint32_t interval_push(const Params& P,
                                                  const Clause* __restrict__ pool_in,
                                                  const TapeRec* __restrict__ recs_in,
                                                  uint32_t tape_id,
                                                  bool active, const float flo[3],
                                                  const float fhi[3], bool push_nan,
                                                  uint16_t* choice, uint8_t* lv,
                                                  uint32_t* need, uint8_t* bytes,
                                                  float& rlo, float& rhi, bool& rnan)
{
    const int lane = threadIdx.x;
    const TapeRec tr = recs_in[tape_id];
    const uint32_t off = __builtin_amdgcn_readfirstlane(tr.off);
    const uint32_t n = __builtin_amdgcn_readfirstlane(tr.len);
    const uint32_t tflags = __builtin_amdgcn_readfirstlane(tr.flags);
    const Clause* __restrict__ t = pool_in + off;
    const uint32_t W = (P.nregs + 31u) / 32u;

    // The reference slot each register's value resolves to (the same in
    // every lane), kept as a register bank so that the a == b test of the
    // keep rule needs no LDS traffic.
    LF_RF_DECL(RS);
    LF_RF_SET(RS, NV, kRegX, fmath::u2f(P.slot_x));
    LF_RF_SET(RS, NV, kRegY, fmath::u2f(P.slot_y));
    LF_RF_SET(RS, NV, kRegZ, fmath::u2f(P.slot_z));
    __syncthreads();

    LF_RF_DECL(L);
    LF_RF_DECL(U);
    LF_RF_DECL(F);
    {
        const Iv ix = make_checked(flo[0], fhi[0]);
        const Iv iy = make_checked(flo[1], fhi[1]);
        const Iv iz = make_checked(flo[2], fhi[2]);
        LF_RF_SET(L, NV, kRegX, ix.lo); LF_RF_SET(U, NV, kRegX, ix.hi); LF_RF_SET(F, NV, kRegX, ix.nan ? 1.0f : 0.0f);
        LF_RF_SET(L, NV, kRegY, iy.lo); LF_RF_SET(U, NV, kRegY, iy.hi); LF_RF_SET(F, NV, kRegY, iy.nan ? 1.0f : 0.0f);
        LF_RF_SET(L, NV, kRegZ, iz.lo); LF_RF_SET(U, NV, kRegZ, iz.hi); LF_RF_SET(F, NV, kRegZ, iz.nan ? 1.0f : 0.0f);
    }
    bool any_ab = false;
    const Batch4* tb = reinterpret_cast<const Batch4*>(t);
    const uint32_t nb = (n + 3u) / 4u;
    Batch4 cur4 = tb[0];
    for (uint32_t b = 0; b < nb; ++b) {
      const Batch4 nxt4 = tb[b + 1];
#pragma unroll
      for (uint32_t q = 0; q < 4; ++q) {
        const uint32_t k = 4u * b + q;
        if (k >= n) break;
        const Clause c = cur4.c[q];
        const uint32_t o = tape::op(c);
        const uint32_t m = mode(c);
        const float im = imm(c);
        Iv a, b = {0.0f, 0.0f, false};
        if (m == MODE_A_IMM || o == OP_LOADI) {
            a = make_checked(im, im);
        } else {
            const uint32_t r = ra(c);
            a = {LF_RF_GET(L, NV, r), LF_RF_GET(U, NV, r), LF_RF_GET(F, NV, r) != 0.0f};
        }
        if (is_binary(c)) {
            if (m == MODE_B_IMM) {
                b = make_checked(im, im);
            } else {
                const uint32_t r = rb(c);
                b = {LF_RF_GET(L, NV, r), LF_RF_GET(U, NV, r), LF_RF_GET(F, NV, r) != 0.0f};
            }
        }
        if (o == OC::OP_MIN || o == OC::OP_MAX) {
            // The interval keep rule (eval_interval.cpp:116-167); operands
            // that resolve to the same reference slot keep a.
            const bool same = (m == MODE_RR) &&
                fmath::f2u(LF_RF_GET(RS, NV, ra(c))) == fmath::f2u(LF_RF_GET(RS, NV, rb(c)));
            uint32_t keep;
            if (same) keep = KEEP_A;
            else if (a.lo > b.hi) keep = (o == OC::OP_MAX) ? KEEP_A : KEEP_B;
            else if (b.lo > a.hi) keep = (o == OC::OP_MAX) ? KEEP_B : KEEP_A;
            else keep = KEEP_BOTH;
            const uint64_t ma = ballot(active && keep == KEEP_A);
            const uint64_t mb = ballot(active && keep == KEEP_B);
            choice[k] = uint16_t((ma & 0xffu) | ((mb & 0xffu) << 8));
            any_ab |= (keep == KEEP_A || keep == KEEP_B);
        }
        const Iv out = iv_apply(o, a, b, im);
        const uint32_t d = dst(c);
        const float rs = (o == OP_COPY) ? LF_RF_GET(RS, NV, ra(c)) : fmath::u2f(slot(c));
        LF_RF_SET(RS, NV, d, rs);
        LF_RF_SET(L, NV, d, out.lo);
        LF_RF_SET(U, NV, d, out.hi);
        LF_RF_SET(F, NV, d, out.nan ? 1.0f : 0.0f);
      }
      cur4 = nxt4;
    }
    const uint32_t root = P.root_reg;
    rlo = LF_RF_GET(L, NV, root);
    rhi = LF_RF_GET(U, NV, root);
    rnan = LF_RF_GET(F, NV, root) != 0.0f;

    // Tape::push for every box that may prune, in two walks per lane.  The
    // first (root first) marks the clauses whose result is still needed, a
    // MIN/MAX resolved to one side needing only that side.  The second
    // (leaves first) writes the kept clauses with their operands renamed
    // through the resolved clauses, as the reference resolves its remap
    // chains (tape.cpp:124-146): a resolved clause or a copy makes its
    // destination an alias of the register that holds the value and writes
    // nothing.  Because registers are reused, an alias is written out as a
    // copy just before its register would be overwritten; an alias always
    // names a register that holds its own value, so copies never chain.
    bool dopush = active && !(rlo > 0.0f) && !(rhi < 0.0f) && (push_nan || !rnan) &&
                  !(tflags & TAPE_TERMINAL) && any_ab;
    uint32_t child_tape = tape_id;
    __syncthreads();
    if (ballot(dopush)) {
        uint32_t* my = need + uint32_t(lane & 7) * W;
        uint8_t* ren = bytes + uint32_t(lane & 7) * P.nregs;
        uint8_t* cnt = ren + 8u * P.nregs;
        uint32_t base = 0;
        if (dopush) {
            for (uint32_t w = 0; w < W; ++w) my[w] = 0;
            my[root >> 5] |= 1u << (root & 31u);
            base = atomicAdd(&P.counters->pool_top, n + kPushExtra);
            if (base + n + kPushExtra > P.pool_cap) {
                atomicOr(&P.counters->error, ERR_POOL);
                dopush = false;
            }
        }
        bool changed = false, nonterm = false;
        Batch4 bcur = tb[nb - 1];
        for (uint32_t b = nb; b-- > 0;) {
          const Batch4 bprv = tb[b > 0 ? b - 1 : 0];
#pragma unroll
          for (int q = 3; q >= 0; --q) {
            const uint32_t k = 4u * b + uint32_t(q);
            if (k >= n) continue;
            const Clause c = bcur.c[q];
            bool live = false;
            const uint32_t d = dst(c);
            if (dopush && ((my[d >> 5] >> (d & 31u)) & 1u)) {
                live = true;
                my[d >> 5] &= ~(1u << (d & 31u));
                const uint32_t o = tape::op(c);
                const uint32_t m = mode(c);
                bool use_a = nargs(o) >= 1 && m != MODE_A_IMM && o != OP_LOADI;
                bool use_b = is_binary(c) && m != MODE_B_IMM;
                if (o == OC::OP_MIN || o == OC::OP_MAX) {
                    const uint32_t ch = choice[k];
                    const bool ka = (ch >> lane) & 1u;
                    const bool kb = (ch >> (8 + lane)) & 1u;
                    if (ka || kb) {
                        changed = true;
                        use_a = use_a && ka;
                        use_b = use_b && kb;
                    } else {
                        nonterm = true;
                    }
                }
                if (use_a) { const uint32_t r = ra(c); my[r >> 5] |= 1u << (r & 31u); }
                if (use_b) { const uint32_t r = rb(c); my[r >> 5] |= 1u << (r & 31u); }
            }
            const uint64_t lm = ballot(live);
            if (lane == 0) lv[k] = uint8_t(lm & 0xffu);
          }
          bcur = bprv;
        }
        __syncthreads();
        uint32_t pos = base, refc = 0;
        const bool emit = dopush && changed;
        if (emit) {
            for (uint32_t r = 0; r < P.nregs; ++r) {
                ren[r] = uint8_t(r);
                cnt[r] = 0;
            }
        }
        // Before register d is written: aliases of d become copies of it,
        // and d stops being an alias itself.
        auto will_write = [&](uint32_t d) {
            for (uint32_t y = 0; cnt[d] != 0 && y < P.nregs; ++y) {
                if (y != d && ren[y] == d) {
                    P.pool[pos++] = make(OP_COPY, MODE_RR, y, d, 0, 0.0f, 0);
                    ren[y] = uint8_t(y);
                    cnt[d]--;
                }
            }
            if (ren[d] != d) {
                cnt[ren[d]]--;
                ren[d] = uint8_t(d);
            }
        };
        Batch4 ecur = tb[0];
        for (uint32_t b = 0; b < nb; ++b) {
          const Batch4 enxt = tb[b + 1];
#pragma unroll
          for (uint32_t qq = 0; qq < 4; ++qq) {
            const uint32_t k = 4u * b + qq;
            if (k >= n) break;
            if (!emit || !((lv[k] >> lane) & 1u)) continue;
            const Clause c = ecur.c[qq];
            const uint32_t o = tape::op(c);
            const uint32_t m = mode(c);
            const uint32_t d = dst(c);
            bool alias = (o == OP_COPY);
            bool ka = true;
            if (o == OC::OP_MIN || o == OC::OP_MAX) {
                const uint32_t ch = choice[k];
                ka = (ch >> lane) & 1u;
                alias = ka || ((ch >> (8 + lane)) & 1u);
            }
            if (alias) {
                const bool from_imm = (o != OP_COPY) &&
                                      (ka ? (m == MODE_A_IMM) : (m == MODE_B_IMM));
                if (from_imm) {
                    will_write(d);
                    P.pool[pos++] = make(OP_LOADI, MODE_A_IMM, d, 0, 0, imm(c), slot(c));
                    continue;
                }
                const uint32_t r = ren[ka ? ra(c) : rb(c)];
                if (r == d) continue;            // d already holds the value
                if (cnt[d] != 0) {               // d's register is aliased: copy
                    will_write(d);
                    P.pool[pos++] = make(OP_COPY, MODE_RR, d, r, 0, 0.0f, slot(c));
                } else {
                    if (ren[d] != d) cnt[ren[d]]--;
                    ren[d] = uint8_t(r);
                    cnt[r]++;
                }
                continue;
            }
            const bool use_a = nargs(o) >= 1 && m != MODE_A_IMM && o != OP_LOADI;
            const bool use_b = is_binary(c) && m != MODE_B_IMM;
            const uint32_t a2 = use_a ? ren[ra(c)] : ra(c);
            const uint32_t b2 = use_b ? ren[rb(c)] : rb(c);
            will_write(d);
            Clause out = c;
            out.w1 = (a2 & 0xffffu) | (b2 << 16);
            P.pool[pos++] = out;
            if (o != OP_LOADI) ++refc;
          }
          ecur = enxt;
        }
        if (emit && ren[root] != root) {
            P.pool[pos++] = make(OP_COPY, MODE_RR, root, ren[root], 0, 0.0f, 0);
        }
        if (emit) {
            for (uint32_t k = 0; k < 8u; ++k) P.pool[pos + k] = Clause{0, 0, 0, 0};
            const uint32_t id = atomicAdd(&P.counters->tape_count, 1u);
            if (id >= P.tape_cap) {
                atomicOr(&P.counters->error, ERR_TAPES);
            } else {
                TapeRec r;
                r.off = base;
                r.len = pos - base;
                r.parent = tape_id;
                r.flags = TAPE_INTERVAL | (nonterm ? 0u : TAPE_TERMINAL);
                r.box[0] = flo[0]; r.box[1] = fhi[0];
                r.box[2] = flo[1]; r.box[3] = fhi[1];
                r.box[4] = flo[2]; r.box[5] = fhi[2];
                r.ref_len = refc;
                r.model = tr.model;
                P.tapes[id] = r;
                child_tape = id;
            }
        }
    }
    return child_tape;
}

// Carves a workgroup's work arrays (see interval_push) out of LDS, or
// choice and lv out of the global workspace ws when `global`.
struct IntervalWork {
    uint16_t* choice;
    uint8_t* lv;
    uint32_t* need;
    uint8_t* bytes;
};
// The two forms are separate instantiations so that the compiler keeps LDS
// accesses as LDS instructions (a pointer that may point to either memory
// would make every access a flat one).
template <bool GLOBAL>
__device__ __forceinline__ IntervalWork interval_work(uint32_t* lds, const Params& P,
                                                      uint32_t max_len, uint8_t* ws) {
    const uint32_t W = (P.nregs + 31u) / 32u;
    IntervalWork w;
    if constexpr (GLOBAL) {
        uint8_t* arrays = ws + size_t(blockIdx.x) * interval_ws_stride(max_len);
        w.choice = reinterpret_cast<uint16_t*>(arrays);
        w.lv = arrays + 2u * size_t(max_len);
        w.need = lds;
    } else {
        // LDS: [choice, lv (3 * max_len bytes, rounded to words)], need, bytes
        w.choice = reinterpret_cast<uint16_t*>(lds);
        w.lv = reinterpret_cast<uint8_t*>(lds) + 2u * size_t(max_len);
        w.need = lds + (3u * max_len + 3u) / 4u;
    }
    w.bytes = reinterpret_cast<uint8_t*>(w.need + 8u * W);
    return w;
}

// One wave per ambiguous parent (8 children), or the root alone when
// groups is null; workgroup i takes group g0 + i.  The global-workspace
// form is launched over at most kIntervalWaves groups at a time, so that
// workgroup i owns slot i of the workspace.  (A loop over groups inside
// the kernel would make the compiler keep the tape reads on the vector
// unit and spill the register files.)
template <int NV, bool GLOBAL>
__global__ __launch_bounds__(64)
void k_interval(Params P, const Clause* __restrict__ pool_in,
                const TapeRec* __restrict__ recs_in, int level,
                const uint32_t* __restrict__ groups, uint32_t g0, LevelView parent,
                LevelView cur, uint32_t max_len, uint8_t* ws, uint32_t first_root)
{
    HIP_DYNAMIC_SHARED(uint32_t, lds)
    const IntervalWork w = interval_work<GLOBAL>(lds, P, max_len, ws);
    const int lane = threadIdx.x;
    const bool at_root = (groups == nullptr);
    auto run = [&](uint32_t g) {
        // At the roots, group g is model first_root + g, whose base tape is
        // record first_root + g.
        uint32_t tape_id = first_root + g, first = 0, pkey[3] = {0, 0, 0};
        if (!at_root) {
            const uint32_t p = groups[g];
            tape_id = parent.tape[p];
            first = uint32_t(parent.child[p]);
            for (int a = 0; a < 3; ++a) pkey[a] = parent.key[3 * p + a];
        }
        tape_id = __builtin_amdgcn_readfirstlane(tape_id);
        const Params Q = model_view(P, __builtin_amdgcn_readfirstlane(recs_in[tape_id].model));
        if (at_root) first = Q.root_cell;
        const uint32_t nchild = at_root ? 1u : 8u;
        const bool active = uint32_t(lane) < nchild;
        const uint32_t octant = active ? uint32_t(lane) : 0u;
        const uint32_t cidx = first + octant;
        uint32_t key[3];
        for (int a = 0; a < 3; ++a) {
            key[a] = pkey[a] + (at_root ? 0u : (((octant >> a) & 1u) << level));
        }
        double lo[3], hi[3];
        cell_bounds(Q, key, level, lo, hi);
        float flo[3], fhi[3];
        for (int a = 0; a < 3; ++a) {
            flo[a] = float(lo[a]);
            fhi[a] = float(hi[a]);
        }
        float rlo, rhi;
        bool rnan;
        const uint32_t child_tape = interval_push<NV>(Q, pool_in, recs_in, tape_id, active,
                                                      flo, fhi, false,
                                                      w.choice, w.lv, w.need, w.bytes,
                                                      rlo, rhi, rnan);
        // Interval::state (interval.hpp:66-77)
        uint8_t state;
        if (rnan) state = AMBIGUOUS;
        else if (rlo > 0.0f) state = EMPTY;
        else if (rhi < 0.0f) state = FILLED;
        else state = AMBIGUOUS;
        // The longest tape and the push budget of the ambiguous children,
        // reduced over the wave's 8 lanes and added by lane 0.
        uint32_t len = (active && state == AMBIGUOUS) ? P.tapes[child_tape].len : 0u;
        unsigned long long need = (active && state == AMBIGUOUS)
            ? 8ull * (len + kPushExtra) : 0ull;
        for (int d = 1; d < 8; d <<= 1) {
            const uint32_t ol = __shfl_xor(len, d);
            len = ol > len ? ol : len;
            need += (unsigned long long)__shfl_xor((uint32_t)need, d) |
                    ((unsigned long long)__shfl_xor((uint32_t)(need >> 32), d) << 32);
        }
        if (lane == 0 && len) {
            atomicMax(&P.counters->max_len, len);
            atomicAdd(&P.counters->push_need, need);
        }
        if (active) {
            cur.type[cidx] = state;
            cur.tape[cidx] = child_tape;
            cur.child[cidx] = -1;
            for (int a = 0; a < 3; ++a) cur.key[3 * cidx + a] = key[a];
        }
    };
    run(g0 + blockIdx.x);
}

// Groups of up to 8 boxes that share a tape, one wave per group.
template <int NV, bool GLOBAL>
__global__ __launch_bounds__(64)
void k_interval_boxes(Params P, const Clause* __restrict__ pool_in,
                      const TapeRec* __restrict__ recs_in,
                      const BoxGroup* __restrict__ groups, uint32_t g0,
                      const float* __restrict__ boxes, float* __restrict__ bounds,
                      uint8_t* __restrict__ nan, uint32_t* __restrict__ tapes,
                      uint32_t push_nan, uint32_t max_len, uint8_t* ws)
{
    HIP_DYNAMIC_SHARED(uint32_t, lds)
    const IntervalWork w = interval_work<GLOBAL>(lds, P, max_len, ws);
    const int lane = threadIdx.x;
    auto run = [&](uint32_t g) {
        const BoxGroup G = groups[g];
        const uint32_t tape_id = __builtin_amdgcn_readfirstlane(G.tape);
        const uint32_t count = __builtin_amdgcn_readfirstlane(G.count);
        const bool active = uint32_t(lane) < count;
        const uint32_t b = G.first + (active ? uint32_t(lane) : 0u);
        const Params Q = model_view(P, __builtin_amdgcn_readfirstlane(recs_in[tape_id].model));
        const float flo[3] = {boxes[6 * b], boxes[6 * b + 2], boxes[6 * b + 4]};
        const float fhi[3] = {boxes[6 * b + 1], boxes[6 * b + 3], boxes[6 * b + 5]};
        float rlo, rhi;
        bool rnan;
        const uint32_t child_tape = interval_push<NV>(Q, pool_in, recs_in, tape_id, active,
                                                      flo, fhi, push_nan != 0, w.choice, w.lv,
                                                      w.need, w.bytes, rlo, rhi, rnan);
        if (active) {
            bounds[2 * b] = rlo;
            bounds[2 * b + 1] = rhi;
            nan[b] = rnan ? 1 : 0;
            tapes[b] = child_tape;
            atomicMax(&P.counters->max_len, P.tapes[child_tape].len);
        }
    };
    run(g0 + blockIdx.x);
}

// Register banks of 32 per file: 1, 2 or 4 stay in VGPRs; 8 (tapes of up
// to 256 registers) leave the compiler to keep part of each file in AGPRs
// or scratch memory, which is slower but has no other limit.
static int banks(uint32_t nregs) {
    const uint32_t b = (nregs + 31) / 32;
    return b <= 1 ? 1 : b <= 2 ? 2 : b <= 4 ? 4 : b <= 8 ? 8 : 0;
}

int register_banks(uint32_t nregs) { return banks(nregs); }

size_t interval_lds_bytes(uint32_t nregs, uint32_t max_len) {
    const uint32_t W = (nregs + 31) / 32;
    return 32u * W + 16u * size_t(nregs) + 4u * ((3u * size_t(max_len) + 3u) / 4u);
}

__host__ __device__ size_t interval_wide_stride(uint32_t max_len);

// The wide kernel serves levels of at most kWideMaxGroups parents whose
// longest tape has at least kWideMinLen clauses (and at most kWideMaxLen,
// the done bits' LDS): there the one-wave walks are few and long, and the
// level's time is the longest walk.  With more parents the walks fill the
// device anyway.  The "interval_wide" decision of the autotuner overrides
// the rule for tests (0: never, 1: for every level up to kWideMaxLen).
constexpr uint32_t kWideMinLen = 1024;
constexpr uint32_t kWideMaxLen = 1u << 16;
constexpr uint32_t kWideMaxGroups = 1024;
constexpr uint32_t kWideChunk = 256;       // workgroups per launch (workspace slots)

bool interval_uses_wide(uint32_t max_len, uint32_t ngroups) {
    if (max_len > kWideMaxLen) return false;
    const int forced = tune::known("interval_wide");
    if (forced == 0) return false;
    if (forced == 1) return true;
    return max_len >= kWideMinLen && ngroups <= kWideMaxGroups;
}

size_t interval_boxes_workspace_bytes(uint32_t nregs, uint32_t max_len) {
    if (interval_lds_bytes(nregs, max_len) <= kIntervalLdsMax) return 0;
    return size_t(kIntervalWaves) * interval_ws_stride(max_len);
}

size_t interval_workspace_bytes(uint32_t nregs, uint32_t max_len, uint32_t ngroups) {
    size_t b = interval_boxes_workspace_bytes(nregs, max_len);
    if (interval_uses_wide(max_len, ngroups)) {
        b = std::max(b, size_t(std::min(ngroups, kWideChunk)) * interval_wide_stride(max_len));
    }
    return b;
}

namespace {

// Launch geometry of one interval pass: the LDS form runs one workgroup
// per group in one launch; the global-workspace form launches at most
// kIntervalWaves workgroups at a time.
struct IntervalLaunch {
    bool global;
    size_t lds;
    uint32_t chunk;
};

hipError_t interval_launch(uint32_t nregs, uint32_t max_len, uint32_t ngroups,
                           uint8_t* ws, size_t ws_bytes, IntervalLaunch& l) {
    if (banks(nregs) == 0) return hipErrorInvalidValue;
    const size_t full = interval_lds_bytes(nregs, max_len);
    l.global = full > kIntervalLdsMax;
    l.lds = l.global ? interval_lds_bytes(nregs, 0) : full;
    l.chunk = ngroups;
    if (l.global) {
        if (ws == nullptr || ws_bytes < interval_boxes_workspace_bytes(nregs, max_len)) {
            return hipErrorInvalidValue;
        }
        l.chunk = kIntervalWaves;
    }
    return hipSuccess;
}

}   // namespace

////////////////////////////////////////////////////////////////////////////////
// Interval classification and pruning of long tapes
//
// k_interval walks a parent's tape with one wave: 8 lanes (the children)
// step through the clauses one at a time, so a tape of n clauses costs n
// dependent steps however few parents a level has.  At the top levels of a
// long tape (a few parents, thousands of clauses) that serial walk is the
// whole cost of the level.  k_interval_wide gives each parent a workgroup
// of kWideWaves waves:
//   1. sources: for every operand, the clause that defines the value it
//      reads (the SSA form of the register-allocated tape); each wave
//      renames one chunk of the tape and the chunks are joined through
//      their last definitions;
//   2. values: the clauses run out of order, 8 lanes per clause (one per
//      child) and 64 clauses in flight, each as soon as the clauses that
//      define its operands are done; every clause's value has its own slot
//      in the workspace (the register file's reuse would order clauses
//      that are independent).  The smallest pending clause can always run,
//      so the pass always progresses.  The keep decisions of MIN/MAX are
//      recorded per child;
//   3. liveness: each pushing child's wave walks the tape backwards with
//      its set of needed registers spread over its lanes, as
//      interval_push does, listing the live clauses;
//   4. emission: the same wave writes the child's tape with interval_push's
//      walk over that list (operands renamed through decided clauses,
//      aliases copied before their register is overwritten).
// Each value is computed by the same operations on the same operands as in
// interval_push, and the liveness and emission walks are the same, so the
// bounds, the decisions and the pushed tapes are the same.  Every wait is
// bounded: a loop that makes no progress for too long sets ERR_WIDE and
// ends.

constexpr int kWideWaves = 8;
constexpr uint32_t kWideThreads = 64u * kWideWaves;
constexpr int32_t kSrcX = -1, kSrcY = -2, kSrcZ = -3;   // registers 0, 1, 2
constexpr int32_t kSrcImm = -4;                         // the clause's immediate
constexpr int32_t kSrcNone = -5;                        // no operand / no definition
constexpr int32_t kSrcUnres = -16;                      // kSrcUnres - r: register r, not yet found
constexpr uint32_t kWideIdleCap = 1u << 22;             // idle iterations before ERR_WIDE

__host__ __device__ inline size_t wide_align(size_t b) { return (b + 255) & ~size_t(255); }

__host__ __device__ size_t interval_wide_stride(uint32_t max_len) {
    const size_t n = max_len;
    return wide_align(64 * n) + wide_align(8 * n) + wide_align(4 * n) + wide_align(32 * n) +
           wide_align(2 * n) + wide_align(n);
}

struct WideWork {
    float2* val;        // [n][8]: interval of each clause's value per child
    int32_t* src;       // [n][2]: source of operands a and b
    uint32_t* rslot;    // [n]: reference slot of the value (the keep rule's a == b)
    uint32_t* list;     // [8][n]: live clauses of each child, last first
    uint16_t* choice;   // [n]: keep-A children | keep-B children << 8
    uint8_t* vnan;      // [n]: children whose value may be NaN
};

__device__ inline WideWork wide_work(uint8_t* base, uint32_t max_len) {
    const size_t n = max_len;
    WideWork w;
    size_t o = 0;
    w.val = reinterpret_cast<float2*>(base + o); o += wide_align(64 * n);
    w.src = reinterpret_cast<int32_t*>(base + o); o += wide_align(8 * n);
    w.rslot = reinterpret_cast<uint32_t*>(base + o); o += wide_align(4 * n);
    w.list = reinterpret_cast<uint32_t*>(base + o); o += wide_align(32 * n);
    w.choice = reinterpret_cast<uint16_t*>(base + o); o += wide_align(2 * n);
    w.vnan = base + o;
    return w;
}

// Tapes of at most this many clauses keep the liveness pass in LDS.
constexpr uint32_t kWideLdsLive = 12288;

// Dynamic LDS of k_interval_wide: the waves' last-definition tables (4 B x
// kWideWaves per register) and the clauses' done bits; later the liveness
// pass's live bits and pending user counts (3 B per clause) in the same
// space.
size_t interval_wide_lds(uint32_t nregs, uint32_t max_len) {
    const size_t a = size_t(nregs) * 4u * kWideWaves + 4u * ((size_t(max_len) + 31u) / 32u);
    const size_t ml = std::min<size_t>(max_len, kWideLdsLive);
    const size_t b = 4u * ((ml + 3u) / 4u + (ml + 1u) / 2u);
    return std::max(a, b);
}

// One table over the registers, spread over the 64 lanes of a wave: entry
// r lives in lane r % 64, slot r / 64 (up to 256 entries).  The slots are
// named registers, never an array: an indexed array would be placed in
// scratch memory.
struct LaneTable {
    int32_t v0, v1, v2, v3;
};
__device__ __forceinline__ LaneTable lt_make(int32_t a, int32_t b, int32_t c, int32_t d) {
    LaneTable t;
    t.v0 = a;
    t.v1 = b;
    t.v2 = c;
    t.v3 = d;
    return t;
}
// this lane's entry in slot q
__device__ __forceinline__ int32_t lt_slot(const LaneTable& t, uint32_t q) {
    return q == 0 ? t.v0 : q == 1 ? t.v1 : q == 2 ? t.v2 : t.v3;
}
__device__ __forceinline__ int32_t lt_get(const LaneTable& t, uint32_t r) {
    const uint32_t l = r & 63u;
    switch (r >> 6) {
        case 0: return __builtin_amdgcn_readlane(t.v0, l);
        case 1: return __builtin_amdgcn_readlane(t.v1, l);
        case 2: return __builtin_amdgcn_readlane(t.v2, l);
        default: return __builtin_amdgcn_readlane(t.v3, l);
    }
}
__device__ __forceinline__ void lt_set(LaneTable& t, uint32_t r, int32_t x, uint32_t lane) {
    if (lane == (r & 63u)) {
        switch (r >> 6) {
            case 0: t.v0 = x; break;
            case 1: t.v1 = x; break;
            case 2: t.v2 = x; break;
            default: t.v3 = x; break;
        }
    }
}

// A set of registers spread the same way: lane r % 64 holds bit r / 64.
__device__ __forceinline__ bool bs_get(uint32_t bits, uint32_t r) {
    return (__builtin_amdgcn_readlane(bits, r & 63u) >> (r >> 6)) & 1u;
}
__device__ __forceinline__ void bs_put(uint32_t& bits, uint32_t r, bool on, uint32_t lane) {
    if (lane == (r & 63u)) bits = on ? (bits | (1u << (r >> 6))) : (bits & ~(1u << (r >> 6)));
}

__device__ __forceinline__ bool wide_done(const uint32_t* done, int32_t s) {
    if (s < 0) return true;
    const uint32_t w = __hip_atomic_load(&done[uint32_t(s) >> 5], __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_WORKGROUP);
    return (w >> (uint32_t(s) & 31u)) & 1u;
}

__global__ __launch_bounds__(512)
void k_interval_wide(Params P, const Clause* __restrict__ pool_in,
                     const TapeRec* __restrict__ recs_in, int level,
                     const uint32_t* __restrict__ groups, uint32_t g0, LevelView parent,
                     LevelView cur, uint32_t max_len, uint8_t* ws, uint32_t first_root)
{
    HIP_DYNAMIC_SHARED(uint32_t, lds)
    __shared__ float s_lo[8][3], s_hi[8][3];
    __shared__ uint32_t s_key[8][3];
    __shared__ uint32_t s_cidx[8], s_base[8], s_child_tape[8], s_nlive[8];
    __shared__ uint8_t s_state[8];
    __shared__ uint32_t s_anyab, s_dopush, s_changed, s_nonterm, s_err;
    __shared__ int32_t s_root;

    const uint32_t t = threadIdx.x, lane = t & 63u;
    const uint32_t wv = __builtin_amdgcn_readfirstlane(t >> 6);
    const uint32_t g = g0 + blockIdx.x;
    const bool at_root = (groups == nullptr);
    const WideWork W = wide_work(ws + size_t(blockIdx.x) * interval_wide_stride(max_len), max_len);
    const uint32_t nregs = P.nregs;
    // LDS carve (see interval_wide_lds)
    int32_t* defs = reinterpret_cast<int32_t*>(lds);                       // [kWideWaves][nregs]
    uint32_t* done = reinterpret_cast<uint32_t*>(defs + kWideWaves * nregs);
    const uint32_t nwords = (max_len + 31u) / 32u;

    uint32_t tape_id = first_root + g, first = 0, pkey[3] = {0, 0, 0};
    if (!at_root) {
        const uint32_t p = groups[g];
        tape_id = parent.tape[p];
        first = uint32_t(parent.child[p]);
        for (int a = 0; a < 3; ++a) pkey[a] = parent.key[3 * p + a];
    }
    tape_id = __builtin_amdgcn_readfirstlane(tape_id);
    const TapeRec tr = recs_in[tape_id];
    const Params Q = model_view(P, __builtin_amdgcn_readfirstlane(tr.model));
    if (at_root) first = Q.root_cell;
    const uint32_t off = __builtin_amdgcn_readfirstlane(tr.off);
    const uint32_t n = __builtin_amdgcn_readfirstlane(tr.len);
    const uint32_t tflags = __builtin_amdgcn_readfirstlane(tr.flags);
    const uint32_t nchild = at_root ? 1u : 8u;
    const uint32_t root = Q.root_reg;

    if (t < 8) {
        const uint32_t octant = t < nchild ? t : 0u;
        uint32_t key[3];
        for (int a = 0; a < 3; ++a) {
            key[a] = pkey[a] + (at_root ? 0u : (((octant >> a) & 1u) << level));
            s_key[t][a] = key[a];
        }
        double lo[3], hi[3];
        cell_bounds(Q, key, level, lo, hi);
        for (int a = 0; a < 3; ++a) {
            s_lo[t][a] = float(lo[a]);
            s_hi[t][a] = float(hi[a]);
        }
        s_cidx[t] = first + octant;
        s_child_tape[t] = tape_id;
        s_base[t] = 0;
        s_nlive[t] = 0;
    }
    if (t == 0) {
        s_anyab = 0;
        s_dopush = 0;
        s_changed = 0;
        s_nonterm = 0;
        s_err = 0;
        s_root = kSrcNone;
    }
    for (uint32_t i = t; i < nwords; i += kWideThreads) done[i] = 0u;
    if (n > max_len && t == 0) atomicOr(&s_err, 8u);   // the host sized the workspace by max_len
    const uint32_t nn = n > max_len ? 0u : n;
    __syncthreads();

    // ---- 1. sources ---------------------------------------------------------
    const uint32_t csz = (nn + kWideWaves - 1) / kWideWaves;
    const uint32_t c0 = wv * csz, c1 = min(nn, c0 + csz);
    {
        LaneTable def = lt_make(-1, -1, -1, -1);
        for (uint32_t kb = c0; kb < c1; kb += 64) {
            const uint32_t k = kb + lane;
            uint32_t w0 = 0, w1 = 0;
            if (k < c1) {
                const Clause cl = pool_in[off + k];
                w0 = cl.w0;
                w1 = cl.w1;
            }
            int32_t my_a = kSrcNone, my_b = kSrcNone;
            const uint32_t cnt = min(64u, c1 - kb);
            for (uint32_t j = 0; j < cnt; ++j) {
                Clause cl;
                cl.w0 = __builtin_amdgcn_readlane(w0, j);
                cl.w1 = __builtin_amdgcn_readlane(w1, j);
                const uint32_t o = tape::op(cl), m = mode(cl);
                auto reg_src = [&](uint32_t r) -> int32_t {
                    if (r < kFirstFreeReg) return -int32_t(r) - 1;
                    const int32_t d = lt_get(def, r);
                    return d >= 0 ? d : kSrcUnres - int32_t(r);
                };
                int32_t sa = kSrcNone, sb = kSrcNone;
                if (m == MODE_A_IMM || o == OP_LOADI) sa = kSrcImm;
                else if (nargs(o) >= 1) sa = reg_src(ra(cl));
                if (is_binary(cl)) sb = (m == MODE_B_IMM) ? kSrcImm : reg_src(rb(cl));
                if (lane == j) {
                    my_a = sa;
                    my_b = sb;
                }
                lt_set(def, dst(cl), int32_t(kb + j), lane);
            }
            if (k < c1) {
                W.src[2 * k] = my_a;
                W.src[2 * k + 1] = my_b;
            }
        }
        for (int q = 0; q < 4; ++q) {
            const uint32_t r = lane + 64u * uint32_t(q);
            if (r < nregs) defs[wv * nregs + r] = lt_slot(def, uint32_t(q));
        }
    }
    __syncthreads();
    // operands defined in an earlier chunk; the root's source
    auto resolve = [&](int32_t s, uint32_t w) -> int32_t {
        if (s > kSrcUnres) return s;
        const uint32_t r = uint32_t(kSrcUnres - s);
        for (uint32_t w2 = w; w2-- > 0;) {
            const int32_t d = defs[w2 * nregs + r];
            if (d >= 0) return d;
        }
        atomicOr(&s_err, 1u);               // a register read before any write
        return kSrcNone;
    };
    for (uint32_t k = t; k < nn; k += kWideThreads) {
        const uint32_t w = k / csz;
        const int32_t a = W.src[2 * k], b = W.src[2 * k + 1];
        if (a <= kSrcUnres) W.src[2 * k] = resolve(a, w);
        if (b <= kSrcUnres) W.src[2 * k + 1] = resolve(b, w);
    }
    if (t == 0) {
        s_root = root < kFirstFreeReg ? -int32_t(root) - 1 : resolve(kSrcUnres - int32_t(root),
                                                                      kWideWaves);
    }
    __syncthreads();
    // ---- 2. values ------------------------------------------------------------
    {
        const uint32_t slot_g = t >> 3, c = t & 7u;
        const bool cact = c < nchild;
        uint32_t k = slot_g;
        bool have = false;
        Clause cl = Clause{0, 0, 0, 0};
        int32_t sa = kSrcNone, sb = kSrcNone;
        uint32_t idle = 0;
        auto input = [&](int a) -> Iv { return make_checked(s_lo[c][a], s_hi[c][a]); };
        auto operand = [&](int32_t s, float im) -> Iv {
            if (s >= 0) {
                const float2 v = W.val[8u * uint32_t(s) + c];
                return {v.x, v.y, ((W.vnan[s] >> c) & 1u) != 0};
            }
            if (s == kSrcX) return input(0);
            if (s == kSrcY) return input(1);
            if (s == kSrcZ) return input(2);
            if (s == kSrcImm) return make_checked(im, im);
            return {0.0f, 0.0f, false};
        };
        auto rslot_of = [&](int32_t s) -> uint32_t {
            if (s >= 0) return W.rslot[s];
            if (s == kSrcX) return Q.slot_x;
            if (s == kSrcY) return Q.slot_y;
            if (s == kSrcZ) return Q.slot_z;
            return 0u;
        };
        for (;;) {
            const bool mine = k < nn;
            if (ballot(mine) == 0) break;
            if (mine && !have) {
                cl = pool_in[off + k];
                sa = W.src[2 * k];
                sb = W.src[2 * k + 1];
                have = true;
            }
            const bool ready = mine && wide_done(done, sa) && wide_done(done, sb);
            uint32_t keep = KEEP_BOTH;
            bool onan = false;
            uint32_t rs = 0;
            if (ready) {
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                const uint32_t o = tape::op(cl), m = mode(cl);
                const float im = imm(cl);
                Iv a = operand(sa, im);
                Iv b = {0.0f, 0.0f, false};
                if (is_binary(cl)) b = operand(sb, im);
                if (o == OC::OP_MIN || o == OC::OP_MAX) {
                    const bool same = (m == MODE_RR) && rslot_of(sa) == rslot_of(sb);
                    if (same) keep = KEEP_A;
                    else if (a.lo > b.hi) keep = (o == OC::OP_MAX) ? KEEP_A : KEEP_B;
                    else if (b.lo > a.hi) keep = (o == OC::OP_MAX) ? KEEP_B : KEEP_A;
                }
                const Iv out = iv_apply(o, a, b, im);
                onan = out.nan;
                rs = (o == OP_COPY) ? rslot_of(sa) : slot(cl);
                W.val[8u * k + c] = make_float2(out.lo, out.hi);
            }
            const uint64_t rm = ballot(ready);
            const uint64_t ka = ballot(ready && cact && keep == KEEP_A);
            const uint64_t kbm = ballot(ready && cact && keep == KEEP_B);
            const uint64_t nm = ballot(ready && onan);
            if (ready && cact && (keep == KEEP_A || keep == KEEP_B)) atomicOr(&s_anyab, 1u << c);
            if (ready && c == 0) {
                const uint32_t sh = lane & ~7u;
                W.choice[k] = uint16_t(((ka >> sh) & 0xffu) | (((kbm >> sh) & 0xffu) << 8));
                W.vnan[k] = uint8_t((nm >> sh) & 0xffu);
                W.rslot[k] = rs;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            if (ready && c == 0) {
                __hip_atomic_fetch_or(&done[k >> 5], 1u << (k & 31u), __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            if (ready) {
                k += 64;
                have = false;
            }
            if (rm == 0) {
                if (++idle > kWideIdleCap) {
                    if (lane == 0) atomicOr(&s_err, 2u);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
    }
    __syncthreads();

    // ---- the children's bounds and states; push decisions -----------------------
    if (t < nchild) {
        Iv r;
        const int32_t rsrc = s_root;
        if (rsrc >= 0) {
            const float2 v = W.val[8u * uint32_t(rsrc) + t];
            r = {v.x, v.y, ((W.vnan[rsrc] >> t) & 1u) != 0};
        } else if (rsrc >= kSrcZ) {
            const int a = -rsrc - 1;
            r = make_checked(s_lo[t][a], s_hi[t][a]);
        } else {
            r = {0.0f, 0.0f, true};
        }
        uint8_t state;
        if (r.nan) state = AMBIGUOUS;
        else if (r.lo > 0.0f) state = EMPTY;
        else if (r.hi < 0.0f) state = FILLED;
        else state = AMBIGUOUS;
        s_state[t] = state;
        bool dopush = !(r.lo > 0.0f) && !(r.hi < 0.0f) && !r.nan &&
                      !(tflags & TAPE_TERMINAL) && ((s_anyab >> t) & 1u) && s_err == 0;
        if (dopush) {
            const uint32_t base = atomicAdd(&P.counters->pool_top, n + kPushExtra);
            if (base + n + kPushExtra > P.pool_cap) {
                atomicOr(&P.counters->error, ERR_POOL);
                dopush = false;
            }
            s_base[t] = base;
        }
        if (dopush) atomicOr(&s_dopush, 1u << t);
    }
    __syncthreads();

    // ---- 3. liveness ----------------------------------------------------------------
    // Tapes of at most kWideLdsLive clauses: all threads, from the root back
    // to the leaves, each clause passing its children's needs to its
    // sources once all of its users are through (live bits and pending user
    // counts packed in LDS, over the sources' and done bits' space).  Longer
    // tapes: each pushing child's wave walks the tape backwards with its set
    // of needed registers spread over its lanes.  Both give interval_push's
    // live clauses, changed and nonterminal flags.
    const bool lds_live = nn <= kWideLdsLive;
    if (lds_live && s_dopush != 0u) {
        uint32_t* lw = lds;                                  // live bits, 4 clauses a word
        uint32_t* pw = lds + (nn + 3u) / 4u;                 // pending users, 2 clauses a word
        const uint32_t nlw = (nn + 3u) / 4u, npw = (nn + 1u) / 2u;
        for (uint32_t i = t; i < nlw + npw; i += kWideThreads) lds[i] = 0u;
        __syncthreads();
        for (uint32_t k = t; k < nn; k += kWideThreads) {
            const int32_t a = W.src[2 * k], b = W.src[2 * k + 1];
            if (a >= 0) atomicAdd(&pw[uint32_t(a) >> 1], 1u << (16u * (uint32_t(a) & 1u)));
            if (b >= 0) atomicAdd(&pw[uint32_t(b) >> 1], 1u << (16u * (uint32_t(b) & 1u)));
        }
        if (t == 0 && s_root >= 0) {
            atomicOr(&lw[uint32_t(s_root) >> 2], s_dopush << (8u * (uint32_t(s_root) & 3u)));
        }
        __syncthreads();
        int32_t k = int32_t(nn) - 1 - int32_t(t);
        bool have = false;
        int32_t sa = kSrcNone, sb = kSrcNone;
        uint32_t o = 0, ch = 0, idle = 0, changed = 0, nonterm = 0;
        for (;;) {
            const bool mine = k >= 0;
            if (ballot(mine) == 0) break;
            if (mine && !have) {
                o = tape::op(pool_in[off + uint32_t(k)]);
                sa = W.src[2 * k];
                sb = W.src[2 * k + 1];
                ch = W.choice[k];
                have = true;
            }
            bool ready = false;
            if (mine) {
                const uint32_t pwv = __hip_atomic_load(&pw[uint32_t(k) >> 1], __ATOMIC_ACQUIRE,
                                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                ready = ((pwv >> (16u * (uint32_t(k) & 1u))) & 0xffffu) == 0u;
            }
            if (ready) {
                const uint32_t lwv = __hip_atomic_load(&lw[uint32_t(k) >> 2], __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                const uint32_t lv = (lwv >> (8u * (uint32_t(k) & 3u))) & 0xffu;
                uint32_t na = lv, nb = lv;
                if (lv && (o == OC::OP_MIN || o == OC::OP_MAX)) {
                    const uint32_t kam = ch & 0xffu, kbm = ch >> 8;
                    na = lv & ~kbm;
                    nb = lv & ~kam;
                    changed |= lv & (kam | kbm);
                    nonterm |= lv & ~(kam | kbm);
                }
                auto pass = [&](int32_t src, uint32_t needs) {
                    const uint32_t u = uint32_t(src);
                    if (needs) {
                        __hip_atomic_fetch_or(&lw[u >> 2], needs << (8u * (u & 3u)),
                                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    }
                    __hip_atomic_fetch_sub(&pw[u >> 1], 1u << (16u * (u & 1u)), __ATOMIC_RELEASE,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                };
                if (sa >= 0) pass(sa, na);
                if (sb >= 0) pass(sb, nb);
                k -= int32_t(kWideThreads);
                have = false;
            }
            if (ballot(ready) == 0) {
                if (++idle > kWideIdleCap) {
                    if (lane == 0) atomicOr(&s_err, 4u);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        if (changed) atomicOr(&s_changed, changed);
        if (nonterm) atomicOr(&s_nonterm, nonterm);
        __syncthreads();
        // each pushing child's live clauses, in tape order
        if (wv < nchild && ((s_dopush >> wv) & 1u)) {
            const uint32_t c = wv;
            uint32_t* lst = W.list + size_t(c) * max_len;
            uint32_t nl = 0;
            for (uint32_t b = 0; b < nn; b += 64) {
                const uint32_t kk = b + lane;
                const bool bit = kk < nn && ((lw[kk >> 2] >> (8u * (kk & 3u) + c)) & 1u);
                const uint64_t m = ballot(bit);
                if (bit) lst[nl + uint32_t(__popcll(m & ((uint64_t(1) << lane) - 1)))] = kk;
                nl += uint32_t(__popcll(m));
            }
            if (lane == 0) s_nlive[c] = nl;
        }
    }

    // ---- 4. emission: wave c for child c -------------------------------------------
    if (wv < nchild && ((s_dopush >> wv) & 1u)) {
        const uint32_t c = wv;
        uint32_t* lst = W.list + size_t(c) * max_len;
        uint32_t nl = 0;
        bool changed = false;
        if (lds_live) {
            nl = s_nlive[c];
            changed = (s_changed >> c) & 1u;
        } else {
            // the backward walk with the needed registers spread over the lanes
            uint32_t need = 0;
            bs_put(need, root, true, lane);
            bool nonterm = false;
            for (uint32_t kb = nn; kb > 0;) {
                const uint32_t cntb = min(64u, kb);
                kb -= cntb;
                const uint32_t k = kb + lane;
                uint32_t w0 = 0, w1 = 0, chv = 0;
                if (lane < cntb) {
                    const Clause cc = pool_in[off + k];
                    w0 = cc.w0;
                    w1 = cc.w1;
                    chv = W.choice[k];
                }
                uint32_t my_live = 0;
                for (uint32_t j = cntb; j-- > 0;) {
                    Clause c2;
                    c2.w0 = __builtin_amdgcn_readlane(w0, j);
                    c2.w1 = __builtin_amdgcn_readlane(w1, j);
                    const uint32_t d = dst(c2);
                    if (!bs_get(need, d)) continue;
                    bs_put(need, d, false, lane);
                    const uint32_t o = tape::op(c2), m = mode(c2);
                    bool use_a = nargs(o) >= 1 && m != MODE_A_IMM && o != OP_LOADI;
                    bool use_b = is_binary(c2) && m != MODE_B_IMM;
                    if (o == OC::OP_MIN || o == OC::OP_MAX) {
                        const uint32_t ch = __builtin_amdgcn_readlane(chv, j);
                        const bool ka = (ch >> c) & 1u, kb2 = (ch >> (8 + c)) & 1u;
                        if (ka || kb2) {
                            changed = true;
                            use_a = use_a && ka;
                            use_b = use_b && kb2;
                        } else {
                            nonterm = true;
                        }
                    }
                    if (use_a) bs_put(need, ra(c2), true, lane);
                    if (use_b) bs_put(need, rb(c2), true, lane);
                    if (lane == j) my_live = 1;
                }
                // this batch's live clauses at the end of the list, last last
                const uint64_t lm = ballot(my_live != 0);
                const uint32_t hi_rank = uint32_t(__popcll(lm & ~((uint64_t(2) << lane) - 1)));
                if (my_live) lst[max_len - 1 - (nl + hi_rank)] = k;
                nl += uint32_t(__popcll(lm));
            }
            lst += max_len - nl;                       // now in tape order
            if (lane == 0) {
                s_nlive[c] = nl;
                if (changed) atomicOr(&s_changed, 1u << c);
                if (nonterm) atomicOr(&s_nonterm, 1u << c);
            }
        }
        __threadfence_block();

        if (changed && s_err == 0) {
            // interval_push's emission walk, uniform over the wave: the
            // renaming (ren) and alias counts (cnt) are tables spread over
            // the lanes; lane 0 stores the clauses
            LaneTable ren = lt_make(int32_t(lane), int32_t(lane + 64u), int32_t(lane + 128u),
                                    int32_t(lane + 192u));
            LaneTable cnt = lt_make(0, 0, 0, 0);
            uint32_t pos = s_base[c], refc = 0;
            auto put = [&](const Clause& x) {
                if (lane == 0) P.pool[pos] = x;
                ++pos;
            };
            auto will_write = [&](uint32_t d) {
                if (lt_get(cnt, d) != 0) {
                    // every register aliased to d becomes a copy of d, in
                    // register order
                    for (int q = 0; q < 4; ++q) {
                        const uint32_t y = lane + 64u * uint32_t(q);
                        uint64_t m = ballot(y < nregs && y != d &&
                                            uint32_t(lt_slot(ren, uint32_t(q))) == d);
                        while (m) {
                            const uint32_t l = uint32_t(__builtin_ctzll(m));
                            m &= m - 1;
                            const uint32_t yy = l + 64u * uint32_t(q);
                            put(make(OP_COPY, MODE_RR, yy, d, 0, 0.0f, 0));
                            lt_set(ren, yy, int32_t(yy), lane);
                        }
                    }
                    lt_set(cnt, d, 0, lane);
                }
                const uint32_t rd = uint32_t(lt_get(ren, d));
                if (rd != d) {
                    lt_set(cnt, rd, lt_get(cnt, rd) - 1, lane);
                    lt_set(ren, d, int32_t(d), lane);
                }
            };
            // registers that are aliases or aliased (S, four 64-bit words over
            // the register numbers): a clause that reads or writes none of
            // them and is no alias itself is emitted as it stands
            uint64_t sp0 = 0, sp1 = 0, sp2 = 0, sp3 = 0;
            auto recompute = [&]() {
                auto word = [&](uint32_t q) {
                    const uint32_t y = lane + 64u * q;
                    return ballot(y < nregs && (uint32_t(lt_slot(ren, q)) != y ||
                                                lt_slot(cnt, q) != 0));
                };
                sp0 = word(0);
                sp1 = word(1);
                sp2 = word(2);
                sp3 = word(3);
            };
            // interval_push's step for one clause (uniform over the wave)
            auto serial = [&](const Clause& c2, uint32_t ch) {
                const uint32_t o = tape::op(c2);
                const uint32_t m = mode(c2);
                const uint32_t d = dst(c2);
                bool alias = (o == OP_COPY);
                bool ka = true;
                if (o == OC::OP_MIN || o == OC::OP_MAX) {
                    ka = (ch >> c) & 1u;
                    alias = ka || ((ch >> (8 + c)) & 1u);
                }
                if (alias) {
                    const bool from_imm = (o != OP_COPY) &&
                                          (ka ? (m == MODE_A_IMM) : (m == MODE_B_IMM));
                    if (from_imm) {
                        will_write(d);
                        put(make(OP_LOADI, MODE_A_IMM, d, 0, 0, imm(c2), slot(c2)));
                        return;
                    }
                    const uint32_t r = uint32_t(lt_get(ren, ka ? ra(c2) : rb(c2)));
                    if (r == d) return;
                    if (lt_get(cnt, d) != 0) {
                        will_write(d);
                        put(make(OP_COPY, MODE_RR, d, r, 0, 0.0f, slot(c2)));
                    } else {
                        const uint32_t rd = uint32_t(lt_get(ren, d));
                        if (rd != d) lt_set(cnt, rd, lt_get(cnt, rd) - 1, lane);
                        lt_set(ren, d, int32_t(r), lane);
                        lt_set(cnt, r, lt_get(cnt, r) + 1, lane);
                    }
                    return;
                }
                const bool use_a = nargs(o) >= 1 && m != MODE_A_IMM && o != OP_LOADI;
                const bool use_b = is_binary(c2) && m != MODE_B_IMM;
                const uint32_t a2 = use_a ? uint32_t(lt_get(ren, ra(c2))) : ra(c2);
                const uint32_t b2 = use_b ? uint32_t(lt_get(ren, rb(c2))) : rb(c2);
                will_write(d);
                Clause outc = c2;
                outc.w1 = (a2 & 0xffffu) | (b2 << 16);
                put(outc);
                if (o != OP_LOADI) ++refc;
            };
            for (uint32_t jb = 0; jb < nl; jb += 64) {
                const uint32_t cntb = min(64u, nl - jb);
                Clause mine = Clause{0, 0, 0, 0};
                uint32_t chv = 0;
                if (lane < cntb) {
                    const uint32_t k = lst[jb + lane];
                    mine = pool_in[off + k];
                    chv = W.choice[k];
                }
                const uint32_t lo = tape::op(mine), lm = mode(mine);
                bool lal = (lo == OP_COPY);
                if (lo == OC::OP_MIN || lo == OC::OP_MAX) {
                    lal = (((chv >> c) | (chv >> (8 + c))) & 1u) != 0;
                }
                const bool lua = nargs(lo) >= 1 && lm != MODE_A_IMM && lo != OP_LOADI;
                const bool lub = is_binary(mine) && lm != MODE_B_IMM;
                for (uint32_t q = 0; q < cntb;) {
                    bool special = lal;
                    if ((sp0 | sp1 | sp2 | sp3) != 0) {
                        auto in_s = [&](uint32_t r) {
                            const uint64_t w = r < 64u ? sp0 : r < 128u ? sp1 : r < 192u ? sp2 : sp3;
                            return ((w >> (r & 63u)) & 1u) != 0;
                        };
                        special = special || (lua && in_s(ra(mine))) || (lub && in_s(rb(mine))) ||
                                  in_s(dst(mine));
                    }
                    const bool ahead = lane >= q && lane < cntb;
                    const uint64_t sm = ballot(ahead && special);
                    const uint32_t e = sm ? uint32_t(__builtin_ctzll(sm)) : cntb;
                    // clauses q .. e-1 as they stand
                    const bool run = lane >= q && lane < e;
                    if (run) P.pool[pos + lane - q] = mine;
                    refc += uint32_t(__popcll(ballot(run && lo != OP_LOADI)));
                    pos += e - q;
                    if (e >= cntb) break;
                    Clause c2;
                    c2.w0 = __builtin_amdgcn_readlane(mine.w0, e);
                    c2.w1 = __builtin_amdgcn_readlane(mine.w1, e);
                    c2.w2 = __builtin_amdgcn_readlane(mine.w2, e);
                    c2.w3 = __builtin_amdgcn_readlane(mine.w3, e);
                    serial(c2, __builtin_amdgcn_readlane(chv, e));
                    recompute();
                    q = e + 1;
                }
            }
            const uint32_t rr = uint32_t(lt_get(ren, root));
            if (rr != root) put(make(OP_COPY, MODE_RR, root, rr, 0, 0.0f, 0));
            if (lane < 8u) P.pool[pos + lane] = Clause{0, 0, 0, 0};
            if (lane == 0) {
                const uint32_t id = atomicAdd(&P.counters->tape_count, 1u);
                if (id >= P.tape_cap) {
                    atomicOr(&P.counters->error, ERR_TAPES);
                } else {
                    TapeRec r;
                    const uint32_t base = s_base[c];
                    r.off = base;
                    r.len = pos - base;
                    r.parent = tape_id;
                    r.flags = TAPE_INTERVAL | (((s_nonterm >> c) & 1u) ? 0u : TAPE_TERMINAL);
                    r.box[0] = s_lo[c][0]; r.box[1] = s_hi[c][0];
                    r.box[2] = s_lo[c][1]; r.box[3] = s_hi[c][1];
                    r.box[4] = s_lo[c][2]; r.box[5] = s_hi[c][2];
                    r.ref_len = refc;
                    r.model = tr.model;
                    P.tapes[id] = r;
                    s_child_tape[c] = id;
                }
            }
        }
    }
    __syncthreads();
    if (t == 0 && s_err) atomicOr(&P.counters->error, ERR_WIDE);

    // ---- the children's cell records, as k_interval --------------------------------
    if (wv == 0) {
        const bool active = lane < nchild;
        const uint32_t ci = active ? lane : 0u;
        const uint32_t child_tape = s_child_tape[ci];
        const uint8_t state = s_state[ci];
        uint32_t len = (active && state == AMBIGUOUS) ? P.tapes[child_tape].len : 0u;
        unsigned long long need = (active && state == AMBIGUOUS)
            ? 8ull * (len + kPushExtra) : 0ull;
        for (int d = 1; d < 8; d <<= 1) {
            const uint32_t ol = __shfl_xor(len, d);
            len = ol > len ? ol : len;
            need += (unsigned long long)__shfl_xor((uint32_t)need, d) |
                    ((unsigned long long)__shfl_xor((uint32_t)(need >> 32), d) << 32);
        }
        if (lane == 0 && len) {
            atomicMax(&P.counters->max_len, len);
            atomicAdd(&P.counters->push_need, need);
        }
        if (active) {
            const uint32_t cidx = s_cidx[ci];
            cur.type[cidx] = state;
            cur.tape[cidx] = child_tape;
            cur.child[cidx] = -1;
            for (int a = 0; a < 3; ++a) cur.key[3 * cidx + a] = s_key[ci][a];
        }
    }
}

#define LF_INTERVAL_SWITCH(KERNEL, ...)                                                   \
    for (uint32_t g0 = 0; g0 < ngroups; g0 += l.chunk) {                                  \
        const dim3 grid(ngroups - g0 < l.chunk ? ngroups - g0 : l.chunk);                 \
        switch (banks(p.nregs) * (l.global ? -1 : 1)) {                                   \
            case 1: hipLaunchKernelGGL((KERNEL<1, false>), grid, dim3(64), l.lds, s, __VA_ARGS__); break; \
            case 2: hipLaunchKernelGGL((KERNEL<2, false>), grid, dim3(64), l.lds, s, __VA_ARGS__); break; \
            case 4: hipLaunchKernelGGL((KERNEL<4, false>), grid, dim3(64), l.lds, s, __VA_ARGS__); break; \
            case 8: hipLaunchKernelGGL((KERNEL<8, false>), grid, dim3(64), l.lds, s, __VA_ARGS__); break; \
            case -1: hipLaunchKernelGGL((KERNEL<1, true>), grid, dim3(64), l.lds, s, __VA_ARGS__); break; \
            case -2: hipLaunchKernelGGL((KERNEL<2, true>), grid, dim3(64), l.lds, s, __VA_ARGS__); break; \
            case -4: hipLaunchKernelGGL((KERNEL<4, true>), grid, dim3(64), l.lds, s, __VA_ARGS__); break; \
            case -8: hipLaunchKernelGGL((KERNEL<8, true>), grid, dim3(64), l.lds, s, __VA_ARGS__); break; \
            default: return hipErrorInvalidValue;                                         \
        }                                                                                 \
        const hipError_t le = hipGetLastError();                                          \
        if (le != hipSuccess) return le;                                                  \
    }

hipError_t launch_interval(const Params& p, int level, const uint32_t* groups,
                           uint32_t ngroups, const LevelView& parent,
                           const LevelView& cur, uint32_t max_len, uint8_t* ws,
                           size_t ws_bytes, hipStream_t s, uint32_t first_root) {
    if (ngroups == 0) return hipSuccess;
    if (banks(p.nregs) != 0 && interval_uses_wide(max_len, ngroups)) {
        const size_t stride = interval_wide_stride(max_len);
        if (ws == nullptr || ws_bytes < size_t(std::min(ngroups, kWideChunk)) * stride) {
            return hipErrorInvalidValue;
        }
        const size_t lds = interval_wide_lds(p.nregs, max_len);
        for (uint32_t g0 = 0; g0 < ngroups; g0 += kWideChunk) {
            const uint32_t ng = std::min(kWideChunk, ngroups - g0);
            hipLaunchKernelGGL(k_interval_wide, dim3(ng), dim3(kWideThreads), lds, s, p, p.pool,
                               p.tapes, level, groups, g0, parent, cur, max_len, ws, first_root);
            const hipError_t le = hipGetLastError();
            if (le != hipSuccess) return le;
        }
        return hipSuccess;
    }
    IntervalLaunch l;
    const hipError_t e = interval_launch(p.nregs, max_len, ngroups, ws, ws_bytes, l);
    if (e != hipSuccess) return e;
    LF_INTERVAL_SWITCH(k_interval, p, p.pool, p.tapes, level, groups, g0, parent, cur,
                       max_len, ws, first_root)
    return hipSuccess;
}

hipError_t launch_interval_boxes(const Params& p, const BoxGroup* groups, uint32_t ngroups,
                                 const float* boxes, float* bounds, uint8_t* nan,
                                 uint32_t* tapes, bool push_nan, uint32_t max_len,
                                 uint8_t* ws, size_t ws_bytes, hipStream_t s) {
    if (ngroups == 0) return hipSuccess;
    IntervalLaunch l;
    const hipError_t e = interval_launch(p.nregs, max_len, ngroups, ws, ws_bytes, l);
    if (e != hipSuccess) return e;
    LF_INTERVAL_SWITCH(k_interval_boxes, p, p.pool, p.tapes, groups, g0, boxes, bounds, nan,
                       tapes, push_nan ? 1u : 0u, max_len, ws)
    return hipSuccess;
}

#undef LF_INTERVAL_SWITCH

////////////////////////////////////////////////////////////////////////////////
// Level linking

__global__ void k_flags(LevelView cur, uint32_t* flags) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cur.n) flags[i] = (cur.type[i] == AMBIGUOUS) ? 1u : 0u;
}

__global__ void k_link(LevelView cur, const uint32_t* base, uint32_t* groups) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cur.n) return;
    if (cur.type[i] == AMBIGUOUS) {
        cur.child[i] = int32_t(8u * base[i]);
        groups[base[i]] = i;
    } else {
        cur.child[i] = -1;
    }
}

hipError_t launch_flags(const LevelView& cur, uint32_t* flags, hipStream_t s) {
    if (cur.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_flags, dim3((cur.n + 255) / 256), dim3(256), 0, s, cur, flags);
    return hipGetLastError();
}

hipError_t launch_link(const LevelView& cur, const uint32_t* base,
                       uint32_t* groups, hipStream_t s) {
    if (cur.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_link, dim3((cur.n + 255) / 256), dim3(256), 0, s, cur, base, groups);
    return hipGetLastError();
}

////////////////////////////////////////////////////////////////////////////////
// Leaf cells

// FEAT = false: a cell that needs the exact-zero feature evaluation is
// appended to `deferred` and left untouched.  FEAT = true: the cells listed
// in `work` are evaluated with that path, each wave with its own workspace.
// With a one-bank register file, at least 4 waves per SIMD: the compiler
// keeps the kernel within 128 VGPRs, spilling rarely used values, which
// measured faster than 2 waves with no spills (gyroid 512^3: 169 -> 152 ms
// per render).  Larger files keep the compiler's choice.
template <int NV, bool FEAT>
__global__ __launch_bounds__(64, NV == 1 ? 4 : 1)
void k_leaf(Params P, const uint32_t* __restrict__ groups, uint32_t ngroups,
            uint32_t per_group, LevelView l1, LevelView l0,
            const uint32_t* __restrict__ work, uint32_t* __restrict__ deferred,
            uint8_t* __restrict__ ws_base, size_t ws_bytes)
{
    __shared__ double s_tf[12][3], s_ts[12][3];
    __shared__ double s_acc[12][14];
    __shared__ int s_rank[12];
    __shared__ float s_v[32], s_g[32][3];
    __shared__ uint32_t s_fl[32];
    __shared__ float s_feat[FEAT ? 24 : 1][feat::kMaxOut][3];
    __shared__ int s_nfeat[FEAT ? 24 : 1];
    const int lane = threadIdx.x;
    const uint32_t total = ngroups * per_group;
    uint32_t* err = &P.counters->error;

    const uint32_t w = FEAT ? work[blockIdx.x] : blockIdx.x;
    if (w >= total) return;
    auto defer = [&]() {
        if (lane == 0) deferred[atomicAdd(&P.counters->deferred, 1u)] = w;
    };
    do {
        const uint32_t p = groups[w / per_group];
        const uint32_t oct = w % per_group;
        const uint32_t cidx = uint32_t(l1.child[p]) + oct;
        uint32_t key[3];
        for (int a = 0; a < 3; ++a) key[a] = l1.key[3 * p + a] + ((oct >> a) & 1u);
        const uint32_t tape_id = __builtin_amdgcn_readfirstlane(l1.tape[p]);
        const TapeRec tr = P.tapes[tape_id];
        const uint32_t n = __builtin_amdgcn_readfirstlane(tr.len);
        const Clause* __restrict__ t = P.pool + __builtin_amdgcn_readfirstlane(tr.off);
        const Params Q = model_view(P, __builtin_amdgcn_readfirstlane(tr.model));
        const uint32_t root = Q.root_reg;
        double lo[3], hi[3];
        cell_bounds(Q, key, 0, lo, hi);
        feat::Workspace W{};
        if (FEAT) W = feat::carve(ws_base + size_t(blockIdx.x) * ws_bytes, n, P.nregs);

        // Corner states (dc_tree.inl:131-239): lane & 7 is the corner.
        const int corner = lane & 7;
        float cp[3];
        for (int a = 0; a < 3; ++a) cp[a] = float(((corner >> a) & 1) ? hi[a] : lo[a]);
        // Values first; partials only when a corner is an unambiguous zero.
        float g[3] = {0.0f, 0.0f, 0.0f};
        uint32_t fl;
        const float v = pe_value<NV>(t, n, root, cp[0], cp[1], cp[2], &fl);
        if (ballot(lane < 8 && v == 0.0f && !(fl & PT_AMBIGUOUS))) {
            uint32_t fl2;
            (void)pe_grad3<NV>(t, n, root, cp[0], cp[1], cp[2], g, &fl2);
        }
        bool inside;
        if (v > 0.0f || fmath::isnan_f(v)) inside = false;
        else if (v < 0.0f) inside = true;
        else if (!(fl & PT_AMBIGUOUS)) inside = (g[0] != 0.0f || g[1] != 0.0f || g[2] != 0.0f);
        else inside = false;
        {
            // Ambiguous zeros: FeatureEvaluator::isInside, one point at a
            // time on the whole wave.
            uint64_t need = ballot(lane < 8 && v == 0.0f && (fl & PT_AMBIGUOUS));
            if (need) {
                if (!FEAT) { defer(); return; }
                while (need) {
                    const int l = __builtin_ctzll(need);
                    need &= need - 1;
                    const float q[3] = {__shfl(cp[0], l), __shfl(cp[1], l), __shfl(cp[2], l)};
                    int r = 0;
                    if (lane == 0) {
                        r = feat::is_inside(t, n, P.nregs, root, Q.slot_x, Q.slot_y, Q.slot_z,
                                            q, W, err) ? 1 : 0;
                    }
                    r = __shfl(r, 0);
                    if (lane == l) inside = r != 0;
                }
            }
        }
        const uint32_t mask = uint32_t(ballot(lane < 8 && inside) & 0xffu);
        if (mask == 0u || mask == 0xffu) {
            if (lane == 0) {
                l0.type[cidx] = mask ? FILLED : EMPTY;
                l0.child[cidx] = -1;
                l0.tape[cidx] = tape_id;
                for (int a = 0; a < 3; ++a) l0.key[3 * cidx + a] = key[a];
            }
            continue;
        }

        LeafRec* leaf = l0.leaf + cidx;
        uint32_t vcount = 0, rank = 0;
        for (int patch = 0; patch < 4; ++patch) {
            if (mv(mask, patch, 0, 0) == -1) break;
            int ec = 0;
            while (ec < 12 && mv(mask, patch, ec, 0) != -1) ++ec;
            if (lane < ec) {
                const int a = mv(mask, patch, lane, 0), b = mv(mask, patch, lane, 1);
                for (int q = 0; q < 3; ++q) {
                    s_tf[lane][q] = ((a >> q) & 1) ? hi[q] : lo[q];
                    s_ts[lane][q] = ((b >> q) & 1) ? hi[q] : lo[q];
                }
            }
            wave_sync();
            // Edge search: 4 rounds of 16 samples (dc_tree.inl:339-410).
            const int passes = (16 * ec + 63) / 64;
            for (int round = 0; round < 4; ++round) {
                for (int pass = 0; pass < passes; ++pass) {
                    const int i = pass * 64 + lane;
                    const int e = i >> 4, j = i & 15;
                    const bool valid = e < ec;
                    const double frac = j / 15.0;
                    double ps[3];
                    for (int q = 0; q < 3; ++q) {
                        ps[q] = s_tf[e][q] * (1 - frac) + s_ts[e][q] * frac;
                    }
                    uint32_t vf;
                    const float sv = pe_value<NV>(t, n, root, float(ps[0]), float(ps[1]),
                                                  float(ps[2]), &vf);
                    // First j >= 1 with v > 0, or v == 0 outside, else j = 15.
                    const bool zero = valid && j >= 1 && sv == 0.0f;
                    bool zero_out = false;
                    const bool fast = zero && !(vf & PT_TIE);
                    if (ballot(fast)) {
                        float sg[3];
                        uint32_t f2;
                        pe_grad3<NV>(t, n, root, float(ps[0]), float(ps[1]), float(ps[2]), sg, &f2);
                        const float nrm = sqrtf(sg[0] * sg[0] + (sg[1] * sg[1] + sg[2] * sg[2]));
                        if (fast) zero_out = !(nrm > 0.0f);
                    }
                    {
                        uint64_t need = ballot(zero && (vf & PT_TIE));
                        if (need) {
                            if (!FEAT) { defer(); return; }
                            const float fp[3] = {float(ps[0]), float(ps[1]), float(ps[2])};
                            while (need) {
                                const int l = __builtin_ctzll(need);
                                need &= need - 1;
                                const float q[3] = {__shfl(fp[0], l), __shfl(fp[1], l),
                                                    __shfl(fp[2], l)};
                                int r = 0;
                                if (lane == 0) {
                                    r = feat::is_inside(t, n, P.nregs, root, Q.slot_x, Q.slot_y,
                                                        Q.slot_z, q, W, err) ? 1 : 0;
                                }
                                r = __shfl(r, 0);
                                if (lane == l) zero_out = (r == 0);
                            }
                        }
                    }
                    const bool take = valid && j >= 1 &&
                        (sv > 0.0f || zero_out || (j == 15 && !(sv == 0.0f)));
                    const uint64_t tm = ballot(take);
                    const uint32_t seg = uint32_t(tm >> ((lane >> 4) * 16)) & 0xffffu;
                    const int first = seg ? __builtin_ctz(seg) : -1;
                    wave_sync();
                    if (valid && first >= 1) {
                        if (j == first) {
                            for (int q = 0; q < 3; ++q) s_ts[e][q] = ps[q];
                        } else if (j == first - 1) {
                            for (int q = 0; q < 3; ++q) s_tf[e][q] = ps[q];
                        }
                    }
                    wave_sync();
                }
            }
            // Derivatives at both bracket ends (dc_tree.inl:412-471): one
            // lane per (axis, end), one partial per lane, so that the 2*ec
            // ends of all three axes share a pass.
            {
                const int m = 2 * ec;
                for (int base = 0; base < 3 * m; base += 64) {
                    const int i = base + lane;
                    const int axis = i / m, k = i - axis * m;
                    const bool valid = axis < 3;
                    const int e = valid ? (k >> 1) : 0;
                    const bool up = valid && (k & 1);
                    float ep[3];
                    for (int q = 0; q < 3; ++q) ep[q] = float(up ? s_ts[e][q] : s_tf[e][q]);
                    float ga;
                    uint32_t ef;
                    const float ev = pe_deriv1<NV>(t, n, root, ep[0], ep[1], ep[2],
                                                   valid ? axis : 0, &ga, &ef);
                    if (valid) {
                        s_g[k][axis] = ga;
                        if (axis == 0) {
                            s_v[k] = ev;
                            s_fl[k] = ef;
                        }
                    }
                }
                wave_sync();
                // Ambiguous ends contribute every distinct gradient of the
                // point (FeatureEvaluator::features).
                uint64_t need = ballot(lane < m && (s_fl[lane < m ? lane : 0] & PT_AMBIGUOUS));
                if (need) {
                    if (!FEAT) { defer(); return; }
                    while (need) {
                        const int l = __builtin_ctzll(need);
                        need &= need - 1;
                        const int e = l >> 1;
                        float q[3];
                        for (int c = 0; c < 3; ++c) q[c] = float((l & 1) ? s_ts[e][c] : s_tf[e][c]);
                        if (lane == 0) {
                            float fo[feat::kMaxOut][3];
                            const int nf = feat::features(t, n, P.nregs, root, Q.slot_x,
                                                          Q.slot_y, Q.slot_z, q, W, fo, err);
                            s_nfeat[l] = nf;
                            for (int f = 0; f < nf; ++f) {
                                s_feat[l][f][0] = fo[f][0];
                                s_feat[l][f][1] = fo[f][1];
                                s_feat[l][f][2] = fo[f][2];
                            }
                        }
                    }
                }
            }
            wave_sync();
            if (lane < ec) {
                double acc[14];
                for (int q = 0; q < 14; ++q) acc[q] = 0.0;
                qef::AccF2 af;
                qef::f2_zero(af);
                const bool f2 = P.qef_f32c != 0;
                bool exists = false;
                for (int end = 0; end < 2; ++end) {
                    const int k = 2 * lane + end;
                    double pos[3];
                    for (int q = 0; q < 3; ++q) pos[q] = end ? s_ts[lane][q] : s_tf[lane][q];
                    const float posf[3] = {float(pos[0]), float(pos[1]), float(pos[2])};
                    if (!(s_fl[k] & PT_AMBIGUOUS)) {
                        const double d[3] = {double(s_g[k][0]), double(s_g[k][1]),
                                             double(s_g[k][2])};
                        if (f2) qef::push_sample_f2(af, posf, s_g[k], s_v[k]);
                        else qef::push_sample(acc, pos, d, double(s_v[k]));
                        exists = true;
                    } else if (FEAT) {
                        for (int f = 0; f < s_nfeat[k]; ++f) {
                            const double d[3] = {double(s_feat[k][f][0]), double(s_feat[k][f][1]),
                                                 double(s_feat[k][f][2])};
                            if (f2) qef::push_sample_f2(af, posf, s_feat[k][f], s_v[k]);
                            else qef::push_sample(acc, pos, d, double(s_v[k]));
                            exists = true;
                        }
                    }
                }
                if (f2) qef::f2_widen(af, acc);
                int r = -1;
                if (exists) {
                    double ata[3][3];
                    qef::unpack_ata(acc, ata);
                    r = qef::small_eigenvalue_count(ata);
                }
                s_rank[lane] = r;
                for (int q = 0; q < 14; ++q) s_acc[lane][q] = acc[q];
            }
            wave_sync();
            if (lane == 0) {
                int max_rank = 0;
                for (int e = 0; e < ec; ++e) max_rank = s_rank[e] > max_rank ? s_rank[e] : max_rank;
                double mass[4] = {0.0, 0.0, 0.0, 0.0};
                double sum[14];
                for (int q = 0; q < 14; ++q) sum[q] = 0.0;
                for (int e = 0; e < ec; ++e) {
                    if (s_rank[e] < 0) continue;
                    if (s_rank[e] == max_rank) {
                        mass[0] = mass[0] + s_acc[e][0] / s_acc[e][3];
                        mass[1] = mass[1] + s_acc[e][1] / s_acc[e][3];
                        mass[2] = mass[2] + s_acc[e][2] / s_acc[e][3];
                        mass[3] = mass[3] + 1.0;
                    }
                }
                if (P.qef_f32c) {
                    qef::AccF2 af;
                    qef::f2_zero(af);
                    for (int e = 0; e < ec; ++e) {
                        if (s_rank[e] < 0) continue;
                        qef::f2_add_acc(af, s_acc[e]);
                    }
                    qef::f2_widen(af, sum);
                } else {
                    for (int e = 0; e < ec; ++e) {
                        if (s_rank[e] < 0) continue;
                        for (int q = 4; q < 14; ++q) sum[q] = sum[q] + s_acc[e][q];
                    }
                }
                double ata[3][3];
                qef::unpack_ata(sum, ata);
                int rk = 0;
                double vert[3];
                qef::find_vertex(ata, sum + 10, sum[13], mass, vert, &rk);
                for (int q = 0; q < 3; ++q) leaf->vert[patch][q] = vert[q];
                for (int q = 0; q < 4; ++q) leaf->mass[q] = mass[q];
                for (int q = 0; q < 6; ++q) leaf->ata[q] = sum[4 + q];
                for (int q = 0; q < 3; ++q) leaf->atb[q] = sum[10 + q];
                leaf->btb = sum[13];
                rank = uint32_t(rk);
            }
            ++vcount;
            wave_sync();
        }
        if (lane == 0) {
            leaf->rank = rank;
            leaf->vcount = vcount;
            leaf->mask = mask;
            leaf->flags = c_manifold[mask] ? 1u : 0u;
            l0.type[cidx] = AMBIGUOUS;
            l0.child[cidx] = -1;
            l0.tape[cidx] = tape_id;
            for (int a = 0; a < 3; ++a) l0.key[3 * cidx + a] = key[a];
        }
    } while (false);
}

#define LF_LEAF_LAUNCH(NV, FEAT, grid)                                             \
    hipLaunchKernelGGL((k_leaf<NV, FEAT>), dim3(grid), dim3(64), 0, s, p, groups,    \
                       ngroups, per_group, l1, l0, work, deferred, ws, ws_bytes)

hipError_t launch_leaf(const Params& p, const uint32_t* groups, uint32_t ngroups,
                       uint32_t per_group, const LevelView& l1, const LevelView& l0,
                       uint32_t* deferred, hipStream_t s) {
    if (ngroups == 0) return hipSuccess;
    const uint32_t grid = ngroups * per_group;
    const uint32_t* work = nullptr;
    uint8_t* ws = nullptr;
    const size_t ws_bytes = 0;
    switch (banks(p.nregs)) {
        case 1: LF_LEAF_LAUNCH(1, false, grid); break;
        case 2: LF_LEAF_LAUNCH(2, false, grid); break;
        case 4: LF_LEAF_LAUNCH(4, false, grid); break;
        case 8: LF_LEAF_LAUNCH(8, false, grid); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_leaf_features(const Params& p, const uint32_t* groups,
                                uint32_t ngroups, uint32_t per_group,
                                const LevelView& l1, const LevelView& l0,
                                const uint32_t* work, uint32_t nwork, uint8_t* ws,
                                size_t ws_bytes, hipStream_t s) {
    if (nwork == 0) return hipSuccess;
    uint32_t* deferred = nullptr;
    switch (banks(p.nregs)) {
        case 1: LF_LEAF_LAUNCH(1, true, nwork); break;
        case 2: LF_LEAF_LAUNCH(2, true, nwork); break;
        case 4: LF_LEAF_LAUNCH(4, true, nwork); break;
        case 8: LF_LEAF_LAUNCH(8, true, nwork); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

size_t leaf_feature_workspace(uint32_t max_len, uint32_t nregs) {
    return (feat::workspace_bytes(max_len, nregs) + 255) & ~size_t(255);
}

////////////////////////////////////////////////////////////////////////////////
// Phase-split leaf pass (see mesh_kernels.hpp)

// The children of a level-1 parent tile its box as a 2 x 2 x 2 block; their
// corners are the 27 points of a lattice, point (x, y, z) (each 0..2) with
// index x + 3y + 9z, at coordinate lower / midpoint / upper of the parent
// along each axis (the midpoint is the one Region::subdivide computes, so
// the coordinates are the children's own bounds).  A lattice edge joins a
// point and its neighbour along one axis: id = 18 axis + 9 c + 3 o1 + o2,
// with c (0 or 1) the lower point's coordinate along the axis and o1, o2
// the other two coordinates in x, y, z order.
__device__ __forceinline__ int lattice_point(uint32_t oct, int corner) {
    const int x = int(oct & 1u) + (corner & 1);
    const int y = int((oct >> 1) & 1u) + ((corner >> 1) & 1);
    const int z = int((oct >> 2) & 1u) + ((corner >> 2) & 1);
    return x + 3 * y + 9 * z;
}

__device__ __forceinline__ int lattice_edge(int pa, int pb) {
    const int lo = pa < pb ? pa : pb;
    const int d = pa < pb ? pb - pa : pa - pb;       // 1, 3 or 9
    const int axis = (d == 1) ? 0 : (d == 3) ? 1 : 2;
    const int x = lo % 3, y = (lo / 3) % 3, z = lo / 9;
    const int c = (axis == 0) ? x : (axis == 1) ? y : z;
    const int o1 = (axis == 0) ? y : x;
    const int o2 = (axis == 2) ? y : z;
    return axis * 18 + c * 9 + o1 * 3 + o2;
}

// The lattice points of edge `id`, lower first.
__device__ __forceinline__ void edge_points(int id, int* pa, int* pb) {
    const int axis = id / 18, r = id % 18;
    const int c = r / 9, o1 = (r / 3) % 3, o2 = r % 3;
    int x, y, z;
    if (axis == 0) { x = c; y = o1; z = o2; }
    else if (axis == 1) { x = o1; y = c; z = o2; }
    else { x = o1; y = o2; z = c; }
    *pa = x + 3 * y + 9 * z;
    *pb = *pa + ((axis == 0) ? 1 : (axis == 1) ? 3 : 9);
}

__device__ __forceinline__ double lattice_coord(const double lo[3], const double hi[3],
                                                int a, int i) {
    return (i == 0) ? lo[a] : (i == 2) ? hi[a] : (lo[a] + hi[a]) / 2;
}

__device__ __forceinline__ void lattice_pos(const double lo[3], const double hi[3],
                                            int li, double out[3]) {
    out[0] = lattice_coord(lo, hi, 0, li % 3);
    out[1] = lattice_coord(lo, hi, 1, (li / 3) % 3);
    out[2] = lattice_coord(lo, hi, 2, li / 9);
}

// Lists the parent's 8 cells for the feature pass, once.
__device__ __forceinline__ void defer_group(const Params& P, LeafGroupRec* grec,
                                            uint32_t g, uint32_t* deferred) {
    if ((atomicOr(&grec[g].flags, 1u) & 1u) == 0u) {
        const uint32_t b = atomicAdd(&P.counters->deferred, 8u);
        for (uint32_t k = 0; k < 8u; ++k) deferred[b + k] = 8u * g + k;
    }
}

// Corners (dc_tree.inl:131-239 for the 8 children at once).  WG waves per
// workgroup, one parent per wave (a launch schedule: the results do not
// depend on it).
template <int NV, bool FAST, int WG>
__global__ __launch_bounds__(64 * WG)
void k_leaf_corners(Params P, const uint32_t* __restrict__ groups, uint32_t g0,
                    uint32_t ngroups, uint32_t nall, LevelView l1, LevelView l0,
                    LeafGroupRec* __restrict__ grec, uint32_t* __restrict__ cnt,
                    uint32_t* __restrict__ deferred)
{
    // This launch covers groups g0 .. g0 + ngroups - 1 of the level (nall
    // in all).  cnt[g], cnt[nall + g], cnt[2 nall + g]: the group's lattice
    // edges, edge-search tasks and ambiguous leaves, turned into list
    // offsets by scans and k_leaf_scatter (no atomics on shared counters:
    // a million waves adding to three addresses serialise there).
    const int lane = int(threadIdx.x & 63u);
    const uint32_t task = blockIdx.x * uint32_t(WG) +
                          __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (task >= ngroups) return;
    const uint32_t g = g0 + task;
    const uint32_t p = groups[g];
    const uint32_t tape_id = __builtin_amdgcn_readfirstlane(l1.tape[p]);
    const TapeRec tr = P.tapes[tape_id];
    const uint32_t n = __builtin_amdgcn_readfirstlane(tr.len);
    const Clause* __restrict__ t = P.pool + __builtin_amdgcn_readfirstlane(tr.off);
    const Params Q = model_view(P, __builtin_amdgcn_readfirstlane(tr.model));
    const uint32_t root = Q.root_reg;
    uint32_t pkey[3];
    for (int a = 0; a < 3; ++a) pkey[a] = l1.key[3 * p + a];
    double plo[3], phi[3];
    cell_bounds(Q, pkey, 1, plo, phi);

    // Lanes 0..26: one lattice point each.
    const bool pt = lane < 27;
    double pd[3];
    lattice_pos(plo, phi, pt ? lane : 0, pd);
    const float cp[3] = {float(pd[0]), float(pd[1]), float(pd[2])};
    uint32_t fl;
    const float v = pe_value<NV, FAST>(t, n, root, cp[0], cp[1], cp[2], &fl);
    float gr[3] = {0.0f, 0.0f, 0.0f};
    if (ballot(pt && v == 0.0f && !(fl & PT_AMBIGUOUS))) {
        uint32_t fl2 = 0;
        (void)pe_grad3<NV, FAST>(t, n, root, cp[0], cp[1], cp[2], gr, &fl2);
        fl |= fl2 & PT_SLOW;
    }
    bool inside;
    if (v > 0.0f || fmath::isnan_f(v)) inside = false;
    else if (v < 0.0f) inside = true;
    else inside = (gr[0] != 0.0f || gr[1] != 0.0f || gr[2] != 0.0f);
    const uint32_t in27 = uint32_t(ballot(pt && inside)) & 0x7ffffffu;
    if (ballot(pt && ((v == 0.0f && (fl & PT_AMBIGUOUS)) || (fl & PT_SLOW)))) {
        // FeatureEvaluator::isInside needed (or an operation the fast
        // evaluator leaves out): the feature pass does the parent's 8
        // cells.
        if (lane == 0) {
            LeafGroupRec r;
            r.edge_mask = 0;
            r.edge_base = 0;
            r.inside = in27;
            r.flags = grec[g].flags & 1u;    // listed already by an earlier run
            r.pad = 0;
            grec[g] = r;
            cnt[g] = 0;
            cnt[nall + g] = 0;
            cnt[2 * nall + g] = 0;
            __threadfence();
            defer_group(P, grec, g, deferred);
        }
        return;
    }

    // Lanes 0..7: one child each.
    const uint32_t oct = uint32_t(lane & 7);
    const uint32_t first = uint32_t(l1.child[p]);
    uint32_t mask = 0;
    for (int c = 0; c < 8; ++c) mask |= ((in27 >> lattice_point(oct, c)) & 1u) << c;
    const bool amb = lane < 8 && mask != 0u && mask != 0xffu;
    uint64_t emask = 0;
    if (lane < 8) {
        const uint32_t cidx = first + oct;
        l0.child[cidx] = -1;
        l0.tape[cidx] = tape_id;
        for (int a = 0; a < 3; ++a) l0.key[3 * cidx + a] = pkey[a] + ((oct >> a) & 1u);
        if (!amb) {
            l0.type[cidx] = mask ? FILLED : EMPTY;
        } else {
            l0.type[cidx] = AMBIGUOUS;
            LeafRec* leaf = l0.leaf + cidx;
            leaf->mask = mask;
            leaf->flags = c_manifold[mask] ? 1u : 0u;
            for (int patch = 0; patch < 4; ++patch) {
                if (mv(mask, patch, 0, 0) == -1) break;
                for (int k = 0; k < 12; ++k) {
                    const int a = mv(mask, patch, k, 0);
                    if (a == -1) break;
                    const int b = mv(mask, patch, k, 1);
                    emask |= uint64_t(1) << lattice_edge(lattice_point(oct, a),
                                                         lattice_point(oct, b));
                }
            }
        }
    }
    // OR over the 8 child lanes.
    for (int d = 1; d < 8; d <<= 1) {
        const uint32_t lo = __shfl_xor(uint32_t(emask), d);
        const uint32_t hi = __shfl_xor(uint32_t(emask >> 32), d);
        emask |= uint64_t(lo) | (uint64_t(hi) << 32);
    }
    const uint32_t ambm = uint32_t(ballot(amb)) & 0xffu;
    if (lane == 0) {
        const uint32_t ne = uint32_t(__popcll(emask));
        LeafGroupRec r;
        r.edge_mask = emask;
        r.edge_base = 0;             // set by k_leaf_scatter
        r.inside = in27;
        r.flags = 0;
        r.pad = ambm;
        grec[g] = r;
        cnt[g] = ne;
        cnt[nall + g] = (ne + 3u) / 4u;
        cnt[2 * nall + g] = uint32_t(__popc(ambm));
    }
}

// Offsets from the scanned counts: each group's first edge record, its
// edge-search tasks and its ambiguous leaves (one thread per group).  The
// thread of the last group stores the totals in the counters.
__global__ __launch_bounds__(256)
void k_leaf_scatter(Params P, uint32_t nall, LeafGroupRec* __restrict__ grec,
                    const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ base,
                    uint64_t* __restrict__ tasks, uint32_t* __restrict__ leaves)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nall) return;
    const uint32_t eb = base[g], tb = base[nall + g], lb = base[2 * nall + g];
    const uint32_t nt = cnt[nall + g];
    grec[g].edge_base = eb;
    for (uint32_t j = 0; j < nt; ++j) tasks[tb + j] = (uint64_t(g) << 32) | j;
    uint32_t ambm = cnt[2 * nall + g] ? grec[g].pad : 0u;
    for (uint32_t k = 0; ambm; ++k) {
        const uint32_t oct = uint32_t(__builtin_ctz(ambm));
        ambm &= ambm - 1u;
        leaves[lb + k] = 8u * g + oct;
    }
    if (g == nall - 1) {
        P.counters->edges = eb + cnt[g];
        P.counters->tasks = tb + nt;
        P.counters->leaves = lb + cnt[2 * nall + g];
    }
}

// Edge search, partials and per-edge QEF (dc_tree.inl:339-502 for one
// lattice edge, shared by the children whose patches cross it).
// WG waves per workgroup, one task per wave, each with its own LDS slots
// (a launch schedule: the results do not depend on it).
template <int NV, bool FAST, int WG>
__global__ __launch_bounds__(64 * WG)
void k_leaf_edges(Params P, const uint32_t* __restrict__ groups, LevelView l1,
                  LeafGroupRec* __restrict__ grec, const uint64_t* __restrict__ tasks,
                  uint32_t ntasks, EdgeRec* __restrict__ edges,
                  uint32_t* __restrict__ deferred)
{
    __shared__ double s_tf_w[WG][4][3], s_ts_w[WG][4][3];
    __shared__ float s_v_w[WG][8], s_g_w[WG][8][3];
    __shared__ uint32_t s_fl_w[WG][8];
    const int lane = int(threadIdx.x & 63u);
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t ti = blockIdx.x * uint32_t(WG) + wv;
    if (ti >= ntasks) return;
    double (&s_tf)[4][3] = s_tf_w[wv];
    double (&s_ts)[4][3] = s_ts_w[wv];
    float (&s_v)[8] = s_v_w[wv];
    float (&s_g)[8][3] = s_g_w[wv];
    uint32_t (&s_fl)[8] = s_fl_w[wv];
    const uint64_t task = tasks[ti];
    const uint32_t g = uint32_t(task >> 32), j = uint32_t(task & 0xffffffffu);
    const uint64_t emask = grec[g].edge_mask;
    const uint32_t ebase = grec[g].edge_base;
    const uint32_t inside = grec[g].inside;
    const uint32_t p = groups[g];
    const uint32_t tape_id = __builtin_amdgcn_readfirstlane(l1.tape[p]);
    const TapeRec tr = P.tapes[tape_id];
    const uint32_t n = __builtin_amdgcn_readfirstlane(tr.len);
    const Clause* __restrict__ t = P.pool + __builtin_amdgcn_readfirstlane(tr.off);
    const Params Q = model_view(P, __builtin_amdgcn_readfirstlane(tr.model));
    const uint32_t root = Q.root_reg;
    uint32_t pkey[3];
    for (int a = 0; a < 3; ++a) pkey[a] = l1.key[3 * p + a];
    double plo[3], phi[3];
    cell_bounds(Q, pkey, 1, plo, phi);

    const uint32_t ne = uint32_t(__popcll(emask));
    const int e = lane >> 4, smp = lane & 15;
    const uint32_t r = 4u * j + uint32_t(e);
    const bool valid = r < ne;
    if (smp == 0) {
        double a[3] = {0.0, 0.0, 0.0}, b[3] = {0.0, 0.0, 0.0};
        if (valid) {
            uint64_t m = emask;
            for (uint32_t k = 0; k < r; ++k) m &= m - 1;
            int pa, pb;
            edge_points(__builtin_ctzll(m), &pa, &pb);
            const bool a_in = (inside >> pa) & 1u;
            lattice_pos(plo, phi, a_in ? pa : pb, a);      // inside end
            lattice_pos(plo, phi, a_in ? pb : pa, b);      // outside end
        }
        for (int q = 0; q < 3; ++q) {
            s_tf[e][q] = a[q];
            s_ts[e][q] = b[q];
        }
    }
    wave_sync();
    // 4 rounds of 16 samples (dc_tree.inl:339-410).
    const double frac = smp / 15.0;
    for (int round = 0; round < 4; ++round) {
        double ps[3];
        for (int q = 0; q < 3; ++q) ps[q] = s_tf[e][q] * (1 - frac) + s_ts[e][q] * frac;
        uint32_t vf;
        const float sv = pe_value<NV, FAST>(t, n, root, float(ps[0]), float(ps[1]),
                                            float(ps[2]), &vf);
        const bool zero = valid && smp >= 1 && sv == 0.0f;
        if (ballot((zero && (vf & PT_TIE)) || (valid && (vf & PT_SLOW)))) {
            if (lane == 0) defer_group(P, grec, g, deferred);
            return;
        }
        bool zero_out = false;
        if (ballot(zero)) {
            float sg[3];
            uint32_t f2 = 0;
            pe_grad3<NV, FAST>(t, n, root, float(ps[0]), float(ps[1]), float(ps[2]), sg, &f2);
            if (ballot(zero && (f2 & PT_SLOW))) {
                if (lane == 0) defer_group(P, grec, g, deferred);
                return;
            }
            const float nrm = sqrtf(sg[0] * sg[0] + (sg[1] * sg[1] + sg[2] * sg[2]));
            if (zero) zero_out = !(nrm > 0.0f);
        }
        const bool take = valid && smp >= 1 &&
            (sv > 0.0f || zero_out || (smp == 15 && !(sv == 0.0f)));
        const uint64_t tm = ballot(take);
        const uint32_t seg = uint32_t(tm >> (e * 16)) & 0xffffu;
        const int first = seg ? __builtin_ctz(seg) : -1;
        wave_sync();
        if (valid && first >= 1) {
            if (smp == first) {
                for (int q = 0; q < 3; ++q) s_ts[e][q] = ps[q];
            } else if (smp == first - 1) {
                for (int q = 0; q < 3; ++q) s_tf[e][q] = ps[q];
            }
        }
        wave_sync();
    }
    // Partials at both bracket ends (dc_tree.inl:412-471), lanes 0..23:
    // edge lane / 6, end (lane / 3) & 1, axis lane % 3.
    {
        const int de = lane / 6, end = (lane / 3) & 1, axis = lane % 3;
        const bool dv = lane < 24 && 4u * j + uint32_t(de) < ne;
        const int ee = dv ? de : 0;
        float ep[3];
        for (int q = 0; q < 3; ++q) ep[q] = float(end ? s_ts[ee][q] : s_tf[ee][q]);
        float ga;
        uint32_t ef;
        const float ev = pe_deriv1<NV, FAST>(t, n, root, ep[0], ep[1], ep[2], axis, &ga, &ef);
        if (dv) {
            s_g[2 * de + end][axis] = ga;
            if (axis == 0) {
                s_v[2 * de + end] = ev;
                s_fl[2 * de + end] = ef;
            }
        }
    }
    wave_sync();
    const bool end_ok = lane < 8 && 4u * j + uint32_t(lane >> 1) < ne;
    if (ballot(end_ok && (s_fl[lane & 7] & (PT_AMBIGUOUS | PT_SLOW)))) {
        if (lane == 0) defer_group(P, grec, g, deferred);
        return;
    }
    if (lane < 4 && 4u * j + uint32_t(lane) < ne) {
        EdgeRec out;
        if (P.qef_f32c) {
            qef::AccF2 a;
            qef::f2_zero(a);
            for (int end = 0; end < 2; ++end) {
                const int k = 2 * lane + end;
                float pos[3];
                for (int q = 0; q < 3; ++q) pos[q] = float(end ? s_ts[lane][q] : s_tf[lane][q]);
                qef::push_sample_f2(a, pos, s_g[k], s_v[k]);
            }
            qef::f2_widen(a, out.acc);
        } else {
            for (int q = 0; q < 14; ++q) out.acc[q] = 0.0;
            for (int end = 0; end < 2; ++end) {
                const int k = 2 * lane + end;
                double pos[3];
                for (int q = 0; q < 3; ++q) pos[q] = end ? s_ts[lane][q] : s_tf[lane][q];
                const double d[3] = {double(s_g[k][0]), double(s_g[k][1]), double(s_g[k][2])};
                qef::push_sample(out.acc, pos, d, double(s_v[k]));
            }
        }
        double ata[3][3];
        qef::unpack_ata(out.acc, ata);
        out.rank = qef::small_eigenvalue_count(ata);
        out.pad = 0;
        edges[ebase + 4u * j + uint32_t(lane)] = out;
    }
}

// Vertex solve per patch (dc_tree.inl:481-552, 740-813), one lane per
// ambiguous leaf.
__global__ __launch_bounds__(64)
void k_leaf_solve(Params P, const uint32_t* __restrict__ groups, LevelView l1,
                  LevelView l0, const LeafGroupRec* __restrict__ grec,
                  const uint32_t* __restrict__ leaves, uint32_t nleaves,
                  const EdgeRec* __restrict__ edges)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nleaves) return;
    const uint32_t w = leaves[i];
    const uint32_t g = w >> 3, oct = w & 7u;
    const LeafGroupRec gr = grec[g];
    if (gr.flags & 1u) return;              // redone by the feature pass
    const uint32_t cidx = uint32_t(l1.child[groups[g]]) + oct;
    LeafRec* leaf = l0.leaf + cidx;
    const uint32_t mask = leaf->mask;
    uint32_t vcount = 0, rank = 0;
    for (int patch = 0; patch < 4; ++patch) {
        if (mv(mask, patch, 0, 0) == -1) break;
        int ec = 0;
        while (ec < 12 && mv(mask, patch, ec, 0) != -1) ++ec;
        uint32_t idx[12];
        int max_rank = 0;
        for (int e = 0; e < ec; ++e) {
            const int le = lattice_edge(lattice_point(oct, mv(mask, patch, e, 0)),
                                        lattice_point(oct, mv(mask, patch, e, 1)));
            idx[e] = gr.edge_base +
                     uint32_t(__popcll(gr.edge_mask & ((uint64_t(1) << le) - 1)));
            const int rk = edges[idx[e]].rank;
            max_rank = rk > max_rank ? rk : max_rank;
        }
        double mass[4] = {0.0, 0.0, 0.0, 0.0};
        double sum[14];
        for (int q = 0; q < 14; ++q) sum[q] = 0.0;
        for (int e = 0; e < ec; ++e) {
            const EdgeRec& er = edges[idx[e]];
            if (er.rank < 0) continue;
            if (er.rank == max_rank) {
                mass[0] = mass[0] + er.acc[0] / er.acc[3];
                mass[1] = mass[1] + er.acc[1] / er.acc[3];
                mass[2] = mass[2] + er.acc[2] / er.acc[3];
                mass[3] = mass[3] + 1.0;
            }
        }
        if (P.qef_f32c) {
            qef::AccF2 a;
            qef::f2_zero(a);
            for (int e = 0; e < ec; ++e) {
                const EdgeRec& er = edges[idx[e]];
                if (er.rank < 0) continue;
                qef::f2_add_acc(a, er.acc);
            }
            qef::f2_widen(a, sum);
        } else {
            for (int e = 0; e < ec; ++e) {
                const EdgeRec& er = edges[idx[e]];
                if (er.rank < 0) continue;
                for (int q = 4; q < 14; ++q) sum[q] = sum[q] + er.acc[q];
            }
        }
        double ata[3][3];
        qef::unpack_ata(sum, ata);
        int rk = 0;
        double vert[3];
        qef::find_vertex(ata, sum + 10, sum[13], mass, vert, &rk);
        for (int q = 0; q < 3; ++q) leaf->vert[patch][q] = vert[q];
        for (int q = 0; q < 4; ++q) leaf->mass[q] = mass[q];
        for (int q = 0; q < 6; ++q) leaf->ata[q] = sum[4 + q];
        for (int q = 0; q < 3; ++q) leaf->atb[q] = sum[10 + q];
        leaf->btb = sum[13];
        rank = uint32_t(rk);
        ++vcount;
    }
    leaf->rank = rank;
    leaf->vcount = vcount;
}

hipError_t launch_leaf_corners(const Params& p, bool fast, const uint32_t* groups,
                               uint32_t g0, uint32_t ngroups, uint32_t nall,
                               const LevelView& l1, const LevelView& l0, LeafGroupRec* grec,
                               uint32_t* cnt, uint32_t* deferred, hipStream_t s, int waves) {
    if (ngroups == 0) return hipSuccess;
    if (waves != 1 && waves != 4) return hipErrorInvalidValue;
    const uint32_t nb = (ngroups + uint32_t(waves) - 1) / uint32_t(waves);
#define LF_CORNERS(NV, F)                                                                      \
    do {                                                                                       \
        if (waves == 4) {                                                                      \
            hipLaunchKernelGGL((k_leaf_corners<NV, F, 4>), dim3(nb), dim3(256), 0, s, p,       \
                               groups, g0, ngroups, nall, l1, l0, grec, cnt, deferred);        \
        } else {                                                                               \
            hipLaunchKernelGGL((k_leaf_corners<NV, F, 1>), dim3(nb), dim3(64), 0, s, p,        \
                               groups, g0, ngroups, nall, l1, l0, grec, cnt, deferred);        \
        }                                                                                      \
    } while (0)
    switch (banks(p.nregs) * (fast ? -1 : 1)) {
        case 1: LF_CORNERS(1, false); break;
        case 2: LF_CORNERS(2, false); break;
        case 4: LF_CORNERS(4, false); break;
        case -1: LF_CORNERS(1, true); break;
        case -2: LF_CORNERS(2, true); break;
        case -4: LF_CORNERS(4, true); break;
        case 8:
        case -8:                                  // calls cost nothing next to spills
            hipLaunchKernelGGL((k_leaf_corners<8, false, 1>), dim3(ngroups), dim3(64), 0, s, p,
                               groups, g0, ngroups, nall, l1, l0, grec, cnt, deferred);
            break;
        default: return hipErrorInvalidValue;
    }
#undef LF_CORNERS
    return hipGetLastError();
}

hipError_t launch_leaf_scatter(const Params& p, uint32_t nall, LeafGroupRec* grec,
                               const uint32_t* cnt, const uint32_t* base, uint64_t* tasks,
                               uint32_t* leaves, hipStream_t s) {
    if (nall == 0) return hipSuccess;
    hipLaunchKernelGGL(k_leaf_scatter, dim3((nall + 255) / 256), dim3(256), 0, s, p, nall,
                       grec, cnt, base, tasks, leaves);
    return hipGetLastError();
}

hipError_t launch_leaf_edges(const Params& p, bool fast, const uint32_t* groups,
                             const LevelView& l1, LeafGroupRec* grec,
                             const uint64_t* tasks, uint32_t ntasks,
                             EdgeRec* edges, uint32_t* deferred, hipStream_t s, int waves) {
    if (ntasks == 0) return hipSuccess;
    if (waves != 1 && waves != 4) return hipErrorInvalidValue;
    const uint32_t nb = (ntasks + uint32_t(waves) - 1) / uint32_t(waves);
#define LF_EDGES(NV, F)                                                                        \
    do {                                                                                       \
        if (waves == 4) {                                                                      \
            hipLaunchKernelGGL((k_leaf_edges<NV, F, 4>), dim3(nb), dim3(256), 0, s, p, groups, \
                               l1, grec, tasks, ntasks, edges, deferred);                      \
        } else {                                                                               \
            hipLaunchKernelGGL((k_leaf_edges<NV, F, 1>), dim3(nb), dim3(64), 0, s, p, groups,  \
                               l1, grec, tasks, ntasks, edges, deferred);                      \
        }                                                                                      \
    } while (0)
    switch (banks(p.nregs) * (fast ? -1 : 1)) {
        case 1: LF_EDGES(1, false); break;
        case 2: LF_EDGES(2, false); break;
        case 4: LF_EDGES(4, false); break;
        case -1: LF_EDGES(1, true); break;
        case -2: LF_EDGES(2, true); break;
        case -4: LF_EDGES(4, true); break;
        case 8:
        case -8:
            hipLaunchKernelGGL((k_leaf_edges<8, false, 1>), dim3(ntasks), dim3(64), 0, s, p, groups,
                               l1, grec, tasks, ntasks, edges, deferred);
            break;
        default: return hipErrorInvalidValue;
    }
#undef LF_EDGES
    return hipGetLastError();
}

hipError_t launch_leaf_solve(const Params& p, const uint32_t* groups,
                             const LevelView& l1, const LevelView& l0,
                             const LeafGroupRec* grec, const uint32_t* leaves,
                             uint32_t nleaves, const EdgeRec* edges,
                             hipStream_t s) {
    if (nleaves == 0) return hipSuccess;
    hipLaunchKernelGGL(k_leaf_solve, dim3((nleaves + 63) / 64), dim3(64), 0, s, p, groups,
                       l1, l0, grec, leaves, nleaves, edges);
    return hipGetLastError();
}

////////////////////////////////////////////////////////////////////////////////
// Collapse (collectChildren, dc_tree.inl:581-735)

// Tape::getBase (tape.cpp:168-205) for a region and for a point.
__device__ uint32_t base_for_region(const Params& P, uint32_t id, const double lo[3],
                                    const double hi[3]) {
    // A tape's parent always has a smaller id, so the walk ends at the
    // base tape; the step limit only guards against corrupted records.
    for (int step = 0; step < 64 && P.tapes[id].parent != kNoTape; ++step) {
        const TapeRec& r = P.tapes[id];
        if ((r.flags & TAPE_INTERVAL) &&
            lo[0] >= double(r.box[0]) && hi[0] <= double(r.box[1]) &&
            lo[1] >= double(r.box[2]) && hi[1] <= double(r.box[3]) &&
            lo[2] >= double(r.box[4]) && hi[2] <= double(r.box[5])) {
            break;
        }
        id = r.parent;
    }
    return id;
}

__device__ __forceinline__ uint8_t corner_state(const LevelView& lv, uint32_t c, int i) {
    const uint8_t ty = lv.type[c];
    if (ty == AMBIGUOUS) return ((lv.leaf[c].mask >> i) & 1u) ? FILLED : EMPTY;
    return ty;
}

template <int NV>
__global__ __launch_bounds__(64)
void k_collapse(Params P, int level, LevelView cur, LevelView below)
{
    const uint32_t cell = blockIdx.x;
    if (cell >= cur.n) return;
    if (cur.type[cell] != AMBIGUOUS || cur.child[cell] < 0) return;
    const uint32_t cb = uint32_t(cur.child[cell]);
    for (int i = 0; i < 8; ++i) {
        if (below.child[cb + i] >= 0) return;       // a branch child
    }
    uint8_t corners[8];
    bool all_empty = true, all_full = true;
    for (int i = 0; i < 8; ++i) {
        corners[i] = corner_state(below, cb + i, i);
        const uint8_t ty = below.type[cb + i];
        all_empty &= (ty == EMPTY);
        all_full &= (ty == FILLED);
    }
    if (all_empty || all_full) {
        if (threadIdx.x == 0) {
            cur.type[cell] = all_empty ? EMPTY : FILLED;
            cur.child[cell] = -1;
        }
        return;
    }
    uint32_t mask = 0;
    for (int i = 0; i < 8; ++i) mask |= uint32_t(corners[i] == FILLED) << i;
    bool manifold = c_manifold[mask] != 0;
    for (int i = 0; i < 8 && manifold; ++i) {
        if (below.type[cb + i] == AMBIGUOUS) manifold = (below.leaf[cb + i].flags & 1u) != 0;
    }
    if (manifold) {
        // leafsAreManifold (dc_tree3.cpp:17-109)
        auto cs = [&](int c, int i) { return corner_state(below, cb + c, i); };
        auto e2 = [&](uint8_t s, int a, int b) { return s == corners[a] || s == corners[b]; };
        auto e4 = [&](uint8_t s, int a, int b, int c, int d) {
            return s == corners[a] || s == corners[b] || s == corners[c] || s == corners[d];
        };
        const int X = 1, Y = 2, Z = 4;
        const bool edges_safe =
            e2(cs(0, Z), 0, Z) && e2(cs(0, X), 0, X) && e2(cs(0, Y), 0, Y) &&
            e2(cs(X, X | Y), X, X | Y) && e2(cs(X, X | Z), X, X | Z) &&
            e2(cs(Y, Y | X), Y, Y | X) && e2(cs(Y, Y | Z), Y, Y | Z) &&
            e2(cs(X | Y, X | Y | Z), X | Y, X | Y | Z) &&
            e2(cs(Z, Z | X), Z, Z | X) && e2(cs(Z, Z | Y), Z, Z | Y) &&
            e2(cs(Z | X, Z | X | Y), Z | X, Z | X | Y) &&
            e2(cs(Z | Y, Z | Y | X), Z | Y, Z | Y | X);
        const bool faces_safe =
            e4(cs(0, X | Z), 0, X, Z, X | Z) &&
            e4(cs(0, Y | Z), 0, Y, Z, Y | Z) &&
            e4(cs(0, Y | X), 0, Y, X, Y | X) &&
            e4(cs(X | Y | Z, X), X, X | Z, X | Y, X | Y | Z) &&
            e4(cs(X | Y | Z, Y), Y, Y | Z, Y | X, X | Y | Z) &&
            e4(cs(X | Y | Z, Z), Z, Z | Y, Z | X, X | Y | Z);
        const uint8_t ctr = cs(0, X | Y | Z);
        bool center_safe = false;
        for (int i = 0; i < 8; ++i) center_safe |= (ctr == corners[i]);
        manifold = edges_safe && faces_safe && center_safe;
    }
    if (!manifold) return;

    uint32_t rank = 0;
    for (int i = 0; i < 8; ++i) {
        if (below.type[cb + i] == AMBIGUOUS) {
            const uint32_t r = below.leaf[cb + i].rank;
            rank = r > rank ? r : rank;
        }
    }
    double mass[4] = {0.0, 0.0, 0.0, 0.0};
    double sum[14];
    for (int q = 0; q < 14; ++q) sum[q] = 0.0;
    for (int i = 0; i < 8; ++i) {
        if (below.type[cb + i] != AMBIGUOUS) continue;
        const LeafRec& l = below.leaf[cb + i];
        if (l.rank == rank) {
            for (int q = 0; q < 4; ++q) mass[q] = mass[q] + l.mass[q];
        }
        for (int q = 0; q < 6; ++q) sum[4 + q] = sum[4 + q] + l.ata[q];
        for (int q = 0; q < 3; ++q) sum[10 + q] = sum[10 + q] + l.atb[q];
        sum[13] = sum[13] + l.btb;
    }
    double ata[3][3];
    qef::unpack_ata(sum, ata);
    double vert[3];
    const double err = qef::find_vertex(ata, sum + 10, sum[13], mass, vert, nullptr);
    uint32_t key[3];
    for (int a = 0; a < 3; ++a) key[a] = cur.key[3 * cell + a];
    const Params Q = model_view(P, __builtin_amdgcn_readfirstlane(P.tapes[cur.tape[cell]].model));
    double lo[3], hi[3];
    cell_bounds(Q, key, level, lo, hi);
    const double eps = 1e-6;
    const bool inside = vert[0] >= lo[0] - eps && vert[1] >= lo[1] - eps &&
                        vert[2] >= lo[2] - eps && vert[0] <= hi[0] + eps &&
                        vert[1] <= hi[1] + eps && vert[2] <= hi[2] + eps;
    if (!(err < P.max_err && inside)) return;
    const float vf[3] = {float(vert[0]), float(vert[1]), float(vert[2])};
    const double vd[3] = {vf[0], vf[1], vf[2]};
    uint32_t tid = base_for_region(P, cur.tape[cell], lo, hi);
    tid = __builtin_amdgcn_readfirstlane(base_for_region(P, tid, vd, vd));
    const TapeRec tr = P.tapes[tid];
    uint32_t fl;
    const float fv = pe_value<NV>(P.pool + __builtin_amdgcn_readfirstlane(tr.off),
                                  __builtin_amdgcn_readfirstlane(tr.len), Q.root_reg,
                                  vf[0], vf[1], vf[2], &fl);
    if (!(fabsf(fv) < P.max_err)) return;
    if (threadIdx.x == 0) {
        LeafRec& l = cur.leaf[cell];
        for (int q = 0; q < 3; ++q) l.vert[0][q] = vert[q];
        for (int q = 0; q < 4; ++q) l.mass[q] = mass[q];
        for (int q = 0; q < 6; ++q) l.ata[q] = sum[4 + q];
        for (int q = 0; q < 3; ++q) l.atb[q] = sum[10 + q];
        l.btb = sum[13];
        l.rank = rank;
        l.vcount = 1;
        l.mask = mask;
        l.flags = 1u | (uint32_t(level) << 8);
        cur.child[cell] = -1;
    }
}

hipError_t launch_collapse(const Params& p, int level, const LevelView& cur,
                           const LevelView& below, hipStream_t s) {
    if (cur.n == 0) return hipSuccess;
    switch (banks(p.nregs)) {
        case 1: hipLaunchKernelGGL(k_collapse<1>, dim3(cur.n), dim3(64), 0, s, p, level, cur, below); break;
        case 2: hipLaunchKernelGGL(k_collapse<2>, dim3(cur.n), dim3(64), 0, s, p, level, cur, below); break;
        case 4: hipLaunchKernelGGL(k_collapse<4>, dim3(cur.n), dim3(64), 0, s, p, level, cur, below); break;
        case 8: hipLaunchKernelGGL(k_collapse<8>, dim3(cur.n), dim3(64), 0, s, p, level, cur, below); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

////////////////////////////////////////////////////////////////////////////////
// Quad emission (Dual<3> walk + DCMesher::load)

// Leaf cell containing a point given in doubled finest-cell units (odd
// coordinates never lie on a cell boundary of a cell larger than the unit).
__device__ bool locate(const Params& P, int64_t px, int64_t py, int64_t pz,
                       int* lvl, uint32_t* idx) {
    const int64_t lim = int64_t(2) << P.top;
    if (px < 0 || py < 0 || pz < 0 || px >= lim || py >= lim || pz >= lim) return false;
    int l = P.top;
    uint32_t i = P.root_cell;
    for (;;) {
        const int32_t ch = P.levels[l].child[i];
        if (ch < 0 || l == 0) break;
        const uint32_t oct = uint32_t((px >> l) & 1) | (uint32_t((py >> l) & 1) << 1) |
                             (uint32_t((pz >> l) & 1) << 2);
        i = uint32_t(ch) + oct;
        --l;
    }
    *lvl = l;
    *idx = i;
    return true;
}

__device__ __forceinline__ int axis_q(int a) { return a == 1 ? 2 : a == 2 ? 4 : 1; }
__device__ __forceinline__ int axis_r(int a) { return a == 1 ? 4 : a == 2 ? 1 : 2; }
__device__ __forceinline__ int bit_index(int a) { return a == 1 ? 0 : a == 2 ? 1 : 2; }

__global__ void k_quads(Params P, LevelView cur, int pass, uint32_t* tri_count,
                        uint32_t* vslot, const uint32_t* tri_off,
                        const uint32_t* vid, uint32_t* tris)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cur.n) return;
    const uint32_t model = P.tapes[cur.tape[c]].model;
    const uint32_t gc = cur.cbase[model] + c;       // global cell number
    if (cur.type[c] != AMBIGUOUS || cur.child[c] >= 0) {
        if (pass == 0) tri_count[gc] = 0;
        return;
    }
    const Params MP = model_view(P, model);
    const int my_level = int((cur.leaf[c].flags >> 8) & 0xffu);
    int64_t key[3];
    for (int a = 0; a < 3; ++a) key[a] = cur.key[3 * c + a];
    const int64_t size = int64_t(1) << my_level;
    {
        // Cells below a collapsed ancestor stay in their level's arrays but
        // are no longer part of the tree: the tree leaf at this cell's
        // centre must be the cell itself.
        int cl;
        uint32_t ci;
        const bool found = locate(MP, 2 * key[0] + size, 2 * key[1] + size,
                                  2 * key[2] + size, &cl, &ci);
        if (!found || cl != my_level || ci != c) {
            if (pass == 0) tri_count[gc] = 0;
            return;
        }
    }
    uint32_t count = 0;
    uint32_t out = (pass == 1) ? tri_off[gc] : 0u;
    // vertex numbers are per model, after the model's dummy vertex 0
    const uint32_t vbase = (pass == 1) ? vid[MP.slot0] : 0u;
    for (int A = 1; A <= 4; A <<= 1) {
        const int Q = axis_q(A), R = axis_r(A);
        const int ia = bit_index(A), iq = bit_index(Q), ir = bit_index(R);
        for (int qr = 0; qr < 4; ++qr) {
            const int q = qr & 1, r = qr >> 1;
            const int kc = (q ? 0 : 1) | (r ? 0 : 2);
            int lv[4];
            uint32_t ix[4];
            bool ok = true;
            for (int k = 0; k < 4 && ok; ++k) {
                if (k == kc) { lv[k] = my_level; ix[k] = c; continue; }
                int64_t pt[3];
                pt[ia] = 2 * key[ia] + size;
                pt[iq] = 2 * (key[iq] + q * size) + ((k & 1) ? 1 : -1);
                pt[ir] = 2 * (key[ir] + r * size) + ((k & 2) ? 1 : -1);
                ok = locate(MP, pt[0], pt[1], pt[2], &lv[k], &ix[k]);
            }
            if (!ok) continue;
            // The segment belongs to the smallest of the four cells, and
            // among equal ones to the first in walk order.
            bool owner = true;
            for (int k = 0; k < 4; ++k) {
                if (lv[k] < my_level || (k < kc && lv[k] == my_level)) owner = false;
            }
            if (!owner) continue;
            bool all_amb = true;
            for (int k = 0; k < 4; ++k) {
                all_amb &= (P.levels[lv[k]].type[ix[k]] == AMBIGUOUS);
            }
            if (!all_amb) continue;
            const int corners4[4] = {Q | R, R, Q, 0};
            const uint32_t m = cur.leaf[c].mask;
            const bool sa = (m >> corners4[kc]) & 1u;
            const bool sb = (m >> (corners4[kc] | A)) & 1u;
            if (sa == sb) continue;
            const bool D = sa;
            const int ev[4][2] = {{Q | R, Q | R | A}, {R, R | A}, {Q, Q | A}, {0, A}};
            uint32_t slots[4];
            float vp[4][3];
            for (int i = 0; i < 4; ++i) {
                const int es = D ? c_me[ev[i][0] * 8 + ev[i][1]] : c_me[ev[i][1] * 8 + ev[i][0]];
                const LevelView& L = P.levels[lv[i]];
                const LeafRec& lf = L.leaf[ix[i]];
                const int vi = lv[i] > 0 ? 0 : c_mp[lf.mask * 24 + es];
                slots[i] = 4u * (L.cbase[model] + ix[i]) + uint32_t(vi < 0 ? 0 : vi);
                for (int q2 = 0; q2 < 3; ++q2) vp[i][q2] = float(lf.vert[vi < 0 ? 0 : vi][q2]);
            }
            if (!D) {
                uint32_t ts = slots[1]; slots[1] = slots[2]; slots[2] = ts;
                for (int q2 = 0; q2 < 3; ++q2) {
                    const float tv = vp[1][q2]; vp[1][q2] = vp[2][q2]; vp[2][q2] = tv;
                }
            }
            float nrm[4][3];
            const int tri_n[4][3] = {{0, 1, 2}, {1, 3, 0}, {2, 0, 3}, {3, 2, 1}};
            for (int s = 0; s < 4; ++s) {
                const int a0 = tri_n[s][0], b0 = tri_n[s][1], c0 = tri_n[s][2];
                float u[3], w[3];
                for (int q2 = 0; q2 < 3; ++q2) {
                    u[q2] = vp[b0][q2] - vp[a0][q2];
                    w[q2] = vp[c0][q2] - vp[a0][q2];
                }
                float nn[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2],
                               u[0] * w[1] - u[1] * w[0]};
                const float sq = nn[0] * nn[0] + (nn[1] * nn[1] + nn[2] * nn[2]);
                if (sq > 0.0f) {
                    const float len = sqrtf(sq);
                    for (int q2 = 0; q2 < 3; ++q2) nn[q2] = nn[q2] / len;
                }
                for (int q2 = 0; q2 < 3; ++q2) nrm[a0][q2] = nn[q2];
            }
            const float d03 = nrm[0][0] * nrm[3][0] + (nrm[0][1] * nrm[3][1] + nrm[0][2] * nrm[3][2]);
            const float d12 = nrm[1][0] * nrm[2][0] + (nrm[1][1] * nrm[2][1] + nrm[1][2] * nrm[2][2]);
            uint32_t tv[2][3];
            if (d03 > d12) {
                tv[0][0] = slots[0]; tv[0][1] = slots[1]; tv[0][2] = slots[2];
                tv[1][0] = slots[2]; tv[1][1] = slots[1]; tv[1][2] = slots[3];
            } else {
                tv[0][0] = slots[0]; tv[0][1] = slots[1]; tv[0][2] = slots[3];
                tv[1][0] = slots[0]; tv[1][1] = slots[3]; tv[1][2] = slots[2];
            }
            for (int s = 0; s < 2; ++s) {
                if (tv[s][0] == tv[s][1] || tv[s][1] == tv[s][2] || tv[s][0] == tv[s][2]) continue;
                if (pass == 0) {
                    for (int q2 = 0; q2 < 3; ++q2) vslot[tv[s][q2]] = 1u;
                } else {
                    for (int q2 = 0; q2 < 3; ++q2) {
                        tris[3 * (out + count) + q2] = vid[tv[s][q2]] - vbase + 1u;
                    }
                }
                ++count;
            }
        }
    }
    if (pass == 0) tri_count[gc] = count;
}

hipError_t launch_quads(const Params& p, const LevelView& cur, int pass,
                        uint32_t* tri_count, uint32_t* vslot,
                        const uint32_t* tri_off, const uint32_t* vid,
                        uint32_t* tris, hipStream_t s) {
    if (cur.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_quads, dim3((cur.n + 63) / 64), dim3(64), 0, s, p, cur, pass, tri_count, vslot,
                                                         tri_off, vid, tris);
    return hipGetLastError();
}

__global__ void k_verts(Params P, LevelView cur, int level, const uint32_t* vslot,
                        const uint32_t* vid, float* verts, uint32_t* keys)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 4u * cur.n) return;
    const uint32_t c = i >> 2, vi = i & 3u;
    const uint32_t m = P.tapes[cur.tape[c]].model;
    const uint32_t slot = 4u * (cur.cbase[m] + c) + vi;
    if (vi == 0 && P.models != nullptr && slot == P.models[m].slot0) {
        // the model's dummy vertex (and key), before its first vertex
        const uint32_t d = vid[slot] + m;
        for (int q = 0; q < 3; ++q) verts[3 * d + q] = 0.0f;
        if (keys) {
            for (int q = 0; q < 5; ++q) keys[5 * d + q] = 0u;
        }
    }
    if (!vslot[slot]) return;
    const uint32_t id = vid[slot] + 1u + m;
    const LeafRec& l = cur.leaf[c];
    for (int q = 0; q < 3; ++q) verts[3 * id + q] = float(l.vert[vi][q]);
    if (keys) {
        for (int q = 0; q < 3; ++q) keys[5 * id + q] = cur.key[3 * c + q];
        keys[5 * id + 3] = uint32_t(level);
        keys[5 * id + 4] = vi;
    }
}

hipError_t launch_verts(const Params& p, const LevelView& cur, int level,
                        const uint32_t* vslot, const uint32_t* vid, float* verts,
                        uint32_t* keys, hipStream_t s) {
    if (cur.n == 0) return hipSuccess;
    const uint32_t n = 4u * cur.n;
    hipLaunchKernelGGL(k_verts, dim3((n + 255) / 256), dim3(256), 0, s, p, cur, level, vslot, vid,
                       verts, keys);
    return hipGetLastError();
}

__global__ void k_model_bases(const ModelRec* models, uint32_t nmodels,
                              const uint32_t* tri_off, const uint32_t* tri_count,
                              uint32_t ncells, const uint32_t* vid, const uint32_t* vslot,
                              uint32_t nslots, uint32_t* out)
{
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m > nmodels) return;
    if (m < nmodels) {
        const uint32_t s0 = models[m].slot0;
        out[m] = tri_off[s0 / 4u];
        out[nmodels + 1 + m] = vid[s0];
    } else {
        out[m] = tri_off[ncells - 1] + tri_count[ncells - 1];
        out[nmodels + 1 + m] = vid[nslots - 1] + vslot[nslots - 1];
    }
}

hipError_t launch_model_bases(const ModelRec* models, uint32_t nmodels,
                              const uint32_t* tri_off, const uint32_t* tri_count,
                              uint32_t ncells, const uint32_t* vid, const uint32_t* vslot,
                              uint32_t nslots, uint32_t* out, hipStream_t s) {
    if (ncells == 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_model_bases, dim3((nmodels + 256) / 256), dim3(256), 0, s, models,
                       nmodels, tri_off, tri_count, ncells, vid, vslot, nslots, out);
    return hipGetLastError();
}

// See launch_child_ranges.  upper is null at the top level of the batch.
__global__ void k_child_ranges(const ModelRec* models, uint32_t nmodels, int level,
                               const uint32_t* upper, const uint32_t* upper_base,
                               const uint32_t* upper_flags, uint32_t upper_n,
                               uint32_t* ranges)
{
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmodels) return;
    uint32_t b = 0, e = 0;
    if (models[m].top == level) {
        b = models[m].root_cell;
        e = b + 1;
    } else if (upper != nullptr && upper_n > 0) {
        // ambiguous cells of the level above before index i, times 8
        auto before = [&](uint32_t i) {
            return i < upper_n ? upper_base[i]
                               : upper_base[upper_n - 1] + upper_flags[upper_n - 1];
        };
        b = 8u * before(upper[m]);
        e = 8u * before(upper[nmodels + m]);
    }
    ranges[m] = b;
    ranges[nmodels + m] = e;
}

hipError_t launch_child_ranges(const ModelRec* models, uint32_t nmodels, int level,
                               const uint32_t* upper, const uint32_t* upper_base,
                               const uint32_t* upper_flags, uint32_t upper_n,
                               uint32_t* ranges, hipStream_t s) {
    if (nmodels == 0) return hipSuccess;
    hipLaunchKernelGGL(k_child_ranges, dim3((nmodels + 255) / 256), dim3(256), 0, s, models,
                       nmodels, level, upper, upper_base, upper_flags, upper_n, ranges);
    return hipGetLastError();
}

// Exclusive prefix sums: each block scans 1024 values (256 threads x 4)
// and records its total; the totals are scanned the same way and added
// back.  No block waits on another, so every launch finishes.
constexpr uint32_t kScanBlock = 1024;

__global__ __launch_bounds__(256)
void k_scan_blocks(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                   uint32_t* __restrict__ sums, size_t n)
{
    __shared__ uint32_t part[256];
    const uint32_t t = threadIdx.x;
    const size_t base = size_t(blockIdx.x) * kScanBlock + 4u * t;
    uint32_t v[4], tot = 0;
    for (int k = 0; k < 4; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        tot += v[k];
    }
    part[t] = tot;
    __syncthreads();
    for (uint32_t off = 1; off < 256; off <<= 1) {
        const uint32_t x = (t >= off) ? part[t - off] : 0u;
        __syncthreads();
        part[t] += x;
        __syncthreads();
    }
    uint32_t run = t ? part[t - 1] : 0u;
    for (int k = 0; k < 4; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (t == 255) sums[blockIdx.x] = part[255];
}

__global__ __launch_bounds__(256)
void k_scan_add(uint32_t* __restrict__ out, const uint32_t* __restrict__ offs, size_t n)
{
    const size_t base = size_t(blockIdx.x) * kScanBlock + 4u * threadIdx.x;
    const uint32_t o = offs[blockIdx.x];
    for (int k = 0; k < 4; ++k) {
        if (base + k < n) out[base + k] += o;
    }
}

size_t scan_scratch_words(size_t n) {
    size_t words = 0;
    while (n > 1) {
        n = (n + kScanBlock - 1) / kScanBlock;
        words += 2 * n;
    }
    return words + 2;
}

hipError_t exclusive_scan(const uint32_t* in, uint32_t* out, size_t n,
                          uint32_t* scratch, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const size_t nb = (n + kScanBlock - 1) / kScanBlock;
    uint32_t* sums = scratch;
    uint32_t* offs = scratch + nb;
    hipLaunchKernelGGL(k_scan_blocks, dim3(unsigned(nb)), dim3(256), 0, s, in, out, sums, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || nb == 1) return e;
    if ((e = exclusive_scan(sums, offs, nb, scratch + 2 * nb, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_scan_add, dim3(unsigned(nb)), dim3(256), 0, s, out, offs, n);
    return hipGetLastError();
}

}   // namespace mesh
}   // namespace amd
}   // namespace libfive
