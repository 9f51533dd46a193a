// Point evaluation kernels: tape interpreters for values
// (ArrayEvaluator, libfive/src/eval/eval_array.cpp:107-118, 221-369) and for
// values plus gradients (DerivArrayEvaluator,
// libfive/src/eval/eval_deriv_array.cpp:102-256).
//
// Mapping: one point per lane, points stored SoA (x[], y[], z[]) so loads
// and stores are fully coalesced; the tape is wave-uniform and is read with
// scalar loads; the register file lives in VGPRs (regfile.hpp).  The
// reference evaluates 256 points per pass through per-clause rows of a
// slot matrix; here a lane walks the whole tape for its point and no
// intermediate value leaves the register file.
//
// Compile with -ffp-contract=off: the derivative rules are several
// roundings in a fixed order and must not be fused.
#include <hip/hip_runtime.h>

#include "eval_kernels.hpp"
#include "opcode.hpp"
#include "fmath.hpp"
#include "regfile.hpp"
#include "tape.hpp"
#include "clause_ops.hpp"
#include "point_eval.hpp"

namespace libfive {
namespace amd {

using namespace tape;
namespace OC = Opcode;

// Register access for point p (a compile-time constant after unrolling).
#define LF_PGET(p, NV, i)                                                    \
    ((p) == 0 ? LF_RF_GET(r0, NV, i) : (p) == 1 ? LF_RF_GET(r1, NV, i)       \
     : (p) == 2 ? LF_RF_GET(r2, NV, i) : LF_RF_GET(r3, NV, i))
#define LF_PSET(p, NV, i, x)                                                 \
    do {                                                                     \
        if ((p) == 0) LF_RF_SET(r0, NV, i, x);                               \
        else if ((p) == 1) LF_RF_SET(r1, NV, i, x);                          \
        else if ((p) == 2) LF_RF_SET(r2, NV, i, x);                          \
        else LF_RF_SET(r3, NV, i, x);                                        \
    } while (0)

// One clause for P points.
#define LF_VALUE_STEP(C)                                                     \
    do {                                                                     \
        const Clause c_ = (C);                                               \
        const uint32_t o_ = tape::op(c_);                                    \
        if (o_ != OP_NOP) {                                                  \
            const uint32_t m_ = mode(c_);                                    \
            const float im_ = imm(c_);                                       \
            float a_[P], b_[P], q_[P];                                       \
            if (m_ == MODE_A_IMM) {                                          \
                _Pragma("unroll") for (int p = 0; p < P; ++p) a_[p] = im_;   \
            } else {                                                         \
                const uint32_t ia_ = ra(c_);                                 \
                _Pragma("unroll") for (int p = 0; p < P; ++p)                \
                    a_[p] = LF_PGET(p, NV, ia_);                             \
            }                                                                \
            if (!is_binary(c_)) {                                            \
                _Pragma("unroll") for (int p = 0; p < P; ++p) b_[p] = 0.0f;  \
            } else if (m_ == MODE_B_IMM) {                                   \
                _Pragma("unroll") for (int p = 0; p < P; ++p) b_[p] = im_;   \
            } else {                                                         \
                const uint32_t ib_ = rb(c_);                                 \
                _Pragma("unroll") for (int p = 0; p < P; ++p)                \
                    b_[p] = LF_PGET(p, NV, ib_);                             \
            }                                                                \
            apply_value<P, FAST>(o_, a_, b_, im_, q_, amax_);                \
            const uint32_t d_ = dst(c_);                                     \
            _Pragma("unroll") for (int p = 0; p < P; ++p)                    \
                LF_PSET(p, NV, d_, q_[p]);                                   \
        }                                                                    \
    } while (0)

struct Clause4 {
    Clause c[kFetch];
};

// P points per lane: lane l of block b handles points b*256*P + l + 256*j.
// The tape is read kFetch clauses at a time, the next batch requested
// before the current one is interpreted.
// FAST: the tape has only the common opcodes and the kernel calls no
// function; a point with a sin / cos argument of 2^17 or more gets
// tape::kRedoMark, for k_eval_values_redo.
template <int NV, int P, bool FAST>
__global__ __launch_bounds__(256)
void k_eval_values(const Clause* __restrict__ t, uint32_t n, uint32_t root,
                   const float* __restrict__ xs, const float* __restrict__ ys,
                   const float* __restrict__ zs, float* __restrict__ out,
                   size_t npts)
{
    const size_t base = size_t(blockIdx.x) * (256u * P) + threadIdx.x;
    float amax_[P];
#pragma unroll
    for (int p = 0; p < P; ++p) amax_[p] = 0.0f;
    LF_RF_DECL(r0);
    LF_RF_DECL(r1);
    LF_RF_DECL(r2);
    LF_RF_DECL(r3);
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const size_t i = base + 256u * p;
        const bool ok = i < npts;
        LF_PSET(p, NV, kRegX, ok ? xs[i] : 0.0f);
        LF_PSET(p, NV, kRegY, ok ? ys[i] : 0.0f);
        LF_PSET(p, NV, kRegZ, ok ? zs[i] : 0.0f);
    }
    const Clause4* t4 = reinterpret_cast<const Clause4*>(t);
    Clause4 cur = t4[0];
    for (uint32_t k = 0; k < n; k += kFetch) {
        const Clause4 nxt = t4[k / kFetch + 1];
        LF_VALUE_STEP(cur.c[0]);
        LF_VALUE_STEP(cur.c[1]);
        LF_VALUE_STEP(cur.c[2]);
        LF_VALUE_STEP(cur.c[3]);
        cur = nxt;
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const size_t i = base + 256u * p;
        const float v = LF_PGET(p, NV, root);
        if (i < npts) {
            out[i] = (!FAST || amax_[p] < fmath::kCwLimit) ? v
                                                            : fmath::u2f(kRedoMark);
        }
    }
}

// Values and one partial derivative per launch row (blockIdx.y is the
// axis): the register file is two files (values, that partial) instead of
// four, so that it stays in VGPRs with several waves per SIMD; the values
// are computed in every row and written by row 0.  Each component is
// bit-identical to a pass carrying all three partials (the rules of
// eval_deriv_array.cpp:119-256 are per component).
//
// FAST: the tape has only the common opcodes, and the kernel calls no
// function (no stack, no registers held around calls).  For a point it
// cannot evaluate (a sin / cos argument of 2^17 or above; every row sees
// the same values, so every row agrees), row 0 writes tape::kRedoMark as
// the value, and k_eval_deriv1_redo evaluates the marked points again with
// the general rules.  With two banks the fast form fits in 168 VGPRs
// (three waves per SIMD) when asked to, without spilling.

template <int NV, bool FAST>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu((FAST && NV == 2) ? 3 : 1)))
void k_eval_deriv1(const Clause* __restrict__ t, uint32_t n, uint32_t root,
                   const float* __restrict__ xs, const float* __restrict__ ys,
                   const float* __restrict__ zs, float* __restrict__ ov_out,
                   float* __restrict__ odx, float* __restrict__ ody,
                   float* __restrict__ odz, size_t npts)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool ok = i < npts;
    const int axis = int(blockIdx.y);
    const float x = ok ? xs[i] : 0.0f;
    const float y = ok ? ys[i] : 0.0f;
    const float z = ok ? zs[i] : 0.0f;
    uint32_t flags = 0;
    float g = 0.0f;
    const float v = pe_deriv1<NV, FAST>(t, n, root, x, y, z, axis, &g, &flags);
    if (ok) {
        if (axis == 0) {
            ov_out[i] = (FAST && (flags & PT_SLOW)) ? fmath::u2f(kRedoMark) : v;
        }
        float* o = (axis == 0) ? odx : (axis == 1) ? ody : odz;
        o[i] = g;
    }
}

// The points k_eval_deriv1<NV, true> marked, all three components, with
// the general rules (one lane reads the marker before it overwrites it).
template <int NV>
__global__ __launch_bounds__(256)
void k_eval_deriv1_redo(const Clause* __restrict__ t, uint32_t n, uint32_t root,
                        const float* __restrict__ xs, const float* __restrict__ ys,
                        const float* __restrict__ zs, float* __restrict__ ov_out,
                        float* __restrict__ odx, float* __restrict__ ody,
                        float* __restrict__ odz, size_t npts)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= npts || fmath::f2u(ov_out[i]) != kRedoMark) return;
    const float x = xs[i], y = ys[i], z = zs[i];
    float v = 0.0f;
#pragma unroll 1
    for (int axis = 0; axis < 3; ++axis) {
        uint32_t flags = 0;
        float g = 0.0f;
        v = pe_deriv1<NV, false>(t, n, root, x, y, z, axis, &g, &flags);
        float* o = (axis == 0) ? odx : (axis == 1) ? ody : odz;
        o[i] = g;
    }
    ov_out[i] = v;
}

// Values of the points a call-free value kernel marked, with the general
// rules.
template <int NV>
__global__ __launch_bounds__(256)
void k_eval_values_redo(const Clause* __restrict__ t, uint32_t n, uint32_t root,
                        const float* __restrict__ xs, const float* __restrict__ ys,
                        const float* __restrict__ zs, float* __restrict__ out,
                        size_t npts)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= npts || fmath::f2u(out[i]) != kRedoMark) return;
    uint32_t flags = 0;
    out[i] = pe_value<NV, false>(t, n, root, xs[i], ys[i], zs[i], &flags);
}

////////////////////////////////////////////////////////////////////////////////
// Ambiguity flags: ArrayEvaluator::getAmbiguous (eval_array.cpp:198-217),
// a MIN/MAX clause whose operands compare equal; and
// DerivArrayEvaluator::getAmbiguousDerivs (eval_deriv_array.cpp:58-81),
// equal operands whose partials differ in some component.  The flags are
// OR-ed over the whole tape, as the reference's loop over its clauses.

template <int NV>
__global__ __launch_bounds__(256)
void k_eval_ambiguous(const Clause* __restrict__ t, uint32_t n, uint32_t root,
                      const float* __restrict__ xs, const float* __restrict__ ys,
                      const float* __restrict__ zs, uint8_t* __restrict__ out,
                      size_t npts)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool ok = i < npts;
    LF_RF_DECL(r0);
    LF_RF_SET(r0, NV, kRegX, ok ? xs[i] : 0.0f);
    LF_RF_SET(r0, NV, kRegY, ok ? ys[i] : 0.0f);
    LF_RF_SET(r0, NV, kRegZ, ok ? zs[i] : 0.0f);
    bool amb = false;
    for (uint32_t k = 0; k < n; ++k) {
        const Clause c = t[k];
        const uint32_t o = tape::op(c);
        if (o == OP_NOP) continue;
        const uint32_t m = mode(c);
        const float im = imm(c);
        const float a = (m == MODE_A_IMM || o == OP_LOADI) ? im : LF_RF_GET(r0, NV, ra(c));
        float b = 0.0f;
        if (is_binary(c)) b = (m == MODE_B_IMM) ? im : LF_RF_GET(r0, NV, rb(c));
        if (o == OC::OP_MIN || o == OC::OP_MAX) amb |= (a == b);
        float q;
        apply_value<1>(o, &a, &b, im, &q);
        LF_RF_SET(r0, NV, dst(c), q);
    }
    if (ok) out[i] = amb ? 1 : 0;
}

template <int NV>
__global__ __launch_bounds__(256)
void k_eval_deriv_ambiguous(const Clause* __restrict__ t, uint32_t n,
                            uint32_t root, const float* __restrict__ xs,
                            const float* __restrict__ ys,
                            const float* __restrict__ zs,
                            uint8_t* __restrict__ out, size_t npts)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool ok = i < npts;
    LF_RF_DECL(v);
    LF_RF_DECL(gx);
    LF_RF_DECL(gy);
    LF_RF_DECL(gz);
    LF_RF_SET(v, NV, kRegX, ok ? xs[i] : 0.0f);
    LF_RF_SET(v, NV, kRegY, ok ? ys[i] : 0.0f);
    LF_RF_SET(v, NV, kRegZ, ok ? zs[i] : 0.0f);
    LF_RF_SET(gx, NV, kRegX, 1.0f);
    LF_RF_SET(gy, NV, kRegY, 1.0f);
    LF_RF_SET(gz, NV, kRegZ, 1.0f);
    bool amb = false;
    for (uint32_t k = 0; k < n; ++k) {
        const Clause c = t[k];
        const uint32_t o = tape::op(c);
        if (o == OP_NOP) continue;
        const uint32_t m = mode(c);
        const float im = imm(c);
        float av = im, ad[3] = {0.0f, 0.0f, 0.0f};
        float bv = 0.0f, bd[3] = {0.0f, 0.0f, 0.0f};
        if (m != MODE_A_IMM && o != OP_LOADI) {
            const uint32_t a = ra(c);
            av = LF_RF_GET(v, NV, a);
            ad[0] = LF_RF_GET(gx, NV, a);
            ad[1] = LF_RF_GET(gy, NV, a);
            ad[2] = LF_RF_GET(gz, NV, a);
        }
        if (is_binary(c)) {
            if (m == MODE_B_IMM) {
                bv = im;
            } else {
                const uint32_t b = rb(c);
                bv = LF_RF_GET(v, NV, b);
                bd[0] = LF_RF_GET(gx, NV, b);
                bd[1] = LF_RF_GET(gy, NV, b);
                bd[2] = LF_RF_GET(gz, NV, b);
            }
        }
        if (o == OC::OP_MIN || o == OC::OP_MAX) {
            amb |= (av == bv) &&
                   (ad[0] != bd[0] || ad[1] != bd[1] || ad[2] != bd[2]);
        }
        float ov, od[3];
        apply_deriv(o, av, ad, bv, bd, im, ov, od);
        const uint32_t d = dst(c);
        LF_RF_SET(v, NV, d, ov);
        LF_RF_SET(gx, NV, d, od[0]);
        LF_RF_SET(gy, NV, d, od[1]);
        LF_RF_SET(gz, NV, d, od[2]);
    }
    if (ok) out[i] = amb ? 1 : 0;
}

////////////////////////////////////////////////////////////////////////////////
// Launchers

static int banks_for(uint32_t nregs) {
    const uint32_t b = (nregs + 31) / 32;
    return b <= 1 ? 1 : b <= 2 ? 2 : b <= 4 ? 4 : b <= 8 ? 8 : 0;
}

static int g_points_per_lane = 0;     // 0: choose by bank count

void set_points_per_lane(int p) { g_points_per_lane = p; }

static int points_for(int nv) {
    int p = g_points_per_lane;
    if (p <= 0) p = (nv == 1) ? 2 : 1;
    if (nv >= 4) p = 1;
    if (nv == 2 && p > 2) p = 2;
    return (p >= 4) ? 4 : (p >= 2) ? 2 : 1;
}

template <int NV, int P>
static void launch_v(const Clause* t, uint32_t n, uint32_t root,
                     const float* x, const float* y, const float* z,
                     float* out, size_t npts, bool fast, hipStream_t s) {
    const dim3 block(256);
    const dim3 grid(unsigned((npts + 256u * P - 1) / (256u * P)));
    if (fast) {
        k_eval_values<NV, P, true><<<grid, block, 0, s>>>(t, n, root, x, y, z, out, npts);
        const dim3 grid1(unsigned((npts + 255) / 256));
        k_eval_values_redo<NV><<<grid1, block, 0, s>>>(t, n, root, x, y, z, out, npts);
    } else {
        k_eval_values<NV, P, false><<<grid, block, 0, s>>>(t, n, root, x, y, z, out, npts);
    }
}

std::vector<int> value_points_options(uint32_t nregs) {
    const int nv = banks_for(nregs);
    if (nv == 1) return {1, 2, 4};
    if (nv == 2) return {1, 2};
    return {1};
}

int default_value_points(uint32_t nregs) { return points_for(banks_for(nregs)); }

bool value_points_forced() { return g_points_per_lane > 0; }

hipError_t launch_eval_values(const Clause* t, uint32_t n, uint32_t root,
                              uint32_t nregs, bool fast, const float* x,
                              const float* y, const float* z, float* out,
                              size_t npts, hipStream_t s)
{
    return launch_eval_values_ppl(t, n, root, nregs, fast, x, y, z, out, npts, s,
                                  points_for(banks_for(nregs)));
}

hipError_t launch_eval_values_ppl(const Clause* t, uint32_t n, uint32_t root,
                                  uint32_t nregs, bool fast, const float* x,
                                  const float* y, const float* z, float* out,
                                  size_t npts, hipStream_t s, int p)
{
    if (npts == 0) return hipSuccess;
    const int nv = banks_for(nregs);
    if (!(p == 1 || (p == 2 && nv <= 2) || (p == 4 && nv == 1))) return hipErrorInvalidValue;
    if (nv == 1 && p == 1) launch_v<1, 1>(t, n, root, x, y, z, out, npts, fast, s);
    else if (nv == 1 && p == 2) launch_v<1, 2>(t, n, root, x, y, z, out, npts, fast, s);
    else if (nv == 1) launch_v<1, 4>(t, n, root, x, y, z, out, npts, fast, s);
    else if (nv == 2 && p == 1) launch_v<2, 1>(t, n, root, x, y, z, out, npts, fast, s);
    else if (nv == 2) launch_v<2, 2>(t, n, root, x, y, z, out, npts, fast, s);
    else if (nv == 4) launch_v<4, 1>(t, n, root, x, y, z, out, npts, fast, s);
    else if (nv == 8) launch_v<8, 1>(t, n, root, x, y, z, out, npts, fast, s);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_eval_derivs(const Clause* t, uint32_t n, uint32_t root,
                              uint32_t nregs, bool fast, const float* x,
                              const float* y, const float* z, float* v,
                              float* dx, float* dy, float* dz, size_t npts,
                              hipStream_t s)
{
    if (npts == 0) return hipSuccess;
    const int nv = banks_for(nregs);
    if (nv == 0) return hipErrorInvalidValue;
    const dim3 block(256);
    const dim3 grid(unsigned((npts + 255) / 256), 3);
    const dim3 grid1(unsigned((npts + 255) / 256));
#define LF_D1(NVV)                                                           \
    do {                                                                     \
        if (fast) {                                                          \
            k_eval_deriv1<NVV, true><<<grid, block, 0, s>>>(                 \
                t, n, root, x, y, z, v, dx, dy, dz, npts);                   \
            k_eval_deriv1_redo<NVV><<<grid1, block, 0, s>>>(                 \
                t, n, root, x, y, z, v, dx, dy, dz, npts);                   \
        } else {                                                             \
            k_eval_deriv1<NVV, false><<<grid, block, 0, s>>>(                \
                t, n, root, x, y, z, v, dx, dy, dz, npts);                   \
        }                                                                    \
    } while (0)
    switch (nv) {
        case 1: LF_D1(1); break;
        case 2: LF_D1(2); break;
        case 4: LF_D1(4); break;
        default: LF_D1(8); break;
    }
#undef LF_D1
    return hipGetLastError();
}

hipError_t launch_eval_redo(const Clause* t, uint32_t n, uint32_t root,
                            uint32_t nregs, const float* x, const float* y,
                            const float* z, float* v, float* dx, float* dy,
                            float* dz, size_t npts, hipStream_t s)
{
    if (npts == 0) return hipSuccess;
    const int nv = banks_for(nregs);
    if (nv == 0) return hipErrorInvalidValue;
    const dim3 block(256);
    const dim3 grid(unsigned((npts + 255) / 256));
#define LF_REDO(NVV)                                                         \
    do {                                                                     \
        if (dx) k_eval_deriv1_redo<NVV><<<grid, block, 0, s>>>(              \
                    t, n, root, x, y, z, v, dx, dy, dz, npts);               \
        else k_eval_values_redo<NVV><<<grid, block, 0, s>>>(                 \
                    t, n, root, x, y, z, v, npts);                           \
    } while (0)
    switch (nv) {
        case 1: LF_REDO(1); break;
        case 2: LF_REDO(2); break;
        case 4: LF_REDO(4); break;
        default: LF_REDO(8); break;
    }
#undef LF_REDO
    return hipGetLastError();
}

hipError_t launch_eval_ambiguous(const Clause* t, uint32_t n, uint32_t root,
                                 uint32_t nregs, bool derivs, const float* x,
                                 const float* y, const float* z, uint8_t* out,
                                 size_t npts, hipStream_t s)
{
    if (npts == 0) return hipSuccess;
    const dim3 block(256);
    const dim3 grid(unsigned((npts + 255) / 256));
    const int nv = banks_for(nregs);
    if (derivs) {
        switch (nv) {
            case 1: k_eval_deriv_ambiguous<1><<<grid, block, 0, s>>>(t, n, root, x, y, z, out, npts); break;
            case 2: k_eval_deriv_ambiguous<2><<<grid, block, 0, s>>>(t, n, root, x, y, z, out, npts); break;
            case 4: k_eval_deriv_ambiguous<4><<<grid, block, 0, s>>>(t, n, root, x, y, z, out, npts); break;
            case 8: k_eval_deriv_ambiguous<8><<<grid, block, 0, s>>>(t, n, root, x, y, z, out, npts); break;
            default: return hipErrorInvalidValue;
        }
    } else {
        switch (nv) {
            case 1: k_eval_ambiguous<1><<<grid, block, 0, s>>>(t, n, root, x, y, z, out, npts); break;
            case 2: k_eval_ambiguous<2><<<grid, block, 0, s>>>(t, n, root, x, y, z, out, npts); break;
            case 4: k_eval_ambiguous<4><<<grid, block, 0, s>>>(t, n, root, x, y, z, out, npts); break;
            case 8: k_eval_ambiguous<8><<<grid, block, 0, s>>>(t, n, root, x, y, z, out, npts); break;
            default: return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

// Staging memory of the allocation-free value entry point
// (libfive_tape_eval_values_noalloc): x, y, z and the values, each
// kNoAllocMaxPoints floats, 16 MiB in all, held by every GPU the library
// runs on until the process ends.  Host code never passes this name to a
// kernel: it stands for a host-side shadow, not the device address, which
// comes from noalloc_scratch_address() once per device at initialisation.
static __device__ __align__(256) float scratch[4 * kNoAllocMaxPoints];

hipError_t noalloc_scratch_address(float** out) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    *out = (e == hipSuccess) ? static_cast<float*>(p) : nullptr;
    return e;
}

int max_value_regs() { return 8 * 32; }
int max_deriv_regs() { return 4 * 32; }

}   // namespace amd
}   // namespace libfive
