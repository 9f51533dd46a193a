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

namespace libfive {
namespace amd {

using namespace tape;
namespace OC = Opcode;

// Value of one clause given its operand values (b is ignored by unary
// opcodes).
__device__ __forceinline__ float apply_value(uint32_t o, float a, float b,
                                             float im) {
    switch (o) {
        case OC::OP_ADD: return a + b;
        case OC::OP_MUL: return a * b;
        case OC::OP_MIN: return (b < a) ? b : a;
        case OC::OP_MAX: return (a < b) ? b : a;
        case OC::OP_SUB: return a - b;
        case OC::OP_NEG: return -a;
        case OC::OP_SQUARE: return a * a;
        case OC::OP_ABS: return fabsf(a);
        case OC::OP_SQRT: return sqrtf(a);
        case OC::OP_SIN: return fmath::sin_f32(a);
        case OC::OP_COS: return fmath::cos_f32(a);
        case OC::OP_DIV: return a / b;
        case OP_LOADI: return im;
        case OP_COPY:
        case OC::CONST_VAR: return a;
        default: return fmath::eval_op(int(o), a, b);
    }
}

template <int NV>
__global__ __launch_bounds__(256)
void k_eval_values(const Clause* __restrict__ t, uint32_t n, uint32_t root,
                   const float* __restrict__ xs, const float* __restrict__ ys,
                   const float* __restrict__ zs, float* __restrict__ out,
                   size_t npts)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    LF_RF_DECL(r);
    LF_RF_SET(r, NV, kRegX, xs[i]);
    LF_RF_SET(r, NV, kRegY, ys[i]);
    LF_RF_SET(r, NV, kRegZ, zs[i]);
    for (uint32_t k = 0; k < n; ++k) {
        const Clause c = t[k];
        const uint32_t o = tape::op(c);
        const uint32_t m = mode(c);
        const float im = imm(c);
        const float a = (m == MODE_A_IMM) ? im : LF_RF_GET(r, NV, ra(c));
        const float b = (nargs(o) != 2) ? 0.0f
                      : (m == MODE_B_IMM) ? im : LF_RF_GET(r, NV, rb(c));
        LF_RF_SET(r, NV, dst(c), apply_value(o, a, b, im));
    }
    out[i] = LF_RF_GET(r, NV, root);
}

// Value and partial derivatives of one clause (eval_deriv_array.cpp:119-256;
// the value is the same as apply_value).
__device__ __forceinline__ void apply_deriv(uint32_t o, float av,
                                            const float ad[3], float bv,
                                            const float bd[3], float im,
                                            float& ov, float od[3]) {
    switch (o) {
        case OC::OP_ADD:
            ov = av + bv;
            for (int j = 0; j < 3; ++j) od[j] = ad[j] + bd[j];
            break;
        case OC::OP_SUB:
            ov = av - bv;
            for (int j = 0; j < 3; ++j) od[j] = ad[j] - bd[j];
            break;
        case OC::OP_MUL:
            ov = av * bv;
            for (int j = 0; j < 3; ++j) {
                const float p = bd[j] * av;
                const float q = ad[j] * bv;
                od[j] = p + q;
            }
            break;
        case OC::OP_MIN: {
            ov = (bv < av) ? bv : av;
            const bool s = av < bv;
            for (int j = 0; j < 3; ++j) od[j] = s ? ad[j] : bd[j];
            break;
        }
        case OC::OP_MAX: {
            ov = (av < bv) ? bv : av;
            const bool s = av < bv;
            for (int j = 0; j < 3; ++j) od[j] = s ? bd[j] : ad[j];
            break;
        }
        case OC::OP_DIV: {
            ov = av / bv;
            const float b2 = bv * bv;
            for (int j = 0; j < 3; ++j) {
                const float p = ad[j] * bv;
                const float q = bd[j] * av;
                od[j] = (p - q) / b2;
            }
            break;
        }
        case OC::OP_ATAN2: {
            ov = fmath::atan2_f32(av, bv);
            const float a2 = av * av;
            const float b2 = bv * bv;
            const float den = a2 + b2;
            for (int j = 0; j < 3; ++j) {
                const float p = ad[j] * bv;
                const float q = bd[j] * av;
                od[j] = (p - q) / den;
            }
            break;
        }
        case OC::OP_POW: {
            ov = fmath::pow_f32(av, bv);
            const float f = bv * fmath::pow_f32(av, bv - 1.0f);
            for (int j = 0; j < 3; ++j) od[j] = ad[j] * f;
            break;
        }
        case OC::OP_NTH_ROOT: {
            ov = fmath::nth_root_f32(av, bv);
            const float f = fmath::pow_f32(av, 1.0f / bv - 1.0f) / bv;
            for (int j = 0; j < 3; ++j) od[j] = (ad[j] == 0.0f) ? 0.0f : ad[j] * f;
            break;
        }
        case OC::OP_MOD:
            ov = fmath::mod_f32(av, bv);
            for (int j = 0; j < 3; ++j) od[j] = ad[j];
            break;
        case OC::OP_NANFILL: {
            const bool n_ = fmath::isnan_f(av);
            ov = n_ ? bv : av;
            for (int j = 0; j < 3; ++j) od[j] = n_ ? bd[j] : ad[j];
            break;
        }
        case OC::OP_COMPARE:
            ov = (av < bv) ? -1.0f : (av > bv) ? 1.0f : 0.0f;
            for (int j = 0; j < 3; ++j) od[j] = 0.0f;
            break;
        case OC::OP_SQUARE:
            ov = av * av;
            for (int j = 0; j < 3; ++j) od[j] = (ad[j] * av) * 2.0f;
            break;
        case OC::OP_SQRT: {
            ov = sqrtf(av);
            const float den = 2.0f * ov;
            for (int j = 0; j < 3; ++j) {
                od[j] = (av < 0.0f || ad[j] == 0.0f) ? 0.0f : ad[j] / den;
            }
            break;
        }
        case OC::OP_NEG:
            ov = -av;
            for (int j = 0; j < 3; ++j) od[j] = -ad[j];
            break;
        case OC::OP_SIN: {
            float s, co;
            fmath::sin_cos_f32(av, &s, &co);
            ov = s;
            for (int j = 0; j < 3; ++j) od[j] = ad[j] * co;
            break;
        }
        case OC::OP_COS: {
            float s, co;
            fmath::sin_cos_f32(av, &s, &co);
            ov = co;
            const float ns = -s;
            for (int j = 0; j < 3; ++j) od[j] = ad[j] * ns;
            break;
        }
        case OC::OP_TAN: {
            ov = fmath::tan_f32(av);
            const float r = 1.0f / fmath::cos_f32(av);
            const float f = r * r;
            for (int j = 0; j < 3; ++j) od[j] = ad[j] * f;
            break;
        }
        case OC::OP_ASIN: {
            ov = fmath::asin_f32(av);
            const float s = sqrtf(1.0f - av * av);
            for (int j = 0; j < 3; ++j) od[j] = ad[j] / s;
            break;
        }
        case OC::OP_ACOS: {
            ov = fmath::acos_f32(av);
            const float s = -sqrtf(1.0f - av * av);
            for (int j = 0; j < 3; ++j) od[j] = ad[j] / s;
            break;
        }
        case OC::OP_ATAN: {
            ov = fmath::atan_f32(av);
            const float s = av * av + 1.0f;
            for (int j = 0; j < 3; ++j) od[j] = ad[j] / s;
            break;
        }
        case OC::OP_LOG:
            ov = fmath::log_f32(av);
            for (int j = 0; j < 3; ++j) od[j] = ad[j] / av;
            break;
        case OC::OP_EXP:
            ov = fmath::exp_f32(av);
            for (int j = 0; j < 3; ++j) od[j] = ad[j] * ov;
            break;
        case OC::OP_ABS:
            ov = fabsf(av);
            for (int j = 0; j < 3; ++j) od[j] = (av > 0.0f) ? ad[j] : -ad[j];
            break;
        case OC::OP_RECIP: {
            ov = 1.0f / av;
            const float s = -(av * av);
            for (int j = 0; j < 3; ++j) od[j] = ad[j] / s;
            break;
        }
        case OP_LOADI:
            ov = im;
            for (int j = 0; j < 3; ++j) od[j] = 0.0f;
            break;
        default:    // COPY, CONST_VAR
            ov = av;
            for (int j = 0; j < 3; ++j) od[j] = ad[j];
            break;
    }
}

template <int NV>
__global__ __launch_bounds__(256)
void k_eval_derivs(const Clause* __restrict__ t, uint32_t n, uint32_t root,
                   const float* __restrict__ xs, const float* __restrict__ ys,
                   const float* __restrict__ zs, float* __restrict__ ov_out,
                   float* __restrict__ odx, float* __restrict__ ody,
                   float* __restrict__ odz, size_t npts)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    LF_RF_DECL(v);
    LF_RF_DECL(gx);
    LF_RF_DECL(gy);
    LF_RF_DECL(gz);
    LF_RF_SET(v, NV, kRegX, xs[i]);
    LF_RF_SET(v, NV, kRegY, ys[i]);
    LF_RF_SET(v, NV, kRegZ, zs[i]);
    LF_RF_SET(gx, NV, kRegX, 1.0f);
    LF_RF_SET(gy, NV, kRegY, 1.0f);
    LF_RF_SET(gz, NV, kRegZ, 1.0f);
    for (uint32_t k = 0; k < n; ++k) {
        const Clause c = t[k];
        const uint32_t o = tape::op(c);
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
        if (nargs(o) == 2) {
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
        float ov, od[3];
        apply_deriv(o, av, ad, bv, bd, im, ov, od);
        const uint32_t d = dst(c);
        LF_RF_SET(v, NV, d, ov);
        LF_RF_SET(gx, NV, d, od[0]);
        LF_RF_SET(gy, NV, d, od[1]);
        LF_RF_SET(gz, NV, d, od[2]);
    }
    ov_out[i] = LF_RF_GET(v, NV, root);
    odx[i] = LF_RF_GET(gx, NV, root);
    ody[i] = LF_RF_GET(gy, NV, root);
    odz[i] = LF_RF_GET(gz, NV, root);
}

////////////////////////////////////////////////////////////////////////////////
// Launchers

static int banks_for(uint32_t nregs) {
    return int((nregs + 31) / 32);
}

hipError_t launch_eval_values(const Clause* t, uint32_t n, uint32_t root,
                              uint32_t nregs, const float* x, const float* y,
                              const float* z, float* out, size_t npts,
                              hipStream_t s)
{
    if (npts == 0) return hipSuccess;
    const dim3 block(256);
    const dim3 grid(unsigned((npts + 255) / 256));
    switch (banks_for(nregs)) {
        case 1: k_eval_values<1><<<grid, block, 0, s>>>(t, n, root, x, y, z, out, npts); break;
        case 2: k_eval_values<2><<<grid, block, 0, s>>>(t, n, root, x, y, z, out, npts); break;
        case 3:
        case 4: k_eval_values<4><<<grid, block, 0, s>>>(t, n, root, x, y, z, out, npts); break;
        case 5: case 6: case 7:
        case 8: k_eval_values<8><<<grid, block, 0, s>>>(t, n, root, x, y, z, out, npts); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_eval_derivs(const Clause* t, uint32_t n, uint32_t root,
                              uint32_t nregs, const float* x, const float* y,
                              const float* z, float* v, float* dx, float* dy,
                              float* dz, size_t npts, hipStream_t s)
{
    if (npts == 0) return hipSuccess;
    const dim3 block(256);
    const dim3 grid(unsigned((npts + 255) / 256));
    switch (banks_for(nregs)) {
        case 1: k_eval_derivs<1><<<grid, block, 0, s>>>(t, n, root, x, y, z, v, dx, dy, dz, npts); break;
        case 2: k_eval_derivs<2><<<grid, block, 0, s>>>(t, n, root, x, y, z, v, dx, dy, dz, npts); break;
        case 3:
        case 4: k_eval_derivs<4><<<grid, block, 0, s>>>(t, n, root, x, y, z, v, dx, dy, dz, npts); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int max_value_regs() { return 8 * 32; }
int max_deriv_regs() { return 4 * 32; }

}   // namespace amd
}   // namespace libfive
