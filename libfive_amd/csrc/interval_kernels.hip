// Interval evaluation over boxes (IntervalEvaluator::eval,
// libfive/src/eval/eval_interval.cpp:62-93 and 189-294): one box per lane,
// SoA box bounds, the same wave-uniform tape walk as the point kernels with
// three register banks (lower, upper, maybe-NaN).
#include <hip/hip_runtime.h>

#include "interval_kernels.hpp"
#include "interval_ops.hpp"
#include "regfile.hpp"
#include "tape.hpp"

namespace libfive {
namespace amd {

using namespace tape;

template <int NV>
__global__ __launch_bounds__(256)
void k_eval_intervals(const Clause* __restrict__ t, uint32_t n, uint32_t root,
                      const float* __restrict__ xlo, const float* __restrict__ xhi,
                      const float* __restrict__ ylo, const float* __restrict__ yhi,
                      const float* __restrict__ zlo, const float* __restrict__ zhi,
                      float* __restrict__ olo, float* __restrict__ ohi,
                      uint8_t* __restrict__ onan, size_t nbox)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nbox) return;
    LF_RF_DECL(L);
    LF_RF_DECL(U);
    LF_RF_DECL(F);
    // X/Y/Z slots hold Interval(lower, upper) (eval_interval.cpp:70-72).
    const Iv ix = make_checked(xlo[i], xhi[i]);
    const Iv iy = make_checked(ylo[i], yhi[i]);
    const Iv iz = make_checked(zlo[i], zhi[i]);
    LF_RF_SET(L, NV, kRegX, ix.lo); LF_RF_SET(U, NV, kRegX, ix.hi); LF_RF_SET(F, NV, kRegX, ix.nan ? 1.0f : 0.0f);
    LF_RF_SET(L, NV, kRegY, iy.lo); LF_RF_SET(U, NV, kRegY, iy.hi); LF_RF_SET(F, NV, kRegY, iy.nan ? 1.0f : 0.0f);
    LF_RF_SET(L, NV, kRegZ, iz.lo); LF_RF_SET(U, NV, kRegZ, iz.hi); LF_RF_SET(F, NV, kRegZ, iz.nan ? 1.0f : 0.0f);
    for (uint32_t k = 0; k < n; ++k) {
        const Clause c = t[k];
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
        if (nargs(o) == 2) {
            if (m == MODE_B_IMM) {
                b = make_checked(im, im);
            } else {
                const uint32_t r = rb(c);
                b = {LF_RF_GET(L, NV, r), LF_RF_GET(U, NV, r), LF_RF_GET(F, NV, r) != 0.0f};
            }
        }
        const Iv out = iv_apply(o, a, b, im);
        const uint32_t d = dst(c);
        LF_RF_SET(L, NV, d, out.lo);
        LF_RF_SET(U, NV, d, out.hi);
        LF_RF_SET(F, NV, d, out.nan ? 1.0f : 0.0f);
    }
    olo[i] = LF_RF_GET(L, NV, root);
    ohi[i] = LF_RF_GET(U, NV, root);
    onan[i] = LF_RF_GET(F, NV, root) != 0.0f;
}

hipError_t launch_eval_intervals(const Clause* t, uint32_t n, uint32_t root,
                                 uint32_t nregs, const float* xlo,
                                 const float* xhi, const float* ylo,
                                 const float* yhi, const float* zlo,
                                 const float* zhi, float* olo, float* ohi,
                                 uint8_t* onan, size_t nbox, hipStream_t s)
{
    if (nbox == 0) return hipSuccess;
    const dim3 block(256);
    const dim3 grid(unsigned((nbox + 255) / 256));
    const int nv = int((nregs + 31) / 32);
    if (nv <= 1) {
        k_eval_intervals<1><<<grid, block, 0, s>>>(t, n, root, xlo, xhi, ylo, yhi, zlo, zhi, olo, ohi, onan, nbox);
    } else if (nv <= 2) {
        k_eval_intervals<2><<<grid, block, 0, s>>>(t, n, root, xlo, xhi, ylo, yhi, zlo, zhi, olo, ohi, onan, nbox);
    } else if (nv <= 4) {
        k_eval_intervals<4><<<grid, block, 0, s>>>(t, n, root, xlo, xhi, ylo, yhi, zlo, zhi, olo, ohi, onan, nbox);
    } else if (nv <= 8) {
        k_eval_intervals<8><<<grid, block, 0, s>>>(t, n, root, xlo, xhi, ylo, yhi, zlo, zhi, olo, ohi, onan, nbox);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}   // namespace amd
}   // namespace libfive
