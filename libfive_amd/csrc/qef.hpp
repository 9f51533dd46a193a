// Symmetric 3x3 eigen-decomposition and the dual-contouring vertex solve,
// in double precision, shared by the HIP kernels and host code so that both
// produce identical bits.
//
// The reference solves the QEF with Eigen's SelfAdjointEigenSolver
// (libfive/src/render/brep/dc/dc_tree.inl:740-813) and ranks intersections
// by the eigenvalues of A^T A (libfive/include/libfive/render/brep/dc/
// intersection.hpp:63-89).  Eigen is not available here; the eigenvalues
// and eigenvectors come from the cyclic Jacobi method (Jacobi 1846; see
// Golub & Van Loan, "Matrix Computations", 4th ed., section 8.5), which is
// exact to rounding for symmetric 3x3 matrices.  The solve itself follows
// the reference: truncated pseudo-inverse with cutoff 0.1 on the
// eigenvalues, minimising towards the mass point, and the QEF error
// v^T A v - 2 v^T b + b^T b.
//
// Compile without floating-point contraction.
#pragma once

#include <cmath>
#include <cstdint>

#include "fmath.hpp"

namespace libfive {
namespace qef {

constexpr double kEigenvalueCutoff = 0.1;    // dc_flags.hpp: EIGENVALUE_CUTOFF

// a: symmetric matrix (read: a[0..2][0..2]); w: eigenvalues; v: columns are
// eigenvectors (v[row][col]).
LF_HD void jacobi3(const double a_in[3][3], double w[3], double v[3][3]) {
    double a[3][3];
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
            a[i][j] = a_in[i][j];
            v[i][j] = (i == j) ? 1.0 : 0.0;
        }
    }
    for (int sweep = 0; sweep < 32; ++sweep) {
        const double off = ::fabs(a[0][1]) + ::fabs(a[0][2]) + ::fabs(a[1][2]);
        if (off == 0.0) break;
        for (int pi = 0; pi < 3; ++pi) {
            const int p = (pi == 2) ? 1 : 0;
            const int q = (pi == 0) ? 1 : 2;
            const double apq = a[p][q];
            if (apq == 0.0) continue;
            // Skip rotations that cannot change the diagonal any more.
            const double app = a[p][p], aqq = a[q][q];
            const double g = 100.0 * ::fabs(apq);
            if (sweep > 3 && ::fabs(app) + g == ::fabs(app) &&
                ::fabs(aqq) + g == ::fabs(aqq)) {
                a[p][q] = 0.0;
                a[q][p] = 0.0;
                continue;
            }
            const double theta = (aqq - app) / (2.0 * apq);
            const double t = ((theta >= 0.0) ? 1.0 : -1.0) /
                             (::fabs(theta) + ::sqrt(theta * theta + 1.0));
            const double c = 1.0 / ::sqrt(t * t + 1.0);
            const double s = t * c;
            // A' = J^T A J with J the rotation in the (p, q) plane
            for (int k = 0; k < 3; ++k) {
                const double akp = a[k][p], akq = a[k][q];
                a[k][p] = c * akp - s * akq;
                a[k][q] = s * akp + c * akq;
            }
            for (int k = 0; k < 3; ++k) {
                const double apk = a[p][k], aqk = a[q][k];
                a[p][k] = c * apk - s * aqk;
                a[q][k] = s * apk + c * aqk;
            }
            a[p][q] = 0.0;
            a[q][p] = 0.0;
            for (int k = 0; k < 3; ++k) {
                const double vkp = v[k][p], vkq = v[k][q];
                v[k][p] = c * vkp - s * vkq;
                v[k][q] = s * vkp + c * vkq;
            }
        }
    }
    for (int i = 0; i < 3; ++i) w[i] = a[i][i];
}

// Number of eigenvalues of A^T A with magnitude below the cutoff
// (Intersection::get_rank, intersection.hpp:63-89).
LF_HD int small_eigenvalue_count(const double ata[3][3]) {
    double w[3], v[3][3];
    jacobi3(ata, w, v);
    int r = 0;
    for (int j = 0; j < 3; ++j) r += (::fabs(w[j]) < kEigenvalueCutoff);
    return r;
}

// DCTree::findVertex (dc_tree.inl:740-813).  mass = (sum x, sum y, sum z,
// count).  Writes the vertex and the number of kept eigenvalues (the
// feature rank); returns the QEF error.
LF_HD double find_vertex(const double ata[3][3], const double atb[3],
                         double btb, const double mass[4], double vert[3],
                         int* rank) {
    double w[3], u[3][3];
    jacobi3(ata, w, u);
    double d[3];
    int r = 0;
    for (int i = 0; i < 3; ++i) {
        d[i] = (::fabs(w[i]) < kEigenvalueCutoff) ? 0.0 : (1.0 / w[i]);
        r += (d[i] != 0.0);
    }
    if (rank) *rank = r;
    // pseudo-inverse U D U^T
    double p[3][3];
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
            double acc = 0.0;
            for (int k = 0; k < 3; ++k) {
                const double t = u[i][k] * d[k];
                acc = acc + t * u[j][k];
            }
            p[i][j] = acc;
        }
    }
    double center[3];
    for (int i = 0; i < 3; ++i) center[i] = mass[i] / mass[3];
    double rhs[3];
    for (int i = 0; i < 3; ++i) {
        double acc = 0.0;
        for (int k = 0; k < 3; ++k) acc = acc + ata[i][k] * center[k];
        rhs[i] = atb[i] - acc;
    }
    for (int i = 0; i < 3; ++i) {
        double acc = 0.0;
        for (int k = 0; k < 3; ++k) acc = acc + p[i][k] * rhs[k];
        vert[i] = acc + center[i];
    }
    // error = v^T A v - 2 v^T b + b^T b
    double vav = 0.0;
    for (int i = 0; i < 3; ++i) {
        double row = 0.0;
        for (int k = 0; k < 3; ++k) row = row + ata[i][k] * vert[k];
        vav = vav + vert[i] * row;
    }
    double vb = 0.0;
    for (int i = 0; i < 3; ++i) vb = vb + vert[i] * atb[i];
    return (vav - 2.0 * vb) + btb;
}

// Intersection::push (intersection.hpp:32-59) on a flat accumulator:
// acc[0..3] mass point (x, y, z, count), acc[4..9] A^T A upper triangle
// (xx, xy, xz, yy, yz, zz), acc[10..12] A^T b, acc[13] b^T b.
LF_HD void push_sample(double acc[14], const double pos[3],
                       const double deriv_in[3], double value) {
    acc[0] = acc[0] + pos[0];
    acc[1] = acc[1] + pos[1];
    acc[2] = acc[2] + pos[2];
    acc[3] = acc[3] + 1.0;
    const double norm = ::sqrt((deriv_in[0] * deriv_in[0] +
                                deriv_in[1] * deriv_in[1]) +
                               deriv_in[2] * deriv_in[2]);
    const double n[3] = {deriv_in[0] / norm, deriv_in[1] / norm,
                         deriv_in[2] / norm};
    const double val = value / norm;
    const bool finite = __builtin_isfinite(n[0]) && __builtin_isfinite(n[1]) &&
                        __builtin_isfinite(n[2]);
    if (norm <= 1e-12 || !finite) return;
    acc[4] = acc[4] + n[0] * n[0];
    acc[5] = acc[5] + n[0] * n[1];
    acc[6] = acc[6] + n[0] * n[2];
    acc[7] = acc[7] + n[1] * n[1];
    acc[8] = acc[8] + n[1] * n[2];
    acc[9] = acc[9] + n[2] * n[2];
    const double b = ((n[0] * pos[0] + n[1] * pos[1]) + n[2] * pos[2]) - val;
    acc[10] = acc[10] + n[0] * b;
    acc[11] = acc[11] + n[1] * b;
    acc[12] = acc[12] + n[2] * b;
    acc[13] = acc[13] + b * b;
}

// Compensated single-precision accumulation (the mesher's precision form,
// libfive_mesh_settings::precision = 1): each of the 14 sums is a pair of
// floats (hi, lo) whose exact sum carries the value.  Terms are added with
// the error-free transformations TwoSum (Knuth, TAOCP vol. 2, 4.2.2) and
// TwoProduct (one fused multiply-add), the errors collected in lo, as in
// the Sum2 / Dot2 algorithms of Ogita, Rump and Oishi ("Accurate sum and
// dot product", SIAM J. Sci. Comput. 26(6), 2005).  The pairs are widened
// to double (exactly, in practice) for the rank test and the eigensolve.
struct AccF2 {
    float hi[14];
    float lo[14];
};

LF_HD void f2_zero(AccF2& a) {
    for (int q = 0; q < 14; ++q) {
        a.hi[q] = 0.0f;
        a.lo[q] = 0.0f;
    }
}

// hi + lo += x (TwoSum: s + e == hi + x exactly)
LF_HD void f2_add(float& hi, float& lo, float x) {
    const float s = hi + x;
    const float bp = s - hi;
    const float e = (hi - (s - bp)) + (x - bp);
    hi = s;
    lo = lo + e;
}

// hi + lo += a * b (TwoProduct: p + pe == a * b exactly)
LF_HD void f2_add_prod(float& hi, float& lo, float a, float b) {
    const float p = a * b;
    const float pe = __builtin_fmaf(a, b, -p);
    f2_add(hi, lo, p);
    lo = lo + pe;
}

// Intersection::push (intersection.hpp:32-59) with the sums as float
// pairs: single-precision normalisation, then A^T A, A^T b and b^T b from
// the normal and b = n . pos - value / |d| (itself a float pair).
LF_HD void push_sample_f2(AccF2& a, const float pos[3], const float deriv[3],
                          float value) {
    f2_add(a.hi[0], a.lo[0], pos[0]);
    f2_add(a.hi[1], a.lo[1], pos[1]);
    f2_add(a.hi[2], a.lo[2], pos[2]);
    a.hi[3] = a.hi[3] + 1.0f;
    const float norm = sqrtf((deriv[0] * deriv[0] + deriv[1] * deriv[1]) +
                             deriv[2] * deriv[2]);
    const float n[3] = {deriv[0] / norm, deriv[1] / norm, deriv[2] / norm};
    const float val = value / norm;
    const bool finite = __builtin_isfinite(n[0]) && __builtin_isfinite(n[1]) &&
                        __builtin_isfinite(n[2]);
    if (norm <= 1e-12f || !finite) return;
    f2_add_prod(a.hi[4], a.lo[4], n[0], n[0]);
    f2_add_prod(a.hi[5], a.lo[5], n[0], n[1]);
    f2_add_prod(a.hi[6], a.lo[6], n[0], n[2]);
    f2_add_prod(a.hi[7], a.lo[7], n[1], n[1]);
    f2_add_prod(a.hi[8], a.lo[8], n[1], n[2]);
    f2_add_prod(a.hi[9], a.lo[9], n[2], n[2]);
    float bh = 0.0f, bl = 0.0f;
    f2_add_prod(bh, bl, n[0], pos[0]);
    f2_add_prod(bh, bl, n[1], pos[1]);
    f2_add_prod(bh, bl, n[2], pos[2]);
    f2_add(bh, bl, -val);
    for (int j = 0; j < 3; ++j) {
        f2_add_prod(a.hi[10 + j], a.lo[10 + j], n[j], bh);
        a.lo[10 + j] = a.lo[10 + j] + n[j] * bl;
    }
    f2_add_prod(a.hi[13], a.lo[13], bh, bh);
    a.lo[13] = a.lo[13] + 2.0f * (bh * bl);
}

// Adds a widened accumulator's A^T A, A^T b and b^T b (entries 4..13),
// each split back into a float pair.
LF_HD void f2_add_acc(AccF2& a, const double acc[14]) {
    for (int q = 4; q < 14; ++q) {
        const float xh = float(acc[q]);
        const float xl = float(acc[q] - double(xh));
        f2_add(a.hi[q], a.lo[q], xh);
        a.lo[q] = a.lo[q] + xl;
    }
}

LF_HD void f2_widen(const AccF2& a, double acc[14]) {
    for (int q = 0; q < 14; ++q) acc[q] = double(a.hi[q]) + double(a.lo[q]);
}

LF_HD void unpack_ata(const double acc[14], double ata[3][3]) {
    ata[0][0] = acc[4]; ata[0][1] = acc[5]; ata[0][2] = acc[6];
    ata[1][0] = acc[5]; ata[1][1] = acc[7]; ata[1][2] = acc[8];
    ata[2][0] = acc[6]; ata[2][1] = acc[8]; ata[2][2] = acc[9];
}

}   // namespace qef
}   // namespace libfive
