// Point passes of the slice renderers (2D contours and the raster):
// batched values, partials, ambiguity, edge search and feature evaluation
// over point lists, and the raster's pixel pass, on the interpreters
// shared with the mesher's leaf kernel.  Each point is evaluated with its
// own interval-pruned tape.
#include <hip/hip_runtime.h>

#include "contour_kernels.hpp"
#include "feature_eval.hpp"
#include "point_eval.hpp"
#include "regfile.hpp"

namespace libfive {
namespace amd {

namespace {

// Register banks of 32 per file (8 banks: part of the file in AGPRs or
// scratch memory, placed by the compiler).
int banks(uint32_t nregs) {
    const uint32_t b = (nregs + 31) / 32;
    return b <= 1 ? 1 : b <= 2 ? 2 : b <= 4 ? 4 : b <= 8 ? 8 : 0;
}

// Runs f(t, n) once for each distinct tape id among the active lanes of
// the wave, with exactly the lanes of that tape active; t and n are
// wave-uniform inside f, so the walk keeps its scalar clause loads.  Lanes
// of one point list usually share few tapes (the corners of sibling cells
// share their parent's tape).
template <class F>
__device__ __forceinline__ void for_each_tape(const PoolTapes& T, uint32_t tid, F&& f) {
    uint64_t todo = __ballot(1);
    while (todo) {
        const int first = __builtin_ctzll(todo);
        const uint32_t t0 = uint32_t(__builtin_amdgcn_readlane(int(tid), first));
        const bool mine = (tid == t0);
        todo &= ~__ballot(mine);
        if (mine) {
            const mesh::TapeRec r = T.recs[t0];
            const uint32_t off = __builtin_amdgcn_readfirstlane(r.off);
            const uint32_t len = __builtin_amdgcn_readfirstlane(r.len);
            f(T.pool + off, len);
        }
    }
}

template <int NV>
__global__ __launch_bounds__(64)
void k_point_values(PoolTapes T, const uint32_t* __restrict__ tid,
                    const float* __restrict__ xyz, float* __restrict__ v,
                    uint8_t* __restrict__ amb, size_t npts)
{
    const size_t i = size_t(blockIdx.x) * 64 + threadIdx.x;
    const size_t k = i < npts ? i : npts - 1;
    float r = 0.0f;
    uint32_t fl = 0;
    for_each_tape(T, tid[k], [&](const tape::Clause* t, uint32_t n) {
        r = pe_value<NV>(t, n, T.root, xyz[3 * k], xyz[3 * k + 1], xyz[3 * k + 2], &fl);
    });
    if (i < npts) {
        v[i] = r;
        amb[i] = (fl & PT_AMBIGUOUS) ? 1 : 0;
    }
}

template <int NV>
__global__ __launch_bounds__(64)
void k_point_grads(PoolTapes T, const uint32_t* __restrict__ tid,
                   const float* __restrict__ xyz, float* __restrict__ v,
                   float* __restrict__ g, uint8_t* __restrict__ amb, size_t npts)
{
    const size_t i = size_t(blockIdx.x) * 64 + threadIdx.x;
    const size_t k = i < npts ? i : npts - 1;
    uint32_t fl = 0;
    float d[3] = {0.0f, 0.0f, 0.0f};
    float r = 0.0f;
    for_each_tape(T, tid[k], [&](const tape::Clause* t, uint32_t n) {
        r = pe_grad3<NV>(t, n, T.root, xyz[3 * k], xyz[3 * k + 1], xyz[3 * k + 2], d, &fl);
    });
    if (i < npts) {
        v[i] = r;
        g[3 * i] = d[0];
        g[3 * i + 1] = d[1];
        g[3 * i + 2] = d[2];
        amb[i] = (fl & PT_AMBIGUOUS) ? 1 : 0;
    }
}

// Edge search (dc_tree.inl:339-410): 16 lanes per edge, four edges per
// wave, the four rounds in registers.  A round whose first candidate
// sample is an exact zero needs FeatureEvaluator::isInside; the kernel
// stops that edge there (stop[e] = round, bracket as at the round's start)
// and the host finishes it.  stop[e] = 4 when all rounds ran here.
template <int NV>
__global__ __launch_bounds__(64)
void k_edge_search(PoolTapes T, const uint32_t* __restrict__ tid, float z,
                   double* __restrict__ ends, uint8_t* __restrict__ stop, size_t nedges)
{
    const int lane = threadIdx.x;
    const int j = lane & 15;
    const size_t e = size_t(blockIdx.x) * 4 + (lane >> 4);
    const size_t ec = e < nedges ? e : nedges - 1;
    double a0 = ends[4 * ec], a1 = ends[4 * ec + 1];
    double b0 = ends[4 * ec + 2], b1 = ends[4 * ec + 3];
    const int shift = lane & 48;
    const double f = j / 15.0;
    int round = 0;
    // The 16 lanes of an edge share its tape, so they run together.
    for_each_tape(T, tid[ec], [&](const tape::Clause* t, uint32_t n) {
        for (; round < 4; ++round) {
            const double p0 = a0 * (1 - f) + b0 * f;
            const double p1 = a1 * (1 - f) + b1 * f;
            uint32_t fl;
            const float v = pe_value<NV>(t, n, T.root, float(p0), float(p1), z, &fl);
            const bool cand = j >= 1 && (v > 0.0f || v == 0.0f || j == 15);
            const uint32_t cm = uint32_t(__ballot(cand) >> shift) & 0xffffu;
            const uint32_t zm = uint32_t(__ballot(v == 0.0f) >> shift) & 0xffffu;
            const int js = __builtin_ctz(cm);          // cm always has bit 15
            if ((zm >> js) & 1u) break;                // exact zero: host finishes
            const double fs = js / 15.0, fp = (js - 1) / 15.0;
            const double n0 = a0 * (1 - fp) + b0 * fp, n1 = a1 * (1 - fp) + b1 * fp;
            const double m0 = a0 * (1 - fs) + b0 * fs, m1 = a1 * (1 - fs) + b1 * fs;
            a0 = n0; a1 = n1;
            b0 = m0; b1 = m1;
        }
    });
    if (e < nedges && j == 0) {
        ends[4 * e] = a0;
        ends[4 * e + 1] = a1;
        ends[4 * e + 2] = b0;
        ends[4 * e + 3] = b1;
        stop[e] = uint8_t(round);
    }
}

// Raster pixel pass (Heightmap::pixels, heightmap.cpp:108-155, one Z
// voxel): one workgroup per view of at most 256 pixels, with the view's
// tape; pixel (i, j) of view v is out[offs[v] + i * sy + j] = (value at
// the voxel centre < 0).
template <int NV>
__global__ __launch_bounds__(256)
void k_raster_views(PoolTapes T, const uint32_t* __restrict__ vtape,
                    const float* __restrict__ px, const float* __restrict__ py, float z,
                    const int32_t* __restrict__ views, const uint32_t* __restrict__ offs,
                    uint8_t* __restrict__ out)
{
    const uint32_t v = blockIdx.x;
    const int32_t cx = views[4 * v], cy = views[4 * v + 1];
    const int32_t sx = views[4 * v + 2], sy = views[4 * v + 3];
    const int32_t k = int32_t(threadIdx.x);
    const bool live = k < sx * sy;
    const int32_t i = live ? k / sy : 0, j = live ? k % sy : 0;
    const mesh::TapeRec r = T.recs[vtape[v]];
    const uint32_t off = __builtin_amdgcn_readfirstlane(r.off);
    const uint32_t n = __builtin_amdgcn_readfirstlane(r.len);
    uint32_t fl;
    const float val = pe_value<NV>(T.pool + off, n, T.root, px[cx + i], py[cy + j], z, &fl);
    if (live) out[offs[v] + uint32_t(k)] = val < 0.0f ? 1 : 0;
}

// One wave per point; lane 0 runs the (serial) feature walk with the
// point's tape.
__global__ __launch_bounds__(64)
void k_point_features(PoolTapes T, const uint32_t* __restrict__ tid, uint32_t sx,
                      uint32_t sy, uint32_t sz, const float* __restrict__ xyz,
                      uint8_t* __restrict__ inside, float* __restrict__ feats,
                      int32_t* __restrict__ nfeat, uint8_t* __restrict__ ws, size_t ws_bytes,
                      uint32_t* err, size_t npts)
{
    if (threadIdx.x != 0) return;
    feat::Workspace W = feat::carve(ws + size_t(blockIdx.x) * ws_bytes, T.max_len, T.nregs);
    for (size_t i = blockIdx.x; i < npts; i += gridDim.x) {
        const mesh::TapeRec r = T.recs[tid[i]];
        const tape::Clause* t = T.pool + r.off;
        const uint32_t n = r.len;
        const float p[3] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
        if (inside) {
            inside[i] = feat::is_inside(t, n, T.nregs, T.root, sx, sy, sz, p, W, err) ? 1 : 0;
        }
        if (feats) {
            float out[feat::kMaxOut][3];
            const int m = feat::features(t, n, T.nregs, T.root, sx, sy, sz, p, W, out, err);
            nfeat[i] = m;
            for (int j = 0; j < m; ++j) {
                for (int q = 0; q < 3; ++q) feats[(i * feat::kMaxOut + j) * 3 + q] = out[j][q];
            }
        }
    }
}

}   // namespace

#define LF_BANK_SWITCH(KERNEL, GRID, BLOCK, ...)                                         \
    switch (banks(T.nregs)) {                                                           \
        case 1: hipLaunchKernelGGL(KERNEL<1>, GRID, BLOCK, 0, s, __VA_ARGS__); break;   \
        case 2: hipLaunchKernelGGL(KERNEL<2>, GRID, BLOCK, 0, s, __VA_ARGS__); break;   \
        case 4: hipLaunchKernelGGL(KERNEL<4>, GRID, BLOCK, 0, s, __VA_ARGS__); break;   \
        case 8: hipLaunchKernelGGL(KERNEL<8>, GRID, BLOCK, 0, s, __VA_ARGS__); break;   \
        default: return hipErrorInvalidValue;                                           \
    }

hipError_t launch_point_values(const PoolTapes& T, const uint32_t* tid, const float* xyz,
                               float* v, uint8_t* amb, size_t npts, hipStream_t s) {
    if (npts == 0) return hipSuccess;
    const dim3 grid(uint32_t((npts + 63) / 64));
    LF_BANK_SWITCH(k_point_values, grid, dim3(64), T, tid, xyz, v, amb, npts)
    return hipGetLastError();
}

hipError_t launch_point_grads(const PoolTapes& T, const uint32_t* tid, const float* xyz,
                              float* v, float* g, uint8_t* amb, size_t npts, hipStream_t s) {
    if (npts == 0) return hipSuccess;
    const dim3 grid(uint32_t((npts + 63) / 64));
    LF_BANK_SWITCH(k_point_grads, grid, dim3(64), T, tid, xyz, v, g, amb, npts)
    return hipGetLastError();
}

hipError_t launch_edge_search(const PoolTapes& T, const uint32_t* tid, float z, double* ends,
                              uint8_t* stop, size_t nedges, hipStream_t s) {
    if (nedges == 0) return hipSuccess;
    const dim3 grid(uint32_t((nedges + 3) / 4));
    LF_BANK_SWITCH(k_edge_search, grid, dim3(64), T, tid, z, ends, stop, nedges)
    return hipGetLastError();
}

hipError_t launch_raster_views(const PoolTapes& T, const uint32_t* vtape, const float* px,
                               const float* py, float z, const int32_t* views,
                               const uint32_t* offs, uint8_t* out, size_t nviews,
                               hipStream_t s) {
    if (nviews == 0) return hipSuccess;
    const dim3 grid(static_cast<uint32_t>(nviews), 1, 1);
    LF_BANK_SWITCH(k_raster_views, grid, dim3(256), T, vtape, px, py, z, views, offs, out)
    return hipGetLastError();
}

#undef LF_BANK_SWITCH

size_t point_feature_workspace(uint32_t max_len, uint32_t nregs) {
    return (feat::workspace_bytes(max_len, nregs) + 255) & ~size_t(255);
}

hipError_t launch_point_features(const PoolTapes& T, const uint32_t* tid, uint32_t sx,
                                 uint32_t sy, uint32_t sz, const float* xyz, uint8_t* inside,
                                 float* feats, int32_t* nfeat, uint8_t* ws, size_t ws_bytes,
                                 uint32_t* err, size_t npts, hipStream_t s) {
    if (npts == 0) return hipSuccess;
    const uint32_t grid = uint32_t(npts < kFeatureWaves ? npts : kFeatureWaves);
    hipLaunchKernelGGL(k_point_features, dim3(grid), dim3(64), 0, s, T, tid, sx, sy, sz, xyz,
                       inside, feats, nfeat, ws, ws_bytes, err, npts);
    return hipGetLastError();
}

}   // namespace amd
}   // namespace libfive
