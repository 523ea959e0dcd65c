// skm_walk.hpp -- sample bricks, control-point tiles and per-sample steps
// shared by the on-the-fly kernels (skm_kernels.hip) and the assembled
// kernel-matrix kernels (skm_operator.hip).
#pragma once

#include <climits>

#include "skm_internal.hpp"

namespace skm {

// ------------------------------------------------------------ tiling helpers

template <int DIM> struct Tile;
template <> struct Tile<2> { static constexpr int X = 16, Y = 16, Z = 1; };
template <> struct Tile<3> { static constexpr int X = 8, Y = 8, Z = 4; };
// One wave's brick of samples (the block tile holds four).
template <int DIM> struct WaveBrick;
template <> struct WaveBrick<2> { static constexpr int X = 16, Y = 4, Z = 1; };
template <> struct WaveBrick<3> { static constexpr int X = 4, Y = 4, Z = 4; };

// Thread -> offset inside the block tile; each wave64 gets a compact brick.
template <int DIM>
__device__ __forceinline__ void tile_coords(int t, int *l)
{
    const int lane = t & 63, wave = t >> 6;
    if (DIM == 3) {
        l[0] = (lane & 3) + (wave & 1) * 4;
        l[1] = ((lane >> 2) & 3) + (wave >> 1) * 4;
        l[2] = lane >> 4;
    } else {
        l[0] = lane & 15;
        l[1] = (lane >> 4) + wave * 4;
        l[2] = 0;
    }
}

__device__ __forceinline__ long long block_linear()
{
    return blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
}

template <int DIM>
__device__ __forceinline__ void sample_position(const Problem &pr, const int *m, long long s, int *pos)
{
#pragma unroll
    for (int d = 0; d < DIM; ++d) pos[d] = m[d];
    if (pr.sub > 1) {
#pragma unroll
        for (int d = 0; d < DIM; ++d) pos[d] = m[d] * pr.sub + pr.offsets[s * DIM + d];
    }
}

template <int DIM>
__device__ __forceinline__ float fixed_value(const Problem &pr, const int *pos)
{
    // A jitter offset equal to `subsample` (nextF() == 1.0) can step past the
    // last voxel; the index is clamped instead of reading out of bounds.
    long long f = 0, stride = 1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        int i = clampi(pos[d], 0, pr.fixed.n[d] - 1);
        f += stride * i;
        stride *= pr.fixed.n[d];
    }
    return pr.fixed_data[f];
}

template <int DIM>
__device__ __forceinline__ void df_lookup(const Problem &pr, const float *x, float *u)
{
#pragma unroll
    for (int d = 0; d < DIM; ++d) u[d] = 0.0f;
    if (!pr.has_df) return;
    long long f;
    if (nearest<DIM>(pr.df, x, f)) {
#pragma unroll
        for (int d = 0; d < DIM; ++d) u[d] = pr.df_data[f * DIM + d];
    }
}

// State of control point j in the weighted modes; returns its lam.
template <int DIM, int MODE>
__device__ __forceinline__ float cp_state_load(const Problem &pr, long long j, PState &st)
{
    if (is_stationary(MODE)) return 0.0f;
    const float4 b = pr.cp_pstate[2 * j + 1];
    pstate_unpack<DIM, MODE>(pr.cp_pstate[2 * j], b, st);
    return b.w;
}

// Control-point tile of the derivative gather: small enough that the
// tile's constants stay in scalar registers.  In 3D the four points share
// x and y, so their chords along a sample row nearly coincide and the lanes
// of a row serve all four (measured faster than 4x1x1, 1x4x1 and 2x2x1).
template <int DIM> struct CpTile;
template <> struct CpTile<2> { static constexpr int X = 2, Y = 2, Z = 1; };
template <> struct CpTile<3> { static constexpr int X = 1, Y = 1, Z = 4; };

// Sample planes (last axis) per slab of the derivative gather: 3D windows
// are walked slab by slab so that the workgroups in flight share one slab of
// samples (a few MB, resident in every XCD's L2); 2D windows in one slab.
template <int DIM>
__host__ __device__ __forceinline__ int gather_slab(const Problem &pr)
{
    return DIM == 3 ? (pr.gather_split == 4 ? 4 : 8) : max(pr.fixed.n[1], 1);
}

// Workgroup -> (slab, control-point tile) of the 3D derivative gather.  With
// regions == 8 the tiles are split into eight regions of the (x, y) tile
// plane, rx x (8 / rx) of cw x ch tiles each, and workgroup b works in region
// b % 8: workgroups reach the eight XCDs in turn, so each XCD's L2 serves one
// region's samples.  Within a region (and without regions) the order is
// slab-major, so the workgroups in flight share a few sample planes.
struct GatherMap {
    int regions, rx, cw, ch, cap;  // cap: tiles per region (padded), cw * ch * tiles along z
};

// Union of the reference windows of one control-point tile along the last
// axis (sample cells); empty when lo > hi.
template <int DIM>
__device__ __forceinline__ void tile_plane_range(const Problem &pr, int t0, int t1, int t2, int &lo, int &hi)
{
    constexpr int TX = CpTile<DIM>::X, TY = CpTile<DIM>::Y, TZ = CpTile<DIM>::Z, T = TX * TY * TZ;
    lo = INT_MAX;
    hi = INT_MIN;
#pragma unroll
    for (int c = 0; c < T; ++c) {
        const int jc[3] = {t0 * TX + c % TX, t1 * TY + (c / TX) % TY, t2 * TZ + c / (TX * TY)};
        if (jc[0] < pr.cp.n[0] && jc[1] < pr.cp.n[1] && (DIM == 2 || jc[2] < pr.cp.n[2])) {
            const long long j = jc[0] + (long long)pr.cp.n[0] * (jc[1] + (long long)pr.cp.n[1] * jc[2]);
            lo = min(lo, pr.cp_window[j * 6 + DIM - 1]);
            hi = max(hi, pr.cp_window[j * 6 + 3 + DIM - 1]);
        }
    }
}
__device__ __forceinline__ int wave_min_i(int v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ int wave_max_i(int v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ float wave_max_f(float v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

// End of a derivative-gather workgroup: the tile points' sums of every lane
// added by a fixed FP32 butterfly per wave, then over the four waves in FP64,
// stored as the (slab, tile) partials that k_gather_finish adds in slab order.
template <int DIM>
__device__ __forceinline__ void gather_tile_store(const Problem &pr,
                                                  float (&acc)[CpTile<DIM>::X * CpTile<DIM>::Y * CpTile<DIM>::Z][3],
                                                  double *lds, int t0, int t1, int t2, int tile, int slab_id,
                                                  int n_tiles)
{
    constexpr int TX = CpTile<DIM>::X, TY = CpTile<DIM>::Y, TZ = CpTile<DIM>::Z, T = TX * TY * TZ;
    // wave sums (fixed butterfly, FP32), then the four waves in FP64
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < T; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            float v = acc[c][d];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if (lane == 0) lds[wave * T * 3 + c * 3 + d] = v;
        }
    __syncthreads();
    if (threadIdx.x < T * 3) {
        const int c = threadIdx.x / 3, d = threadIdx.x % 3;
        double sum = 0.0;
        for (int w = 0; w < kThreads / 64; ++w) sum += lds[w * T * 3 + threadIdx.x];
        if (d < DIM) {
            const int jc[3] = {t0 * TX + c % TX, t1 * TY + (c / TX) % TY, t2 * TZ + c / (TX * TY)};
            const bool inside = jc[0] < pr.cp.n[0] && jc[1] < pr.cp.n[1] && (DIM == 2 || jc[2] < pr.cp.n[2]);
            if (inside) pr.gather_partials[((long long)slab_id * n_tiles + tile) * (T * 3) + c * 3 + d] = sum;
        }
    }
}

// The value pass's tail for one sample (evaluateValueMSE, GpuEvaluatorMSE.cu:81-95;
// evaluateValueNCC, GpuEvaluatorNCC.cu:83-109): the moving image and its
// gradient at x + u, the stores (gradM, a) and b with a = M - F (MSE) or a = F,
// b = M (NCC), zeros outside, and the sample's share of the FP64 sums.
template <int DIM, int MEASURE>
__device__ __forceinline__ void value_sample(const Problem &pr, long long s, const int *pos, const float *x,
                                             const float *u, float4 *__restrict__ sg, float *__restrict__ sb,
                                             double *acc)
{
    float y[3];
#pragma unroll
    for (int d = 0; d < DIM; ++d) y[d] = x[d] + u[d];
    float val = 0.0f, g[3] = {0.0f, 0.0f, 0.0f};
    bool in = bspline_sample<DIM>(pr.coeffs, pr.moving, y, val, g);
    float a = 0.0f, b = 0.0f;
    if (in) {
        float f = fixed_value<DIM>(pr, pos);
        if (MEASURE == SKM_MEASURE_MSE) a = val - f;
        else { a = f; b = val; }
    } else {
        g[0] = g[1] = g[2] = 0.0f;
    }
    sg[s] = make_float4(g[0], g[1], DIM == 3 ? g[2] : 0.0f, a);
    if (MEASURE == SKM_MEASURE_NCC) sb[s] = b;
    double da = a, db = b;
    if (MEASURE == SKM_MEASURE_MSE) {
        acc[0] = da * da;
    } else {
        acc[0] = da; acc[1] = db; acc[2] = da * da; acc[3] = db * db; acc[4] = da * db;
    }
    acc[6] = in ? 1.0 : 0.0;
}

}  // namespace skm
