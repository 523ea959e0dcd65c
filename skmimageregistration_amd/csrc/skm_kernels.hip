// skm_kernels.hip -- gfx950 kernels of the SKM kernel-sum / MSE / NCC /
// regularizer path and the device-resident optimizer step.
//
// Reference counterparts (CUDA, one thread per voxel or per control point on
// 4x4x4 / 16x16 blocks): include/cuda/GpuEvaluatorMSE.cu, GpuEvaluatorNCC.cu,
// GpuEvaluatorRegularizers.cu, GpuEvaluatorDisplacementField.cu and the host
// loops of GpuEvaluator.cu.  Here the value pass runs on 256-lane workgroups
// whose four wave64 each cover a 4x4x4 (3D) or 16x4 (2D) sample brick, the
// derivative is a per-control-point gather by a whole workgroup over the
// sample window with rows pruned to the kernel ball, and every reduction is a
// fixed-order FP64 tree (per-block partials, then one finishing block).
#include <cfloat>
#include <climits>
#include <type_traits>

#include "skm_internal.hpp"
#include "skm_walk.hpp"

#include "skm.h"

namespace skm {

// Kernel sum u(x) = sum_j w_j k(x, p_j) c_j over the control points of the
// reference window of x (GpuEvaluatorMSE.cu:61-80).  Rows farther from x than
// any partner's kernel reaches (the support in the stationary modes,
// pair_reach2 in the weighted modes) only hold k == 0 and are skipped.
template <int DIM, int MODE>
__device__ __forceinline__ void kernel_sum(const Problem &pr, const float *x, const PState &st, float *u,
                                           int split = 1, int part = 0)
{
    float ci[3];
    to_local<DIM>(pr.cp, x, ci);
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        float r = st.support * pr.cp.ih[d];
        lo[d] = clampi((int)(ci[d] - r + 1.0f), 0, pr.cp.n[d] - 1);
        hi[d] = clampi((int)(ci[d] + r), 0, pr.cp.n[d] - 1);
    }
    const float s2 = is_stationary(MODE) ? st.support * st.support
                                         : pair_reach2(pr.kc.smin, pstate_lam<DIM, MODE>(st), pr.lam_cp_max);
    const int z0 = DIM == 3 ? lo[2] : 0, z1 = DIM == 3 ? hi[2] : 0;
    int row_index = 0;
    for (int cz = z0; cz <= z1; ++cz) {
        float dz = DIM == 3 ? x[2] - (pr.cp.o[2] + pr.cp.h[2] * cz) : 0.0f;
        for (int cy = lo[1]; cy <= hi[1]; ++cy) {
            // with split > 1 this call takes every split-th row of the window
            if (split > 1 && row_index++ % split != part) continue;
            float dy = x[1] - (pr.cp.o[1] + pr.cp.h[1] * cy);
            float rem = s2 - dy * dy - dz * dz;
            if (rem <= 0.0f) continue;
            // lo >= 0 makes truncation act as floor; +1 covers the ceiling
            float rx = fast_sqrt(rem) * pr.cp.ih[0];
            int xa = max(lo[0], (int)(ci[0] - rx));
            int xb = min(hi[0], (int)(ci[0] + rx) + 1);
            long long row = (long long)pr.cp.n[0] * (cy + (long long)pr.cp.n[1] * cz);
            if constexpr (is_stationary(MODE)) {
                const float c_row = fmaf(dy, dy, dz * dz);
                float dx = x[0] - (pr.cp.o[0] + pr.cp.h[0] * (float)xa);
                const float4 *src = pr.cpc + row;
                for (int cx = xa; cx <= xb; ++cx, dx -= pr.cp.h[0]) {
                    float k = stationary_k<MODE>(fmaf(dx, dx, c_row), pr.kc.inv_s4, pr.kc.inv_s0);
                    float4 c = src[cx];
                    float wk = c.w * k;
                    u[0] = fmaf(wk, c.x, u[0]);
                    u[1] = fmaf(wk, c.y, u[1]);
                    if (DIM == 3) u[2] = fmaf(wk, c.z, u[2]);
                }
                continue;
            }
            for (int cx = xa; cx <= xb; ++cx) {
                float v[3] = {x[0] - (pr.cp.o[0] + pr.cp.h[0] * cx), dy, dz};
                PState cst;
                cp_state_load<DIM, MODE>(pr, row + cx, cst);
                float k = kernel_pair<DIM, MODE>(pr.kc, v, st, cst);
                float4 c = pr.cpc[row + cx];
                float wk = c.w * k;
                u[0] = fmaf(wk, c.x, u[0]);
                u[1] = fmaf(wk, c.y, u[1]);
                if (DIM == 3) u[2] = fmaf(wk, c.z, u[2]);
            }
        }
    }
}

// Kernel sum of the weighted modes with a wave-uniform control-point loop:
// the wave walks the union of its lanes' reference windows row by row, so
// every control point's state and coefficients are scalar loads shared by
// the 64 lanes; a lane adds a pair only inside its own window and reach.
// Every lane of the wave must call this (inactive lanes pass active = false).
template <int DIM, int MODE>
__device__ __forceinline__ void kernel_sum_wave(const Problem &pr, const float *x, const PState &st, float lam_x,
                                                bool active, float *u, int split = 1, int part = 0)
{
    float ci[3];
    to_local<DIM>(pr.cp, x, ci);
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        float r = st.support * pr.cp.ih[d];
        lo[d] = active ? clampi((int)(ci[d] - r + 1.0f), 0, pr.cp.n[d] - 1) : INT_MAX;
        hi[d] = active ? clampi((int)(ci[d] + r), 0, pr.cp.n[d] - 1) : INT_MIN;
    }
    int ulo[3] = {0, 0, 0}, uhi[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        ulo[d] = __builtin_amdgcn_readfirstlane(wave_min_i(lo[d]));
        uhi[d] = __builtin_amdgcn_readfirstlane(wave_max_i(hi[d]));
    }
    // pair reach: |v|^2 < smin^2 (lam_x + lam_c) / 2 (x 1.0001), as the sum of
    // the sample's share and the control point's pre-scaled one (cp_reach)
    const float qx = pair_reach2(pr.kc.smin, lam_x, 0.0f);
    // the wave's samples: bounding box and largest reach share (once per wave;
    // each row's range of control points follows from them without per-row
    // cross-lane reductions)
    float bmin[3] = {0.0f, 0.0f, 0.0f}, bmax[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        bmin[d] = -wave_max_f(active ? -x[d] : -FLT_MAX);
        bmax[d] = wave_max_f(active ? x[d] : -FLT_MAX);
    }
    const float qx_max = wave_max_f(active ? qx : 0.0f);
    int row_index = 0;
    for (int cz = ulo[2]; cz <= uhi[2]; ++cz) {
        const float pz = DIM == 3 ? pr.cp.o[2] + pr.cp.h[2] * cz : 0.0f;
        const float dz = DIM == 3 ? x[2] - pz : 0.0f;
        const float dzb = DIM == 3 ? fmaxf(0.0f, fmaxf(bmin[2] - pz, pz - bmax[2])) : 0.0f;
        const bool zin = DIM == 2 || (cz >= lo[2] && cz <= hi[2]);
        for (int cy = ulo[1]; cy <= uhi[1]; ++cy) {
            // with split > 1 this wave takes every split-th row of the union
            if (split > 1 && row_index++ % split != part) continue;
            const float py = pr.cp.o[1] + pr.cp.h[1] * cy;
            const float dyb = fmaxf(0.0f, fmaxf(bmin[1] - py, py - bmax[1]));
            // largest pair reach in this row, from the box: control points
            // farther along x than rx from the box are out of every lane's reach
            const float rem = qx_max + pair_reach2(pr.kc.smin, 0.0f, pr.cp_row_lam[cy + (long long)pr.cp.n[1] * cz]) -
                              dyb * dyb - dzb * dzb;
            if (!(rem > 0.0f)) continue;
            const float rx = fast_sqrt(rem);
            const int uxa = __builtin_amdgcn_readfirstlane(
                max(ulo[0], (int)floorf((bmin[0] - rx - pr.cp.o[0]) * pr.cp.ih[0])));
            const int uxb = __builtin_amdgcn_readfirstlane(
                min(uhi[0], (int)ceilf((bmax[0] + rx - pr.cp.o[0]) * pr.cp.ih[0])));
            // this lane's window along x, empty outside its rows
            const bool rin = active && zin && cy >= lo[1] && cy <= hi[1];
            const int xa = rin ? lo[0] : INT_MAX, xb = rin ? hi[0] : INT_MIN;
            const float dy = x[1] - py;
            const long long row = (long long)pr.cp.n[0] * (cy + (long long)pr.cp.n[1] * cz);
            const RowTerms rt = row_terms<DIM>(dy, dz);
            float fcx = (float)uxa;
            for (int cx = uxa; cx <= uxb; ++cx, fcx += 1.0f) {
                const float qc = pr.cp_reach[row + cx];
                const float dx = x[0] - fmaf(pr.cp.h[0], fcx, pr.cp.o[0]);
                // window and this pair's own reach (both lams, tighter than the
                // row's) as one predicate: one branch per control point
                if ((cx >= xa) & (cx <= xb) & (fmaf(dx, dx, rt.yz2) < qx + qc)) {
                    PState cst;
                    cp_state_load<DIM, MODE>(pr, row + cx, cst);
                    const float4 c = pr.cpc[row + cx];
                    const float wk = c.w * weighted_k_row<DIM, MODE>(pr.kc, dx, rt, st, cst);
                    u[0] = fmaf(wk, c.x, u[0]);
                    u[1] = fmaf(wk, c.y, u[1]);
                    if (DIM == 3) u[2] = fmaf(wk, c.z, u[2]);
                }
            }
        }
    }
}

// ------------------------------------------------------------ value pass

// MSE / NCC value pass (evaluateValueMSE, GpuEvaluatorMSE.cu:29-95, 174-237;
// evaluateValueNCC, GpuEvaluatorNCC.cu:37-109, 194-263).  Per sample it stores
// (gradM, a) and b with a = M - F (MSE) or a = F, b = M (NCC), zeros outside,
// and adds its share of the FP64 sums to the block partials.
// Eight waves per SIMD: the per-pair loop is latency bound (measured faster
// than the compiler's own register budget, which stops at seven).
// SPLIT == 1: 256 threads, each wave a brick of its own.  SPLIT > 1 (small
// sample grids, too few waves to fill the chip): SPLIT waves share one brick,
// each takes every SPLIT-th row of control points, and their partial sums
// are added in wave order before wave 0 finishes the samples.
template <int DIM, int MODE, int MEASURE, int SPLIT>
__global__ __launch_bounds__(SPLIT > 1 ? 64 * SPLIT : kThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_value(
    Problem pr, float4 *__restrict__ sg, float *__restrict__ sb, double *__restrict__ partials)
{
    constexpr int WAVES = SPLIT > 1 ? SPLIT : kThreads / 64;
    __shared__ double lds[kPartialsK * WAVES];
    __shared__ float lds_u[SPLIT > 1 ? SPLIT * 64 * 3 : 1];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int l[3];
    tile_coords<DIM>(SPLIT > 1 ? lane : (int)threadIdx.x, l);
    constexpr int BX = SPLIT > 1 ? WaveBrick<DIM>::X : Tile<DIM>::X, BY = SPLIT > 1 ? WaveBrick<DIM>::Y : Tile<DIM>::Y,
                  BZ = SPLIT > 1 ? WaveBrick<DIM>::Z : Tile<DIM>::Z;
    int m[3] = {(int)blockIdx.x * BX + l[0], (int)blockIdx.y * BY + l[1], DIM == 3 ? (int)blockIdx.z * BZ + l[2] : 0};
    bool valid = m[0] < pr.nsub[0] && m[1] < pr.nsub[1] && (DIM == 2 || m[2] < pr.nsub[2]);
    double acc[kPartialsK];
#pragma unroll
    for (int k = 0; k < kPartialsK; ++k) acc[k] = 0.0;
    // lanes past the sample grid run the kernel sum of sample 0 (the weighted
    // modes' wave-uniform loop needs every lane) and store nothing
    if (!valid) m[0] = m[1] = m[2] = 0;
    const long long s = m[0] + (long long)pr.nsub[0] * (m[1] + (long long)pr.nsub[1] * m[2]);
    int pos[3] = {0, 0, 0};
    sample_position<DIM>(pr, m, s, pos);
    float fp[3] = {(float)pos[0], (float)pos[1], (float)pos[2]}, x[3];
    to_global<DIM>(pr.fixed, fp, x);
    PState st;
    point_state<DIM, MODE>(pr.kc, pr.weight, pr.weight_data, x, st);
    float u[3];
    df_lookup<DIM>(pr, x, u);
    // with a split the previous displacement is wave 0's share only
    if (SPLIT > 1 && wave > 0) u[0] = u[1] = u[2] = 0.0f;
    float lam_x = 0.0f;
    if constexpr (is_stationary(MODE)) {
        kernel_sum<DIM, MODE>(pr, x, st, u, SPLIT, wave);
    } else {
        lam_x = pstate_lam<DIM, MODE>(st);
        kernel_sum_wave<DIM, MODE>(pr, x, st, lam_x, valid, u, SPLIT, wave);
    }
    if constexpr (SPLIT > 1) {
#pragma unroll
        for (int d = 0; d < 3; ++d) lds_u[(wave * 64 + lane) * 3 + d] = u[d];
        __syncthreads();
        if (wave == 0) {
#pragma unroll
            for (int w = 1; w < SPLIT; ++w)
#pragma unroll
                for (int d = 0; d < 3; ++d) u[d] += lds_u[(w * 64 + lane) * 3 + d];
        }
        valid = valid && wave == 0;
    }
    if (valid) {
        if (!is_stationary(MODE)) {
            // the derivative gather reads the sample's state from here
            float4 a, b;
            pstate_pack<DIM, MODE>(st, lam_x, a, b);
            pr.smp_pstate[2 * s] = a;
            pr.smp_pstate[2 * s + 1] = b;  // .w = lam, read by the gather's brick test
        }
        value_sample<DIM, MEASURE>(pr, s, pos, x, u, sg, sb, acc);
    }
    block_sum<kPartialsK, WAVES>(acc, lds);
    if (threadIdx.x == 0) {
        long long b = block_linear();
#pragma unroll
        for (int k = 0; k < kPartialsK; ++k) partials[b * kPartialsK + k] = acc[k];
    }
}

// Host reduction of GpuEvaluator.cu:192-200 (MSE), 229-262 / 501-534 (NCC) and
// 287-297 (LCC), done by one block in a fixed order.
__global__ __launch_bounds__(kThreads) void k_reduce(int measure, const double *__restrict__ partials, int blocks,
                                                     Scalars *sc, uint64_t *rng, uint64_t rng_delta, uint64_t rng_inc,
                                                     int keep_reg)
{
    __shared__ double lds[kPartialsK * 4];
    double acc[kPartialsK];
#pragma unroll
    for (int k = 0; k < kPartialsK; ++k) acc[k] = 0.0;
    for (int b = threadIdx.x; b < blocks; b += kThreads)
#pragma unroll
        for (int k = 0; k < kPartialsK; ++k) acc[k] += partials[(long long)b * kPartialsK + k];
    block_sum<kPartialsK>(acc, lds);
    if (threadIdx.x != 0) return;
    double n = acc[6];
    sc->count = n;
    for (int k = 0; k < 6; ++k) sc->sums[k] = acc[k];
    sc->mean_f = sc->mean_m = sc->rho = sc->inv_denom = 0.0;
    if (measure == SKM_MEASURE_MSE) {
        sc->measure = acc[0] / n;
        sc->coef_scale = 2.0 / n;
    } else if (measure == SKM_MEASURE_NCC) {
        double sf = acc[0], sm = acc[1], sff = acc[2], smm = acc[3], sfm = acc[4];
        sff -= sf * sf / n;
        smm -= sm * sm / n;
        sfm -= sf * sm / n;
        double denom = -1.0 * sqrt(sff * smm);
        sc->measure = denom != 0.0 ? sfm / denom : 0.0;
        sc->mean_f = sf / n;
        sc->mean_m = sm / n;
        if (denom != 0.0) {
            sc->rho = sfm / smm;
            sc->inv_denom = 1.0 / denom;
        }
        sc->coef_scale = 0.0;
    } else {
        sc->measure = acc[0] / n;
        sc->coef_scale = 1.0 / n;
    }
    if (!keep_reg) sc->reg[0] = sc->reg[1] = sc->reg[2] = sc->reg[3] = 0.0;
    // past the offsets this evaluation drew (k_offsets has finished: the
    // value pass ran in between)
    if (rng_delta) *rng = lcg_advance(*rng, rng_delta, rng_inc);
}

// Per-sample derivative vector v = coef * gradM with coef folding the host
// normalisation (GpuEvaluator.cu:602-603) or the NCC combination
// (dF - sfm/smm dM)/denom (GpuEvaluator.cu:628-633; GpuEvaluatorNCC.cu:177-179).
// The jitter offset rides in .w as packed bytes.
template <int DIM, int MEASURE>
__global__ __launch_bounds__(kThreads) void k_prepare(Problem pr, const float4 *__restrict__ sg,
                                                      const float *__restrict__ sb, const Scalars *__restrict__ sc,
                                                      float4 *__restrict__ sv)
{
    long long s = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (s >= pr.n_sub) return;
    float4 g = sg[s];
    float coef;
    if (MEASURE == SKM_MEASURE_MSE) {
        coef = (float)(sc->coef_scale * (double)g.w);
    } else {
        double a = g.w, b = sb[s];
        coef = (float)(((a - sc->mean_f) - sc->rho * (b - sc->mean_m)) * sc->inv_denom);
    }
    unsigned packed = 0;
    if (pr.sub > 1) {
#pragma unroll
        for (int d = 0; d < DIM; ++d) packed |= (unsigned)pr.offsets[s * DIM + d] << (8 * d);
    }
    sv[s] = make_float4(coef * g.x, coef * g.y, coef * g.z, __uint_as_float(packed));
}

// ------------------------------------------------------------ derivative gather

// Derivative gather, d_j = w_j * sum_{samples m in window(j)} k(p_j, x_m) v_m
// (evaluateDerivativeMSE, GpuEvaluatorMSE.cu:97-170, 239-309;
// evaluateDerivativeNCC, GpuEvaluatorNCC.cu:111-190, 265-340; the windows,
// reference or exact, computed on the host in FP64).  One workgroup per
// (slab of sample planes, tile of control points), slab-major.  The
// workgroup walks the rows of the union of the tile's windows inside its
// slab; in 3D each wave takes 4 x 4 (y, z) rows with 4 lanes along x (a
// 4 x 4 x 4 block of samples per step), in 2D 8 rows with 8 lanes along x.
// Each lane loads a sample (v, and in the weighted modes the state the value
// pass stored) once and adds its pair with every tile control point whose
// window and reach hold it (weighted modes: both points' lam; stationary
// modes: the support).  The tile constants are wave-uniform (scalar
// registers).  Six waves per SIMD (measured faster than five with the next
// sample's loads issued early, and than seven or eight, which spill).  Each
// workgroup writes its FP64 sums; k_gather_finish adds the slabs in order.
template <int DIM, int MODE>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(6, 6))) void k_gather_tile(Problem pr, const float4 *__restrict__ sv, int n_tiles, GatherMap gm)
{
    constexpr int TX = CpTile<DIM>::X, TY = CpTile<DIM>::Y, TZ = CpTile<DIM>::Z, T = TX * TY * TZ;
    __shared__ double lds[T * 3 * 4];
    const int ntx = (pr.cp.n[0] + TX - 1) / TX, nty = (pr.cp.n[1] + TY - 1) / TY;
    int tile, slab_id, t0, t1, t2;
    if (DIM == 3 && gm.regions == 8) {
        const int region = (int)(blockIdx.x % 8u);
        const unsigned k = blockIdx.x / 8u;
        slab_id = (int)(k / (unsigned)gm.cap);
        const int li = (int)(k % (unsigned)gm.cap);
        t0 = (region % gm.rx) * gm.cw + li % gm.cw;
        t1 = (region / gm.rx) * gm.ch + (li / gm.cw) % gm.ch;
        t2 = li / (gm.cw * gm.ch);
        // padding of the last regions (uniform over the workgroup)
        if (t0 >= ntx || t1 >= nty) return;
        tile = t0 + ntx * (t1 + nty * t2);
    } else {
        tile = (int)(blockIdx.x % (unsigned)n_tiles);
        slab_id = (int)(blockIdx.x / (unsigned)n_tiles);
        t0 = tile % ntx;
        t1 = (tile / ntx) % nty;
        t2 = tile / (ntx * nty);
    }
    const int slab = gather_slab<DIM>(pr);
    int plo, phi;
    tile_plane_range<DIM>(pr, t0, t1, t2, plo, phi);
    // slabs outside the tile's windows write nothing (k_gather_finish skips them)
    if (plo > phi || slab_id < plo / slab || slab_id > phi / slab) return;
    // tile control points (uniform)
    long long jj[T];
    bool ok[T];
    float px[T], py[T], pz[T], pc0[T], lam[T];
    int lo[T][3], hi[T][3];
    PState st[T];
    int ulo[3] = {INT_MAX, INT_MAX, DIM == 3 ? INT_MAX : 0}, uhi[3] = {INT_MIN, INT_MIN, DIM == 3 ? INT_MIN : 0};
#pragma unroll
    for (int c = 0; c < T; ++c) {
        const int jc[3] = {t0 * TX + c % TX, t1 * TY + (c / TX) % TY, t2 * TZ + c / (TX * TY)};
        ok[c] = jc[0] < pr.cp.n[0] && jc[1] < pr.cp.n[1] && (DIM == 2 || jc[2] < pr.cp.n[2]);
        jj[c] = ok[c] ? jc[0] + (long long)pr.cp.n[0] * (jc[1] + (long long)pr.cp.n[1] * jc[2]) : 0;
        const float fj[3] = {(float)jc[0], (float)jc[1], (float)jc[2]};
        float pos[3] = {0.0f, 0.0f, 0.0f}, pcl[3] = {0.0f, 0.0f, 0.0f};
        to_global<DIM>(pr.cp, fj, pos);
        to_local<DIM>(pr.fixed, pos, pcl);
        px[c] = pos[0];
        py[c] = pos[1];
        pz[c] = pos[2];
        pc0[c] = pcl[0];
        // stationary modes: a reach share of 1 for every point makes the
        // reach tests |v|^2 < smin^2 (x 1.0001), where their kernel ends
        lam[c] = is_stationary(MODE) ? 1.0f : cp_state_load<DIM, MODE>(pr, jj[c], st[c]);
        const int *win = pr.cp_window + jj[c] * 6;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            lo[c][d] = d < DIM ? win[d] : 0;
            hi[c][d] = d < DIM ? win[d + 3] : 0;
        }
        if (ok[c]) {
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                ulo[d] = min(ulo[d], lo[c][d]);
                uhi[d] = max(uhi[d], hi[c][d]);
            }
        }
    }
    // this slab's planes of the union
    ulo[DIM - 1] = max(ulo[DIM - 1], slab_id * slab);
    uhi[DIM - 1] = min(uhi[DIM - 1], slab_id * slab + slab - 1);
    const int by = uhi[1] - ulo[1] + 1, bz = uhi[2] - ulo[2] + 1;
    // (a window can be empty along an axis: then there is nothing to gather)
    const int rows = (uhi[0] >= ulo[0] && by > 0 && bz > 0) ? by * bz : 0;
    // lanes: LX along x, the rest over rows; in 3D a wave takes 4 x 4 (y, z)
    // rows at a time, so its 64 samples form a 4 x 4 x 4 block
    constexpr int LX = DIM == 3 ? 4 : 8;
    const int tx = threadIdx.x % LX;
    const int wv = threadIdx.x >> 6, lr = (threadIdx.x & 63) / LX;
    const int sub = pr.sub;
    const float span = sub > 1 ? (float)sub : 0.0f;
    const float inv_sub = 1.0f / (float)sub;
    const float smin2h = 0.5f * pr.kc.smin * pr.kc.smin * 1.0001f;
    float acc[T][3];
#pragma unroll
    for (int c = 0; c < T; ++c) acc[c][0] = acc[c][1] = acc[c][2] = 0.0f;
    constexpr int RY = DIM == 3 ? 4 : 64 / LX, RZ = 64 / LX / RY;
    const int nby = rows ? (by + RY - 1) / RY : 0, nbz = rows ? (bz + RZ - 1) / RZ : 0;
    for (int blk = wv; blk < nby * nbz; blk += kThreads / 64) {
        const int my = ulo[1] + (blk % nby) * RY + lr % RY, mz = ulo[2] + (blk / nby) * RZ + lr / RY;
        if (my > uhi[1] || mz > uhi[2]) continue;
        const float y0 = pr.fixed.o[1] + pr.fixed.h[1] * (float)(my * sub);
        const float y1 = y0 + pr.fixed.h[1] * span;
        const float z0 = DIM == 3 ? pr.fixed.o[2] + pr.fixed.h[2] * (float)(mz * sub) : 0.0f;
        const float z1 = z0 + (DIM == 3 ? pr.fixed.h[2] * span : 0.0f);
        const float lam_row = is_stationary(MODE) ? 1.0f : pr.smp_row_lam[my + (long long)pr.nsub[1] * mz];
        // per tile control point: x cells of this row inside its window and reach
        int xa[T], xb[T];
        int rxa = INT_MAX, rxb = INT_MIN;
#pragma unroll
        for (int c = 0; c < T; ++c) {
            xa[c] = INT_MAX;
            xb[c] = INT_MIN;
            const bool rin = ok[c] && my >= lo[c][1] && my <= hi[c][1] &&
                             (DIM == 2 || (mz >= lo[c][2] && mz <= hi[c][2]));
            const float dy = fmaxf(0.0f, fmaxf(y0 - py[c], py[c] - y1));
            const float dz = DIM == 3 ? fmaxf(0.0f, fmaxf(z0 - pz[c], pz[c] - z1)) : 0.0f;
            // reach of the control point over this row's samples
            const float rem = smin2h * (lam[c] + lam_row) - dy * dy - dz * dz;
            if (rin && rem > 0.0f) {
                const float rx = fast_sqrt(rem) * pr.fixed.ih[0];
                xa[c] = max(lo[c][0], (int)((pc0[c] - rx - span) * inv_sub) - 1);
                xb[c] = min(hi[c][0], (int)((pc0[c] + rx) * inv_sub) + 1);
                rxa = min(rxa, xa[c]);
                rxb = max(rxb, xb[c]);
            }
        }
        // no tile control point reaches this row (rxa is still INT_MAX: rxa + tx
        // would overflow)
        if (rxa > rxb) continue;
        const long long rowbase = (long long)pr.nsub[0] * (my + (long long)pr.nsub[1] * mz);
        for (int mx = rxa + tx; mx <= rxb; mx += LX) {
            const float4 cv = sv[rowbase + mx];
            float4 cb = make_float4(0.0f, 0.0f, 0.0f, 1.0f);  // .w: the sample's reach share
            PState sst;
            if constexpr (!is_stationary(MODE)) {
                const float4 ca = pr.smp_pstate[2 * (rowbase + mx)];
                cb = pr.smp_pstate[2 * (rowbase + mx) + 1];
                pstate_unpack<DIM, MODE>(ca, cb, sst);
            }
            int pos[3] = {mx, my, mz};
            if (sub > 1) {
                const unsigned pk = __float_as_uint(cv.w);
                pos[0] = mx * sub + (int)(pk & 0xff);
                pos[1] = my * sub + (int)((pk >> 8) & 0xff);
                pos[2] = mz * sub + (int)((pk >> 16) & 0xff);
            }
            const float fpos[3] = {(float)pos[0], (float)pos[1], (float)pos[2]};
            float xs[3];
            to_global<DIM>(pr.fixed, fpos, xs);
#pragma unroll
            for (int c = 0; c < T; ++c) {
                const float v[3] = {px[c] - xs[0], py[c] - xs[1], DIM == 3 ? pz[c] - xs[2] : 0.0f};
                const float d2 = v[0] * v[0] + v[1] * v[1] + (DIM == 3 ? v[2] * v[2] : 0.0f);
                // the pair's own reach: both points' lam
                if ((mx >= xa[c]) & (mx <= xb[c]) & (d2 < smin2h * (lam[c] + cb.w))) {
                    const float k = kernel_pair<DIM, MODE>(pr.kc, v, st[c], sst);
                    acc[c][0] = fmaf(k, cv.x, acc[c][0]);
                    acc[c][1] = fmaf(k, cv.y, acc[c][1]);
                    if (DIM == 3) acc[c][2] = fmaf(k, cv.z, acc[c][2]);
                }
            }
        }
    }
    gather_tile_store<DIM>(pr, acc, lds, t0, t1, t2, tile, slab_id, n_tiles);
}

// deriv_j = w_j * the sum over the slabs of j's tile, in slab order.
template <int DIM>
__global__ __launch_bounds__(kThreads) void k_gather_finish(Problem pr, int n_tiles, float *__restrict__ deriv)
{
    constexpr int TX = CpTile<DIM>::X, TY = CpTile<DIM>::Y, TZ = CpTile<DIM>::Z, T = TX * TY * TZ;
    const long long j = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (j >= pr.n_cp) return;
    const int jc[3] = {(int)(j % pr.cp.n[0]), (int)((j / pr.cp.n[0]) % pr.cp.n[1]),
                       (int)(j / ((long long)pr.cp.n[0] * pr.cp.n[1]))};
    const int ntx = (pr.cp.n[0] + TX - 1) / TX, nty = (pr.cp.n[1] + TY - 1) / TY;
    const int t0 = jc[0] / TX, t1 = jc[1] / TY, t2 = jc[2] / TZ;
    const long long tile = t0 + (long long)ntx * (t1 + (long long)nty * t2);
    const int c = jc[0] % TX + TX * (jc[1] % TY + TY * (jc[2] % TZ));
    const int slab = gather_slab<DIM>(pr);
    int plo, phi;
    tile_plane_range<DIM>(pr, t0, t1, t2, plo, phi);
    double sum[3] = {0.0, 0.0, 0.0};
    if (plo <= phi) {
        for (int s = plo / slab; s <= phi / slab; ++s) {
            const double *p = pr.gather_partials + ((long long)s * n_tiles + tile) * (T * 3) + c * 3;
#pragma unroll
            for (int d = 0; d < DIM; ++d) sum[d] += p[d];
        }
    }
    const double w = pr.cpc[j].w;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const float g = (float)(w * sum[d]);
        deriv[j + d * pr.n_cp] = pr.reg_add ? (float)((double)g + pr.reg_add[j + d * pr.n_cp]) : g;
    }
}

// ------------------------------------------------------------ regularizers

// RKHS, RD and PG in one pass over the cp window (regularizerRKHS / RD / PG,
// GpuEvaluatorRegularizers.cu:24-255, 259-472).  Values are unweighted per-cp
// sums; derivatives are weighted and added to the normalised metric
// derivative (GpuEvaluator.cu:652-711).  One wave per control point, lanes
// over the (x, y) plane of its window, z looped; wave sums in FP64.
//
// The bilateral term (no reference counterpart; the reference's closest term
// is RD above, and "radiometric" appears only in a comment,
// itkRegularizedImageToImageMetricv4.h:204): per control point j over the
// points i within 3 sigma_s,
//   B_ij = exp(-r_ij^2 / 2 sigma_s^2) exp(-(I_j - I_i)^2 / 2 sigma_i^2),
//   value_j = sum_i w_j B_ij |c_j - c_i|^2,
//   d/dc_j = sum_i 2 B_ij (w_j + w_i)(c_j - c_i)
// with r_ij from the index difference times the spacing and I the fixed
// image at the control points; its value goes to the fourth partial slot.
template <int DIM, int MODE>
__global__ __launch_bounds__(kThreads) void k_regularizers(Problem pr, double w_rkhs, double w_rd, double w_pg,
                                                           double pg_scale, BilateralArgs bl, int with_derivative,
                                                           float *__restrict__ deriv, double *__restrict__ partials)
{
    __shared__ double lds[4 * 4];
    const int lane = threadIdx.x & 63;
    long long j = blockIdx.x * (long long)(kThreads / 64) + (threadIdx.x >> 6);
    double vals[4] = {0.0, 0.0, 0.0, 0.0};
    if (j < pr.n_cp) {
        int jc[3] = {0, 0, 0};
        {
            long long r = j;
            jc[0] = (int)(r % pr.cp.n[0]); r /= pr.cp.n[0];
            jc[1] = (int)(r % pr.cp.n[1]); r /= pr.cp.n[1];
            jc[2] = (int)r;
        }
        float fj[3] = {(float)jc[0], (float)jc[1], (float)jc[2]}, p[3];
        to_global<DIM>(pr.cp, fj, p);
        PState st;
        cp_state_load<DIM, MODE>(pr, j, st);
        float4 cj4 = pr.cpc[j];
        float cj[3] = {cj4.x, cj4.y, cj4.z};
        float wj = cj4.w;
        const int *win = pr.reg_window + j * 6;
        float s = (float)pg_scale;
        float vr = 0.0f, vd = 0.0f, vp = 0.0f, vb = 0.0f;
        float dr[3] = {0, 0, 0}, dd[3] = {0, 0, 0}, dp[3] = {0, 0, 0}, db[3] = {0, 0, 0};
        int z0 = DIM == 3 ? win[2] : 0, z1 = DIM == 3 ? win[5] : 0;
        const int wx = win[3] - win[0] + 1, plane = (w_rkhs > 0 || w_rd > 0 || w_pg > 0) ? wx * (win[4] - win[1] + 1) : 0;
        for (int t = lane; t < plane; t += 64)
            for (int iz = z0; iz <= z1; ++iz) {
                    const int ix = win[0] + t % wx, iy = win[1] + t / wx;
                    long long i = ix + (long long)pr.cp.n[0] * (iy + (long long)pr.cp.n[1] * iz);
                    float fi[3] = {(float)ix, (float)iy, (float)iz}, q[3];
                    to_global<DIM>(pr.cp, fi, q);
                    float v[3] = {p[0] - q[0], p[1] - q[1], DIM == 3 ? p[2] - q[2] : 0.0f};
                    PState ist;
                    cp_state_load<DIM, MODE>(pr, i, ist);
                    float k = kernel_pair<DIM, MODE>(pr.kc, v, st, ist);
                    if (k == 0.0f) continue;
                    float4 ci4 = pr.cpc[i];
                    float ci[3] = {ci4.x, ci4.y, ci4.z};
                    float kw = k * wj;
                    if (w_rkhs > 0) {
                        float dot = ci[0] * cj[0] + ci[1] * cj[1] + (DIM == 3 ? ci[2] * cj[2] : 0.0f);
                        vr = fmaf(kw, dot, vr);
#pragma unroll
                        for (int d = 0; d < DIM; ++d) dr[d] = fmaf(2.0f * kw, ci[d], dr[d]);
                    }
                    if (w_rd > 0) {
                        float e[3] = {cj[0] - ci[0], cj[1] - ci[1], DIM == 3 ? cj[2] - ci[2] : 0.0f};
                        float l2 = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
                        vd = fmaf(kw, l2, vd);
#pragma unroll
                        for (int d = 0; d < DIM; ++d) dd[d] = fmaf(4.0f * kw, e[d], dd[d]);
                    }
                    if (w_pg > 0) {
                        if (DIM == 3) {
                            float c0 = ci[1] * cj[2] - ci[2] * cj[1];
                            float c1 = ci[2] * cj[0] - ci[0] * cj[2];
                            float c2 = ci[0] * cj[1] - ci[1] * cj[0];
                            float n2 = c0 * c0 + c1 * c1 + c2 * c2;
                            vp = fmaf(kw, s * s * 0.5f * logf(1.0f + n2 / s), vp);
                            float den = 2.0f * (n2 / s + 1.0f);
                            float q0 = 2.0f * ci[2] * (ci[2] * cj[0] - ci[0] * cj[2]) - 2.0f * ci[1] * (ci[0] * cj[1] - ci[1] * cj[0]);
                            float q1 = 2.0f * ci[0] * (ci[0] * cj[1] - ci[1] * cj[0]) - 2.0f * ci[2] * (ci[1] * cj[2] - ci[2] * cj[1]);
                            float q2 = 2.0f * ci[1] * (ci[1] * cj[2] - ci[2] * cj[1]) - 2.0f * ci[0] * (ci[2] * cj[0] - ci[0] * cj[2]);
                            dp[0] = fmaf(2.0f * kw, s * q0 / den, dp[0]);
                            dp[1] = fmaf(2.0f * kw, s * q1 / den, dp[1]);
                            dp[2] = fmaf(2.0f * kw, s * q2 / den, dp[2]);
                        } else {
                            float val = ci[0] * cj[1] - ci[1] * cj[0];
                            float val2 = val * val;
                            vp = fmaf(kw, s * s * 0.5f * logf(1.0f + val2 / s), vp);
                            float factor = val2 / s + 1.0f;
                            dp[0] = fmaf(2.0f * kw, -s * ci[1] * val / factor, dp[0]);
                            dp[1] = fmaf(2.0f * kw, s * ci[0] * val / factor, dp[1]);
                        }
                    }
                }
        if (bl.weight > 0) {
            int blo[3] = {0, 0, 0}, bhi[3] = {0, 0, 0};
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                blo[d] = clampi((int)((float)jc[d] - bl.rad[d] + 1.0f), 0, pr.cp.n[d] - 1);
                bhi[d] = clampi((int)((float)jc[d] + bl.rad[d]), 0, pr.cp.n[d] - 1);
            }
            const float ij = pr.cp_intensity[j];
            const int bx = bhi[0] - blo[0] + 1, bplane = bx * (bhi[1] - blo[1] + 1);
            const int bz0 = DIM == 3 ? blo[2] : 0, bz1 = DIM == 3 ? bhi[2] : 0;
            for (int t = lane; t < bplane; t += 64)
                for (int iz = bz0; iz <= bz1; ++iz) {
                    const int ix = blo[0] + t % bx, iy = blo[1] + t / bx;
                    const float v0 = (float)(jc[0] - ix) * pr.cp.h[0], v1 = (float)(jc[1] - iy) * pr.cp.h[1];
                    const float v2 = DIM == 3 ? (float)(jc[2] - iz) * pr.cp.h[2] : 0.0f;
                    const float r2 = v0 * v0 + v1 * v1 + v2 * v2;
                    if (!(r2 < bl.cut2)) continue;
                    const long long i = ix + (long long)pr.cp.n[0] * (iy + (long long)pr.cp.n[1] * iz);
                    const float di = ij - pr.cp_intensity[i];
                    const float b = expf(-r2 * bl.inv2ss) * expf(-di * di * bl.inv2si);
                    const float4 ci4 = pr.cpc[i];
                    const float e[3] = {cj[0] - ci4.x, cj[1] - ci4.y, DIM == 3 ? cj[2] - ci4.z : 0.0f};
                    vb = fmaf(wj * b, e[0] * e[0] + e[1] * e[1] + e[2] * e[2], vb);
                    const float f = 2.0f * b * (wj + ci4.w);
#pragma unroll
                    for (int d = 0; d < DIM; ++d) db[d] = fmaf(f, e[d], db[d]);
                }
        }
        vals[0] = vr; vals[1] = vd; vals[2] = vp; vals[3] = vb;
        if (with_derivative) {
            double add[3];
#pragma unroll
            for (int d = 0; d < DIM; ++d)
                add[d] = wave_sum(w_rkhs * (double)dr[d] + w_rd * (double)dd[d] + w_pg * (double)dp[d] +
                                  bl.weight * (double)db[d]);
            if (lane == 0)
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    if (pr.reg_add) pr.reg_add[j + d * pr.n_cp] = add[d];
                    else deriv[j + d * pr.n_cp] = (float)((double)deriv[j + d * pr.n_cp] + add[d]);
                }
        }
    }
    block_sum<4>(vals, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) partials[blockIdx.x * 4LL + k] = vals[k];
}

__global__ __launch_bounds__(kThreads) void k_reg_reduce(const double *__restrict__ partials, int blocks,
                                                         Scalars *sc)
{
    __shared__ double lds[4 * 4];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < blocks; b += kThreads)
        for (int k = 0; k < 4; ++k) acc[k] += partials[b * 4LL + k];
    block_sum<4>(acc, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) sc->reg[k] = acc[k];
}

// ------------------------------------------------------------ displacement

// Dense u(x) + previous field on the full fixed grid, AoS
// (evaluateDisplacement, GpuEvaluatorDisplacementField.cu:24-127).
template <int DIM, int MODE>
__global__ __launch_bounds__(kThreads) void k_displacement(Problem pr, float *__restrict__ dst)
{
    int l[3];
    tile_coords<DIM>(threadIdx.x, l);
    int m[3] = {(int)blockIdx.x * Tile<DIM>::X + l[0], (int)blockIdx.y * Tile<DIM>::Y + l[1],
                DIM == 3 ? (int)blockIdx.z * Tile<DIM>::Z + l[2] : 0};
    if (m[0] >= pr.fixed.n[0] || m[1] >= pr.fixed.n[1] || (DIM == 3 && m[2] >= pr.fixed.n[2])) return;
    float fm[3] = {(float)m[0], (float)m[1], (float)m[2]}, x[3];
    to_global<DIM>(pr.fixed, fm, x);
    PState st;
    point_state<DIM, MODE>(pr.kc, pr.weight, pr.weight_data, x, st);
    float u[3];
    df_lookup<DIM>(pr, x, u);
    kernel_sum<DIM, MODE>(pr, x, st, u);
    long long v = m[0] + (long long)pr.fixed.n[0] * (m[1] + (long long)pr.fixed.n[1] * m[2]);
#pragma unroll
    for (int d = 0; d < DIM; ++d) dst[v * DIM + d] = u[d];
}

// Packed states (pstate_pack) of the weighted modes at every fixed voxel,
// for gathers over the dense grid (the exact LCC derivative).
template <int DIM, int MODE>
__global__ __launch_bounds__(kThreads) void k_voxel_states(Problem pr, float4 *__restrict__ out)
{
    const long long v = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (v >= pr.n_vox) return;
    if (is_stationary(MODE)) return;
    int m[3] = {(int)(v % pr.fixed.n[0]), (int)((v / pr.fixed.n[0]) % pr.fixed.n[1]),
                (int)(v / ((long long)pr.fixed.n[0] * pr.fixed.n[1]))};
    float fm[3] = {(float)m[0], (float)m[1], (float)m[2]}, x[3];
    to_global<DIM>(pr.fixed, fm, x);
    PState st;
    point_state<DIM, MODE>(pr.kc, pr.weight, pr.weight_data, x, st);
    float4 a, b;
    pstate_pack<DIM, MODE>(st, pstate_lam<DIM, MODE>(st), a, b);
    out[2 * v] = a;
    out[2 * v + 1] = b;
}

// ------------------------------------------------------------ small kernels

// PCG32 cube offsets: cell i draws numbers i*D .. i*D+D-1 of the stream, so
// every thread jumps ahead in O(log n) and the result is bit-identical to the
// serial host loop (GpuEvaluator.cu:107-113; UniformRng.h:34-55).  The stream
// state lives in device memory (k_reduce of the same evaluation moves it past
// the n*D draws), so every launch of a recorded evaluation draws anew.
template <int DIM>
__global__ __launch_bounds__(kThreads) void k_offsets(long long n, int sub, const uint64_t *__restrict__ state,
                                                      uint64_t inc, int *__restrict__ out)
{
    long long i = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (i >= n) return;
    uint64_t st = lcg_advance(*state, (uint64_t)i * DIM, inc);
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        uint32_t r = pcg32_output(st);
        st = st * kPcgMult + inc;
        double f = (double)r / 4294967295.0;
        out[i * DIM + d] = (int)(f * (double)sub);
    }
}

__global__ __launch_bounds__(kThreads) void k_pack_cp(long long n, int dim, const float *__restrict__ cp,
                                                      const float *__restrict__ cpw, int soa,
                                                      float4 *__restrict__ cpc)
{
    long long j = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (j >= n) return;
    float c[3] = {0.0f, 0.0f, 0.0f};
    for (int d = 0; d < dim; ++d) c[d] = soa ? cp[j + d * n] : cp[j * dim + d];
    cpc[j] = make_float4(c[0], c[1], c[2], cpw[j]);
}

__global__ void k_result(const Scalars *__restrict__ sc, skm_result *__restrict__ out)
{
    out->measure = sc->measure;
    out->reg_value_rkhs = sc->reg[0];
    out->reg_value_rd = sc->reg[1];
    out->reg_value_pg = sc->reg[2];
    out->num_pixels_counted = (uint32_t)sc->count;
    out->reserved = 0;
    out->reg_value_bilateral = sc->reg[3];
}

// ------------------------------------------------------------ launchers

template <int DIM>
static dim3 tile_grid(const int *n)
{
    return dim3((n[0] + Tile<DIM>::X - 1) / Tile<DIM>::X, (n[1] + Tile<DIM>::Y - 1) / Tile<DIM>::Y,
                DIM == 3 ? (n[2] + Tile<DIM>::Z - 1) / Tile<DIM>::Z : 1);
}

static unsigned blocks_for(long long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

#define SKM_DISPATCH_DIM_MODE(pr, FN, ...)                                                                         \
    do {                                                                                                           \
        if ((pr).dim == 2) {                                                                                       \
            if ((pr).kc.mode == 0) FN<2, 0>(__VA_ARGS__);                                                          \
            else if ((pr).kc.mode == 1) FN<2, 1>(__VA_ARGS__);                                                     \
            else if ((pr).kc.mode == 2) FN<2, 2>(__VA_ARGS__);                                                     \
            else FN<2, kModeStationaryEqual>(__VA_ARGS__);                                                         \
        } else {                                                                                                   \
            if ((pr).kc.mode == 0) FN<3, 0>(__VA_ARGS__);                                                          \
            else if ((pr).kc.mode == 1) FN<3, 1>(__VA_ARGS__);                                                     \
            else if ((pr).kc.mode == 2) FN<3, 2>(__VA_ARGS__);                                                     \
            else FN<3, kModeStationaryEqual>(__VA_ARGS__);                                                         \
        }                                                                                                          \
    } while (0)

void launch_offsets(const Problem &pr, int *offsets, uint64_t *state, uint64_t inc, hipStream_t s)
{
    if (pr.dim == 2) k_offsets<2><<<blocks_for(pr.n_sub), kThreads, 0, s>>>(pr.n_sub, pr.sub, state, inc, offsets);
    else k_offsets<3><<<blocks_for(pr.n_sub), kThreads, 0, s>>>(pr.n_sub, pr.sub, state, inc, offsets);
}

void launch_pack_cp(const Problem &pr, const float *cp, const float *cpw, int soa, float4 *cpc, hipStream_t s)
{
    k_pack_cp<<<blocks_for(pr.n_cp), kThreads, 0, s>>>(pr.n_cp, pr.dim, cp, cpw, soa, cpc);
}

template <int DIM, int MODE>
static void value_impl(const Problem &pr, int measure, float4 *sg, float *sb, double *partials, int *blocks,
                       hipStream_t s)
{
    dim3 grid = tile_grid<DIM>(pr.nsub);
    // fewer than four waves per SIMD (1024 SIMDs on the chip): the waves of a
    // workgroup share a brick and split its control-point rows, 4, 8 or 16
    // ways, the fewest that reach 4096 waves
    constexpr long long kWavesWanted = 4096;
    const bool few = (long long)grid.x * grid.y * grid.z * (kThreads / 64) < kWavesWanted;
    if ((pr.value_split == 0 && few) || pr.value_split > 1) {
        grid = dim3((pr.nsub[0] + WaveBrick<DIM>::X - 1) / WaveBrick<DIM>::X,
                    (pr.nsub[1] + WaveBrick<DIM>::Y - 1) / WaveBrick<DIM>::Y,
                    DIM == 3 ? (pr.nsub[2] + WaveBrick<DIM>::Z - 1) / WaveBrick<DIM>::Z : 1);
        const long long bricks = (long long)grid.x * grid.y * grid.z;
        *blocks = (int)bricks;
        const int split = pr.value_split > 1 ? pr.value_split
                          : bricks * 4 >= kWavesWanted ? 4 : bricks * 8 >= kWavesWanted ? 8 : 16;
        auto go = [&](auto split_c) {
            constexpr int S = decltype(split_c)::value;
            if (measure == SKM_MEASURE_MSE)
                k_value<DIM, MODE, SKM_MEASURE_MSE, S><<<grid, 64 * S, 0, s>>>(pr, sg, sb, partials);
            else
                k_value<DIM, MODE, SKM_MEASURE_NCC, S><<<grid, 64 * S, 0, s>>>(pr, sg, sb, partials);
        };
        if (split == 4) go(std::integral_constant<int, 4>());
        else if (split == 8) go(std::integral_constant<int, 8>());
        else go(std::integral_constant<int, 16>());
        return;
    }
    *blocks = (int)(grid.x * grid.y * grid.z);
    if (measure == SKM_MEASURE_MSE)
        k_value<DIM, MODE, SKM_MEASURE_MSE, 1><<<grid, kThreads, 0, s>>>(pr, sg, sb, partials);
    else
        k_value<DIM, MODE, SKM_MEASURE_NCC, 1><<<grid, kThreads, 0, s>>>(pr, sg, sb, partials);
}

void launch_value(const Problem &pr, int measure, float4 *sg, float *sb, double *partials, int *blocks_out,
                  hipStream_t s)
{
    SKM_DISPATCH_DIM_MODE(pr, value_impl, pr, measure, sg, sb, partials, blocks_out, s);
}

void launch_reduce(const Problem &pr, int measure, const double *partials, int blocks, Scalars *sc, hipStream_t s,
                   uint64_t *rng, uint64_t rng_delta, uint64_t rng_inc, bool keep_reg)
{
    k_reduce<<<1, kThreads, 0, s>>>(measure, partials, blocks, sc, rng, rng_delta, rng_inc, keep_reg ? 1 : 0);
}

void launch_prepare(const Problem &pr, int measure, const float4 *sg, const float *sb, const Scalars *sc,
                    float4 *sv, hipStream_t s)
{
    unsigned g = blocks_for(pr.n_sub);
    if (pr.dim == 2) {
        if (measure == SKM_MEASURE_MSE) k_prepare<2, SKM_MEASURE_MSE><<<g, kThreads, 0, s>>>(pr, sg, sb, sc, sv);
        else k_prepare<2, SKM_MEASURE_NCC><<<g, kThreads, 0, s>>>(pr, sg, sb, sc, sv);
    } else {
        if (measure == SKM_MEASURE_MSE) k_prepare<3, SKM_MEASURE_MSE><<<g, kThreads, 0, s>>>(pr, sg, sb, sc, sv);
        else k_prepare<3, SKM_MEASURE_NCC><<<g, kThreads, 0, s>>>(pr, sg, sb, sc, sv);
    }
}

template <int DIM, int MODE>
static void gather_impl(const Problem &pr, const float4 *sv, float *deriv, hipStream_t s)
{
    {
        const long long tiles = (long long)((pr.cp.n[0] + CpTile<DIM>::X - 1) / CpTile<DIM>::X) *
                                ((pr.cp.n[1] + CpTile<DIM>::Y - 1) / CpTile<DIM>::Y) *
                                (DIM == 3 ? (pr.cp.n[2] + CpTile<DIM>::Z - 1) / CpTile<DIM>::Z : 1);
        const int slab = gather_slab<DIM>(pr);
        const long long slabs = (pr.nsub[DIM - 1] + slab - 1) / slab;
        GatherMap gm{1, 1, 1, 1, 1};
        long long blocks = tiles * slabs;
        if (DIM == 3 && pr.gather_regions == 8) {
            const int ntx = (pr.cp.n[0] + CpTile<DIM>::X - 1) / CpTile<DIM>::X;
            const int nty = (pr.cp.n[1] + CpTile<DIM>::Y - 1) / CpTile<DIM>::Y;
            const int ntz = (pr.cp.n[2] + CpTile<DIM>::Z - 1) / CpTile<DIM>::Z;
            gm.regions = 8;
            gm.rx = ntx >= nty ? 4 : 2;
            gm.cw = (ntx + gm.rx - 1) / gm.rx;
            gm.ch = (nty + 8 / gm.rx - 1) / (8 / gm.rx);
            gm.cap = gm.cw * gm.ch * ntz;
            blocks = 8LL * slabs * gm.cap;
        }
        k_gather_tile<DIM, MODE><<<(unsigned)blocks, kThreads, 0, s>>>(pr, sv, (int)tiles, gm);
        k_gather_finish<DIM><<<blocks_for(pr.n_cp), kThreads, 0, s>>>(pr, (int)tiles, deriv);
    }
}

template <int DIM>
static long long partials_size_impl(const Problem &pr, long long planes)
{
    constexpr int T = CpTile<DIM>::X * CpTile<DIM>::Y * CpTile<DIM>::Z;
    const long long tiles = (long long)((pr.cp.n[0] + CpTile<DIM>::X - 1) / CpTile<DIM>::X) *
                            ((pr.cp.n[1] + CpTile<DIM>::Y - 1) / CpTile<DIM>::Y) *
                            (DIM == 3 ? (pr.cp.n[2] + CpTile<DIM>::Z - 1) / CpTile<DIM>::Z : 1);
    const int slab = gather_slab<DIM>(pr);
    return tiles * ((planes + slab - 1) / slab) * T * 3;
}

template <int DIM>
static void gather_finish_impl(const Problem &pr, float *deriv, hipStream_t s)
{
    const long long tiles = (long long)((pr.cp.n[0] + CpTile<DIM>::X - 1) / CpTile<DIM>::X) *
                            ((pr.cp.n[1] + CpTile<DIM>::Y - 1) / CpTile<DIM>::Y) *
                            (DIM == 3 ? (pr.cp.n[2] + CpTile<DIM>::Z - 1) / CpTile<DIM>::Z : 1);
    k_gather_finish<DIM><<<blocks_for(pr.n_cp), kThreads, 0, s>>>(pr, (int)tiles, deriv);
}

void launch_gather_finish(const Problem &pr, float *deriv_soa, hipStream_t s)
{
    if (pr.dim == 2) gather_finish_impl<2>(pr, deriv_soa, s);
    else gather_finish_impl<3>(pr, deriv_soa, s);
}

long long gather_partials_size(const Problem &pr, long long planes)
{
    return pr.dim == 2 ? partials_size_impl<2>(pr, planes) : partials_size_impl<3>(pr, planes);
}

void launch_gather(const Problem &pr, const float4 *sv, float *deriv_soa, hipStream_t s)
{
    SKM_DISPATCH_DIM_MODE(pr, gather_impl, pr, sv, deriv_soa, s);
}

template <int DIM, int MODE>
static void reg_impl(const Problem &pr, const double *w, double pg, const BilateralArgs &bl, int with_derivative,
                     float *deriv, double *partials, hipStream_t s)
{
    k_regularizers<DIM, MODE><<<reg_blocks(pr.n_cp), kThreads, 0, s>>>(pr, w[0], w[1], w[2], pg, bl,
                                                                       with_derivative, deriv, partials);
}

BilateralArgs make_bilateral_args(const Problem &pr, double weight, double sigma_s, double sigma_i)
{
    BilateralArgs a{};
    a.weight = weight;
    const double cut = 3.0 * sigma_s;
    a.cut2 = (float)(cut * cut);
    a.inv2ss = (float)(1.0 / (2.0 * sigma_s * sigma_s));
    a.inv2si = (float)(1.0 / (2.0 * sigma_i * sigma_i));
    for (int d = 0; d < 3; ++d) a.rad[d] = d < pr.dim ? (float)(cut / (double)pr.cp.h[d]) : 0.0f;
    return a;
}

// one wave per control point: 4 control points per workgroup
unsigned reg_blocks(long long n_cp) { return (unsigned)((n_cp + kThreads / 64 - 1) / (kThreads / 64)); }

void launch_regularizers(const Problem &pr, const double *weights, double pg_scaling, const BilateralArgs &bl,
                         int with_derivative, float *deriv_soa, double *partials, Scalars *sc, hipStream_t s)
{
    SKM_DISPATCH_DIM_MODE(pr, reg_impl, pr, weights, pg_scaling, bl, with_derivative, deriv_soa, partials, s);
    k_reg_reduce<<<1, kThreads, 0, s>>>(partials, (int)reg_blocks(pr.n_cp), sc);
}

template <int DIM, int MODE>
static void displacement_impl(const Problem &pr, float *dst, hipStream_t s)
{
    k_displacement<DIM, MODE><<<tile_grid<DIM>(pr.fixed.n), kThreads, 0, s>>>(pr, dst);
}

void launch_displacement(const Problem &pr, float *dst, hipStream_t s)
{
    SKM_DISPATCH_DIM_MODE(pr, displacement_impl, pr, dst, s);
}

template <int DIM, int MODE>
static void voxel_states_impl(const Problem &pr, float4 *out, hipStream_t s)
{
    k_voxel_states<DIM, MODE><<<(unsigned)((pr.n_vox + kThreads - 1) / kThreads), kThreads, 0, s>>>(pr, out);
}

void launch_voxel_states(const Problem &pr, float4 *out, hipStream_t s)
{
    SKM_DISPATCH_DIM_MODE(pr, voxel_states_impl, pr, out, s);
}

void launch_result(const Scalars *sc, void *result, hipStream_t s)
{
    k_result<<<1, 1, 0, s>>>(sc, (skm_result *)result);
}

// ------------------------------------------------------------ optimizer

// L1 / L2 / L21 values and subgradients and the sign flip of
// RegularizedImageToImageMetricv4::GetValueAndDerivative
// (itkRegularizedImageToImageMetricv4.hxx:430-529), one thread per cp.
__global__ __launch_bounds__(kThreads) void k_host_regs(long long ncp, int dim, const float *__restrict__ th,
                                                        double l1w, double l21w, double l2w, int negate,
                                                        float *__restrict__ der, const float *__restrict__ prev_g,
                                                        double *__restrict__ partials)
{
    __shared__ double lds[4 * 4];
    long long j = blockIdx.x * (long long)kThreads + threadIdx.x;
    double vals[4] = {0.0, 0.0, 0.0, 0.0};
    if (j < ncp) {
        double t[3] = {0, 0, 0}, g[3] = {0, 0, 0};
        for (int d = 0; d < dim; ++d) { t[d] = th[j + d * ncp]; g[d] = der[j + d * ncp]; }
        if (l1w > 0) {
            for (int d = 0; d < dim; ++d) {
                vals[0] += fabs(t[d]);
                if (fabs(t[d]) > 0) g[d] += l1w * (t[d] < 0 ? -1.0 : 1.0);
                else if (g[d] < -l1w) g[d] += l1w;
                else if (g[d] > l1w) g[d] -= l1w;
                else g[d] = 0.0;
            }
        }
        double n2 = 0.0;
        for (int d = 0; d < dim; ++d) n2 += t[d] * t[d];
        if (l2w > 0) {
            vals[2] = n2;
            for (int d = 0; d < dim; ++d) g[d] += l2w * 2.0 * t[d];
        }
        if (l21w > 0) {
            double nrm = sqrt(n2);
            vals[1] = nrm;
            for (int d = 0; d < dim; ++d) {
                if (nrm > 0) g[d] += l21w * t[d] / nrm;
                else if (g[d] < -l21w) g[d] += l21w;
                else if (g[d] > l21w) g[d] -= l21w;
                else g[d] = 0.0;
            }
        }
        for (int d = 0; d < dim; ++d) {
            const float gd = (float)(negate ? -g[d] : g[d]);
            der[j + d * ncp] = gd;
            // <g_k, g_{k-1}> of the adaptive step (CalculateScalarProduct,
            // itkAdaptiveStepGradientDescentOptimizerv4.hxx:100-116)
            vals[3] += (double)gd * (double)prev_g[j + d * ncp];
        }
    }
    block_sum<4>(vals, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) partials[blockIdx.x * 4LL + k] = vals[k];
}

void launch_host_regs_device(const Problem &pr, const float *params_soa, double l1, double l21, double l2,
                             int negate, float *deriv_soa, const float *prev_grad, double *partials, hipStream_t s)
{
    k_host_regs<<<blocks_for(pr.n_cp), kThreads, 0, s>>>(pr.n_cp, pr.dim, params_soa, l1, l21, l2, negate,
                                                         deriv_soa, prev_grad, partials);
}

struct StepParams {
    int kind;
    double a, A, alpha, fmax, fmin, omega;
    double l1w, l21w, l2w;
    double rw[4];
    int reg_blocks;
};

// Learning rate for this step and the adaptive time update
// (EstimateLearningRate, Standard...hxx:159-166, Adaptive...hxx:86-97;
// ModifyCurrentTime, Adaptive...hxx:118-128), plus the iteration record.
// One wave: the lanes add the block partials of k_host_regs (L1, L2,1, L2
// and the scalar product) in a fixed order, lane 0 does the rest.
__global__ __launch_bounds__(64) void k_step_scalars(StepParams sp, const double *__restrict__ reg_partials,
                                                     OptState *st, const Scalars *sc, skm_iteration_record *rec)
{
    double part[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < sp.reg_blocks; b += 64)
#pragma unroll
        for (int k = 0; k < 4; ++k) part[k] += reg_partials[b * 4LL + k];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) part[k] += __shfl_xor(part[k], off, 64);
    if (threadIdx.x != 0) return;
    const double dot = part[3];
    double rate;
    if (sp.kind == SKM_OPT_STANDARD) rate = pow(sp.a / (sp.A + (double)st->iteration + 1.0), sp.alpha);
    else if (sp.kind == SKM_OPT_ADAPTIVE) rate = sp.a / pow(sp.A + st->time + 1.0, sp.alpha);
    else rate = sp.a;
    st->rate = rate;
    if (sp.kind == SKM_OPT_ADAPTIVE) {
        double x = -dot;
        double f = sp.fmin + (sp.fmax - sp.fmin) / (1.0 - (sp.fmax / sp.fmin) * exp(-x / sp.omega));
        st->time = fmax(0.0, st->time + f);
    }
    st->prev_dot = dot;
    const double l1 = part[0], l21 = part[1], l2 = part[2];
    double reg = 0.0;
    if (sp.l1w > 0) reg += sp.l1w * l1;
    if (sp.l21w > 0) reg += sp.l21w * l21;
    if (sp.l2w > 0) reg += sp.l2w * l2;
    for (int k = 0; k < 4; ++k)
        if (sp.rw[k] > 0) reg += sp.rw[k] * sc->reg[k];
    const double value = sc->measure + reg;
    st->improved = st->iteration == 0 || value < st->best_value;
    if (st->improved) st->best_value = value;
    if (rec) {
        rec[st->iteration].value = value;
        rec[st->iteration].measure = sc->measure;
        rec[st->iteration].learning_rate = rate;
        rec[st->iteration].num_pixels_counted = (uint32_t)sc->count;
        rec[st->iteration].reserved = 0;
    }
    st->iteration += 1;
}

// theta <- theta + rate * d with orthant projection; d is kept for the next
// scalar product (ModifyGradientByLearningRateOverSubRange, Standard...hxx:
// 133-143).  One thread per control point: first the copy of the position
// just evaluated when its value is the lowest so far (ITK's
// ReturnBestParametersAndValue; best may be null), then the step, then the
// new position packed as the next evaluation's control points.
__global__ __launch_bounds__(kThreads) void k_update(long long ncp, int dim, float *__restrict__ th,
                                                     const float *__restrict__ g, float *__restrict__ prev_g,
                                                     const OptState *st, int orthant, float *__restrict__ best,
                                                     const float *__restrict__ cpw, float4 *__restrict__ cpc)
{
    const long long j = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (j >= ncp) return;
    const double rate = st->rate;
    const bool keep = best && st->improved;
    float c[3] = {0.0f, 0.0f, 0.0f};
    for (int d = 0; d < dim; ++d) {
        const long long i = j + d * ncp;
        const double t = th[i];
        if (keep) best[i] = th[i];
        double v = t + (double)g[i] * rate;
        if (orthant && v * t < 0) v = 0.0;
        th[i] = (float)v;
        c[d] = (float)v;
        prev_g[i] = g[i];
    }
    cpc[j] = make_float4(c[0], c[1], c[2], cpw[j]);
}

void launch_opt_step(const Problem &pr, int kind, double a, double A, double alpha, double fmax, double fmin,
                     double omega, int orthant, float *params_soa, const float *grad_soa, float *prev_grad,
                     OptState *st, const Scalars *sc, const double *reg_partials, double l1w, double l21w,
                     double l2w, const double *gpu_reg_weights, void *record, float *best_params, const float *cpw,
                     float4 *cpc, hipStream_t s)
{
    StepParams sp{kind, a, A, alpha, fmax, fmin, omega, l1w, l21w, l2w,
                  {gpu_reg_weights[0], gpu_reg_weights[1], gpu_reg_weights[2], gpu_reg_weights[3]},
                  (int)blocks_for(pr.n_cp)};
    k_step_scalars<<<1, 64, 0, s>>>(sp, reg_partials, st, sc, (skm_iteration_record *)record);
    k_update<<<blocks_for(pr.n_cp), kThreads, 0, s>>>(pr.n_cp, pr.dim, params_soa, grad_soa, prev_grad, st, orthant,
                                                      best_params, cpw, cpc);
}

void launch_soa_to_cpc(const Problem &pr, const float *params_soa, const float *cpw, float4 *cpc, hipStream_t s)
{
    k_pack_cp<<<blocks_for(pr.n_cp), kThreads, 0, s>>>(pr.n_cp, pr.dim, params_soa, cpw, 1, cpc);
}

}  // namespace skm
