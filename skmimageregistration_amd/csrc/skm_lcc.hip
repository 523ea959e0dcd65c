// skm_lcc.hip -- local cross-correlation (LCC) value and derivative on gfx950.
//
// The reference (include/cuda/GpuEvaluatorLCC.cu) recomputes the kernel sum
// u(y) for every neighbour y of every centre, inside every control point's
// derivative thread (NOAPPROX, include/CommonConfig.h:41).  Here u(y), the
// warped value M(y + u(y)) and its gradient are computed once per fixed voxel
// (dense warp), the per-centre correlation sums are taken from that field,
// and each centre's derivative of -v^2/u is folded into three coefficients
// A, B, C with  d cc_x / d(kw M(y')) = A kw F(y) + B kw M(y') + C.
// The derivative then only needs k(p_j, y) per (cp, centre, neighbour).
#include "skm_internal.hpp"

#include <vector>

#include "skm.h"

namespace skm {

template <int DIM> struct LTile;
template <> struct LTile<2> { static constexpr int X = 16, Y = 16, Z = 1; };
template <> struct LTile<3> { static constexpr int X = 8, Y = 8, Z = 4; };

template <int DIM>
__device__ __forceinline__ void ltile_coords(int t, int *l)
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

template <int DIM, int MODE, int QUIRK>
__device__ __forceinline__ void lcc_kernel_sum(const Problem &pr, const float *y, const PState &st, float *u)
{
    float ci[3];
    to_local<DIM>(pr.cp, y, ci);
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        float r = st.support * pr.cp.ih[d];
        lo[d] = clampi((int)(ci[d] - r + 1.0f), 0, pr.cp.n[d] - 1);
        hi[d] = clampi((int)(ci[d] + r), 0, pr.cp.n[d] - 1);
    }
    // rows beyond the kernel's reach hold k == 0 (with sigma(y) twice the
    // pair's lam sum is 2 lam(y))
    float s2 = st.support * st.support;
    if (MODE != 0) {
        const float lam_y = pstate_lam<DIM, MODE>(st);
        s2 = pair_reach2(pr.kc.smin, lam_y, QUIRK ? lam_y : pr.lam_cp_max);
    }
    const int z0 = DIM == 3 ? lo[2] : 0, z1 = DIM == 3 ? hi[2] : 0;
    for (int cz = z0; cz <= z1; ++cz) {
        float dz = DIM == 3 ? y[2] - (pr.cp.o[2] + pr.cp.h[2] * cz) : 0.0f;
        for (int cy = lo[1]; cy <= hi[1]; ++cy) {
            float dy = y[1] - (pr.cp.o[1] + pr.cp.h[1] * cy);
            float rem = s2 - dy * dy - dz * dz;
            if (rem <= 0.0f) continue;
            float rx = sqrtf(rem) * pr.cp.ih[0];
            int xa = max(lo[0], (int)floorf(ci[0] - rx));
            int xb = min(hi[0], (int)ceilf(ci[0] + rx));
            long long row = (long long)pr.cp.n[0] * (cy + (long long)pr.cp.n[1] * cz);
            for (int cx = xa; cx <= xb; ++cx) {
                float v[3] = {y[0] - (pr.cp.o[0] + pr.cp.h[0] * cx), dy, dz};
                float k;
                if (QUIRK) {
                    // evaluate(cp_i, y, sigma(y)) looks the second sigma up at y too
                    // (GpuEvaluatorLCC.cu:428-443, 3D derivative kernel)
                    k = kernel_pair<DIM, MODE>(pr.kc, v, st, st);
                } else {
                    PState cst;
                    if (MODE != 0)
                        pstate_unpack<DIM, MODE>(pr.cp_pstate[2 * (row + cx)], pr.cp_pstate[2 * (row + cx) + 1], cst);
                    k = kernel_pair<DIM, MODE>(pr.kc, v, st, cst);
                }
                float4 c = pr.cpc[row + cx];
                float wk = c.w * k;
                u[0] = fmaf(wk, c.x, u[0]);
                u[1] = fmaf(wk, c.y, u[1]);
                if (DIM == 3) u[2] = fmaf(wk, c.z, u[2]);
            }
        }
    }
}

// Dense warp: for every fixed voxel y, (grad M, M) at y + df(y) + u(y); .w is
// NaN where y + u(y) leaves the moving image (insideGlobal, BsplineImage.h:285-296).
template <int DIM, int MODE, int QUIRK>
__global__ __launch_bounds__(kThreads) void k_dense_warp(Problem pr, float4 *__restrict__ out)
{
    int l[3];
    ltile_coords<DIM>(threadIdx.x, l);
    int m[3] = {(int)blockIdx.x * LTile<DIM>::X + l[0], (int)blockIdx.y * LTile<DIM>::Y + l[1],
                DIM == 3 ? (int)blockIdx.z * LTile<DIM>::Z + l[2] : 0};
    if (m[0] >= pr.fixed.n[0] || m[1] >= pr.fixed.n[1] || (DIM == 3 && m[2] >= pr.fixed.n[2])) return;
    float fm[3] = {(float)m[0], (float)m[1], (float)m[2]}, y[3];
    to_global<DIM>(pr.fixed, fm, y);
    PState st;
    point_state<DIM, MODE>(pr.kc, pr.weight, pr.weight_data, y, st);
    float u[3] = {0.0f, 0.0f, 0.0f};
    if (pr.has_df) {
        long long f;
        if (nearest<DIM>(pr.df, y, f))
#pragma unroll
            for (int d = 0; d < DIM; ++d) u[d] = pr.df_data[f * DIM + d];
    }
    lcc_kernel_sum<DIM, MODE, QUIRK>(pr, y, st, u);
    float w[3];
#pragma unroll
    for (int d = 0; d < DIM; ++d) w[d] = y[d] + u[d];
    float val = 0.0f, g[3] = {0.0f, 0.0f, 0.0f};
    bool in = bspline_sample<DIM>(pr.coeffs, pr.moving, w, val, g);
    long long v = m[0] + (long long)pr.fixed.n[0] * (m[1] + (long long)pr.fixed.n[1] * m[2]);
    out[v] = in ? make_float4(g[0], g[1], DIM == 3 ? g[2] : 0.0f, val) : make_float4(0.0f, 0.0f, 0.0f, __int_as_float(0x7fc00000));
}

// neighborhoodWeight(p, q, support, 10) (GpuEvaluatorLCC.cu:28-35).
__device__ __forceinline__ float neighborhood_weight(float r2, float support)
{
    float q = r2 / (support * support);
    float q2 = q * q;
    float r = q2 * q2 * q;  // (|p-q|/support)^10
    float f = fmaxf(0.0f, 1.0f - r);
    return f * f;
}

// Neighbour box of a centre: tight window with the centre's support on the
// fixed grid, FP64 so the stride lattice starts where the reference's does
// (GpuEvaluatorLCC.cu:95-97).
template <int DIM>
__device__ __forceinline__ void neighbour_box(const Problem &pr, const int *pos, double support, int *lo, int *hi)
{
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        double idx = (double)pos[d];  // centre index: toLocal(toGlobal(pos)) == pos
        double r = support / (double)pr.fixed.h[d];
        lo[d] = clampi((int)(idx - r + 1.0), 0, pr.fixed.n[d] - 1);
        hi[d] = clampi((int)(idx + r), 0, pr.fixed.n[d] - 1);
    }
}

template <int DIM>
__device__ __forceinline__ void centre_position(const Problem &pr, long long s, int *pos)
{
    long long r = s;
    int m[3] = {0, 0, 0};
    m[0] = (int)(r % pr.nsub[0]); r /= pr.nsub[0];
    m[1] = (int)(r % pr.nsub[1]); r /= pr.nsub[1];
    m[2] = (int)r;
#pragma unroll
    for (int d = 0; d < DIM; ++d) pos[d] = pr.sub > 1 ? m[d] * pr.sub + pr.offsets[s * DIM + d] : m[d];
}

template <int DIM>
__device__ __forceinline__ long long voxel_index(const Problem &pr, const int *pos)
{
    long long f = 0, stride = 1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        f += stride * clampi(pos[d], 0, pr.fixed.n[d] - 1);
        stride *= pr.fixed.n[d];
    }
    return f;
}

// Per-centre LCC (evaluateValueLCC, GpuEvaluatorLCC.cu:38-204, 536-694) from the
// dense warp, one thread per centre.  Writes the derivative coefficients
// (A, B, C) of the centre and adds cc and the centre's inside flag to the
// block partials.  The coefficients are zero where the centre contributes no
// derivative: the reference's |u| > 1e-16 guard (:517-518) in both modes, and
// in exact mode also d1 > 0 && d2 > 0 (the condition for cc, :186).  Without
// the guard, a neighbourhood whose moving values are the B-spline's ringing
// inside a zero background (d2 ~ 1e-40) gives B = 2 d1 v^2 / u^2 <= 2 / d2
// beyond the FP32 range; with it |A| < 2e8 and B < 2e16 d1.
template <int DIM, int MODE>
__global__ __launch_bounds__(kThreads) void k_lcc_centres(Problem pr, const float4 *__restrict__ dense,
                                                          float4 *__restrict__ coef, double *__restrict__ partials,
                                                          double *__restrict__ bound_blk, unsigned *__restrict__ cpk)
{
    __shared__ double lds[kPartialsK * 4];
    __shared__ double lds_max[kThreads / 64];
    double bound = 0.0;
    long long s = blockIdx.x * (long long)kThreads + threadIdx.x;
    double acc[kPartialsK];
#pragma unroll
    for (int k = 0; k < kPartialsK; ++k) acc[k] = 0.0;
    if (s < pr.n_sub) {
        int pos[3] = {0, 0, 0};
        centre_position<DIM>(pr, s, pos);
        float fp[3] = {(float)pos[0], (float)pos[1], (float)pos[2]}, x[3];
        to_global<DIM>(pr.fixed, fp, x);
        PState st;
        point_state<DIM, MODE>(pr.kc, pr.weight, pr.weight_data, x, st);
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        neighbour_box<DIM>(pr, pos, (double)st.support, lo, hi);
        double sf = 0, sm = 0, sff = 0, smm = 0, sfm = 0;
        float fk = 0.0f, mk = 0.0f, gk = 0.0f;  // largest |kw F|, |kw M|, |gradM| component
        int count = 0;
        const int sn = pr.subn;
        const int z0 = DIM == 3 ? lo[2] : 0, z1 = DIM == 3 ? hi[2] : 0;
        for (int fz = z0; fz <= z1; fz += sn)
            for (int fy = lo[1]; fy <= hi[1]; fy += sn)
                for (int fx = lo[0]; fx <= hi[0]; fx += sn) {
                    long long v = fx + (long long)pr.fixed.n[0] * (fy + (long long)pr.fixed.n[1] * fz);
                    float4 g4 = dense[v];
                    if (isnan(g4.w)) continue;
                    float fy3[3] = {(float)fx, (float)fy, (float)fz}, y[3];
                    to_global<DIM>(pr.fixed, fy3, y);
                    float sup = pr.kc.smin;
                    if (MODE != 0) {
                        PState sty;
                        point_state<DIM, MODE>(pr.kc, pr.weight, pr.weight_data, y, sty);
                        sup = sty.support;
                    }
                    float r2 = 0.0f;
#pragma unroll
                    for (int d = 0; d < DIM; ++d) r2 += (x[d] - y[d]) * (x[d] - y[d]);
                    float kw = neighborhood_weight(r2, sup);
                    if (!(kw > 0.0f)) continue;
                    double mv = (double)(kw * g4.w);
                    double fv = (double)(kw * pr.fixed_data[v]);
                    sf += fv; sm += mv; sff += fv * fv; smm += mv * mv; sfm += fv * mv;
                    ++count;
                    if (bound_blk) {
                        fk = fmaxf(fk, fabsf((float)fv));
                        mk = fmaxf(mk, fabsf((float)mv));
                        gk = fmaxf(gk, fmaxf(fabsf(g4.x), fmaxf(fabsf(g4.y), fabsf(g4.z))));
                    }
                }
        double cc = 0.0;
        float A = 0.0f, B = 0.0f, C = 0.0f;
        if (count > 0) {
            double msf = sf / count, msm = sm / count;
            double d1 = sff - 2.0 * msf * sf + count * msf * msf;
            double d2 = smm - 2.0 * msm * sm + count * msm * msm;
            double vv = sfm - msm * sf - msf * sm + count * msf * msm;
            if (d1 > 0 && d2 > 0) cc = -vv * vv / (d1 * d2);
            double u = d1 * d2;
            bool use = fabs(u) > 1e-16 && (!pr.exact || (d1 > 0 && d2 > 0));
            if (use) {
                // beta(y) = -(2 v (fv - msf) u - 2 d1 (mv - msm) v^2) / u^2
                double inv_u = 1.0 / u;
                A = (float)(-2.0 * vv * inv_u);
                B = (float)(2.0 * d1 * vv * vv * inv_u * inv_u);
                C = (float)(2.0 * vv * msf * inv_u - 2.0 * d1 * msm * vv * vv * inv_u * inv_u);
                // |beta kw g| <= (|A| max|kw F| + |B| max|kw M| + |C|) max|g|, with
                // kw <= 1; the factor covers the FP32 roundings of the scatter
                bound = ((double)fabsf(A) * fk + (double)fabsf(B) * mk + (double)fabsf(C)) * gk * 1.001;
            }
        }
        unsigned w = 0u;
        if (cpk) {
            w = (A != 0.0f || B != 0.0f || C != 0.0f) ? 0x80000000u : 0u;
            if (pr.sub > 1)
#pragma unroll
                for (int d = 0; d < DIM; ++d) w |= (unsigned)pr.offsets[s * DIM + d] << (8 * d);
            cpk[s] = w;
        }
        // .w carries the packed word too, so that one load gives both
        coef[s] = make_float4(A, B, C, __uint_as_float(w));
        acc[0] = cc;
        long long cv = voxel_index<DIM>(pr, pos);
        acc[6] = isnan(dense[cv].w) ? 0.0 : 1.0;
    }
    block_sum<kPartialsK>(acc, lds);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < kPartialsK; ++k) partials[blockIdx.x * (long long)kPartialsK + k] = acc[k];
    if (bound_blk) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) bound = fmax(bound, __shfl_xor(bound, off, 64));
        if ((threadIdx.x & 63) == 0) lds_max[threadIdx.x >> 6] = bound;
        __syncthreads();
        if (threadIdx.x == 0) {
            double m = 0.0;
#pragma unroll
            for (int w = 0; w < kThreads / 64; ++w) m = fmax(m, lds_max[w]);
            bound_blk[blockIdx.x] = m;
        }
    }
}

// Control points per workgroup of the reference-mode derivative: the tile's
// points share every loaded neighbour (its dense warp, fixed value and
// neighbourhood weight), which cuts the re-reads of the dense field that
// bound a one-point-per-workgroup walk (each point's centres and neighbours
// cover ~80^3 voxels, more than an XCD's L2).  In 3D 16 points per tile
// (measured faster than 8 and 32 on the 256x256x128 sliding-motion workload:
// 137, 90 and 97 ms).
template <int DIM> struct LcTile;
template <> struct LcTile<2> { static constexpr int X = 2, Y = 2, Z = 1; };
template <> struct LcTile<3> { static constexpr int X = 2, Y = 2, Z = 4; };

// Reference-mode derivative (evaluateDerivativeLCC, GpuEvaluatorLCC.cu:206-530,
// 696-961): for cp j, centres in the tight window of p_j, all their
// neighbours, ka = k(p_j, y) with the neighbour's sigma / covariance for both
// points (:446-460), no w_j, divided by N on the host (:645-646).
// One workgroup per tile of control points: each thread takes a centre of
// the union of the tile's windows (256 at a time), the wave walks the
// neighbour offsets in step, and each pair is added to every tile point whose
// window holds the centre.  Sums: FP32 per lane, FP32 butterfly per wave,
// FP64 across the waves in order.
// With chunks > 1 the centres of a tile are split over `chunks` workgroups
// (workgroup b: tile b / chunks, every chunks-th block of kThreads centres
// from b % chunks), which write FP64 partials per (workgroup, point, axis) for
// k_lcc_ref_finish: a grid of few control points has too few tiles to fill
// the GPU (13x13x7 points: 98 tiles).
template <int DIM, int MODE>
__global__ __launch_bounds__(kThreads) void k_lcc_gather_ref(Problem pr, const float4 *__restrict__ dense,
                                                             const float4 *__restrict__ coef,
                                                             const Scalars *__restrict__ sc,
                                                             float *__restrict__ deriv, int chunks,
                                                             double *__restrict__ partials)
{
    constexpr int TX = LcTile<DIM>::X, TY = LcTile<DIM>::Y, TZ = LcTile<DIM>::Z, T = TX * TY * TZ;
    __shared__ double lds[(kThreads / 64) * T * 3];
    const int ntx = (pr.cp.n[0] + TX - 1) / TX, nty = (pr.cp.n[1] + TY - 1) / TY;
    const int tile = (int)blockIdx.x / chunks, chunk = (int)blockIdx.x % chunks;
    const int t0 = tile % ntx, t1 = (tile / ntx) % nty, t2 = tile / (ntx * nty);
    float p[T][3];
    int lo[T][3], hi[T][3];
    bool ok[T];
    int ulo[3] = {INT_MAX, INT_MAX, DIM == 3 ? INT_MAX : 0}, uhi[3] = {INT_MIN, INT_MIN, DIM == 3 ? INT_MIN : 0};
#pragma unroll
    for (int c = 0; c < T; ++c) {
        const int jc[3] = {t0 * TX + c % TX, t1 * TY + (c / TX) % TY, t2 * TZ + c / (TX * TY)};
        ok[c] = jc[0] < pr.cp.n[0] && jc[1] < pr.cp.n[1] && (DIM == 2 || jc[2] < pr.cp.n[2]);
        const long long j = ok[c] ? jc[0] + (long long)pr.cp.n[0] * (jc[1] + (long long)pr.cp.n[1] * jc[2]) : 0;
        const float fj[3] = {(float)jc[0], (float)jc[1], (float)jc[2]};
        p[c][0] = p[c][1] = p[c][2] = 0.0f;
        to_global<DIM>(pr.cp, fj, p[c]);
        const int *win = pr.cp_window + j * 6;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            lo[c][d] = d < DIM ? win[d] : 0;
            hi[c][d] = d < DIM ? win[d + 3] : 0;
            if (ok[c] && d < DIM) {
                ulo[d] = min(ulo[d], lo[c][d]);
                uhi[d] = max(uhi[d], hi[c][d]);
            }
        }
    }
    const int bx = uhi[0] - ulo[0] + 1, by = uhi[1] - ulo[1] + 1, bz = uhi[2] - ulo[2] + 1;
    const int ncent = (bx > 0 && by > 0 && bz > 0) ? bx * by * bz : 0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    float acc[T][3];
#pragma unroll
    for (int c = 0; c < T; ++c) acc[c][0] = acc[c][1] = acc[c][2] = 0.0f;
    const int sn = pr.subn;
    // one centre per thread (consecutive threads: consecutive centres along
    // x, so a step's neighbour loads of a wave fall on few cache lines); the
    // neighbour offsets are walked in step by the whole wave
    for (int cb = chunk * kThreads; cb < ncent; cb += chunks * kThreads) {
        const int ci = cb + (int)threadIdx.x;
        unsigned mask = 0;
        float4 abc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float x[3] = {0.0f, 0.0f, 0.0f};
        int nlo[3] = {0, 0, 0}, cnt[3] = {0, 0, 0};
        if (ci < ncent) {
            const int mx = ulo[0] + ci % bx, my = ulo[1] + (ci / bx) % by, mz = ulo[2] + ci / (bx * by);
            // tile points whose window holds this centre
#pragma unroll
            for (int c = 0; c < T; ++c)
                if (ok[c] && mx >= lo[c][0] && mx <= hi[c][0] && my >= lo[c][1] && my <= hi[c][1] &&
                    (DIM == 2 || (mz >= lo[c][2] && mz <= hi[c][2])))
                    mask |= 1u << c;
            if (mask) {
                const long long s = mx + (long long)pr.nsub[0] * (my + (long long)pr.nsub[1] * mz);
                abc = coef[s];
                if (abc.x == 0.0f && abc.y == 0.0f && abc.z == 0.0f) mask = 0;
            }
            if (mask) {
                const long long s = mx + (long long)pr.nsub[0] * (my + (long long)pr.nsub[1] * mz);
                int pos[3] = {0, 0, 0};
                centre_position<DIM>(pr, s, pos);
                float fp[3] = {(float)pos[0], (float)pos[1], (float)pos[2]};
                to_global<DIM>(pr.fixed, fp, x);
                PState st;
                point_state<DIM, MODE>(pr.kc, pr.weight, pr.weight_data, x, st);
                int nhi[3] = {0, 0, 0};
                neighbour_box<DIM>(pr, pos, (double)st.support, nlo, nhi);
                cnt[0] = (nhi[0] - nlo[0]) / sn + 1;
                cnt[1] = (nhi[1] - nlo[1]) / sn + 1;
                cnt[2] = DIM == 3 ? (nhi[2] - nlo[2]) / sn + 1 : 1;
            }
        }
        // the wave's largest neighbour box
        int wn[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            int m = cnt[d];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off, 64));
            wn[d] = __builtin_amdgcn_readfirstlane(m);
        }
        for (int t = 0; t < wn[0] * wn[1] * wn[2]; ++t) {
            const int kx = t % wn[0], ky = (t / wn[0]) % wn[1], kz = t / (wn[0] * wn[1]);
            if (kx >= cnt[0] || ky >= cnt[1] || kz >= cnt[2]) continue;
            int fx = nlo[0] + kx * sn, fy = nlo[1] + ky * sn;
            int fz = DIM == 3 ? nlo[2] + kz * sn : 0;
            long long v = fx + (long long)pr.fixed.n[0] * (fy + (long long)pr.fixed.n[1] * fz);
            float4 g4 = dense[v];
            if (isnan(g4.w)) continue;
            float fy3[3] = {(float)fx, (float)fy, (float)fz}, y[3];
            to_global<DIM>(pr.fixed, fy3, y);
            PState sty;
            point_state<DIM, MODE>(pr.kc, pr.weight, pr.weight_data, y, sty);
            float r2 = 0.0f;
#pragma unroll
            for (int d = 0; d < DIM; ++d) r2 += (x[d] - y[d]) * (x[d] - y[d]);
            float kw = neighborhood_weight(r2, MODE == 0 ? pr.kc.smin : sty.support);
            if (!(kw > 0.0f)) continue;
            float fv = kw * pr.fixed_data[v];
            float mv = kw * g4.w;
            float beta = fmaf(abc.x, fv, fmaf(abc.y, mv, abc.z));
            float bk = beta * kw;
#pragma unroll
            for (int c = 0; c < T; ++c) {
                if (!((mask >> c) & 1u)) continue;
                float vv[3] = {p[c][0] - y[0], p[c][1] - y[1], DIM == 3 ? p[c][2] - y[2] : 0.0f};
                float ka = kernel_pair<DIM, MODE>(pr.kc, vv, sty, sty);
                float w = bk * ka;
                acc[c][0] = fmaf(w, g4.x, acc[c][0]);
                acc[c][1] = fmaf(w, g4.y, acc[c][1]);
                if (DIM == 3) acc[c][2] = fmaf(w, g4.z, acc[c][2]);
            }
        }
    }
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
        if (chunks > 1) {
            partials[(size_t)blockIdx.x * (T * 3) + threadIdx.x] = sum;
        } else if (ok[c] && d < DIM) {
            const int jc[3] = {t0 * TX + c % TX, t1 * TY + (c / TX) % TY, t2 * TZ + c / (TX * TY)};
            const long long j = jc[0] + (long long)pr.cp.n[0] * (jc[1] + (long long)pr.cp.n[1] * jc[2]);
            deriv[j + d * pr.n_cp] = (float)(sum * sc->coef_scale);
        }
    }
}

// Sum of a control point's chunk partials in chunk order (k_lcc_gather_ref
// with chunks > 1); one thread per (point, axis).
template <int DIM>
__global__ __launch_bounds__(kThreads) void k_lcc_ref_finish(Problem pr, int chunks,
                                                             const double *__restrict__ partials,
                                                             const Scalars *__restrict__ sc,
                                                             float *__restrict__ deriv)
{
    constexpr int TX = LcTile<DIM>::X, TY = LcTile<DIM>::Y, TZ = LcTile<DIM>::Z, T = TX * TY * TZ;
    const long long i = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (i >= pr.n_cp * DIM) return;
    const long long j = i % pr.n_cp;
    const int d = (int)(i / pr.n_cp);
    const int ntx = (pr.cp.n[0] + TX - 1) / TX, nty = (pr.cp.n[1] + TY - 1) / TY;
    const int j0 = (int)(j % pr.cp.n[0]), j1 = (int)((j / pr.cp.n[0]) % pr.cp.n[1]);
    const int j2 = DIM == 3 ? (int)(j / ((long long)pr.cp.n[0] * pr.cp.n[1])) : 0;
    const long long tile = j0 / TX + (long long)ntx * (j1 / TY + (long long)nty * (j2 / TZ));
    const int c = j0 % TX + TX * (j1 % TY + TY * (j2 % TZ));
    double sum = 0.0;
    for (int k = 0; k < chunks; ++k) sum += partials[((size_t)(tile * chunks + k) * T + c) * 3 + d];
    deriv[i] = (float)(sum * sc->coef_scale);
}

// Exact mode: every valid centre scatters beta kw gradM onto its neighbours,
// pre-scaled by 1/N; the control points then gather sum_y w_j k(p_j, y) gamma(y).
// The fixed-point step of the exact scatter: with every term bounded by
// 2^e (the largest centre bound times 1/N) and at most n_sub terms per voxel
// (each centre adds at most one term to a voxel), a step of
// 2^(e - 62 + ceil(log2(n_sub + 1)) + 1) keeps every voxel's sum below 2^62.
__global__ __launch_bounds__(kThreads) void k_lcc_fx_unit(const double *__restrict__ bound_blk, int blocks,
                                                          long long n_sub, Scalars *__restrict__ sc)
{
    __shared__ double lds_max[kThreads / 64];
    double m = 0.0;
    for (int b = threadIdx.x; b < blocks; b += kThreads) m = fmax(m, bound_blk[b]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) lds_max[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 0; w < kThreads / 64; ++w) m = fmax(m, lds_max[w]);
        m *= sc->coef_scale;
        int e = 0;
        if (m > 0.0 && isfinite(m)) frexp(m, &e);  // m < 2^e
        int h = 1;
        while ((1LL << h) <= n_sub) ++h;            // 2^h > n_sub
        const int shift = e - 62 + h + 1;
        sc->fx_unit = ldexp(1.0, shift);
        sc->fx_inv_unit = ldexp(1.0, -shift);
    }
}

__global__ __launch_bounds__(kThreads) void k_lcc_fx_finish(const unsigned long long *__restrict__ fx, long long n,
                                                            const Scalars *__restrict__ sc, float4 *__restrict__ gamma)
{
    const long long v = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (v >= n) return;
    const double u = sc->fx_unit;
    gamma[v] = make_float4((float)((double)(long long)fx[3 * v] * u), (float)((double)(long long)fx[3 * v + 1] * u),
                           (float)((double)(long long)fx[3 * v + 2] * u), 0.0f);
}

template <int DIM, int MODE>
__global__ __launch_bounds__(kThreads) void k_lcc_scatter(Problem pr, const float4 *__restrict__ dense,
                                                          const float4 *__restrict__ coef,
                                                          const Scalars *__restrict__ sc,
                                                          unsigned long long *__restrict__ sums)
{
    // one centre per thread (consecutive threads: consecutive centres along
    // x); the wave walks the neighbour offsets in step, so that a step's
    // loads and atomic adds fall on few cache lines
    const long long s = blockIdx.x * (long long)kThreads + threadIdx.x;
    const float scale = (float)sc->coef_scale;
    const double inv_unit = sc->fx_inv_unit;
    float4 abc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float x[3] = {0.0f, 0.0f, 0.0f};
    int nlo[3] = {0, 0, 0}, cnt[3] = {0, 0, 0};
    if (s < pr.n_sub) abc = coef[s];
    if (s < pr.n_sub && !(abc.x == 0.0f && abc.y == 0.0f && abc.z == 0.0f)) {
        int pos[3] = {0, 0, 0};
        centre_position<DIM>(pr, s, pos);
        float fp[3] = {(float)pos[0], (float)pos[1], (float)pos[2]};
        to_global<DIM>(pr.fixed, fp, x);
        PState st;
        point_state<DIM, MODE>(pr.kc, pr.weight, pr.weight_data, x, st);
        int nhi[3] = {0, 0, 0};
        neighbour_box<DIM>(pr, pos, (double)st.support, nlo, nhi);
        cnt[0] = (nhi[0] - nlo[0]) / pr.subn + 1;
        cnt[1] = (nhi[1] - nlo[1]) / pr.subn + 1;
        cnt[2] = DIM == 3 ? (nhi[2] - nlo[2]) / pr.subn + 1 : 1;
    }
    const int sn = pr.subn;
    // the wave's largest neighbour box
    int wn[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        int m = cnt[d];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off, 64));
        wn[d] = __builtin_amdgcn_readfirstlane(m);
    }
    for (int t = 0; t < wn[0] * wn[1] * wn[2]; ++t) {
        const int kx = t % wn[0], ky = (t / wn[0]) % wn[1], kz = t / (wn[0] * wn[1]);
        if (kx >= cnt[0] || ky >= cnt[1] || kz >= cnt[2]) continue;
        int fx = nlo[0] + kx * sn, fy = nlo[1] + ky * sn;
        int fz = DIM == 3 ? nlo[2] + kz * sn : 0;
        long long v = fx + (long long)pr.fixed.n[0] * (fy + (long long)pr.fixed.n[1] * fz);
        float4 g4 = dense[v];
        if (isnan(g4.w)) continue;
        float fy3[3] = {(float)fx, (float)fy, (float)fz}, y[3];
        to_global<DIM>(pr.fixed, fy3, y);
        float sup = pr.kc.smin;
        if (MODE != 0) {
            PState sty;
            point_state<DIM, MODE>(pr.kc, pr.weight, pr.weight_data, y, sty);
            sup = sty.support;
        }
        float r2 = 0.0f;
#pragma unroll
        for (int d = 0; d < DIM; ++d) r2 += (x[d] - y[d]) * (x[d] - y[d]);
        float kw = neighborhood_weight(r2, sup);
        if (!(kw > 0.0f)) continue;
        float fv = kw * pr.fixed_data[v];
        float mv = kw * g4.w;
        float beta = fmaf(abc.x, fv, fmaf(abc.y, mv, abc.z)) * kw * scale;
        // integer adds commute: the sums do not depend on the order of the atomics
        atomicAdd(&sums[3 * v], (unsigned long long)__double2ll_rn((double)(beta * g4.x) * inv_unit));
        atomicAdd(&sums[3 * v + 1], (unsigned long long)__double2ll_rn((double)(beta * g4.y) * inv_unit));
        if (DIM == 3)
            atomicAdd(&sums[3 * v + 2], (unsigned long long)__double2ll_rn((double)(beta * g4.z) * inv_unit));
    }
}

// Exact mode with a stationary kernel: gamma(y) gathered per fixed voxel y
// from the centres whose neighbour lattice holds y -- the transpose of
// k_lcc_scatter, without atomics.  With one support for every centre, the
// centres that reach y along an axis lie within [y - e - 1, y - c + 1]
// (c = floor(1 - r), e = floor(r), r = support / spacing); each position p of
// that range is tested with neighbour_box's own arithmetic (y inside the box
// and on its stride lattice) and kept as a bit of a 64-bit mask, so the launch
// needs e - c + 3 <= 64 on every axis.  A position triple is a centre
// when its cell's jitter offsets equal p - sub floor(p / sub).  Then
//   gamma(y) = scale gradM(y) (F(y) SA + M(y) SB + SC),
//   SA = sum A kw^2, SB = sum B kw^2, SC = sum C kw
// over the centres, in FP64 and in a fixed order (z, y, x positions
// ascending), which is the scatter's sum of beta kw gradM regrouped.
template <int DIM>
__device__ __forceinline__ unsigned long long reach_mask(const Problem &pr, int d, int yi, double r, int lo_c,
                                                         int &base)
{
    base = yi - (int)floor(r) - 1;
    unsigned long long m = 0;
    const int n = pr.fixed.n[d], sn = pr.subn;
    for (int i = 0; i < 64; ++i) {
        const int p = base + i;
        if (p > yi - lo_c + 1) break;
        if (p < 0 || p >= n) continue;
        const int nlo = clampi((int)((double)p - r + 1.0), 0, n - 1);
        const int nhi = clampi((int)((double)p + r), 0, n - 1);
        if (yi >= nlo && yi <= nhi && (yi - nlo) % sn == 0) m |= 1ull << i;
    }
    return m;
}

template <int DIM>
__global__ __launch_bounds__(kThreads) void k_lcc_gamma(Problem pr, const float4 *__restrict__ dense,
                                                        const float4 *__restrict__ coef,
                                                        const unsigned *__restrict__ cpk,
                                                        const Scalars *__restrict__ sc, float4 *__restrict__ gamma)
{
    const long long v = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (v >= pr.n_vox) return;
    const int yi[3] = {(int)(v % pr.fixed.n[0]), (int)((v / pr.fixed.n[0]) % pr.fixed.n[1]),
                       DIM == 3 ? (int)(v / ((long long)pr.fixed.n[0] * pr.fixed.n[1])) : 0};
    const float4 g4 = dense[v];
    if (isnan(g4.w)) {
        gamma[v] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const float fy3[3] = {(float)yi[0], (float)yi[1], (float)yi[2]};
    float y[3];
    to_global<DIM>(pr.fixed, fy3, y);
    unsigned long long mask[3] = {1ull, 1ull, 1ull};
    int base[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const double r = (double)pr.kc.smin / (double)pr.fixed.h[d];
        mask[d] = reach_mask<DIM>(pr, d, yi[d], r, (int)floor(1.0 - r), base[d]);
    }
    const int sub = pr.sub;
    const float inv_sub = 1.0f / (float)sub;
    // cell and jitter offset of a position p >= 0: (p + 0.5) / sub is at least
    // 0.5 / sub away from an integer, far more than the FP32 product's error
    // for any image size below 2^22
    auto cell_of = [&](int p, int &off) {
        const int m = sub > 1 ? (int)(((float)p + 0.5f) * inv_sub) : p;
        off = p - m * sub;
        return m;
    };
    // bits i of a mask whose position base + i lies within R of c
    auto bits_within = [](int base_i, int c, int R) -> unsigned long long {
        const int a = max(c - R - base_i, 0), b = min(c + R - base_i, 63);
        if (a > b) return 0ull;
        const unsigned long long upto = b == 63 ? ~0ull : ((1ull << (b + 1)) - 1ull);
        return upto & (~0ull << a);
    };
    // candidates beyond the support (plus a margin) are skipped before their
    // weight is computed; the weight of the others decides
    const float lim = pr.kc.smin * pr.kc.smin * 1.001f;
    double sa = 0.0, sb = 0.0, scc = 0.0;
    for (unsigned long long mz = mask[2]; mz; mz &= mz - 1) {
        const int pz = DIM == 3 ? base[2] + __builtin_ctzll(mz) : 0;
        const float dz = DIM == 3 ? fmaf(pr.fixed.h[2], (float)pz, pr.fixed.o[2]) - y[2] : 0.0f;
        const float remz = lim - dz * dz;
        if (!(remz > 0.0f)) continue;
        int oz = 0;
        const int mzc = DIM == 3 ? cell_of(pz, oz) : 0;
        if (DIM == 3 && mzc >= pr.nsub[2]) continue;
        const unsigned long long ymask =
            mask[1] & bits_within(base[1], yi[1], (int)(sqrtf(remz) * pr.fixed.ih[1]) + 1);
        for (unsigned long long my = ymask; my; my &= my - 1) {
            const int py = base[1] + __builtin_ctzll(my);
            const float dy = fmaf(pr.fixed.h[1], (float)py, pr.fixed.o[1]) - y[1];
            const float remy = remz - dy * dy;
            if (!(remy > 0.0f)) continue;
            int oy = 0;
            const int myc = cell_of(py, oy);
            if (myc >= pr.nsub[1]) continue;
            const long long rowbase = (long long)pr.nsub[0] * (myc + (long long)pr.nsub[1] * mzc);
            const unsigned wyz = 0x80000000u | ((unsigned)oy << 8) | (DIM == 3 ? (unsigned)oz << 16 : 0u);
            const unsigned long long xmask =
                mask[0] & bits_within(base[0], yi[0], (int)(sqrtf(remy) * pr.fixed.ih[0]) + 1);
            for (unsigned long long mx = xmask; mx; mx &= mx - 1) {
                const int px = base[0] + __builtin_ctzll(mx);
                int ox = 0;
                const int mxc = cell_of(px, ox);
                if (mxc >= pr.nsub[0]) continue;
                const long long s = rowbase + mxc;
                // a centre at this position whose coefficients are not all zero
                if (cpk[s] != (wyz | (unsigned)ox)) continue;
                const float fp[3] = {(float)px, (float)py, (float)pz};
                float x[3];
                to_global<DIM>(pr.fixed, fp, x);
                float r2 = 0.0f;
#pragma unroll
                for (int d = 0; d < DIM; ++d) r2 += (x[d] - y[d]) * (x[d] - y[d]);
                const float kw = neighborhood_weight(r2, pr.kc.smin);
                if (!(kw > 0.0f)) continue;
                const float4 abc = coef[s];
                const double kw2 = (double)kw * (double)kw;
                sa += (double)abc.x * kw2;
                sb += (double)abc.y * kw2;
                scc += (double)abc.z * (double)kw;
            }
        }
    }
    const double gam = sc->coef_scale * ((double)pr.fixed_data[v] * sa + (double)g4.w * sb + scc);
    gamma[v] = make_float4((float)(gam * g4.x), (float)(gam * g4.y), DIM == 3 ? (float)(gam * g4.z) : 0.0f, 0.0f);
}

// ------------------------------------------------------------ reference windows by voxel tables

// The reference-window derivative (k_lcc_gather_ref's sum) regrouped per
// neighbour voxel y, for stationary kernels in 3D:
//   d_j = sum_y ka(p_j, y) gradM(y) (F(y) SA_j(y) + M(y) SB_j(y) + SC_j(y)),
//   SA_j = sum A kw^2, SB_j = sum B kw^2, SC_j = sum C kw
// over the centres whose neighbour lattice holds y (found as k_lcc_gamma
// finds them) and whose cell lies in j's window.  F(y) and M(y) are fixed for
// the voxel, so one sum T_j = sum kw (kw (F A + M B) + C) stands for the
// bracket.  Along an axis, the centres
// reaching y occupy a few sample cells (their "ranks"), and j's window, an
// interval of cells, keeps a contiguous run of those ranks.  So one 3D prefix
// sum of the centres' terms over (x, y, z) ranks per voxel gives every
// control point's restricted sum as an 8-corner difference: the kernel pair
// is evaluated once per (control point, voxel) instead of once per (control
// point, centre, neighbour).
//
// One wave per block of voxels: the voxels whose position lies in one cell
// [p_k, p_k+1) of the control-point grid.  Every control point with
// |p_j - y| < support for such a voxel lies in the box k - A .. k - A + NB - 1
// per axis (NB = 2 ceil(support / h), A = NB / 2 - 1), whose points have one
// FP32 sum slot each for the whole block: lane l holds the slots of box
// column (l % NB0, l / NB0) in registers, one per layer bz.  Per voxel: the
// reach masks and ranks; each column's active layers (points within the
// support whose windows keep a centre on every axis, about half the box); the
// table of the centres' terms kw (kw (F A + M B) + C) in chunks of at most
// kRtCh ranks per axis (prefix along x in registers, then y and z in LDS,
// each line's loads issued together; FP32); and per active slot an 8-corner
// difference of the table.  The LDS left (4.3 KB per wave) lets 4 waves run
// per SIMD, which the registers are held to.
// Each block writes its slots, k_lcc_reft_finish adds the blocks in order in
// FP64.  A control-point cell wider than ~5 voxels is split into pieces along
// each axis, one wave each, so that coarse grids still fill the GPU.
constexpr int kRtCh = 7;              // ranks per axis of one table chunk (the sliding-motion
                                      // workload's voxels away from the faces reach 7 cells)
constexpr int kRtT = kRtCh + 1;       // table extent per axis; index 0 is the zero border
constexpr int kRtBox = 8;             // control points per axis of a block's box, at most

struct RefTableGeom {
    int kmin[3], nk[3];  // control-point cells of the blocks, per axis
    int A[3], NB[3];     // the box of block k: control points k - A .. k - A + NB - 1
    int np[3];           // voxel pieces per cell and axis (one wave each)
};

// control-point cell of voxel index i along axis d: floor((y - o_cp) / h_cp)
__device__ __forceinline__ int rt_cell(const Problem &pr, int d, int i)
{
    const float y = pr.fixed.o[d] + pr.fixed.h[d] * (float)i;
    return (int)floorf((y - pr.cp.o[d]) * pr.cp.ih[d]);
}

// bits [0, t) of a 64-bit mask (t clamped to 0..64)
__device__ __forceinline__ unsigned long long rt_below(int t)
{
    return t <= 0 ? 0ull : (t >= 64 ? ~0ull : ((1ull << t) - 1ull));
}

// Reach of a voxel index yi along axis d, for every index of every axis:
// the centre cells (relative to cbase, one bit each) whose neighbour lattice
// holds yi, and per rank the cell and the jitter offsets allowed in it.  It
// depends on the geometry only; one wave per (axis, index).
struct RefReach {
    unsigned long long *cm;  // per row (off[d] + yi): the cell mask
    int *cbase;              // its first cell
    int *cell;               // 64 per row: the cell of each rank
    unsigned *allowed;       // 64 per row: its allowed offsets (bit o: offset o)
    int off[3];
};

__global__ __launch_bounds__(64) void k_lcc_reach(Problem pr, RefReach rr)
{
    const int lane = threadIdx.x;
    const int row = (int)blockIdx.x;
    const int d = row < rr.off[1] ? 0 : (row < rr.off[2] ? 1 : 2);
    const int yi = row - rr.off[d];
    const double r = (double)pr.kc.smin / (double)pr.fixed.h[d];
    // centre positions that can reach yi: [base, lim] (k_lcc_gamma's range)
    const int base = yi - (int)floor(r) - 1, lim = yi - (int)floor(1.0 - r) + 1;
    const int n = pr.fixed.n[d], sn = pr.subn, sub = pr.sub;
    const int p0 = max(base, 0);
    const int cbase = sub > 1 ? (int)(((float)p0 + 0.5f) / (float)sub) : p0;
    // lane l: cell cbase + l; offset o is allowed when position sub (cbase + l)
    // + o holds yi on its lattice (neighbour_box's arithmetic)
    const int c = cbase + lane;
    unsigned al = 0;
    if (c < pr.nsub[d]) {
        for (int o = 0; o < sub; ++o) {
            const int p = sub * c + o;
            if (p < p0 || p > lim || p >= n) continue;
            const int nlo = clampi((int)((double)p - r + 1.0), 0, n - 1);
            const int nhi = clampi((int)((double)p + r), 0, n - 1);
            if (yi >= nlo && yi <= nhi && (yi - nlo) % sn == 0) al |= 1u << o;
        }
    }
    const unsigned long long cm = __ballot(al != 0u);
    if (al) {
        const int rank = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(cm >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((unsigned)cm, 0u));
        rr.cell[row * 64 + rank] = c;
        rr.allowed[row * 64 + rank] = al;
    }
    if (lane == 0) {
        rr.cm[row] = cm;
        rr.cbase[row] = cbase;
    }
}

// A voxel index's reach into LDS (cells, allowed) and registers (cm, cbase).
__device__ __forceinline__ void rt_axis(const RefReach &rr, int d, int yi, int lane, int (*cells)[64],
                                        unsigned (*allowed)[64], unsigned long long &cm, int &cbase)
{
    const int row = rr.off[d] + yi;
    cm = rr.cm[row];
    cbase = rr.cbase[row];
    if (lane < (int)__popcll(cm)) {
        cells[d][lane] = rr.cell[row * 64 + lane];
        allowed[d][lane] = rr.allowed[row * 64 + lane];
    }
}

// table index of ranks (z, y, x), each 0..kRtCh with 0 the zero border; rows
// and planes padded so that lanes reading different rows fall on different
// LDS banks
constexpr int kRtRow = kRtT + 1, kRtPlane = kRtT * kRtRow + 3, kRtSize = kRtT * kRtPlane;
__device__ __forceinline__ int rt_at(int z, int y, int x) { return z * kRtPlane + y * kRtRow + x; }

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_lcc_ref_table(Problem pr, const float4 *__restrict__ dense,
                                                      const float4 *__restrict__ coef, RefReach rr,
                                                      RefTableGeom g, float *__restrict__ partials)
{
    __shared__ float tT[kRtSize];
    __shared__ int cells[3][64];
    __shared__ unsigned allowed[3][64];
    __shared__ int wlo[3][kRtBox], whi[3][kRtBox];  // box control points' windows (sample cells)
    __shared__ int rlo[3][kRtBox], rhi[3][kRtBox];  // their rank ranges at the current voxel
    const int lane = threadIdx.x;
    const int b = blockIdx.x;
    // block b: piece (q0, q1, q2) of control-point cell (k0, k1, k2), pieces fastest
    const int npc = g.np[0] * g.np[1] * g.np[2];
    const int cell = b / npc, piece = b % npc;
    const int k[3] = {g.kmin[0] + cell % g.nk[0], g.kmin[1] + (cell / g.nk[0]) % g.nk[1],
                      g.kmin[2] + cell / (g.nk[0] * g.nk[1])};
    const int q[3] = {piece % g.np[0], (piece / g.np[0]) % g.np[1], piece / (g.np[0] * g.np[1])};
    for (int i = lane; i < kRtSize; i += 64) tT[i] = 0.0f;  // borders stay zero
    // the sums of this lane's box column, slot bz * 64 + lane, per axis
    float acc[kRtBox][3];
#pragma unroll
    for (int bz = 0; bz < kRtBox; ++bz) acc[bz][0] = acc[bz][1] = acc[bz][2] = 0.0f;
    if (lane < 3 * kRtBox) {
        // windows are separable for stationary kernels: axis d's from a
        // control point on that axis
        const int d = lane / kRtBox, bj = lane % kRtBox;
        const int j = k[d] - g.A[d] + bj;
        int lo = 1, hi = 0;
        if (bj < g.NB[d] && j >= 0 && j < pr.cp.n[d]) {
            const long long jj = d == 0 ? j : (d == 1 ? (long long)j * pr.cp.n[0]
                                                      : (long long)j * pr.cp.n[0] * pr.cp.n[1]);
            lo = pr.cp_window[jj * 6 + d];
            hi = pr.cp_window[jj * 6 + 3 + d];
        }
        wlo[d][bj] = lo;
        whi[d][bj] = hi;
    }
    // this lane's box column (bx, by); slot bz * 64 + lane
    const int bx = lane % g.NB[0], by = lane / g.NB[0];
    const bool lane_ok = lane < g.NB[0] * g.NB[1];
    const int jx = k[0] - g.A[0] + bx, jy = k[1] - g.A[1] + by;
    const bool jxy_ok = lane_ok && jx >= 0 && jx < pr.cp.n[0] && jy >= 0 && jy < pr.cp.n[1];
    const float px = pr.cp.o[0] + pr.cp.h[0] * (float)jx, py = pr.cp.o[1] + pr.cp.h[1] * (float)jy;
    const float s2 = pr.kc.smin * pr.kc.smin * 1.0001f;
    const float inv_s2 = 1.0f / (pr.kc.smin * pr.kc.smin);
    // voxel range of the block per axis (one voxel of margin; rt_cell decides)
    int vlo[3], vhi[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float a = (pr.cp.o[d] + pr.cp.h[d] * (float)k[d] - pr.fixed.o[d]) * pr.fixed.ih[d];
        const float e = (pr.cp.o[d] + pr.cp.h[d] * (float)(k[d] + 1) - pr.fixed.o[d]) * pr.fixed.ih[d];
        const int lo = max(0, (int)floorf(a) - 1), hi = min(pr.fixed.n[d] - 1, (int)ceilf(e) + 1);
        // this block's piece of it
        const int len = hi - lo + 1;
        vlo[d] = lo + (len * q[d]) / g.np[d];
        vhi[d] = lo + (len * (q[d] + 1)) / g.np[d] - 1;
    }
    __syncthreads();
    const int sub = pr.sub;
    unsigned long long cm[3] = {0, 0, 0};
    int cbase[3] = {0, 0, 0};
    for (int iz = vlo[2]; iz <= vhi[2]; ++iz) {
        if (rt_cell(pr, 2, iz) != k[2]) continue;
        __syncthreads();  // the previous voxel's readers of cells / allowed
        rt_axis(rr, 2, iz, lane, cells, allowed, cm[2], cbase[2]);
        for (int iy = vlo[1]; iy <= vhi[1]; ++iy) {
            if (rt_cell(pr, 1, iy) != k[1]) continue;
            __syncthreads();
            rt_axis(rr, 1, iy, lane, cells, allowed, cm[1], cbase[1]);
            // the row's dense values and fixed intensities, one voxel ahead
            const long long row = (long long)pr.fixed.n[0] * (iy + (long long)pr.fixed.n[1] * iz);
            float4 g4n = dense[row + vlo[0]];
            float fn = pr.fixed_data[row + vlo[0]];
            for (int ix = vlo[0]; ix <= vhi[0]; ++ix) {
                const long long v = row + ix;
                const float4 g4 = g4n;
                const float F = fn;
                if (ix < vhi[0]) {
                    g4n = dense[v + 1];
                    fn = pr.fixed_data[v + 1];
                }
                if (rt_cell(pr, 0, ix) != k[0]) continue;
                if (isnan(g4.w)) continue;
                __syncthreads();
                rt_axis(rr, 0, ix, lane, cells, allowed, cm[0], cbase[0]);
                const int R[3] = {(int)__popcll(cm[0]), (int)__popcll(cm[1]), (int)__popcll(cm[2])};
                if (R[0] == 0 || R[1] == 0 || R[2] == 0) continue;
                // rank ranges of the box control points' windows
                if (lane < 3 * kRtBox) {
                    const int d = lane / kRtBox, bj = lane % kRtBox;
                    const unsigned long long m = d == 0 ? cm[0] : (d == 1 ? cm[1] : cm[2]);
                    const int cb = d == 0 ? cbase[0] : (d == 1 ? cbase[1] : cbase[2]);
                    int a = 1, e = 0;
                    if (wlo[d][bj] <= whi[d][bj]) {
                        a = (int)__popcll(m & rt_below(wlo[d][bj] - cb));
                        e = (int)__popcll(m & rt_below(whi[d][bj] + 1 - cb)) - 1;
                    }
                    rlo[d][bj] = a;
                    rhi[d][bj] = e;
                }
                const float fyv[3] = {(float)ix, (float)iy, (float)iz};
                float y[3];
                to_global<3>(pr.fixed, fyv, y);
                const float M = g4.w;
                __syncthreads();
                // the active slots of this lane's column (bit bz): control
                // points within the support whose windows keep some centre on
                // every axis
                const int x0 = lane_ok ? rlo[0][bx] : 1, x1 = lane_ok ? rhi[0][bx] : 0;
                const int y0 = lane_ok ? rlo[1][by] : 1, y1 = lane_ok ? rhi[1][by] : 0;
                unsigned amask = 0u;
                {
                    const bool xy_ok = jxy_ok && x0 <= x1 && y0 <= y1;
                    const float dxy = (px - y[0]) * (px - y[0]) + (py - y[1]) * (py - y[1]);
#pragma unroll
                    for (int bz = 0; bz < kRtBox; ++bz) {
                        const int jz = k[2] - g.A[2] + bz;
                        const float dz = pr.cp.o[2] + pr.cp.h[2] * (float)jz - y[2];
                        const int z0 = rlo[2][bz], z1 = rhi[2][bz];
                        if (bz < g.NB[2] && xy_ok && jz >= 0 && jz < pr.cp.n[2] && z0 <= z1 && dxy + dz * dz < s2)
                            amask |= 1u << bz;
                    }
                }
                if (__ballot(amask != 0u) == 0ull) continue;
                for (int c0z = 0; c0z < R[2]; c0z += kRtCh)
                    for (int c0y = 0; c0y < R[1]; c0y += kRtCh)
                        for (int c0x = 0; c0x < R[0]; c0x += kRtCh) {
                            const int nx = min(kRtCh, R[0] - c0x), ny = min(kRtCh, R[1] - c0y);
                            const int nz = min(kRtCh, R[2] - c0z);
                            // the chunk's terms, prefix along x: lane = (y rank, z rank); the
                            // loads of a lane's row are issued together (indices clamped into
                            // the row, the extra entries unused)
                            {
                                const int ly = lane % kRtCh, lz = lane / kRtCh;
                                if (lz < kRtCh && ly < ny && lz < nz) {
                                    const int cy = cells[1][c0y + ly], cz = cells[2][c0z + lz];
                                    const unsigned ay = allowed[1][c0y + ly], az = allowed[2][c0z + lz];
                                    const unsigned rowc = (unsigned)pr.nsub[0] * (unsigned)(cy + pr.nsub[1] * cz);
                                    unsigned cxs[kRtCh];
#pragma unroll
                                    for (int lx = 0; lx < kRtCh; ++lx) cxs[lx] = (unsigned)cells[0][c0x + min(lx, nx - 1)];
                                    float st = 0.0f;
                                    const int tb = rt_at(lz + 1, ly + 1, 1);
                                    // the candidates in groups of four (four, then three), each
                                    // group's coefficients loaded together: fewer registers held
                                    // at once than with all seven (16 bytes of spill per lane at
                                    // four waves per SIMD instead of 60)
                                    auto term = [&](int lx, const float4 &a4) {
                                        const unsigned wv = __float_as_uint(a4.w);
                                        const unsigned ax = allowed[0][c0x + min(lx, nx - 1)];
                                        const unsigned ox = wv & 0xffu, oy = (wv >> 8) & 0xffu;
                                        const unsigned oz = (wv >> 16) & 0xffu;
                                        if ((wv & 0x80000000u) && ((ax >> ox) & 1u) && ((ay >> oy) & 1u) &&
                                            ((az >> oz) & 1u)) {
                                            const float fp[3] = {(float)((int)cxs[lx] * sub + (sub > 1 ? (int)ox : 0)),
                                                                 (float)(cy * sub + (sub > 1 ? (int)oy : 0)),
                                                                 (float)(cz * sub + (sub > 1 ? (int)oz : 0))};
                                            float x[3];
                                            to_global<3>(pr.fixed, fp, x);
                                            float r2 = 0.0f;
#pragma unroll
                                            for (int d = 0; d < 3; ++d) r2 += (x[d] - y[d]) * (x[d] - y[d]);
                                            // neighborhood_weight with the reciprocal of support^2
                                            const float q = r2 * inv_s2, q2 = q * q;
                                            const float f = fmaxf(0.0f, 1.0f - q2 * q2 * q);
                                            const float kw = f * f;
                                            st = fmaf(kw, fmaf(kw, fmaf(F, a4.x, M * a4.y), a4.z), st);
                                        }
                                        if (lx < nx) tT[tb + lx] = st;
                                    };
                                    constexpr int kG = 4;  // candidates per group
#pragma unroll
                                    for (int g0 = 0; g0 < kRtCh; g0 += kG) {
                                        float4 abc[kG];
#pragma unroll
                                        for (int i = 0; i < kG; ++i)
                                            if (g0 + i < kRtCh) abc[i] = coef[rowc + cxs[g0 + i]];
#pragma unroll
                                        for (int i = 0; i < kG; ++i)
                                            if (g0 + i < kRtCh) term(g0 + i, abc[i]);
                                        __builtin_amdgcn_sched_barrier(0);
                                    }
                                }
                            }
                            __syncthreads();
                            // prefix along y: lane = (x rank, z rank), a line's loads together
                            {
                                const int lx = lane % kRtCh, lz = lane / kRtCh;
                                if (lz < kRtCh && lx < nx && lz < nz) {
                                    float va[kRtCh];
#pragma unroll
                                    for (int ly = 0; ly < kRtCh; ++ly) va[ly] = tT[rt_at(lz + 1, ly + 1, lx + 1)];
#pragma unroll
                                    for (int ly = 1; ly < kRtCh; ++ly) va[ly] += va[ly - 1];
#pragma unroll
                                    for (int ly = 1; ly < kRtCh; ++ly)
                                        if (ly < ny) tT[rt_at(lz + 1, ly + 1, lx + 1)] = va[ly];
                                }
                            }
                            __syncthreads();
                            // prefix along z: lane = (x rank, y rank)
                            {
                                const int lx = lane % kRtCh, ly = lane / kRtCh;
                                if (ly < kRtCh && lx < nx && ly < ny) {
                                    float va[kRtCh];
#pragma unroll
                                    for (int lz = 0; lz < kRtCh; ++lz) va[lz] = tT[rt_at(lz + 1, ly + 1, lx + 1)];
#pragma unroll
                                    for (int lz = 1; lz < kRtCh; ++lz) va[lz] += va[lz - 1];
#pragma unroll
                                    for (int lz = 1; lz < kRtCh; ++lz)
                                        if (lz < nz) tT[rt_at(lz + 1, ly + 1, lx + 1)] = va[lz];
                                }
                            }
                            __syncthreads();
                            // the column's active slots: 8-corner sums of their boxes
                            const int xa = max(x0, c0x) - c0x, xb = min(x1, c0x + nx - 1) - c0x + 1;
                            const int ya = max(y0, c0y) - c0y, yb = min(y1, c0y + ny - 1) - c0y + 1;
                            if (amask != 0u && xa < xb && ya < yb) {
                                const float qx = pr.cp.o[0] + pr.cp.h[0] * (float)(k[0] - g.A[0] + bx);
                                const float qy = pr.cp.o[1] + pr.cp.h[1] * (float)(k[1] - g.A[1] + by);
#pragma unroll
                                for (int bz = 0; bz < kRtBox; ++bz) {
                                    if (!((amask >> bz) & 1u)) continue;
                                    const int za = max(rlo[2][bz], c0z) - c0z;
                                    const int zb = min(rhi[2][bz], c0z + nz - 1) - c0z + 1;
                                    if (za >= zb) continue;
                                    const float qz = pr.cp.o[2] + pr.cp.h[2] * (float)(k[2] - g.A[2] + bz);
                                    const float vv[3] = {qx - y[0], qy - y[1], qz - y[2]};
                                    PState sty;
                                    sty.support = pr.kc.smin;
                                    const float ka = kernel_pair<3, 0>(pr.kc, vv, sty, sty);
                                    if (ka == 0.0f) continue;
                                    const int c11 = rt_at(zb, yb, 0), c10 = rt_at(zb, ya, 0);
                                    const int c01 = rt_at(za, yb, 0), c00 = rt_at(za, ya, 0);
                                    const float ST = ((tT[c11 + xb] - tT[c11 + xa]) - (tT[c10 + xb] - tT[c10 + xa])) -
                                                     ((tT[c01 + xb] - tT[c01 + xa]) - (tT[c00 + xb] - tT[c00 + xa]));
                                    const float wgt = ka * ST;
                                    acc[bz][0] = fmaf(wgt, g4.x, acc[bz][0]);
                                    acc[bz][1] = fmaf(wgt, g4.y, acc[bz][1]);
                                    acc[bz][2] = fmaf(wgt, g4.z, acc[bz][2]);
                                }
                            }
                            __syncthreads();  // the table's readers, before the next chunk or voxel
                        }
            }
        }
    }
    // this block's sums: [block][slot][axis]
    float *out = partials + (long long)b * (kRtBox * 64 * 3);
#pragma unroll
    for (int bz = 0; bz < kRtBox; ++bz) {
        out[(bz * 64 + lane) * 3] = acc[bz][0];
        out[(bz * 64 + lane) * 3 + 1] = acc[bz][1];
        out[(bz * 64 + lane) * 3 + 2] = acc[bz][2];
    }
}

// deriv_j = coef_scale * the sum over the blocks whose box holds j (cells,
// then their pieces), in block order.
__global__ __launch_bounds__(kThreads) void k_lcc_reft_finish(Problem pr, RefTableGeom g,
                                                              const float *__restrict__ partials,
                                                              const Scalars *__restrict__ sc,
                                                              float *__restrict__ deriv)
{
    const long long i = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (i >= pr.n_cp * 3) return;
    const long long j = i % pr.n_cp;
    const int d = (int)(i / pr.n_cp);
    const int jc[3] = {(int)(j % pr.cp.n[0]), (int)((j / pr.cp.n[0]) % pr.cp.n[1]),
                       (int)(j / ((long long)pr.cp.n[0] * pr.cp.n[1]))};
    int klo[3], khi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        klo[a] = max(jc[a] + g.A[a] - g.NB[a] + 1, g.kmin[a]);
        khi[a] = min(jc[a] + g.A[a], g.kmin[a] + g.nk[a] - 1);
    }
    const int npc = g.np[0] * g.np[1] * g.np[2];
    double sum = 0.0;
    for (int kz = klo[2]; kz <= khi[2]; ++kz)
        for (int ky = klo[1]; ky <= khi[1]; ++ky)
            for (int kx = klo[0]; kx <= khi[0]; ++kx) {
                const long long cell = (kx - g.kmin[0]) + (long long)g.nk[0] * ((ky - g.kmin[1]) +
                                                                                (long long)g.nk[1] * (kz - g.kmin[2]));
                const int bx = jc[0] - kx + g.A[0], by = jc[1] - ky + g.A[1], bz = jc[2] - kz + g.A[2];
                const long long slot = (long long)(bz * 64 + bx + g.NB[0] * by) * 3 + d;
                for (int pc = 0; pc < npc; ++pc)
                    sum += (double)partials[(cell * npc + pc) * (kRtBox * 64 * 3) + slot];
            }
    deriv[i] = (float)(sum * sc->coef_scale);
}

// Geometry of the voxel-table launch; false when it does not apply.
static bool ref_table_geom(const Problem &pr, RefTableGeom &g)
{
    if (pr.dim != 3 || !is_stationary(pr.kc.mode) || pr.sub < 1 || pr.sub > 32 || pr.subn < 1) return false;
    for (int d = 0; d < 3; ++d) {
        const double r = (double)pr.kc.smin / (double)pr.fixed.h[d];
        // the centre positions reaching a voxel, [y - floor(r) - 1, y - floor(1 - r) + 1],
        // must lie within 64 cells (the reach rows' masks)
        const double positions = std::floor(r) - std::floor(1.0 - r) + 3.0;
        if (std::ceil(positions / pr.sub) + 1.0 > 64.0) return false;
        const double sh = (double)pr.kc.smin / (double)pr.cp.h[d];
        const int nb = 2 * (int)std::ceil(sh);
        if (nb < 2 || nb > kRtBox) return false;
        g.NB[d] = nb;
        g.A[d] = nb / 2 - 1;
        // blocks: the cells of the first and last voxel, one more each side
        const double y0 = (double)pr.fixed.o[d], y1 = y0 + (double)pr.fixed.h[d] * (pr.fixed.n[d] - 1);
        g.kmin[d] = (int)std::floor((y0 - (double)pr.cp.o[d]) / (double)pr.cp.h[d]) - 1;
        const int kmax = (int)std::floor((y1 - (double)pr.cp.o[d]) / (double)pr.cp.h[d]) + 1;
        g.nk[d] = kmax - g.kmin[d] + 1;
        // pieces of about 5 voxels per axis, so that large cells fill the GPU
        const double vox = (double)pr.cp.h[d] / (double)pr.fixed.h[d];
        g.np[d] = std::max(1, std::min(16, (int)std::lround(vox / 5.0)));
    }
    if (g.NB[0] * g.NB[1] > 64) return false;
    return (long long)g.nk[0] * g.nk[1] * g.nk[2] * g.np[0] * g.np[1] * g.np[2] < (1ll << 31);
}

// The centre walk's (control point, centre, neighbour) triples, counted per
// axis with each centre at its cell origin: the product over the axes of the
// sum over the control points of the neighbours along the axis of the
// centres in the point's window.
static double walk_triples(const Problem &pr)
{
    const auto clampi = [](int v, int lo, int hi) { return std::min(std::max(v, lo), hi); };
    double t = 1.0;
    for (int d = 0; d < pr.dim; ++d) {
        const int n = pr.fixed.n[d], ns = pr.nsub[d], sub = pr.sub;
        const double r = (double)pr.kc.smin / (double)pr.fixed.h[d];
        std::vector<double> pre((size_t)ns + 1, 0.0);
        for (int c = 0; c < ns; ++c) {
            const int pos = c * sub;
            const int nlo = clampi((int)((double)pos - r + 1.0), 0, n - 1), nhi = clampi((int)((double)pos + r), 0, n - 1);
            pre[(size_t)c + 1] = pre[(size_t)c] + (double)((nhi - nlo) / pr.subn + 1);
        }
        double sum = 0.0;
        for (int j = 0; j < pr.cp.n[d]; ++j) {
            const double idx = ((double)pr.cp.o[d] + (double)pr.cp.h[d] * j - (double)pr.fixed.o[d]) / (double)pr.fixed.h[d];
            const int lo = clampi((int)((idx - r) / sub + 1.0), 0, ns - 1), hi = clampi((int)((idx + r) / sub), 0, ns - 1);
            if (lo <= hi) sum += pre[(size_t)hi + 1] - pre[(size_t)lo];
        }
        t *= sum;
    }
    return t;
}

// Which form of the reference-window derivative runs: lcc_ref_walk 1 the
// walk, 2 the voxel tables where their geometry allows, 0 the one with the
// smaller expected time.  Rates measured on MI355X at the sliding-motion
// registration's three levels (256x256x128, supports 80 / 40 / 20): the walk
// 0.9-1.7e12 of walk_triples' triples per second (the lower one here), the
// tables 4.2e8 voxels per second at support 20 (fewer at support 80, where
// the walk wins by far).
bool lcc_ref_table_applies(const Problem &pr)
{
    RefTableGeom g;
    if (pr.lcc_ref_walk == 1 || !ref_table_geom(pr, g)) return false;
    if (pr.lcc_ref_walk == 2) return true;
    constexpr double kWalkTriplesPerS = 0.9e12, kTableVoxelsPerS = 4.2e8;
    return walk_triples(pr) / kWalkTriplesPerS > (double)pr.n_vox / kTableVoxelsPerS;
}

// block sums (FP32), then the reach rows (cm, cell, allowed, cbase) of every
// index of every axis
static long long ref_table_blocks(const RefTableGeom &g)
{
    return (long long)g.nk[0] * g.nk[1] * g.nk[2] * g.np[0] * g.np[1] * g.np[2];
}
static long long ref_reach_rows(const Problem &pr) { return (long long)pr.fixed.n[0] + pr.fixed.n[1] + pr.fixed.n[2]; }

long long lcc_ref_table_partials(const Problem &pr)
{
    RefTableGeom g;
    if (!ref_table_geom(pr, g)) return 0;
    const long long rows = ref_reach_rows(pr);
    const long long bytes = ref_table_blocks(g) * kRtBox * 64 * 3 * 4 + rows * (8 + 64 * 4 + 64 * 4 + 4);
    return (bytes + 7) / 8;
}

void launch_lcc_ref_table(const Problem &pr, const float4 *dense_g, const float4 *centre_coef, const Scalars *sc,
                          float *deriv_soa, double *partials, hipStream_t s)
{
    RefTableGeom g;
    if (!ref_table_geom(pr, g)) return;
    const long long rows = ref_reach_rows(pr), blocks = ref_table_blocks(g);
    float *sums = reinterpret_cast<float *>(partials);
    RefReach rr;
    // (the block sums are a multiple of 8 bytes: 1536 floats per block)
    rr.cm = reinterpret_cast<unsigned long long *>(sums + blocks * kRtBox * 64 * 3);
    rr.cell = reinterpret_cast<int *>(rr.cm + rows);
    rr.allowed = reinterpret_cast<unsigned *>(rr.cell + rows * 64);
    rr.cbase = reinterpret_cast<int *>(rr.allowed + rows * 64);
    rr.off[0] = 0;
    rr.off[1] = pr.fixed.n[0];
    rr.off[2] = pr.fixed.n[0] + pr.fixed.n[1];
    k_lcc_reach<<<(unsigned)rows, 64, 0, s>>>(pr, rr);
    k_lcc_ref_table<<<(unsigned)blocks, 64, 0, s>>>(pr, dense_g, centre_coef, rr, g, sums);
    const unsigned fb = (unsigned)((pr.n_cp * 3 + kThreads - 1) / kThreads);
    k_lcc_reft_finish<<<fb, kThreads, 0, s>>>(pr, g, sums, sc, deriv_soa);
}

// ------------------------------------------------------------ launchers

template <int DIM>
static dim3 ltile_grid(const int *n)
{
    return dim3((n[0] + LTile<DIM>::X - 1) / LTile<DIM>::X, (n[1] + LTile<DIM>::Y - 1) / LTile<DIM>::Y,
                DIM == 3 ? (n[2] + LTile<DIM>::Z - 1) / LTile<DIM>::Z : 1);
}

template <int DIM, int MODE>
static void dense_impl(const Problem &pr, int quirk, float4 *out, hipStream_t s)
{
    dim3 g = ltile_grid<DIM>(pr.fixed.n);
    if (quirk) k_dense_warp<DIM, MODE, 1><<<g, kThreads, 0, s>>>(pr, out);
    else k_dense_warp<DIM, MODE, 0><<<g, kThreads, 0, s>>>(pr, out);
}

// Both stationary modes (0 and 3) run the general stationary instantiation.
#define SKM_LCC_DISPATCH(pr, FN, ...)                                                                              \
    do {                                                                                                           \
        if ((pr).dim == 2) {                                                                                       \
            if (is_stationary((pr).kc.mode)) FN<2, 0>(__VA_ARGS__);                                                \
            else if ((pr).kc.mode == 1) FN<2, 1>(__VA_ARGS__);                                                     \
            else FN<2, 2>(__VA_ARGS__);                                                                            \
        } else {                                                                                                   \
            if (is_stationary((pr).kc.mode)) FN<3, 0>(__VA_ARGS__);                                                \
            else if ((pr).kc.mode == 1) FN<3, 1>(__VA_ARGS__);                                                     \
            else FN<3, 2>(__VA_ARGS__);                                                                            \
        }                                                                                                          \
    } while (0)

void launch_dense_warp(const Problem &pr, int quirk3d, float4 *dense_g, hipStream_t s)
{
    SKM_LCC_DISPATCH(pr, dense_impl, pr, quirk3d, dense_g, s);
}

template <int DIM, int MODE>
static void centres_impl(const Problem &pr, const float4 *dense, float4 *coef, double *partials, int *blocks,
                         double *bound_blk, unsigned *cpk, hipStream_t s)
{
    unsigned g = (unsigned)((pr.n_sub + kThreads - 1) / kThreads);
    *blocks = (int)g;
    k_lcc_centres<DIM, MODE><<<g, kThreads, 0, s>>>(pr, dense, coef, partials, bound_blk, cpk);
}

void launch_lcc_centres(const Problem &pr, const float4 *dense_g, float4 *centre_coef, double *partials,
                        int *blocks_out, hipStream_t s, double *bound_blk, unsigned *cpk)
{
    SKM_LCC_DISPATCH(pr, centres_impl, pr, dense_g, centre_coef, partials, blocks_out, bound_blk, cpk, s);
}

template <int DIM>
static long long ref_tiles(const Problem &pr)
{
    constexpr int TX = LcTile<DIM>::X, TY = LcTile<DIM>::Y, TZ = LcTile<DIM>::Z;
    return (long long)((pr.cp.n[0] + TX - 1) / TX) * ((pr.cp.n[1] + TY - 1) / TY) *
           (DIM == 3 ? (pr.cp.n[2] + TZ - 1) / TZ : 1);
}

// Workgroups per tile of the reference-mode derivative: enough for ~8192
// workgroups in all, at most 64.
static int ref_chunks(long long tiles)
{
    constexpr long long kWanted = 8192;
    return (int)std::min<long long>(64, std::max<long long>(1, (kWanted + tiles - 1) / tiles));
}

static int chunks_of(const Problem &pr, long long tiles)
{
    return pr.lcc_chunks > 0 ? std::min(pr.lcc_chunks, 64) : ref_chunks(tiles);
}

long long lcc_reference_partials(const Problem &pr)
{
    const long long tiles = pr.dim == 3 ? ref_tiles<3>(pr) : ref_tiles<2>(pr);
    const int chunks = chunks_of(pr, tiles);
    const int t = pr.dim == 3 ? LcTile<3>::X * LcTile<3>::Y * LcTile<3>::Z : LcTile<2>::X * LcTile<2>::Y;
    return chunks > 1 ? tiles * chunks * t * 3 : 0;
}

template <int DIM, int MODE>
static void gather_ref_impl(const Problem &pr, const float4 *dense, const float4 *coef, const Scalars *sc,
                            float *deriv, double *partials, hipStream_t s)
{
    const long long tiles = ref_tiles<DIM>(pr);
    const int chunks = partials ? chunks_of(pr, tiles) : 1;
    k_lcc_gather_ref<DIM, MODE><<<(unsigned)(tiles * chunks), kThreads, 0, s>>>(pr, dense, coef, sc, deriv, chunks,
                                                                                partials);
    if (chunks > 1) {
        const unsigned g = (unsigned)((pr.n_cp * DIM + kThreads - 1) / kThreads);
        k_lcc_ref_finish<DIM><<<g, kThreads, 0, s>>>(pr, chunks, partials, sc, deriv);
    }
}

void launch_lcc_gather_reference(const Problem &pr, const float4 *dense_g, const float4 *centre_coef,
                                 const Scalars *sc, float *deriv_soa, double *partials, hipStream_t s)
{
    SKM_LCC_DISPATCH(pr, gather_ref_impl, pr, dense_g, centre_coef, sc, deriv_soa, partials, s);
}

template <int DIM, int MODE>
static void scatter_impl(const Problem &pr, const float4 *dense, const float4 *coef, const Scalars *sc,
                         unsigned long long *fx, hipStream_t s)
{
    unsigned g = (unsigned)((pr.n_sub + kThreads - 1) / kThreads);
    k_lcc_scatter<DIM, MODE><<<g, kThreads, 0, s>>>(pr, dense, coef, sc, fx);
}

bool lcc_gamma_gather_applies(const Problem &pr)
{
    if (!is_stationary(pr.kc.mode)) return false;
    for (int d = 0; d < pr.dim; ++d) {
        const double r = (double)pr.kc.smin / (double)pr.fixed.h[d];
        if ((std::floor(r) - std::floor(1.0 - r)) + 3.0 > 64.0) return false;
    }
    return true;
}

void launch_lcc_gamma(const Problem &pr, const float4 *dense_g, const float4 *centre_coef, const unsigned *cpk,
                      const Scalars *sc, float4 *gamma, hipStream_t s)
{
    const unsigned g = (unsigned)((pr.n_vox + kThreads - 1) / kThreads);
    if (pr.dim == 2) k_lcc_gamma<2><<<g, kThreads, 0, s>>>(pr, dense_g, centre_coef, cpk, sc, gamma);
    else k_lcc_gamma<3><<<g, kThreads, 0, s>>>(pr, dense_g, centre_coef, cpk, sc, gamma);
}

void launch_lcc_scatter(const Problem &pr, const float4 *dense_g, const float4 *centre_coef, Scalars *sc,
                        const double *bound_blk, int bound_blocks, unsigned long long *fx, float4 *gamma,
                        hipStream_t s)
{
    k_lcc_fx_unit<<<1, kThreads, 0, s>>>(bound_blk, bound_blocks, pr.n_sub, sc);
    SKM_LCC_DISPATCH(pr, scatter_impl, pr, dense_g, centre_coef, sc, fx, s);
    const unsigned g = (unsigned)((pr.n_vox + kThreads - 1) / kThreads);
    k_lcc_fx_finish<<<g, kThreads, 0, s>>>(fx, pr.n_vox, sc, gamma);
}

}  // namespace skm
