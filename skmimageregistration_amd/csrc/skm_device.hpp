// skm_device.hpp -- device-side math of the SKM path for gfx950.
//
// Geometry, Wendland kernels in the three modes, cubic B-spline sampling and
// PCG32 jump-ahead.  All per-pair arithmetic is FP32; sums that leave a
// kernel are FP64.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace skm {

constexpr int kThreads = 256;   // 4 wave64 per workgroup

// Image geometry in FP32 (origin + spacing * index), x fastest.
struct Grid {
    int n[3];
    float o[3];
    float h[3];
    float ih[3];   // 1 / spacing
};

struct KernelConst {
    int mode;            // 0 stationary, 1 weight image, 2 weight tensor
    float inv_s4, inv_s0;
    float smin;          // min(sigma_C4, sigma_C0)
    float max_weight;
    float max_eig_cov[9];
};

// Kernel-side state of one point: sigma (mode 1) or the symmetric part of
// the covariance (mode 2, upper triangle xx, xy, (xz,) yy, (yz, zz)),
// sdet = |sigma|^(1/2) or det(cov)^(1/2) (the square of the F factor's
// per-point part, Kernel.h:107-109, 121-123), and the region support.
struct PState {
    float sig;
    float sdet;
    float up[6];
    float support;
};

// Number of upper-triangle entries of a symmetric DIM x DIM matrix.
template <int DIM> constexpr int kUpper = DIM * (DIM + 1) / 2;

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

template <int DIM>
__device__ __forceinline__ void to_global(const Grid &g, const float *idx, float *p)
{
#pragma unroll
    for (int d = 0; d < DIM; ++d) p[d] = g.o[d] + g.h[d] * idx[d];
}

template <int DIM>
__device__ __forceinline__ void to_local(const Grid &g, const float *p, float *idx)
{
#pragma unroll
    for (int d = 0; d < DIM; ++d) idx[d] = g.ih[d] * (p[d] - g.o[d]);
}

// Nearest-neighbour index, toNearest = int(idx + 0.5) (ImageBase.h:64-67).
template <int DIM>
__device__ __forceinline__ bool nearest(const Grid &g, const float *p, long long &flat)
{
    float idx[3];
    to_local<DIM>(g, p, idx);
    long long f = 0, stride = 1;
    bool in = true;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        int i = (int)(idx[d] + 0.5f);
        in = in && i >= 0 && i < g.n[d];
        f += stride * (long long)i;
        stride *= g.n[d];
    }
    flat = f;
    return in;
}

// ---------------------------------------------------------------- kernels

// Kernel modes inside the library: the three reference modes plus mode 3,
// the stationary kernel with sigma_C4 == sigma_C0, where C4 * C0 collapses
// to (1-t)^8 (1 + 6t + 35/3 t^2).
constexpr int kModeStationaryEqual = 3;
__host__ __device__ constexpr bool is_stationary(int mode) { return mode == 0 || mode == kModeStationaryEqual; }

// Hardware square root (v_sqrt_f32, ~1 ulp); the IEEE-rounded sqrtf expands
// to a dozen instructions per pair.
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// Wendland C4 with t = r / sigma_C4 (Kernel.h:72-77): (1-t)_+^6 (3 + 18t + 35t^2) / 3.
__device__ __forceinline__ float wendland_c4(float t)
{
    float f = fmaxf(0.0f, 1.0f - t);
    float f2 = f * f;
    return f2 * f2 * f2 * fmaf(fmaf(35.0f / 3.0f, t, 6.0f), t, 1.0f);
}

// Stationary k(r) from r^2 (Kernel.h:96-103, 173-176).
template <int MODE>
__device__ __forceinline__ float stationary_k(float r2, float inv_s4, float inv_s0)
{
    float r = fast_sqrt(r2);
    float t = r * inv_s4;
    float f = fmaxf(0.0f, fmaf(-inv_s4, r, 1.0f));
    float f2 = f * f;
    float f4 = f2 * f2;
    float poly = fmaf(fmaf(35.0f / 3.0f, t, 6.0f), t, 1.0f);
    if (MODE == kModeStationaryEqual) return f4 * f4 * poly;
    float g = fmaxf(0.0f, fmaf(-inv_s0, r, 1.0f));
    return f4 * f2 * poly * (g * g);
}

// Wendland C0 (Kernel.h:48-53): (1-t)_+^2.
__device__ __forceinline__ float wendland_c0(float t)
{
    float f = fmaxf(0.0f, 1.0f - t);
    return f * f;
}

// C4(r / sigma_C4) * C0(r / sigma_C0) of the weighted modes; with equal
// supports the two share the (1 - t) factor: (1 - t)^8 (1 + 6t + 35/3 t^2).
__device__ __forceinline__ float wendland_c4c0(float rs, const KernelConst &kc)
{
    const float t = rs * kc.inv_s4;
    const float f = fmaxf(0.0f, 1.0f - t);
    const float f2 = f * f, f4 = f2 * f2;
    const float poly = fmaf(fmaf(35.0f / 3.0f, t, 6.0f), t, 1.0f);
    if (kc.inv_s4 == kc.inv_s0) return f4 * f4 * poly;
    const float g = fmaxf(0.0f, fmaf(-kc.inv_s0, rs, 1.0f));
    return f4 * f2 * poly * (g * g);
}

template <int DIM>
__device__ __forceinline__ float det_m(const float *m)
{
    if (DIM == 2) return m[0] * m[3] - m[2] * m[1];
    return m[0] * (m[4] * m[8] - m[5] * m[7]) + m[1] * (m[5] * m[6] - m[3] * m[8]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
}

// Largest eigenvalue, trigonometric closed form for 3x3 (Mat.h:259-281).
template <int DIM>
__device__ __forceinline__ float max_eigenvalue(const float *m)
{
    if (DIM == 2) {
        float tr = m[0] + m[3];
        float det = det_m<2>(m);
        return 0.5f * (tr + sqrtf(fmaxf(tr * tr - 4.0f * det, 0.0f)));
    }
    float tri = m[1] * m[1] + m[2] * m[2] + m[5] * m[5];
    if (tri == 0.0f) return fmaxf(m[0], fmaxf(m[4], m[8]));
    float q = (1.0f / 3.0f) * (m[0] + m[4] + m[8]);
    float a = m[0] - q, b = m[4] - q, c = m[8] - q;
    float psq = 2.0f * tri + a * a + b * b + c * c;
    float p = sqrtf(psq * (1.0f / 6.0f));
    float s[9] = {a, m[1], m[2], m[3], b, m[5], m[6], m[7], c};
    float r = det_m<3>(s) * (0.5f / (p * p * p));
    float phi = (1.0f / 3.0f) * acosf(fminf(1.0f, fmaxf(-1.0f, r)));
    return q + 2.0f * p * cosf(phi);
}

// Symmetric part of a row-major DIM x DIM matrix, upper triangle.
template <int DIM>
__device__ __forceinline__ void sym_upper(const float *m, float *up)
{
    int k = 0;
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int j = i; j < DIM; ++j) up[k++] = i == j ? m[i * DIM + i] : 0.5f * (m[i * DIM + j] + m[j * DIM + i]);
}

// Largest eigenvalue of the symmetric matrix held as its upper triangle.
template <int DIM>
__device__ __forceinline__ float max_eig_upper(const float *up)
{
    float m[9];
    if (DIM == 2) {
        m[0] = up[0]; m[1] = m[2] = up[1]; m[3] = up[2];
    } else {
        m[0] = up[0]; m[1] = m[3] = up[1]; m[2] = m[6] = up[2];
        m[4] = up[3]; m[5] = m[7] = up[4]; m[8] = up[5];
    }
    return max_eigenvalue<DIM>(m);
}

// Packed per-point state of the weighted modes, two float4 per point:
//   mode 1: {sigma, sdet, 0, 0}, {0, 0, 0, lam}
//   mode 2: {up0, up1, up2, sdet}, {up3, up4, up5, lam}
// lam (sigma, or the largest covariance eigenvalue) bounds the distance at
// which a pair with this point can still have k != 0.
template <int DIM, int MODE>
__device__ __forceinline__ void pstate_unpack(const float4 a, const float4 b, PState &st)
{
    if (MODE == 1) {
        st.sig = a.x;
        st.sdet = a.y;
    } else {
        st.up[0] = a.x; st.up[1] = a.y; st.up[2] = a.z;
        st.up[3] = b.x; st.up[4] = b.y; st.up[5] = b.z;
        st.sdet = a.w;
    }
}

template <int DIM, int MODE>
__device__ __forceinline__ void pstate_pack(const PState &st, float lam, float4 &a, float4 &b)
{
    if (MODE == 1) {
        a = make_float4(st.sig, st.sdet, 0.0f, 0.0f);
        b = make_float4(0.0f, 0.0f, 0.0f, lam);
    } else {
        a = make_float4(st.up[0], st.up[1], st.up[2], st.sdet);
        b = make_float4(DIM == 3 ? st.up[3] : 0.0f, DIM == 3 ? st.up[4] : 0.0f, DIM == 3 ? st.up[5] : 0.0f, lam);
    }
}

// lam of a point state: sigma (mode 1) or the largest covariance eigenvalue.
template <int DIM, int MODE>
__device__ __forceinline__ float pstate_lam(const PState &st)
{
    return MODE == 1 ? st.sig : max_eig_upper<DIM>(st.up);
}

// Squared distance beyond which k(p, q) == 0 for every partner q: the kernel
// vanishes for r >= min(sigma_C4, sigma_C0), and r^2 >= |v|^2 2 / (lam_p +
// lam_q) in both weighted modes (mode 2: lambda_max((S_p + S_q) / 2) <=
// (lam_p + lam_q) / 2).  lam_other bounds lam_q over all partners; the
// relative margin covers FP32 rounding of lam.
__device__ __forceinline__ float pair_reach2(float smin, float lam_self, float lam_other)
{
    return smin * smin * 0.5f * (lam_self + lam_other) * 1.0001f;
}

// State of a point from the weight image / tensor field (getSigmaAtPoint /
// getCovarianceAtPoint, Kernel.h:148-156, NN with 0 / identity outside) and
// its region support (Kernel.h:158-171).
template <int DIM, int MODE>
__device__ __forceinline__ void point_state(const KernelConst &kc, const Grid &wg, const float *wdata,
                                            const float *p, PState &st)
{
    if (is_stationary(MODE)) {
        st.support = kc.smin;
        return;
    }
    long long f;
    bool in = nearest<DIM>(wg, p, f);
    if (MODE == 1) {
        float s = in ? wdata[f] : 0.0f;
        st.sig = s;
        st.sdet = sqrtf(fabsf(s));
        st.support = kc.smin / sqrtf(2.0f / (kc.max_weight + s));
    } else {
        constexpr int NN = DIM * DIM;
        float cov[9];
#pragma unroll
        for (int i = 0; i < NN; ++i) cov[i] = in ? wdata[f * NN + i] : ((i % (DIM + 1)) == 0 ? 1.0f : 0.0f);
        st.sdet = sqrtf(det_m<DIM>(cov));
        sym_upper<DIM>(cov, st.up);
        float sum[9];
#pragma unroll
        for (int i = 0; i < NN; ++i) sum[i] = kc.max_eig_cov[i] + cov[i];
        st.support = kc.smin / sqrtf(2.0f / max_eigenvalue<DIM>(sum));
    }
}

// Terms of v = (dx, dy, dz) that stay fixed along a row of control points
// or samples (dy, dz fixed): 2dy, 2dz, dy^2, dz^2, 2 dy dz, dy^2 + dz^2.
struct RowTerms {
    float ty, tz, yy, zz, tyz, yz2;
};

template <int DIM>
__device__ __forceinline__ RowTerms row_terms(float dy, float dz)
{
    RowTerms r;
    r.ty = 2.0f * dy;
    r.yy = dy * dy;
    if (DIM == 3) {
        r.tz = 2.0f * dz;
        r.zz = dz * dz;
        r.tyz = r.ty * dz;
        r.yz2 = fmaf(dz, dz, r.yy);
    } else {
        r.tz = r.zz = r.tyz = 0.0f;
        r.yz2 = r.yy;
    }
    return r;
}

// k(p, q) of the weighted modes for v = (dx, dy, dz) given as dx and the row
// terms (Kernel.h:96-131); both modes carry F twice, F^2 = sdet_p * sdet_q /
// |avg|.  Mode 2 works on S = S_p + S_q: with m = S / 2, v^T m^-1 v =
// 2 v^T adj(S) v / det S and det m = det S / 2^D, so one reciprocal serves
// both terms; v^T adj(S) v = dx (c00 dx + 2 c01 dy + 2 c02 dz) + c11 dy^2 +
// c22 dz^2 + 2 c12 dy dz.  k = 0 unless r < smin: lanes outside return before
// the reciprocal (the bound is loosened by 1e-4, the exact zero comes from
// the polynomial).
template <int DIM, int MODE>
__device__ __forceinline__ float weighted_k_row(const KernelConst &kc, float dx, const RowTerms &r, const PState &a,
                                                const PState &b)
{
    const float s2 = kc.smin * kc.smin * 1.0001f;
    if (MODE == 1) {
        const float r2 = fmaf(dx, dx, r.yz2);
        const float sum = a.sig + b.sig;
        if (!(2.0f * r2 < s2 * sum)) return 0.0f;
        const float inv = __builtin_amdgcn_rcpf(sum);
        const float F2 = 2.0f * a.sdet * b.sdet * fabsf(inv);
        const float rs = fast_sqrt(2.0f * inv * r2);
        return F2 * wendland_c4c0(rs, kc);
    } else {
        float quad, det;
        if (DIM == 2) {
            const float s00 = a.up[0] + b.up[0], s01 = a.up[1] + b.up[1], s11 = a.up[2] + b.up[2];
            det = fmaf(s00, s11, -s01 * s01);
            // dx (s11 dx - 2 s01 dy) + s00 dy^2
            quad = fmaf(dx, fmaf(s11, dx, -s01 * r.ty), s00 * r.yy);
        } else {
            const float s00 = a.up[0] + b.up[0], s01 = a.up[1] + b.up[1], s02 = a.up[2] + b.up[2];
            const float s11 = a.up[3] + b.up[3], s12 = a.up[4] + b.up[4], s22 = a.up[5] + b.up[5];
            const float c00 = fmaf(s11, s22, -s12 * s12), c11 = fmaf(s00, s22, -s02 * s02);
            const float c22 = fmaf(s00, s11, -s01 * s01);
            const float c01 = fmaf(s02, s12, -s01 * s22), c02 = fmaf(s01, s12, -s02 * s11);
            const float c12 = fmaf(s01, s02, -s00 * s12);
            det = fmaf(s00, c00, fmaf(s01, c01, s02 * c02));
            quad = fmaf(dx, fmaf(c00, dx, fmaf(c01, r.ty, c02 * r.tz)), fmaf(c11, r.yy, fmaf(c22, r.zz, c12 * r.tyz)));
        }
        if (!(2.0f * quad < s2 * det)) return 0.0f;
        const float inv = __builtin_amdgcn_rcpf(det);
        const float F2 = (DIM == 3 ? 8.0f : 4.0f) * a.sdet * b.sdet * inv;
        const float rs = fast_sqrt(fmaxf(0.0f, 2.0f * quad * inv));
        return F2 * wendland_c4c0(rs, kc);
    }
}

// k(p, q) for v = p - q with the states of both points (Kernel.h:96-131).
template <int DIM, int MODE>
__device__ __forceinline__ float kernel_pair(const KernelConst &kc, const float *v, const PState &a, const PState &b)
{
    if (is_stationary(MODE)) {
        float r2 = v[0] * v[0] + v[1] * v[1];
        if (DIM == 3) r2 = fmaf(v[2], v[2], r2);
        return stationary_k<MODE>(r2, kc.inv_s4, kc.inv_s0);
    } else {
        return weighted_k_row<DIM, MODE>(kc, v[0], row_terms<DIM>(v[1], DIM == 3 ? v[2] : 0.0f), a, b);
    }
}

// ---------------------------------------------------------------- B-spline

// Value and physical gradient of the cubic B-spline at p, coefficient grid
// padded by 3 and shifted by one; false outside [-0.5, n-0.5)
// (BsplineImage.h:247-296).  `g` is the raw image geometry.
template <int DIM>
__device__ __forceinline__ bool bspline_sample(const float *__restrict__ coef, const Grid &g, const float *p,
                                               float &val, float *grad)
{
    float idx[3];
    to_local<DIM>(g, p, idx);
    bool in = true;
    // written so that a NaN position is outside (the reference's negated test
    // would let it through and index with an undefined integer)
#pragma unroll
    for (int d = 0; d < DIM; ++d) in = in && (idx[d] >= -0.5f && idx[d] + 3.5f < (float)(g.n[d] + 3));
    if (!in) return false;
    int base[3] = {0, 0, 0};
    float w[3][4], dw[3][4];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        base[d] = (int)idx[d];
        float t = idx[d] - (float)base[d];
        float t2 = t * t;
        w[d][3] = (1.0f / 6.0f) * t2 * t;
        w[d][0] = (1.0f / 6.0f) + 0.5f * t * (t - 1.0f) - w[d][3];
        w[d][2] = t + w[d][0] - 2.0f * w[d][3];
        w[d][1] = 1.0f - w[d][0] - w[d][2] - w[d][3];
        dw[d][3] = 0.5f * t2;
        dw[d][0] = t - 0.5f - dw[d][3];
        dw[d][2] = 1.0f + dw[d][0] - 2.0f * dw[d][3];
        dw[d][1] = -dw[d][0] - dw[d][2] - dw[d][3];
    }
    const long long sx = g.n[0] + 3;
    if (DIM == 2) {
        float v = 0.0f, gx = 0.0f, gy = 0.0f;
        const float *row = coef + (long long)base[0] + (long long)base[1] * sx;
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            const float *c = row + y * sx;
            float r = 0.0f, rd = 0.0f;
#pragma unroll
            for (int x = 0; x < 4; ++x) {
                r = fmaf(w[0][x], c[x], r);
                rd = fmaf(dw[0][x], c[x], rd);
            }
            v = fmaf(w[1][y], r, v);
            gx = fmaf(w[1][y], rd, gx);
            gy = fmaf(dw[1][y], r, gy);
        }
        val = v;
        grad[0] = gx * g.ih[0];
        grad[1] = gy * g.ih[1];
    } else {
        const long long sy = g.n[1] + 3;
        float v = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f;
        const float *blk = coef + (long long)base[0] + (long long)base[1] * sx + (long long)base[2] * sx * sy;
#pragma unroll
        for (int z = 0; z < 4; ++z) {
            float az = 0.0f, axz = 0.0f, ayz = 0.0f;
#pragma unroll
            for (int y = 0; y < 4; ++y) {
                const float *c = blk + z * sx * sy + y * sx;
                float r = 0.0f, rd = 0.0f;
#pragma unroll
                for (int x = 0; x < 4; ++x) {
                    r = fmaf(w[0][x], c[x], r);
                    rd = fmaf(dw[0][x], c[x], rd);
                }
                az = fmaf(w[1][y], r, az);
                axz = fmaf(w[1][y], rd, axz);
                ayz = fmaf(dw[1][y], r, ayz);
            }
            v = fmaf(w[2][z], az, v);
            gx = fmaf(w[2][z], axz, gx);
            gy = fmaf(w[2][z], ayz, gy);
            gz = fmaf(dw[2][z], az, gz);
        }
        val = v;
        grad[0] = gx * g.ih[0];
        grad[1] = gy * g.ih[1];
        grad[2] = gz * g.ih[2];
    }
    return true;
}

// ---------------------------------------------------------------- PCG32

constexpr uint64_t kPcgMult = 6364136223846793005ULL;

// Output permutation XSH-RR of the PCG32 generator (UniformRng.h:34-41).
__host__ __device__ __forceinline__ uint32_t pcg32_output(uint64_t old)
{
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((uint32_t)(-(int32_t)rot) & 31));
}

// State after `delta` LCG steps, O(log delta) (O'Neill 2014, "PCG", sec. 4.3.1).
__host__ __device__ __forceinline__ uint64_t lcg_advance(uint64_t state, uint64_t delta, uint64_t inc)
{
    uint64_t acc_mult = 1, acc_plus = 0, cur_mult = kPcgMult, cur_plus = inc;
    while (delta > 0) {
        if (delta & 1) {
            acc_mult *= cur_mult;
            acc_plus = acc_plus * cur_mult + cur_plus;
        }
        cur_plus = (cur_mult + 1) * cur_plus;
        cur_mult *= cur_mult;
        delta >>= 1;
    }
    return acc_mult * state + acc_plus;
}

// ---------------------------------------------------------------- reductions

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block sum of K doubles in a fixed order; result valid in thread 0.
template <int K, int WAVES = kThreads / 64>
__device__ __forceinline__ void block_sum(double (&v)[K], double *lds /* K * WAVES */)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) lds[wave * K + k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double s = 0.0;
            for (int w = 0; w < WAVES; ++w) s += lds[w * K + k];
            v[k] = s;
        }
    }
}

}  // namespace skm
