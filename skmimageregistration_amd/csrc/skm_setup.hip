// skm_setup.hip -- evaluator set-up work on the device: the cubic B-spline
// prefilter of the moving image (BsplineImage.h:37-108, 174-221) and the
// per-voxel scan of the weight image / tensor field (validation as in
// itkTransformAdapter.hxx:310-366, and the lam bounds of the pair reach).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "skm_internal.hpp"

namespace skm {
namespace {

// Unser-Thevenaz cubic B-spline prefilter of one line in FP64, pole
// sqrt(3)-2, mirror boundary (BsplineImage.h:37-70): gain, causal pass with
// the mirrored-signal initial value, anti-causal pass.
__device__ void prefilter_line(double *c, int n, long long stride, double z)
{
    if (n < 2) return;
    const double lambda = 2.0 - z - 1.0 / z;
    for (int i = 0; i < n; ++i) c[i * stride] *= lambda;
    double zn = z, iz = 1.0 / z, z2n = pow(z, (double)(n - 1));
    double sum = c[0] + z2n * c[(long long)(n - 1) * stride];
    z2n *= z2n * iz;
    for (int i = 1; i < n - 1; ++i) {
        sum += (zn + z2n) * c[i * stride];
        zn *= z;
        z2n *= iz;
    }
    c[0] = sum / (1.0 - zn * zn);
    for (int i = 1; i < n; ++i) c[i * stride] += z * c[(i - 1) * stride];
    c[(long long)(n - 1) * stride] =
        (z / (z * z - 1.0)) * (z * c[(long long)(n - 2) * stride] + c[(long long)(n - 1) * stride]);
    for (int i = n - 2; i >= 0; --i) c[i * stride] = z * (c[(i + 1) * stride] - c[i * stride]);
}

__global__ __launch_bounds__(kThreads) void k_to_double(const float *__restrict__ in, double *__restrict__ out,
                                                        long long n)
{
    const long long i = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (i < n) out[i] = in[i];
}

// One thread per line along `axis` of an n[0] x n[1] x n[2] volume (x fastest).
__global__ __launch_bounds__(kThreads) void k_prefilter_axis(double *__restrict__ c, int n0, int n1, int n2, int axis,
                                                             double z)
{
    const long long line = blockIdx.x * (long long)kThreads + threadIdx.x;
    long long base, stride;
    int len;
    if (axis == 0) {
        if (line >= (long long)n1 * n2) return;
        base = line * n0;
        stride = 1;
        len = n0;
    } else if (axis == 1) {
        if (line >= (long long)n0 * n2) return;
        base = line % n0 + (long long)n0 * n1 * (line / n0);
        stride = n0;
        len = n1;
    } else {
        if (line >= (long long)n0 * n1) return;
        base = line;
        stride = (long long)n0 * n1;
        len = n2;
    }
    prefilter_line(c + base, len, stride, z);
}

// Coefficients shifted by one and padded to (n+3)^D with mirrored borders
// (BsplineImage.h:72-108).
__device__ __forceinline__ int mirror(int x, int len) { return (len - 1) - abs((len - 1) - abs(x - 1)); }

__global__ __launch_bounds__(kThreads) void k_pad_mirror(const double *__restrict__ c, float *__restrict__ out,
                                                         int dim, int n0, int n1, int n2)
{
    const int p0 = n0 + 3, p1 = n1 + 3, p2 = dim == 3 ? n2 + 3 : 1;
    const long long i = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (i >= (long long)p0 * p1 * p2) return;
    const int x = (int)(i % p0), y = (int)((i / p0) % p1), zz = (int)(i / ((long long)p0 * p1));
    const int xi = mirror(x, n0), yi = mirror(y, n1), zi = dim == 3 ? mirror(zz, n2) : 0;
    out[i] = (float)c[xi + (long long)n0 * (yi + (long long)n1 * zi)];
}

// Largest eigenvalue of the symmetric part of a DIM x DIM row-major matrix
// in FP64 (trigonometric closed form for 3x3).
template <int DIM>
__device__ double max_eig_sym(const float *m)
{
    double a[9];
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int j = 0; j < DIM; ++j) a[i * DIM + j] = 0.5 * ((double)m[i * DIM + j] + (double)m[j * DIM + i]);
    if (DIM == 2) {
        const double tr = a[0] + a[3], det = a[0] * a[3] - a[1] * a[2];
        return 0.5 * (tr + sqrt(fmax(tr * tr - 4.0 * det, 0.0)));
    }
    const double p1 = a[1] * a[1] + a[2] * a[2] + a[5] * a[5];
    if (p1 == 0.0) return fmax(a[0], fmax(a[4], a[8]));
    const double q = (a[0] + a[4] + a[8]) / 3.0;
    const double p2 = (a[0] - q) * (a[0] - q) + (a[4] - q) * (a[4] - q) + (a[8] - q) * (a[8] - q) + 2.0 * p1;
    const double p = sqrt(p2 / 6.0);
    double b[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) b[i] = (a[i] - (i % 4 == 0 ? q : 0.0)) / p;
    const double r = 0.5 * (b[0] * (b[4] * b[8] - b[5] * b[7]) - b[1] * (b[3] * b[8] - b[5] * b[6]) +
                            b[2] * (b[3] * b[7] - b[4] * b[6]));
    const double phi = acos(fmin(1.0, fmax(-1.0, r))) / 3.0;
    return q + 2.0 * p * cos(phi);
}

template <int DIM>
__device__ double det_full(const float *m)
{
    if (DIM == 2) return (double)m[0] * m[3] - (double)m[2] * m[1];
    const double a0 = m[0], a1 = m[1], a2 = m[2], a3 = m[3], a4 = m[4], a5 = m[5], a6 = m[6], a7 = m[7], a8 = m[8];
    return a0 * (a4 * a8 - a5 * a7) + a1 * (a5 * a6 - a3 * a8) + a2 * (a3 * a7 - a4 * a6);
}

// Per weight voxel: lam (sigma, or the largest eigenvalue of the symmetric
// part of the covariance) and the validation of SetWeightingImage /
// skm_create (sigma > 1e-8; det > 0).  Block partials: max lam, invalid count.
template <int DIM>
__global__ __launch_bounds__(kThreads) void k_weight_lam(int mode, long long nw, const float *__restrict__ w,
                                                         float *__restrict__ lam, double *__restrict__ blk_max,
                                                         unsigned *__restrict__ blk_bad)
{
    __shared__ double s_max[kThreads / 64];
    __shared__ unsigned s_bad[kThreads / 64];
    const long long i = blockIdx.x * (long long)kThreads + threadIdx.x;
    double l = 0.0;
    unsigned bad = 0;
    if (i < nw) {
        if (mode == 1) {
            l = w[i];
            bad = !(w[i] > 1e-8f);
        } else {
            const float *m = w + i * (DIM * DIM);
            l = max_eig_sym<DIM>(m);
            bad = !(det_full<DIM>(m) > 0.0);
        }
        lam[i] = (float)(l * (1.0 + 1e-6));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        l = fmax(l, __shfl_xor(l, off, 64));
        bad += __shfl_xor(bad, off, 64);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        s_max[wave] = l;
        s_bad[wave] = bad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kThreads / 64; ++k) {
            l = fmax(l, s_max[k]);
            bad += s_bad[k];
        }
        blk_max[blockIdx.x] = l;
        blk_bad[blockIdx.x] = bad;
    }
}

// Largest lam over the points every row of sample cells (my, mz) can take:
// voxel rows my*sub .. my*sub+jit in y and z (jit = sub when sub > 1, the
// jitter offset reaches sub), every x; positions off the weight grid take
// lam_out.  One thread per row.
__global__ __launch_bounds__(kThreads) void k_sample_row_lams(int dim, Grid fixed, Grid weight, int sub, int ns0,
                                                              int ns1, int ns2, const float *__restrict__ wlam,
                                                              float lam_out, float *__restrict__ out)
{
    const long long r = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (r >= (long long)ns1 * ns2) return;
    const int my = (int)(r % ns1), mz = (int)(r / ns1);
    const int jit = sub > 1 ? sub : 0;
    // the nearest weight voxel exactly as the kernels look it up (nearest())
    auto widx = [&](int d, int i) {
        const float p = fixed.o[d] + fixed.h[d] * (float)i;
        const int k = (int)(weight.ih[d] * (p - weight.o[d]) + 0.5f);
        return (k >= 0 && k < weight.n[d]) ? k : -1;
    };
    float m = 0.0f;
    const int z0 = dim == 3 ? mz * sub : 0, z1 = dim == 3 ? mz * sub + jit : 0;
    for (int zz = z0; zz <= z1; ++zz) {
        const int c = dim == 3 ? widx(2, zz) : 0;
        for (int y = my * sub; y <= my * sub + jit; ++y) {
            const int b = widx(1, y);
            for (int x = 0; x < ns0 * sub + jit; ++x) {
                const int a = widx(0, x);
                const float v = (a < 0 || b < 0 || c < 0)
                                    ? lam_out
                                    : wlam[a + (long long)weight.n[0] * (b + (long long)weight.n[1] * c)];
                m = fmaxf(m, v);
            }
        }
    }
    out[r] = m;
}

// dst[i * per + k] = src[idx[i] * per + k] (idx >= 0), fill[k] otherwise.
__global__ __launch_bounds__(kThreads) void k_gather_values(const float *__restrict__ src,
                                                            const long long *__restrict__ idx, int per, long long n,
                                                            float fill_diag, float *__restrict__ dst)
{
    const long long t = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (t >= n * per) return;
    const long long i = t / per;
    const int k = (int)(t % per);
    const long long f = idx[i];
    // outside the grid: identity (tensor, diagonal entries) or 0 (weight image)
    const int d = per == 4 ? 2 : 3;
    const float outside = per == 1 ? 0.0f : ((k % (d + 1)) == 0 ? fill_diag : 0.0f);
    dst[t] = f >= 0 ? src[f * per + k] : outside;
}

// Zeroes n words with vector stores (a kernel rather than a memset node, so
// that a stream capture of the caller records only kernels).
__global__ __launch_bounds__(kThreads) void k_zero_words(unsigned *__restrict__ p, long long n)
{
    const long long stride = (long long)gridDim.x * kThreads;
    for (long long t = blockIdx.x * (long long)kThreads + threadIdx.x; t < n; t += stride) p[t] = 0u;
}

unsigned blocks(long long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

}  // namespace

void launch_zero_words(void *d_ptr, long long n_words, hipStream_t s)
{
    if (n_words <= 0) return;
    const unsigned g = (unsigned)std::min<long long>(blocks(n_words), 8192);
    k_zero_words<<<g, kThreads, 0, s>>>(static_cast<unsigned *>(d_ptr), n_words);
}

void launch_bspline_prefilter(int dim, const int *size, const float *d_src, double *d_work, hipStream_t s)
{
    const int n0 = size[0], n1 = size[1], n2 = dim == 3 ? size[2] : 1;
    const long long n = (long long)n0 * n1 * n2;
    const double pole = sqrt(3.0) - 2.0;
    k_to_double<<<blocks(n), kThreads, 0, s>>>(d_src, d_work, n);
    // the order of the host restatement: 3D z, y, x; 2D x, y
    if (dim == 3) {
        k_prefilter_axis<<<blocks((long long)n0 * n1), kThreads, 0, s>>>(d_work, n0, n1, n2, 2, pole);
        k_prefilter_axis<<<blocks((long long)n0 * n2), kThreads, 0, s>>>(d_work, n0, n1, n2, 1, pole);
        k_prefilter_axis<<<blocks((long long)n1 * n2), kThreads, 0, s>>>(d_work, n0, n1, n2, 0, pole);
    } else {
        k_prefilter_axis<<<blocks((long long)n1), kThreads, 0, s>>>(d_work, n0, n1, 1, 0, pole);
        k_prefilter_axis<<<blocks((long long)n0), kThreads, 0, s>>>(d_work, n0, n1, 1, 1, pole);
    }
}

void launch_bspline_coefficients(int dim, const int *size, const float *d_src, double *d_work, float *d_coeffs,
                                 hipStream_t s)
{
    const int n0 = size[0], n1 = size[1], n2 = dim == 3 ? size[2] : 1;
    launch_bspline_prefilter(dim, size, d_src, d_work, s);
    const long long np = (long long)(n0 + 3) * (n1 + 3) * (dim == 3 ? n2 + 3 : 1);
    k_pad_mirror<<<blocks(np), kThreads, 0, s>>>(d_work, d_coeffs, dim, n0, n1, n2);
}

void launch_gather_values(const float *d_src, const long long *d_idx, int per, long long n, float *d_dst,
                          hipStream_t s)
{
    if (n <= 0) return;
    k_gather_values<<<blocks(n * per), kThreads, 0, s>>>(d_src, d_idx, per, n, 1.0f, d_dst);
}

unsigned weight_lam_blocks(long long nw) { return blocks(nw); }

void launch_weight_lam(int dim, int mode, long long nw, const float *d_w, float *d_lam, double *d_blk_max,
                       unsigned *d_blk_bad, hipStream_t s)
{
    if (dim == 2) k_weight_lam<2><<<blocks(nw), kThreads, 0, s>>>(mode, nw, d_w, d_lam, d_blk_max, d_blk_bad);
    else k_weight_lam<3><<<blocks(nw), kThreads, 0, s>>>(mode, nw, d_w, d_lam, d_blk_max, d_blk_bad);
}

void launch_sample_row_lams(int dim, const Grid &fixed, const Grid &weight, int sub, const int *ncell,
                            const float *d_wlam, float lam_out, float *d_out, hipStream_t s)
{
    const long long rows = (long long)ncell[1] * ncell[2];
    k_sample_row_lams<<<blocks(rows), kThreads, 0, s>>>(dim, fixed, weight, sub, ncell[0], ncell[1], ncell[2], d_wlam,
                                                        lam_out, d_out);
}

}  // namespace skm
