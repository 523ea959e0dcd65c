// skm_warp.hip -- warped images from a dense displacement field, on the GPU:
//   * the final registered image of SKMReg: itk::ResampleImageFilter with an
//     itk::DisplacementFieldTransform and the default linear interpolator,
//     default value 0 (src/SKMReg.cxx:507-525);
//   * WarpImage: itk::WarpImageFilter with a cubic B-spline interpolator over
//     the displacement's own grid (include/itkUtils.h:163-179).
//
// One thread per output voxel; geometry and interpolation weights in double
// (ITK's interpolators compute in double), images and fields in float.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "skm.h"
#include "skm_internal.hpp"

namespace skm {
namespace {

// Geometry in double: size, origin, spacing and its reciprocal.
struct Geo {
    int n[3];
    double o[3], h[3], ih[3];
};

Geo to_geo(const skm_grid &g, int dim)
{
    Geo r;
    for (int d = 0; d < 3; ++d) {
        const bool used = d < dim;
        r.n[d] = used ? g.size[d] : 1;
        r.o[d] = used ? g.origin[d] : 0.0;
        r.h[d] = used ? g.spacing[d] : 1.0;
        r.ih[d] = 1.0 / r.h[d];
    }
    return r;
}

// Physical point -> continuous index, (p - origin) times the inverse spacing
// (identity direction).
template <int DIM>
__device__ __forceinline__ void to_cindex(const Geo &g, const double *p, double *ci)
{
#pragma unroll
    for (int d = 0; d < DIM; ++d) ci[d] = (p[d] - g.o[d]) * g.ih[d];
}

// ImageFunction::IsInsideBuffer for a continuous index: [-0.5, n - 0.5) on
// every axis.  NaN is outside.
template <int DIM>
__device__ __forceinline__ bool inside_buffer(const Geo &g, const double *ci)
{
    bool in = true;
#pragma unroll
    for (int d = 0; d < DIM; ++d) in = in && ci[d] >= -0.5 && ci[d] < (double)g.n[d] - 0.5;
    return in;
}

// Linear interpolation with the neighbours clamped to the buffer: the lower
// neighbour of an index in [-0.5, 0) and the upper one of an index in
// (n - 1, n - 0.5) are the edge voxel (LinearInterpolateImageFunction and
// VectorLinearInterpolateImageFunction).  C interleaved components.
template <int DIM, int C>
__device__ void interp_linear(const float *__restrict__ data, const Geo &g, const double *ci, double *out)
{
    int base[3] = {0, 0, 0};
    double frac[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const double fl = floor(ci[d]);
        base[d] = (int)fl;
        frac[d] = ci[d] - fl;
    }
#pragma unroll
    for (int c = 0; c < C; ++c) out[c] = 0.0;
#pragma unroll
    for (int corner = 0; corner < (1 << DIM); ++corner) {
        double w = 1.0;
        long long idx = 0, stride = 1;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const int up = (corner >> d) & 1;
            const int i = min(max(base[d] + up, 0), g.n[d] - 1);
            w *= up ? frac[d] : 1.0 - frac[d];
            idx += stride * i;
            stride *= g.n[d];
        }
        if (w == 0.0) continue;
#pragma unroll
        for (int c = 0; c < C; ++c) out[c] += w * (double)data[idx * C + c];
    }
}

// Index of a cubic B-spline tap reflected at the edges (mirror boundary of
// BSplineInterpolateImageFunction: -i below 0, 2(n-1) - i above n-1); kept
// inside the buffer for lines shorter than the stencil.
__device__ __forceinline__ int mirror_tap(int i, int n)
{
    if (n == 1) return 0;
    if (i < 0) i = -i;
    if (i > n - 1) i = 2 * (n - 1) - i;
    return min(max(i, 0), n - 1);
}

// Cubic B-spline weights of the taps floor(x) - 1 .. floor(x) + 2 for the
// fraction t = x - floor(x): the uniform cubic B-spline
// beta3 evaluated at t + 1, t, 1 - t, 2 - t.
__device__ __forceinline__ void cubic_weights(double t, double *w)
{
    const double t2 = t * t, t3 = t2 * t, s = 1.0 - t;
    w[0] = s * s * s / 6.0;
    w[1] = (3.0 * t3 - 6.0 * t2 + 4.0) / 6.0;
    w[2] = (-3.0 * t3 + 3.0 * t2 + 3.0 * t + 1.0) / 6.0;
    w[3] = t3 / 6.0;
}

// ResampleImageFilter + DisplacementFieldTransform: x -> x + u(x) where u is
// interpolated linearly inside the field's buffer and zero outside it
// (DisplacementFieldTransform::TransformPoint returns the input point there);
// the image is interpolated linearly inside its buffer, default value outside.
template <int DIM>
__global__ __launch_bounds__(kThreads) void k_warp_linear(const float *__restrict__ image, Geo gi,
                                                          const float *__restrict__ df, Geo gd, Geo go,
                                                          double default_value, float *__restrict__ out,
                                                          long long n_out)
{
    const long long i = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (i >= n_out) return;
    const int idx[3] = {(int)(i % go.n[0]), (int)((i / go.n[0]) % go.n[1]), (int)(i / ((long long)go.n[0] * go.n[1]))};
    double p[3], ci[3];
#pragma unroll
    for (int d = 0; d < DIM; ++d) p[d] = go.o[d] + go.h[d] * (double)idx[d];
    to_cindex<DIM>(gd, p, ci);
    if (inside_buffer<DIM>(gd, ci)) {
        double u[DIM];
        interp_linear<DIM, DIM>(df, gd, ci, u);
#pragma unroll
        for (int d = 0; d < DIM; ++d) p[d] += u[d];
    }
    to_cindex<DIM>(gi, p, ci);
    double v = default_value;
    if (inside_buffer<DIM>(gi, ci)) interp_linear<DIM, 1>(image, gi, ci, &v);
    out[i] = (float)v;
}

// WarpImageFilter with the output grid of the displacement field: x -> x +
// u[x] at every voxel; cubic B-spline interpolation (double coefficients of
// the mirror-boundary prefilter) inside the image buffer, default value
// outside.
template <int DIM>
__global__ __launch_bounds__(kThreads) void k_warp_bspline3(const double *__restrict__ coeffs, Geo gi,
                                                            const float *__restrict__ df, Geo go,
                                                            double default_value, float *__restrict__ out,
                                                            long long n_out)
{
    const long long i = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (i >= n_out) return;
    const int idx[3] = {(int)(i % go.n[0]), (int)((i / go.n[0]) % go.n[1]), (int)(i / ((long long)go.n[0] * go.n[1]))};
    double p[3], ci[3];
#pragma unroll
    for (int d = 0; d < DIM; ++d) p[d] = go.o[d] + go.h[d] * (double)idx[d] + (double)df[i * DIM + d];
    to_cindex<DIM>(gi, p, ci);
    if (!inside_buffer<DIM>(gi, ci)) {
        out[i] = (float)default_value;
        return;
    }
    double w[3][4];
    int tap[3][4];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (d < DIM) {
            const double fl = floor(ci[d]);
            cubic_weights(ci[d] - fl, w[d]);
#pragma unroll
            for (int k = 0; k < 4; ++k) tap[d][k] = mirror_tap((int)fl - 1 + k, gi.n[d]);
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                w[d][k] = k == 0 ? 1.0 : 0.0;
                tap[d][k] = 0;
            }
        }
    }
    double v = 0.0;
    const int kz = DIM == 3 ? 4 : 1;
    for (int c = 0; c < kz; ++c) {
        double vy = 0.0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const long long row = (long long)gi.n[0] * (tap[1][b] + (long long)gi.n[1] * tap[2][c]);
            double vx = 0.0;
#pragma unroll
            for (int a = 0; a < 4; ++a) vx += w[0][a] * coeffs[row + tap[0][a]];
            vy += w[1][b] * vx;
        }
        v += w[2][c] * vy;
    }
    out[i] = (float)v;
}

#define SKM_W_HIP(expr)                                                                                            \
    do {                                                                                                           \
        hipError_t e_ = (expr);                                                                                    \
        if (e_ != hipSuccess) {                                                                                    \
            (void)hipStreamSynchronize(nullptr); /* temporaries are released on return */                          \
            return set_last_error(e_ == hipErrorOutOfMemory ? SKM_ERROR_OUT_OF_MEMORY : SKM_ERROR_HIP,             \
                                  std::string(#expr) + ": " + hipGetErrorString(e_));                              \
        }                                                                                                          \
    } while (0)

template <typename T>
struct Buf {
    T *p = nullptr;
    ~Buf() { device_release(p); }
    hipError_t alloc(long long count)
    {
        return device_alloc((void **)&p, (size_t)std::max<long long>(count, 1) * sizeof(T));
    }
};

skm_status check_grid(const skm_grid *g, int dim, const char *what)
{
    if (!g) return set_last_error(SKM_ERROR_INVALID_ARGUMENT, std::string("null ") + what + " grid");
    for (int d = 0; d < dim; ++d) {
        if (g->size[d] < 1) return set_last_error(SKM_ERROR_INVALID_ARGUMENT, std::string(what) + " size must be positive");
        if (!(g->spacing[d] > 0.0))
            return set_last_error(SKM_ERROR_INVALID_ARGUMENT, std::string(what) + " spacing must be positive");
    }
    return SKM_SUCCESS;
}

long long count(const skm_grid &g, int dim)
{
    long long n = 1;
    for (int d = 0; d < dim; ++d) n *= g.size[d];
    return n;
}

bool same_grid(const skm_grid &a, const skm_grid &b, int dim)
{
    for (int d = 0; d < dim; ++d)
        if (a.size[d] != b.size[d] || a.origin[d] != b.origin[d] || a.spacing[d] != b.spacing[d]) return false;
    return true;
}

unsigned blocks_of(long long n) { return (unsigned)std::max<long long>(1, (n + kThreads - 1) / kThreads); }

// ---- synthetic registration pairs (random ellipsoids, smooth displacement)

struct Ellipsoid {
    float c[3], inv_ax[3], value;
};
struct Rbf {
    float c[3], w[3];
};
constexpr int kMaxEllipsoids = 64, kMaxRbf = 32;
struct SynthShapes {
    Ellipsoid e[kMaxEllipsoids];
    int ne;
};
struct SynthField {
    Rbf r[kMaxRbf];
    int nr;
    float inv2s2;
};

// value of the last ellipsoid (in order) holding the voxel, 0 outside all
__global__ __launch_bounds__(kThreads) void k_synth_ellipsoids(SynthShapes sh, int nx, int ny, int nz,
                                                               float *__restrict__ out)
{
    const long long i = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (i >= (long long)nx * ny * nz) return;
    const float p[3] = {(float)(i % nx), (float)((i / nx) % ny), (float)(i / ((long long)nx * ny))};
    float v = 0.0f;
    for (int k = 0; k < sh.ne; ++k) {
        const Ellipsoid &e = sh.e[k];
        float q = 0.0f;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const float t = (p[d] - e.c[d]) * e.inv_ax[d];
            q += t * t;
        }
        if (q <= 1.0f) v = e.value;
    }
    out[i] = v;
}

// sum of Gaussian RBFs per voxel (AoS), with the block maxima of |d|^2
__global__ __launch_bounds__(kThreads) void k_synth_field(SynthField f, int nx, int ny, int nz, float *__restrict__ out,
                                                          float *__restrict__ blk_max)
{
    __shared__ float lds[kThreads / 64];
    const long long i = blockIdx.x * (long long)kThreads + threadIdx.x;
    float m = 0.0f;
    if (i < (long long)nx * ny * nz) {
        const float p[3] = {(float)(i % nx), (float)((i / nx) % ny), (float)(i / ((long long)nx * ny))};
        double u[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k < f.nr; ++k) {
            const Rbf &r = f.r[k];
            const float dx = p[0] - r.c[0], dy = p[1] - r.c[1], dz = p[2] - r.c[2];
            const double g = exp(-(double)(dx * dx + dy * dy + dz * dz) * (double)f.inv2s2);
#pragma unroll
            for (int d = 0; d < 3; ++d) u[d] += g * (double)r.w[d];
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) out[i * 3 + d] = (float)u[d];
        m = (float)(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    }
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.0f;
        for (int w = 0; w < kThreads / 64; ++w) t = fmaxf(t, lds[w]);
        blk_max[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(kThreads) void k_scale(float *__restrict__ a, long long n, float s)
{
    const long long i = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (i < n) a[i] *= s;
}

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// adds N(0, sigma^2) per voxel: Box-Muller on two uniforms hashed from (seed, voxel)
__global__ __launch_bounds__(kThreads) void k_synth_noise(float *__restrict__ img, long long n, uint64_t seed,
                                                          float sigma)
{
    const long long i = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (i >= n) return;
    const uint64_t h = mix64(seed ^ mix64((uint64_t)i));
    const double u1 = ((double)(h >> 11) + 1.0) * (1.0 / 9007199254740993.0);  // (0, 1]
    const double u2 = (double)(mix64(h) >> 11) * (1.0 / 9007199254740992.0);
    img[i] += (float)((double)sigma * sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
}

// host stream of uniform doubles in [0, 1) (splitmix64)
struct HostRng {
    uint64_t s;
    double next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    }
    double uniform(double a, double b) { return a + (b - a) * next(); }
};

}  // namespace
}  // namespace skm

using namespace skm;

extern "C" {

SKM_API skm_status skm_warp_image(int32_t dim, const skm_grid *image_grid, const float *image, const skm_grid *df_grid,
                                  const float *df_aos, const skm_grid *out_grid, int32_t interpolation,
                                  double default_value, float *out)
{
    if (dim != 2 && dim != 3) return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "dimension must be 2 or 3");
    if (!image || !df_aos || !out) return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "null argument");
    skm_status st;
    if ((st = check_grid(image_grid, dim, "image")) != SKM_SUCCESS) return st;
    if ((st = check_grid(df_grid, dim, "displacement")) != SKM_SUCCESS) return st;
    if ((st = check_grid(out_grid, dim, "output")) != SKM_SUCCESS) return st;
    if (interpolation != SKM_INTERPOLATION_LINEAR && interpolation != SKM_INTERPOLATION_BSPLINE3)
        return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "interpolation must be linear (1) or cubic B-spline (3)");
    if (interpolation == SKM_INTERPOLATION_BSPLINE3 && !same_grid(*out_grid, *df_grid, dim))
        return set_last_error(SKM_ERROR_INVALID_ARGUMENT,
                              "the cubic warp writes the displacement field's grid: out_grid must equal df_grid");
    const long long ni = count(*image_grid, dim), nd = count(*df_grid, dim), no = count(*out_grid, dim);
    const Geo gi = to_geo(*image_grid, dim), gd = to_geo(*df_grid, dim), go = to_geo(*out_grid, dim);

    Buf<float> d_img, d_df, d_out;
    SKM_W_HIP(d_img.alloc(ni));
    SKM_W_HIP(d_df.alloc(nd * dim));
    SKM_W_HIP(d_out.alloc(no));
    SKM_W_HIP(hipMemcpy(d_img.p, image, sizeof(float) * ni, hipMemcpyHostToDevice));
    SKM_W_HIP(hipMemcpy(d_df.p, df_aos, sizeof(float) * nd * dim, hipMemcpyHostToDevice));
    if (interpolation == SKM_INTERPOLATION_LINEAR) {
        if (dim == 2)
            k_warp_linear<2><<<blocks_of(no), kThreads>>>(d_img.p, gi, d_df.p, gd, go, default_value, d_out.p, no);
        else
            k_warp_linear<3><<<blocks_of(no), kThreads>>>(d_img.p, gi, d_df.p, gd, go, default_value, d_out.p, no);
    } else {
        Buf<double> coeffs;
        SKM_W_HIP(coeffs.alloc(ni));
        launch_bspline_prefilter(dim, image_grid->size, d_img.p, coeffs.p, nullptr);
        if (dim == 2)
            k_warp_bspline3<2><<<blocks_of(no), kThreads>>>(coeffs.p, gi, d_df.p, go, default_value, d_out.p, no);
        else
            k_warp_bspline3<3><<<blocks_of(no), kThreads>>>(coeffs.p, gi, d_df.p, go, default_value, d_out.p, no);
        SKM_W_HIP(hipGetLastError());
        SKM_W_HIP(hipDeviceSynchronize());
    }
    SKM_W_HIP(hipGetLastError());
    SKM_W_HIP(hipMemcpy(out, d_out.p, sizeof(float) * no, hipMemcpyDeviceToHost));
    return SKM_SUCCESS;
}

}  // extern "C"

extern "C" SKM_API skm_status skm_synthetic_pair(const skm_synthetic_desc *desc, float *reference, float *target,
                                                float *displacement_aos)
{
    if (!desc || !reference || !target) return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "null argument");
    const int nx = desc->size[0], ny = desc->size[1], nz = desc->size[2];
    if (nx < 4 || ny < 4 || nz < 4) return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "volume too small");
    if (desc->n_ellipsoids < 0 || desc->n_ellipsoids > kMaxEllipsoids || desc->n_rbf < 1 || desc->n_rbf > kMaxRbf)
        return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "ellipsoid or RBF count out of range");
    if (!(desc->rbf_sigma > 0.0) || !(desc->smoothing_sigma >= 0.0) || !(desc->noise_sigma >= 0.0))
        return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "sigmas must be positive");
    const long long n = (long long)nx * ny * nz;
    const int size[3] = {nx, ny, nz};
    HostRng rng{desc->seed};
    SynthShapes sh{};
    sh.ne = desc->n_ellipsoids;
    for (int k = 0; k < sh.ne; ++k) {
        for (int d = 0; d < 3; ++d) sh.e[k].c[d] = (float)(rng.uniform(0.15, 0.85) * size[d]);
        for (int d = 0; d < 3; ++d) sh.e[k].inv_ax[d] = (float)(1.0 / (rng.uniform(0.05, 0.2) * size[d]));
        sh.e[k].value = (float)rng.next();
    }
    SynthField f{};
    f.nr = desc->n_rbf;
    f.inv2s2 = (float)(1.0 / (2.0 * desc->rbf_sigma * desc->rbf_sigma));
    for (int k = 0; k < f.nr; ++k) {
        for (int d = 0; d < 3; ++d) f.r[k].c[d] = (float)rng.uniform(0.0, size[d] - 1);
        for (int d = 0; d < 3; ++d) f.r[k].w[d] = (float)rng.uniform(-1.0, 1.0);
    }
    skm_grid g{};
    for (int d = 0; d < 3; ++d) {
        g.size[d] = size[d];
        g.spacing[d] = 1.0;
    }
    const Geo geo = to_geo(g, 3);
    const unsigned nb = blocks_of(n);
    Buf<float> img, df, out, bmax;
    Buf<double> coeffs;
    SKM_W_HIP(img.alloc(n));
    SKM_W_HIP(df.alloc(n * 3));
    SKM_W_HIP(out.alloc(n));
    SKM_W_HIP(bmax.alloc(nb));
    SKM_W_HIP(coeffs.alloc(n));
    k_synth_ellipsoids<<<nb, kThreads>>>(sh, nx, ny, nz, img.p);
    SKM_W_HIP(hipGetLastError());
    if (desc->smoothing_sigma > 0.0) {
        const skm_status st =
            gaussian_smooth_device(3, g, img.p, desc->smoothing_sigma * desc->smoothing_sigma, img.p, nullptr);
        if (st != SKM_SUCCESS) return st;
    }
    k_synth_field<<<nb, kThreads>>>(f, nx, ny, nz, df.p, bmax.p);
    std::vector<float> hmax(nb);
    SKM_W_HIP(hipMemcpy(hmax.data(), bmax.p, sizeof(float) * nb, hipMemcpyDeviceToHost));
    float m2 = 0.0f;
    for (float v : hmax) m2 = std::max(m2, v);
    const double scale = desc->max_displacement / std::max(1e-12, std::sqrt((double)m2)) * desc->displacement_scale;
    k_scale<<<blocks_of(n * 3), kThreads>>>(df.p, n * 3, (float)scale);
    launch_bspline_prefilter(3, size, img.p, coeffs.p, nullptr);
    k_warp_bspline3<3><<<nb, kThreads>>>(coeffs.p, geo, df.p, geo, 0.0, out.p, n);
    if (desc->noise_sigma > 0.0) k_synth_noise<<<nb, kThreads>>>(out.p, n, desc->seed, (float)desc->noise_sigma);
    SKM_W_HIP(hipGetLastError());
    SKM_W_HIP(hipMemcpy(reference, img.p, sizeof(float) * n, hipMemcpyDeviceToHost));
    SKM_W_HIP(hipMemcpy(target, out.p, sizeof(float) * n, hipMemcpyDeviceToHost));
    if (displacement_aos) SKM_W_HIP(hipMemcpy(displacement_aos, df.p, sizeof(float) * n * 3, hipMemcpyDeviceToHost));
    return SKM_SUCCESS;
}
