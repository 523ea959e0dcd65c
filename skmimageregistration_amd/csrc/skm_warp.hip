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
