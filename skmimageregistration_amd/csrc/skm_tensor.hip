// skm_tensor.hip -- image preprocessing of the anisotropic kernel on the GPU:
// the discrete Gaussian of GaussianSmoothing (include/itkUtils.h:140-151) and
// the covariance field of ComputeAnisotropicTensor
// (include/AnisotropicTensorUtils.h:37-355) with the maximum-tilt logic of
// TransformAdapter::SetWeightingTensor (include/itkTransformAdapter.hxx:335-366).
//
// Pipeline, on one stream, float images in HBM and double arithmetic per tap:
//   image * weight -> separable discrete Gaussian (variance sigma), last axis first
//   -> central-difference gradient and its outer product, D(D+1)/2 float planes
//   -> the same separable Gaussian on every plane (variance sigma / 4)
//   -> per voxel: FP64 Jacobi eigen-decomposition, re-weighting of the
//      largest-magnitude direction, covariance a e1 e1^T + c (I - e1 e1^T)
//   -> per-block maximum of the tilt, reduced on the host.
//
// The reference assembles the covariance as U S S U^T from re-scaled eigenvector
// rows (AnisotropicTensorUtils.h:177-355).  The columns of U S are e1/(1+s),
// c v_mid and c v_min with c = 1 + s/(1+s) and s = (|lambda_max| factor)^1.5,
// so the product is a e1 e1^T + c (I - e1 e1^T) with a = 1/(1+s): only the
// largest-magnitude eigenvector is needed, which makes the result independent
// of the basis chosen inside a degenerate eigenspace.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "skm.h"
#include "skm_internal.hpp"

namespace skm {
namespace {

// ------------------------------------------------------------ host: kernel

// Modified Bessel I_0 and I_1 by the polynomial approximations of Abramowitz &
// Stegun 9.8.1-9.8.4, the functions itk::GaussianOperator evaluates.
double bessel_i0(double y)
{
    const double d = std::fabs(y);
    if (d < 3.75) {
        const double m = (y / 3.75) * (y / 3.75);
        return 1.0 + m * (3.5156229 + m * (3.0899424 + m * (1.2067492 + m * (0.2659732 + m * (0.360768e-1 +
                                                                                            m * 0.45813e-2)))));
    }
    const double m = 3.75 / d;
    const double poly =
        0.39894228 +
        m * (0.1328592e-1 +
             m * (0.225319e-2 +
                  m * (-0.157565e-2 +
                       m * (0.916281e-2 + m * (-0.2057706e-1 + m * (0.2635537e-1 + m * (-0.1647633e-1 +
                                                                                         m * 0.392377e-2)))))));
    return std::exp(d) / std::sqrt(d) * poly;
}

double bessel_i1(double y)
{
    const double d = std::fabs(y);
    double acc;
    if (d < 3.75) {
        const double m = (y / 3.75) * (y / 3.75);
        acc = d * (0.5 + m * (0.87890594 + m * (0.51498869 + m * (0.15084934 + m * (0.2658733e-1 +
                                                                                  m * (0.301532e-2 +
                                                                                       m * 0.32411e-3))))));
    } else {
        const double m = 3.75 / d;
        acc = 0.2282967e-1 + m * (-0.2895312e-1 + m * (0.1787654e-1 - m * 0.420059e-2));
        acc = 0.39894228 +
              m * (-0.3988024e-1 + m * (-0.362018e-2 + m * (0.163801e-2 + m * (-0.1031555e-1 + m * acc))));
        acc *= std::exp(d) / std::sqrt(d);
    }
    return y < 0.0 ? -acc : acc;
}

constexpr double kMaximumError = 0.01;   // DiscreteGaussianImageFilter default
constexpr size_t kMaximumWidth = 64;     // GaussianSmoothing (itkUtils.h:148)

// itk::GaussianOperator::GenerateCoefficients for a variance in pixels: the
// discrete Gaussian exp(-t) I_n(t), grown by the Bessel recurrence until the
// two-sided sum reaches 1 - maximum error, a coefficient is not positive or
// the half kernel exceeds the maximum width; normalised by that sum.
// Returns the full symmetric kernel (2r+1 taps).
std::vector<double> gaussian_kernel(double variance)
{
    const double et = std::exp(-variance);
    const double cap = 1.0 - kMaximumError;
    std::vector<double> half{et * bessel_i0(variance), et * bessel_i1(variance)};
    double total = half[0] + 2.0 * half[1];
    for (size_t i = 2; total < cap; ++i) {
        half.push_back(half[i - 2] - 2.0 * (double)(i - 1) * half[i - 1] / variance);
        total += 2.0 * half[i];
        if (half[i] <= 0.0) break;
        if (half.size() > kMaximumWidth) break;
    }
    for (double &c : half) c /= total;
    std::vector<double> full(2 * half.size() - 1);
    const size_t r = half.size() - 1;
    for (size_t i = 0; i <= r; ++i) full[r + i] = full[r - i] = half[i];
    return full;
}

// ------------------------------------------------------------ device

struct Dims {
    int n[3];          // size per axis (1 beyond the dimension)
    long long s[3];    // element stride per axis
    long long n_vox;
};

Dims make_dims(int dim, const skm_grid &g)
{
    Dims d;
    long long stride = 1;
    for (int a = 0; a < 3; ++a) {
        d.n[a] = a < dim ? g.size[a] : 1;
        d.s[a] = stride;
        stride *= d.n[a];
    }
    d.n_vox = stride;
    return d;
}

__device__ __forceinline__ int axis_coord(long long v, const Dims &d, int axis)
{
    return (int)((v / d.s[axis]) % d.n[axis]);
}

// One separable pass of the discrete Gaussian along `axis` over `planes`
// planes: edge-replicated taps, double sums, float result
// (NeighborhoodOperatorImageFilter with a zero-flux Neumann boundary).
// Reads are multiplied by `prescale` in float first (MultiplyImageFilter).
__global__ void __launch_bounds__(kThreads) k_smooth_axis(const float *__restrict__ in, float *__restrict__ out,
                                                          const double *__restrict__ coeffs, int radius, int axis,
                                                          Dims d, int planes, float prescale)
{
    const long long total = d.n_vox * planes;
    const long long stride = d.s[axis];
    const int m = d.n[axis];
    for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < total;
         i += (long long)gridDim.x * kThreads) {
        const long long v = i % d.n_vox;
        const int c = axis_coord(v, d, axis);
        const float *line = in + (i - (long long)c * stride);
        double acc = 0.0;
        for (int k = -radius; k <= radius; ++k) {
            const int j = min(max(c + k, 0), m - 1);
            acc += coeffs[k + radius] * (double)(line[(long long)j * stride] * prescale);
        }
        out[i] = (float)acc;
    }
}

// Central-difference gradient (0.5 f[i+1] - 0.5 f[i-1]) / spacing with edge
// replication (itk::GradientImageFilter) and its outer product, stored as the
// D(D+1)/2 upper-triangle planes xx, xy, (xz,) yy, (yz, zz) in float
// (AnisotropicTensorUtils.h:63-102).
template <int DIM>
__global__ void __launch_bounds__(kThreads) k_structure(const float *__restrict__ sm, float *__restrict__ comps,
                                                        Dims d, double ihx, double ihy, double ihz)
{
    const double ih[3] = {ihx, ihy, ihz};
    for (long long v = blockIdx.x * (long long)kThreads + threadIdx.x; v < d.n_vox;
         v += (long long)gridDim.x * kThreads) {
        double g[DIM];
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
            const int c = axis_coord(v, d, a);
            const double up = sm[v + (c + 1 < d.n[a] ? d.s[a] : 0)];
            const double dn = sm[v - (c > 0 ? d.s[a] : 0)];
            g[a] = (0.5 * up - 0.5 * dn) * ih[a];
        }
        int k = 0;
#pragma unroll
        for (int i = 0; i < DIM; ++i)
#pragma unroll
            for (int j = i; j < DIM; ++j) comps[(long long)(k++) * d.n_vox + v] = (float)(g[i] * g[j]);
    }
}

// Cyclic Jacobi eigen-decomposition of a symmetric DIMxDIM matrix in double:
// on return the diagonal of a holds the eigenvalues and the columns of v the
// orthonormal eigenvectors.
template <int DIM>
__host__ __device__ inline void jacobi_eigen(double (&a)[DIM][DIM], double (&v)[DIM][DIM])
{
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int j = 0; j < DIM; ++j) v[i][j] = i == j ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 12; ++sweep) {
        double off = 0.0, diag = 0.0;
#pragma unroll
        for (int p = 0; p < DIM; ++p) {
            diag += a[p][p] * a[p][p];
#pragma unroll
            for (int q = p + 1; q < DIM; ++q) off += a[p][q] * a[p][q];
        }
        if (!(off > 1e-40 * diag)) break;
#pragma unroll
        for (int p = 0; p < DIM - 1; ++p)
#pragma unroll
            for (int q = p + 1; q < DIM; ++q) {
                const double apq = a[p][q];
                if (apq == 0.0) continue;
                const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
                double t = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
                if (theta < 0.0) t = -t;
                const double c = 1.0 / sqrt(t * t + 1.0);
                const double s = t * c;
                a[p][p] -= t * apq;
                a[q][q] += t * apq;
                a[p][q] = a[q][p] = 0.0;
#pragma unroll
                for (int r = 0; r < DIM; ++r) {
                    if (r == p || r == q) continue;
                    const double arp = a[r][p], arq = a[r][q];
                    a[r][p] = a[p][r] = c * arp - s * arq;
                    a[r][q] = a[q][r] = s * arp + c * arq;
                }
#pragma unroll
                for (int r = 0; r < DIM; ++r) {
                    const double vrp = v[r][p], vrq = v[r][q];
                    v[r][p] = c * vrp - s * vrq;
                    v[r][q] = s * vrp + c * vrq;
                }
            }
    }
}

// Covariance of one voxel from its smoothed structure tensor (upper triangle
// xx, xy, (xz,) yy, (yz, zz)): the re-weighting of
// AnisotropicTensorUtils.h:177-300 and the covariance of :311-355 in closed
// form.  Returns 1 if the voxel was re-weighted (all |lambda| > 1e-8) and
// then stores its tilt |lambda_max| * factor in *tilt; 0 otherwise (identity).
template <int DIM>
__host__ __device__ inline int tensor_covariance(const double *upper, int alignment, double (&out)[DIM][DIM],
                                                 double *tilt)
{
    double a[DIM][DIM], vec[DIM][DIM];
    int k = 0;
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int j = i; j < DIM; ++j) {
            a[i][j] = a[j][i] = upper[k++];
        }
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int j = 0; j < DIM; ++j) out[i][j] = i == j ? 1.0 : 0.0;
    jacobi_eigen<DIM>(a, vec);
    int top = 0;
    bool all_large = true;
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        const double m = fabs(a[i][i]);
        all_large = all_large && m > 1e-8;
        if (m > fabs(a[top][top])) top = i;
    }
    if (!all_large) return 0;
    double e[DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i) e[i] = vec[i][top];
    double factor = 1.0;
    if (alignment >= 0 && alignment < DIM) {
        // |e x m| for the unit axis m (3D); |e_0 m_1 - e_1 m_0| (2D)
        double cr;
        if (DIM == 3) {
            double c2 = 0.0;
#pragma unroll
            for (int i = 0; i < DIM; ++i)
                if (i != alignment) c2 += e[i] * e[i];
            cr = sqrt(c2);
        } else {
            cr = fabs(alignment == 1 ? e[0] : e[1]);
        }
        factor = (cr + 1e-10) * (cr + 1e-10) / sqrt((double)DIM);
    }
    const double rs = fabs(a[top][top]) * factor;
    const double s = rs * sqrt(rs);
    const double wa = 1.0 / (1.0 + s);
    const double wc = 1.0 + s / (1.0 + s);
    *tilt = rs;
    if (wa >= 1e-8) {
#pragma unroll
        for (int i = 0; i < DIM; ++i)
#pragma unroll
            for (int j = 0; j < DIM; ++j) out[i][j] = (i == j ? wc : 0.0) + (wa - wc) * e[i] * e[j];
    }
    return 1;
}

__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ unsigned wave_add(unsigned v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Per voxel: eigen-decomposition of the smoothed structure tensor, the
// re-weighting of AnisotropicTensorUtils.h:177-300 and the covariance of
// :311-355 in closed form.  Writes DIM*DIM floats per voxel, and per block the
// largest tilt |lambda_max| * factor over re-weighted voxels, their number and
// the number of voxels with a non-finite tensor.
template <int DIM>
__global__ void __launch_bounds__(kThreads) k_tensor_finish(const float *__restrict__ comps, long long n_vox,
                                                            int alignment, float *__restrict__ tensor,
                                                            float *__restrict__ blk_max,
                                                            unsigned *__restrict__ blk_count,
                                                            unsigned *__restrict__ blk_bad)
{
    float tilt = -1.0f;
    unsigned count = 0, bad = 0;
    for (long long v = blockIdx.x * (long long)kThreads + threadIdx.x; v < n_vox;
         v += (long long)gridDim.x * kThreads) {
        double upper[DIM * (DIM + 1) / 2];
        bool finite = true;
#pragma unroll
        for (int k = 0; k < DIM * (DIM + 1) / 2; ++k) {
            upper[k] = comps[(long long)k * n_vox + v];
            finite = finite && isfinite(upper[k]);
        }
        double out[DIM][DIM];
        double rs = 0.0;
        if (!finite) {
            ++bad;
#pragma unroll
            for (int i = 0; i < DIM; ++i)
#pragma unroll
                for (int j = 0; j < DIM; ++j) out[i][j] = i == j ? 1.0 : 0.0;
        } else if (tensor_covariance<DIM>(upper, alignment, out, &rs)) {
            tilt = fmaxf(tilt, (float)rs);
            ++count;
        }
        float *dst = tensor + v * (DIM * DIM);
#pragma unroll
        for (int i = 0; i < DIM; ++i)
#pragma unroll
            for (int j = 0; j < DIM; ++j) dst[i * DIM + j] = (float)out[i][j];
    }
    __shared__ float s_max[kThreads / 64];
    __shared__ unsigned s_cnt[kThreads / 64], s_bad[kThreads / 64];
    tilt = wave_max(tilt);
    count = wave_add(count);
    bad = wave_add(bad);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        s_max[wave] = tilt;
        s_cnt[wave] = count;
        s_bad[wave] = bad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kThreads / 64; ++w) {
            tilt = fmaxf(tilt, s_max[w]);
            count += s_cnt[w];
            bad += s_bad[w];
        }
        blk_max[blockIdx.x] = tilt;
        blk_count[blockIdx.x] = count;
        blk_bad[blockIdx.x] = bad;
    }
}

// ------------------------------------------------------------ host helpers

#define SKM_T_HIP(expr)                                                                                            \
    do {                                                                                                           \
        hipError_t e_ = (expr);                                                                                    \
        if (e_ != hipSuccess) {                                                                                    \
            (void)hipStreamSynchronize(nullptr); /* temporaries are released on return */                          \
            return set_last_error(e_ == hipErrorOutOfMemory ? SKM_ERROR_OUT_OF_MEMORY : SKM_ERROR_HIP,             \
                                  std::string(#expr) + ": " + hipGetErrorString(e_));                              \
        }                                                                                                          \
    } while (0)

// Device buffer released on every return path.
template <typename T>
struct DevBuf {
    T *p = nullptr;
    ~DevBuf() { device_release(p); }
    hipError_t alloc(size_t count) { return device_alloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T)); }
};

int launch_blocks(long long work)
{
    const long long b = (work + kThreads - 1) / kThreads;
    return (int)std::max<long long>(1, std::min<long long>(b, 1LL << 20));
}

// Waits for a stream when it goes out of scope: declared after a function's
// temporaries, it runs before they are released, on every return path.
struct StreamWait {
    hipStream_t s;
    ~StreamWait() { (void)hipStreamSynchronize(s); }
};

// Discrete Gaussian coefficients of every axis (variance / spacing^2) on the device.
struct AxisKernels {
    DevBuf<double> coeffs;
    int offset[3] = {0, 0, 0};
    int radius[3] = {0, 0, 0};
};

skm_status make_axis_kernels(int dim, const skm_grid &g, double variance, AxisKernels &k, hipStream_t stream)
{
    std::vector<double> all;
    for (int a = 0; a < dim; ++a) {
        const double t = variance / (g.spacing[a] * g.spacing[a]);
        std::vector<double> c = gaussian_kernel(t);
        for (double x : c)
            if (!std::isfinite(x))
                return set_last_error(SKM_ERROR_INVALID_ARGUMENT,
                                      "Gaussian variance " + std::to_string(t) +
                                          " pixels^2 is too large for the discrete kernel (exp overflows)");
        k.offset[a] = (int)all.size();
        k.radius[a] = (int)(c.size() - 1) / 2;
        all.insert(all.end(), c.begin(), c.end());
    }
    SKM_T_HIP(k.coeffs.alloc(all.size()));
    SKM_T_HIP(hipMemcpyAsync(k.coeffs.p, all.data(), all.size() * sizeof(double), hipMemcpyHostToDevice, stream));
    SKM_T_HIP(hipStreamSynchronize(stream));
    return SKM_SUCCESS;
}

// Separable passes, last axis first, ping-ponging between a and b (both
// overwritten); *result receives the buffer that holds the smoothed planes.
skm_status gaussian_passes(int dim, const Dims &d, int planes, const AxisKernels &k, float prescale, float *a,
                           float *b, float **result, hipStream_t stream)
{
    float *cur = a, *other = b;
    const int blocks = launch_blocks(d.n_vox * planes);
    for (int axis = dim - 1; axis >= 0; --axis) {
        hipLaunchKernelGGL(k_smooth_axis, dim3(blocks), dim3(kThreads), 0, stream, cur, other,
                           k.coeffs.p + k.offset[axis], k.radius[axis], axis, d, planes,
                           axis == dim - 1 ? prescale : 1.0f);
        SKM_T_HIP(hipGetLastError());
        std::swap(cur, other);
    }
    *result = cur;
    return SKM_SUCCESS;
}

skm_status validate_image(int32_t dim, const skm_grid *grid, const void *in, const void *out)
{
    if (dim != 2 && dim != 3) return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "dimension must be 2 or 3");
    if (!grid || !in || !out) return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "null argument");
    for (int a = 0; a < dim; ++a) {
        if (grid->size[a] < 1) return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "image size must be positive");
        if (!(grid->spacing[a] > 0.0))
            return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "Pixel spacing cannot be zero or negative");
    }
    return SKM_SUCCESS;
}

}  // namespace
}  // namespace skm

using namespace skm;

extern "C" {

SKM_API skm_status skm_gaussian_smooth(int32_t dim, const skm_grid *grid, const float *in, double variance,
                                       float *out)
{
    skm_status st = validate_image(dim, grid, in, out);
    if (st != SKM_SUCCESS) return st;
    const Dims d = make_dims(dim, *grid);
    DevBuf<float> a;
    SKM_T_HIP(a.alloc(d.n_vox));
    SKM_T_HIP(hipMemcpy(a.p, in, d.n_vox * sizeof(float), hipMemcpyHostToDevice));
    if ((st = gaussian_smooth_device(dim, *grid, a.p, variance, a.p)) != SKM_SUCCESS) return st;
    SKM_T_HIP(hipMemcpy(out, a.p, d.n_vox * sizeof(float), hipMemcpyDeviceToHost));
    return SKM_SUCCESS;
}

SKM_API skm_status skm_anisotropic_tensor(int32_t dim, const skm_grid *grid, const float *image,
                                          const skm_tensor_options *opt, float *tensor_out, double *maximum_tilt,
                                          double *maximum_weight, double *maximum_eigenvalue_covariance)
{
    skm_status st = validate_image(dim, grid, image, tensor_out);
    if (st != SKM_SUCCESS) return st;
    if (!opt) return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "null options");
    const Dims d = make_dims(dim, *grid);
    DevBuf<float> img, ten;
    SKM_T_HIP(img.alloc(d.n_vox));
    SKM_T_HIP(ten.alloc(d.n_vox * dim * dim));
    SKM_T_HIP(hipMemcpy(img.p, image, d.n_vox * sizeof(float), hipMemcpyHostToDevice));
    if ((st = anisotropic_tensor_device(dim, *grid, img.p, *opt, ten.p, maximum_tilt, maximum_weight,
                                        maximum_eigenvalue_covariance)) != SKM_SUCCESS)
        return st;
    SKM_T_HIP(hipMemcpy(tensor_out, ten.p, d.n_vox * dim * dim * sizeof(float), hipMemcpyDeviceToHost));
    return SKM_SUCCESS;
}

}  // extern "C"

// GaussianSmoothing on device arrays; d_in and d_out may be the same array.
skm_status skm::gaussian_smooth_device(int dim, const skm_grid &grid, const float *d_in, double variance,
                                       float *d_out, hipStream_t stream)
{
    skm_status st = validate_image(dim, &grid, d_in, d_out);
    if (st != SKM_SUCCESS) return st;
    if (!(variance >= 0.0)) return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "variance must not be negative");
    const Dims d = make_dims(dim, grid);
    AxisKernels k;
    DevBuf<float> a, b;
    StreamWait wait{stream};
    if ((st = make_axis_kernels(dim, grid, variance, k, stream)) != SKM_SUCCESS) return st;
    SKM_T_HIP(a.alloc(d.n_vox));
    SKM_T_HIP(b.alloc(d.n_vox));
    SKM_T_HIP(hipMemcpyAsync(a.p, d_in, d.n_vox * sizeof(float), hipMemcpyDeviceToDevice, stream));
    float *res = nullptr;
    if ((st = gaussian_passes(dim, d, 1, k, 1.0f, a.p, b.p, &res, stream)) != SKM_SUCCESS) return st;
    SKM_T_HIP(hipMemcpyAsync(d_out, res, d.n_vox * sizeof(float), hipMemcpyDeviceToDevice, stream));
    SKM_T_HIP(hipStreamSynchronize(stream));  // complete for callers on other streams
    return SKM_SUCCESS;
}

skm_status skm::anisotropic_tensor_device(int dim, const skm_grid &grid_ref, const float *d_image,
                                          const skm_tensor_options &opt_ref, float *d_tensor_out,
                                          double *maximum_tilt, double *maximum_weight,
                                          double *maximum_eigenvalue_covariance, hipStream_t stream)
{
    const skm_grid *grid = &grid_ref;
    const skm_tensor_options *opt = &opt_ref;
    skm_status st = validate_image(dim, grid, d_image, d_tensor_out);
    if (st != SKM_SUCCESS) return st;
    if (opt->anisotropic_filtering)
        return set_last_error(SKM_ERROR_NOT_SUPPORTED,
                              "anisotropic pre-filtering (curvature diffusion) is not supported; SKMReg passes false");
    if (!(opt->sigma > 0.0) || !std::isfinite(opt->weight))
        return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "sigma must be positive and the weight finite");
    const Dims d = make_dims(dim, *grid);
    const int nc = dim * (dim + 1) / 2;
    AxisKernels k_img, k_ten;
    DevBuf<float> a, b, ca, cb;
    DevBuf<float> bmax;
    DevBuf<unsigned> bcnt, bbad;
    StreamWait wait{stream};
    if ((st = make_axis_kernels(dim, *grid, opt->sigma, k_img, stream)) != SKM_SUCCESS) return st;
    if ((st = make_axis_kernels(dim, *grid, opt->sigma / 4.0, k_ten, stream)) != SKM_SUCCESS) return st;

    SKM_T_HIP(a.alloc(d.n_vox));
    SKM_T_HIP(b.alloc(d.n_vox));
    SKM_T_HIP(ca.alloc(d.n_vox * nc));
    SKM_T_HIP(cb.alloc(d.n_vox * nc));
    const int fin_blocks = std::min(launch_blocks(d.n_vox), 4096);
    SKM_T_HIP(bmax.alloc(fin_blocks));
    SKM_T_HIP(bcnt.alloc(fin_blocks));
    SKM_T_HIP(bbad.alloc(fin_blocks));
    SKM_T_HIP(hipMemcpyAsync(a.p, d_image, d.n_vox * sizeof(float), hipMemcpyDeviceToDevice, stream));

    float *smoothed = nullptr, *comps = nullptr;
    if ((st = gaussian_passes(dim, d, 1, k_img, (float)opt->weight, a.p, b.p, &smoothed, stream)) != SKM_SUCCESS)
        return st;
    const double ih[3] = {1.0 / grid->spacing[0], 1.0 / grid->spacing[1], dim == 3 ? 1.0 / grid->spacing[2] : 1.0};
    if (dim == 3)
        hipLaunchKernelGGL(k_structure<3>, dim3(launch_blocks(d.n_vox)), dim3(kThreads), 0, stream, smoothed, ca.p,
                           d, ih[0], ih[1], ih[2]);
    else
        hipLaunchKernelGGL(k_structure<2>, dim3(launch_blocks(d.n_vox)), dim3(kThreads), 0, stream, smoothed, ca.p,
                           d, ih[0], ih[1], ih[2]);
    SKM_T_HIP(hipGetLastError());
    if ((st = gaussian_passes(dim, d, nc, k_ten, 1.0f, ca.p, cb.p, &comps, stream)) != SKM_SUCCESS) return st;
    const int align = opt->alignment_direction;
    if (dim == 3)
        hipLaunchKernelGGL(k_tensor_finish<3>, dim3(fin_blocks), dim3(kThreads), 0, stream, comps, d.n_vox, align,
                           d_tensor_out, bmax.p, bcnt.p, bbad.p);
    else
        hipLaunchKernelGGL(k_tensor_finish<2>, dim3(fin_blocks), dim3(kThreads), 0, stream, comps, d.n_vox, align,
                           d_tensor_out, bmax.p, bcnt.p, bbad.p);
    SKM_T_HIP(hipGetLastError());

    std::vector<float> hmax(fin_blocks);
    std::vector<unsigned> hcnt(fin_blocks), hbad(fin_blocks);
    SKM_T_HIP(hipMemcpyAsync(hmax.data(), bmax.p, fin_blocks * sizeof(float), hipMemcpyDeviceToHost, stream));
    SKM_T_HIP(hipMemcpyAsync(hcnt.data(), bcnt.p, fin_blocks * sizeof(unsigned), hipMemcpyDeviceToHost, stream));
    SKM_T_HIP(hipMemcpyAsync(hbad.data(), bbad.p, fin_blocks * sizeof(unsigned), hipMemcpyDeviceToHost, stream));
    SKM_T_HIP(hipStreamSynchronize(stream));
    unsigned long long count = 0, bad = 0;
    float tilt = -1.0f;
    for (int i = 0; i < fin_blocks; ++i) {
        count += hcnt[i];
        bad += hbad[i];
        tilt = std::max(tilt, hmax[i]);
    }
    if (bad)
        return set_last_error(SKM_ERROR_INVALID_ARGUMENT,
                              std::to_string(bad) + " voxels have a non-finite structure tensor (non-finite image?)");

    // ComputeAnisotropicTensor returns the largest tilt, or the lowest double
    // when no voxel was re-weighted; SetWeightingTensor then takes the tilt as
    // the maximum weight, or (all tensors are the identity) the largest
    // eigenvalue 1 with the identity as maximum-eigenvalue covariance.
    const double t = count ? (double)tilt : std::numeric_limits<double>::lowest();
    if (maximum_tilt) *maximum_tilt = t;
    if (maximum_weight) *maximum_weight = count ? t : 1.0;
    if (maximum_eigenvalue_covariance)
        for (int i = 0; i < 9; ++i) maximum_eigenvalue_covariance[i] = (i % 4 == 0 && i / 3 < dim) ? 1.0 : 0.0;
    return SKM_SUCCESS;
}


// ------------------------------------------------------------ allocation-free smoothing
//
// skm_gaussian_smooth_device_noalloc: the smoothing of skm_gaussian_smooth on
// device arrays with every temporary in `scratch`, a file-scope device array
// of fixed size, so that the call allocates nothing (graph capture, serving
// loops).  Each GPU holds its own copy of the array at its own address,
// looked up once per device with hipGetSymbolAddress; host code uses that
// address, never the array's name.  All calls on one GPU share the array:
// calls that overlap on a GPU race.  Nothing in it needs to start at zero
// (every element is written before it is read).

namespace skm {
namespace {

constexpr long long kNoallocMaxVoxels = 8388608;     // 256 x 256 x 128 floats per plane (32 MiB)
constexpr int kNoallocTapsPerAxis = 2 * (int)kMaximumWidth + 1;
constexpr size_t kNoallocTapBytes = 4096 * sizeof(double);
constexpr size_t kNoallocScratchBytes = kNoallocTapBytes + 2 * (size_t)kNoallocMaxVoxels * sizeof(float);
constexpr int kNoallocMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

std::mutex g_scratch_mutex;
void *g_scratch_ptr[kNoallocMaxDevices] = {};

// device address of `scratch` on the current device, looked up once per device
skm_status scratch_address(void **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_last_error(SKM_ERROR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kNoallocMaxDevices) return set_last_error(SKM_ERROR_NOT_SUPPORTED, "device index too large");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_last_error(SKM_ERROR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return SKM_SUCCESS;
}

// one axis's taps, passed by value so that no host-to-device copy is needed
struct TapArg {
    int n;
    double c[kNoallocTapsPerAxis];
};

__global__ void __launch_bounds__(kThreads) k_write_taps(TapArg taps, double *__restrict__ dst)
{
    if ((int)threadIdx.x < taps.n) dst[threadIdx.x] = taps.c[threadIdx.x];
}

}  // namespace
}  // namespace skm

extern "C" {

SKM_API skm_status skm_gaussian_smooth_device_noalloc(int32_t dim, const skm_grid *grid, const float *d_in,
                                                      double variance, float *d_out, void *hip_stream)
{
    skm_status st = validate_image(dim, grid, d_in, d_out);
    if (st != SKM_SUCCESS) return st;
    if (!(variance >= 0.0)) return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "variance must not be negative");
    const Dims d = make_dims(dim, *grid);
    if (d.n_vox > kNoallocMaxVoxels)
        return set_last_error(SKM_ERROR_INVALID_ARGUMENT,
                              "image of " + std::to_string(d.n_vox) + " voxels exceeds the allocation-free limit of " +
                                  std::to_string(kNoallocMaxVoxels) + " voxels");
    void *base = nullptr;
    if ((st = scratch_address(&base)) != SKM_SUCCESS) return st;
    double *taps = static_cast<double *>(base);
    float *plane_a = reinterpret_cast<float *>(static_cast<unsigned char *>(base) + kNoallocTapBytes);
    float *plane_b = plane_a + kNoallocMaxVoxels;
    hipStream_t stream = static_cast<hipStream_t>(hip_stream);
    // taps of every axis (the coefficients of make_axis_kernels)
    int offset[3] = {0, 0, 0}, radius[3] = {0, 0, 0}, used = 0;
    for (int a = 0; a < dim; ++a) {
        const double t = variance / (grid->spacing[a] * grid->spacing[a]);
        const std::vector<double> c = gaussian_kernel(t);
        TapArg arg{};
        arg.n = (int)c.size();
        if (arg.n > kNoallocTapsPerAxis) return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "Gaussian kernel too wide");
        for (int i = 0; i < arg.n; ++i) {
            if (!std::isfinite(c[i]))
                return set_last_error(SKM_ERROR_INVALID_ARGUMENT,
                                      "Gaussian variance " + std::to_string(t) +
                                          " pixels^2 is too large for the discrete kernel (exp overflows)");
            arg.c[i] = c[i];
        }
        offset[a] = used;
        radius[a] = (arg.n - 1) / 2;
        hipLaunchKernelGGL(k_write_taps, dim3(1), dim3(kThreads), 0, stream, arg, taps + used);
        used += arg.n;
    }
    // separable passes, last axis first: in -> a (-> b) -> out; out may be in
    const int blocks = launch_blocks(d.n_vox);
    const float *src = d_in;
    for (int axis = dim - 1; axis >= 0; --axis) {
        float *dst = axis == 0 ? d_out : ((dim - 1 - axis) % 2 == 0 ? plane_a : plane_b);
        hipLaunchKernelGGL(k_smooth_axis, dim3(blocks), dim3(kThreads), 0, stream, src, dst, taps + offset[axis],
                           radius[axis], axis, d, 1, 1.0f);
        src = dst;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_last_error(SKM_ERROR_HIP, std::string("smoothing launch: ") + hipGetErrorString(e));
    return SKM_SUCCESS;
}

}  // extern "C"
