// skm_register.hip -- the scale-level loop of SKMReg (src/SKMReg.cxx:84-525)
// as a library call over this library's own C ABI: preprocessing, per-level
// smoothing and weight tensor, a new evaluator per level whose previous
// displacement is the sum of the coarser levels, the device-resident
// adaptive-step optimizer, and the dense displacement of all levels.
//
// Images, tensors and displacement fields stay in device memory from level to
// level: the preprocessed images (and a given weight image, weight tensor or
// guidance image) cross PCIe once, and only result files, landmark errors and
// the final field are copied back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <condition_variable>
#include <mutex>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <new>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "skm.h"
#include "skm_internal.hpp"

using skm::set_last_error;

namespace {

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

long long voxels(const skm_grid &g, int dim)
{
    long long n = 1;
    for (int d = 0; d < dim; ++d) n *= g.size[d];
    return n;
}

// ---- preprocessing (SKMReg.cxx:84-91) on the device

constexpr int kPreThreads = 256;
constexpr int kPreBlocks = 512;

// ThresholdImage's four bounds, then RescaleImage's map.
struct PreStats {
    double below_max, above_min, lo, hi;
    double scale, shift;
};

// std::min / std::max of the host restatement (NaN compares false either way).
__device__ __forceinline__ double min_as_std(double a, double b) { return b < a ? b : a; }
__device__ __forceinline__ double max_as_std(double a, double b) { return a < b ? b : a; }

// Block partials of the image's minimum and maximum and, with thresholds, of
// the largest value below thresh_max and the smallest above thresh_min.
__global__ __launch_bounds__(kPreThreads) void k_pre_partials(const float *__restrict__ v, long long n,
                                                              double thresh_max, double thresh_min,
                                                              PreStats *__restrict__ part)
{
    __shared__ double s[4][kPreThreads];
    double bm = -DBL_MAX, am = DBL_MAX, lo = DBL_MAX, hi = -DBL_MAX;
    for (long long i = blockIdx.x * (long long)kPreThreads + threadIdx.x; i < n;
         i += (long long)gridDim.x * kPreThreads) {
        const double x = v[i];
        if (bm < x && x < thresh_max) bm = x;
        if (am > x && x > thresh_min) am = x;
        lo = min_as_std(lo, x);
        hi = max_as_std(hi, x);
    }
    s[0][threadIdx.x] = bm;
    s[1][threadIdx.x] = am;
    s[2][threadIdx.x] = lo;
    s[3][threadIdx.x] = hi;
    __syncthreads();
    for (int w = kPreThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s[0][threadIdx.x] = max_as_std(s[0][threadIdx.x], s[0][threadIdx.x + w]);
            s[1][threadIdx.x] = min_as_std(s[1][threadIdx.x], s[1][threadIdx.x + w]);
            s[2][threadIdx.x] = min_as_std(s[2][threadIdx.x], s[2][threadIdx.x + w]);
            s[3][threadIdx.x] = max_as_std(s[3][threadIdx.x], s[3][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = PreStats{s[0][0], s[1][0], s[2][0], s[3][0], 0.0, 0.0};
}

// One thread: the bounds of the whole image into part[blocks], and the
// rescale map [lo, hi] -> [0, 1] (itkUtils.h:103-112), the product and the
// sum rounded one at a time as on the host.
__global__ void k_pre_final(PreStats *part, int blocks)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    PreStats r{-DBL_MAX, DBL_MAX, DBL_MAX, -DBL_MAX, 0.0, 0.0};
    for (int b = 0; b < blocks; ++b) {
        r.below_max = max_as_std(r.below_max, part[b].below_max);
        r.above_min = min_as_std(r.above_min, part[b].above_min);
        r.lo = min_as_std(r.lo, part[b].lo);
        r.hi = max_as_std(r.hi, part[b].hi);
    }
    if (r.hi != r.lo) r.scale = 1.0 / (r.hi - r.lo);
    else if (r.hi != 0.0) r.scale = 1.0 / r.hi;
    {
#pragma clang fp contract(off)
        r.shift = 0.0 - r.lo * r.scale;
    }
    part[blocks] = r;
}

// ThresholdImage's two ThresholdOutside passes (include/itkUtils.h:45-100):
// values outside [min, largest below thresh_max] become that largest value,
// then values outside [smallest above thresh_min, max] that smallest value.
__global__ __launch_bounds__(kPreThreads) void k_pre_threshold(float *__restrict__ v, long long n,
                                                               const PreStats *__restrict__ st)
{
    const PreStats r = *st;
    for (long long i = blockIdx.x * (long long)kPreThreads + threadIdx.x; i < n;
         i += (long long)gridDim.x * kPreThreads) {
        float x = v[i];
        if (x < r.lo || x > r.below_max) x = (float)r.below_max;
        if (x < r.above_min || x > r.hi) x = (float)r.above_min;
        v[i] = x;
    }
}

// RescaleImage(image, 0, 1): out = clamp(in * scale + shift, 0, 1).
__global__ __launch_bounds__(kPreThreads) void k_pre_rescale(float *__restrict__ v, long long n,
                                                             const PreStats *__restrict__ st)
{
#pragma clang fp contract(off)  // a product and a sum, each rounded, as on the host
    const double scale = st->scale, shift = st->shift;
    for (long long i = blockIdx.x * (long long)kPreThreads + threadIdx.x; i < n;
         i += (long long)gridDim.x * kPreThreads)
        v[i] = (float)min_as_std(1.0, max_as_std(0.0, (double)v[i] * scale + shift));
}

// ThresholdImage (with thresholds) then RescaleImage(0, 1) of a device image;
// stats holds kPreBlocks + 1 entries.
skm_status preprocess_device(float *d_img, long long n, bool thresholds, double thresh_max, double thresh_min,
                             PreStats *stats, hipStream_t s)
{
    const int blocks = (int)std::min<long long>(kPreBlocks, (n + kPreThreads - 1) / kPreThreads);
    if (thresholds) {
        k_pre_partials<<<blocks, kPreThreads, 0, s>>>(d_img, n, thresh_max, thresh_min, stats);
        k_pre_final<<<1, 64, 0, s>>>(stats, blocks);
        k_pre_threshold<<<blocks, kPreThreads, 0, s>>>(d_img, n, stats + blocks);
    }
    k_pre_partials<<<blocks, kPreThreads, 0, s>>>(d_img, n, 0.0, 0.0, stats);
    k_pre_final<<<1, 64, 0, s>>>(stats, blocks);
    k_pre_rescale<<<blocks, kPreThreads, 0, s>>>(d_img, n, stats + blocks);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return set_last_error(SKM_ERROR_HIP, std::string("preprocessing: ") + hipGetErrorString(e));
    return SKM_SUCCESS;
}

// Largest eigenvalue of the symmetric part of a dim x dim matrix (FP64).
double max_eig_sym(int dim, const float *m)
{
    double a[9];
    for (int i = 0; i < dim; ++i)
        for (int j = 0; j < dim; ++j) a[i * dim + j] = 0.5 * ((double)m[i * dim + j] + (double)m[j * dim + i]);
    if (dim == 2) {
        const double tr = a[0] + a[3], det = a[0] * a[3] - a[1] * a[2];
        return 0.5 * (tr + std::sqrt(std::max(tr * tr - 4.0 * det, 0.0)));
    }
    const double p1 = a[1] * a[1] + a[2] * a[2] + a[5] * a[5];
    if (p1 == 0.0) return std::max(a[0], std::max(a[4], a[8]));
    const double q = (a[0] + a[4] + a[8]) / 3.0;
    const double p2 = (a[0] - q) * (a[0] - q) + (a[4] - q) * (a[4] - q) + (a[8] - q) * (a[8] - q) + 2.0 * p1;
    const double p = std::sqrt(p2 / 6.0);
    double b[9];
    for (int i = 0; i < 9; ++i) b[i] = (a[i] - (i % 4 == 0 ? q : 0.0)) / p;
    const double r = 0.5 * (b[0] * (b[4] * b[8] - b[5] * b[7]) - b[1] * (b[3] * b[8] - b[5] * b[6]) +
                            b[2] * (b[3] * b[7] - b[4] * b[6]));
    const double phi = std::acos(std::min(1.0, std::max(-1.0, r))) / 3.0;
    return q + 2.0 * p * std::cos(phi);
}

// DisplacementFieldJacobianDeterminantFilter (include/itkUtils.h:153-161) on
// a D-vector image: det(I + J) with J_ij = (v_j[x + e_i] - v_j[x - e_i]) *
// 0.5 / spacing_i, neighbours clamped to the image (zero-flux Neumann).
std::vector<float> jacobian_determinant(int dim, const skm_grid &g, const float *v)
{
    const int n[3] = {g.size[0], g.size[1], dim == 3 ? g.size[2] : 1};
    const long long nv = (long long)n[0] * n[1] * n[2];
    std::vector<float> out((size_t)nv);
    for (long long i = 0; i < nv; ++i) {
        const int idx[3] = {(int)(i % n[0]), (int)((i / n[0]) % n[1]), (int)(i / ((long long)n[0] * n[1]))};
        double J[3][3] = {{0}};
        for (int a = 0; a < dim; ++a) {
            int lo[3] = {idx[0], idx[1], idx[2]}, hi[3] = {idx[0], idx[1], idx[2]};
            lo[a] = std::max(idx[a] - 1, 0);
            hi[a] = std::min(idx[a] + 1, n[a] - 1);
            const long long il = lo[0] + (long long)n[0] * (lo[1] + (long long)n[1] * lo[2]);
            const long long ih = hi[0] + (long long)n[0] * (hi[1] + (long long)n[1] * hi[2]);
            for (int c = 0; c < dim; ++c)
                J[a][c] = 0.5 / g.spacing[a] * ((double)v[ih * dim + c] - (double)v[il * dim + c]);
            J[a][a] += 1.0;
        }
        out[i] = (float)(dim == 2 ? J[0][0] * J[1][1] - J[0][1] * J[1][0]
                                  : J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                                        J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                                        J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]));
    }
    return out;
}

// Result files of SKMReg under temp_directory.
struct Output {
    std::string dir;
    int dim;
    bool on() const { return !dir.empty(); }
    skm_status write(const std::string &name, const skm_grid &g, int components, const float *data) const
    {
        if (!on()) return SKM_SUCCESS;
        skm_image_header h;
        std::memset(&h, 0, sizeof(h));
        h.dim = dim;
        h.grid = g;
        h.components = components;
        return skm_write_image((dir + "/" + name).c_str(), &h, data);
    }
};

// A device array owned by the driver.
template <typename T>
struct DeviceArray {
    T *p = nullptr;
    size_t n = 0;
    DeviceArray() = default;
    DeviceArray(const DeviceArray &) = delete;
    DeviceArray &operator=(const DeviceArray &) = delete;
    ~DeviceArray() { skm::device_release(p); }
    skm_status alloc(size_t count)
    {
        if (p && n == count) return SKM_SUCCESS;
        skm::device_release(p);
        p = nullptr;
        n = 0;
        hipError_t e = skm::device_alloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            return set_last_error(SKM_ERROR_HIP, std::string("registration buffer: ") + hipGetErrorString(e));
        }
        n = count;
        return SKM_SUCCESS;
    }
    // copies on stream s, complete on return (the evaluators' streams read
    // these buffers without waiting for s)
    skm_status upload(const T *src, size_t count, hipStream_t s)
    {
        skm_status st = alloc(count);
        if (st != SKM_SUCCESS) return st;
        hipError_t e = hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return set_last_error(SKM_ERROR_HIP, std::string("upload: ") + hipGetErrorString(e));
        return SKM_SUCCESS;
    }
    skm_status download(std::vector<T> &dst, hipStream_t s) const
    {
        dst.resize(n);
        hipError_t e = hipMemcpyAsync(dst.data(), p, n * sizeof(T), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return set_last_error(SKM_ERROR_HIP, std::string("download: ") + hipGetErrorString(e));
        return SKM_SUCCESS;
    }
};

// The registration's own stream for its set-up work and copies, so that
// registrations in flight together do not wait for each other's set-up on
// the null stream; returned (synchronised) at the end.
struct PoolStream {
    hipStream_t s = nullptr;
    int dev = 0;
    PoolStream() = default;
    PoolStream(const PoolStream &) = delete;
    PoolStream &operator=(const PoolStream &) = delete;
    ~PoolStream()
    {
        if (s) skm::return_pool_stream(dev, s);
    }
};

// Waits for a stream when it goes out of scope (before the arrays declared
// ahead of it are released).
struct StreamWait {
    hipStream_t s;
    ~StreamWait() { (void)hipStreamSynchronize(s); }
};

struct EvaluatorGuard {
    skm_evaluator *ev = nullptr;
    ~EvaluatorGuard()
    {
        if (ev) skm_destroy(ev);
    }
};

#define SKM_R_TRY(expr)                                                                                            \
    do {                                                                                                           \
        skm_status s_ = (expr);                                                                                    \
        if (s_ != SKM_SUCCESS) return s_;                                                                          \
    } while (0)

// The metric value a level's optimizer reports at SoA parameters `params`
// (RegularizedImageToImageMetricv4::GetValue, .hxx:160-268): the measure plus
// the weighted device and host regularizers.  `resample` draws new jitter
// offsets first.
skm_status level_value(skm_evaluator *ev, int metric, const skm_scale_params &P, const std::vector<float> &params,
                       const std::vector<float> &cpw, int resample, double *value)
{
    const size_t n_cp = cpw.size(), dim = params.size() / n_cp;
    std::vector<float> aos(params.size());
    for (size_t j = 0; j < n_cp; ++j)
        for (size_t d = 0; d < dim; ++d) aos[j * dim + d] = params[d * n_cp + j];
    skm_result r;
    SKM_R_TRY(skm_get_value(ev, metric, aos.data(), cpw.data(), resample, &r));
    double v = r.measure;
    if (P.reg_rkhs > 0) v += P.reg_rkhs * r.reg_value_rkhs;
    if (P.reg_rd > 0) v += P.reg_rd * r.reg_value_rd;
    if (P.reg_pg > 0) v += P.reg_pg * r.reg_value_pg;
    if (P.reg_bilateral > 0) v += P.reg_bilateral * r.reg_value_bilateral;
    if (P.reg_l1 > 0 || P.reg_l21 > 0 || P.reg_l2 > 0) {
        std::vector<double> pd(params.begin(), params.end());
        double reg = 0.0, l1 = 0.0, l21 = 0.0, l2 = 0.0;
        SKM_R_TRY(skm_host_regularizers((int32_t)dim, (int64_t)n_cp, pd.data(), P.reg_l1, P.reg_l21, P.reg_l2, 0,
                                        nullptr, &reg, &l1, &l21, &l2));
        v += reg;
    }
    *value = v;
    return SKM_SUCCESS;
}

// Registrations in lockstep (skm_register_batch_lockstep): each one, when it
// reaches a level's optimisation, waits until every registration still
// running has reached its own; the last to arrive runs all their optimisers
// at once (skm_optimize_batch: one captured graph launch per step for all),
// grouped by identical options.  A registration that ends, or fails, leaves
// the group, so that the others never wait for it.
struct Lockstep {
    struct Job {
        skm_evaluator *ev;
        skm_metric_options mo;
        skm_optimizer_options oo;
        const float *cpw;
        float *params;
        skm_iteration_record *rec;
        skm_status st = SKM_SUCCESS;
        std::string msg;
    };
    std::mutex m;
    std::condition_variable cv;
    int active = 0;
    long long generation = 0;
    std::vector<Job *> waiting;

    static bool same(const Job &a, const Job &b)
    {
        const skm_metric_options &x = a.mo, &y = b.mo;
        const skm_optimizer_options &p = a.oo, &q = b.oo;
        return x.measure == y.measure && x.reg_l1 == y.reg_l1 && x.reg_l21 == y.reg_l21 && x.reg_l2 == y.reg_l2 &&
               x.use_negative_gradient == y.use_negative_gradient && x.do_resampling == y.do_resampling &&
               p.kind == q.kind && p.a == q.a && p.A == q.A && p.alpha == q.alpha && p.fmax == q.fmax &&
               p.fmin == q.fmin && p.omega == q.omega && p.orthant_projection == q.orthant_projection &&
               p.iterations == q.iterations && p.return_best == q.return_best && p.graph == q.graph;
    }

    // runs every waiting job; called with the lock held, releases it meanwhile
    void run(std::unique_lock<std::mutex> &lk)
    {
        std::vector<Job *> jobs;
        jobs.swap(waiting);
        lk.unlock();
        std::vector<bool> done(jobs.size(), false);
        for (size_t i = 0; i < jobs.size(); ++i) {
            if (done[i]) continue;
            std::vector<Job *> group;
            for (size_t k = i; k < jobs.size(); ++k)
                if (!done[k] && same(*jobs[i], *jobs[k])) {
                    group.push_back(jobs[k]);
                    done[k] = true;
                }
            std::vector<skm_evaluator *> evs;
            std::vector<const float *> cpw;
            std::vector<float *> params;
            std::vector<skm_iteration_record *> rec;
            for (Job *j : group) {
                evs.push_back(j->ev);
                cpw.push_back(j->cpw);
                params.push_back(j->params);
                rec.push_back(j->rec);
            }
            const skm_status st = skm_optimize_batch(evs.data(), (int32_t)evs.size(), &group[0]->mo, &group[0]->oo,
                                                     cpw.data(), params.data(), rec.data());
            const std::string msg = st != SKM_SUCCESS ? std::string(skm_last_error()) : std::string();
            for (Job *j : group) {
                j->st = st;
                j->msg = msg;
            }
        }
        lk.lock();
        ++generation;
        cv.notify_all();
    }

    skm_status optimize(Job &job)
    {
        std::unique_lock<std::mutex> lk(m);
        waiting.push_back(&job);
        const long long gen = generation;
        if ((int)waiting.size() == active) run(lk);
        else cv.wait(lk, [&] { return generation != gen; });
        lk.unlock();
        if (job.st != SKM_SUCCESS) return set_last_error(job.st, job.msg);
        return SKM_SUCCESS;
    }

    void leave()
    {
        std::unique_lock<std::mutex> lk(m);
        --active;
        if (!waiting.empty() && (int)waiting.size() == active) run(lk);
    }
};

struct LockstepMember {
    Lockstep *g;
    ~LockstepMember()
    {
        if (g) g->leave();
    }
};

skm_status register_impl(const skm_registration_desc *desc, float *displacement_out, skm_scale_report *reports,
                         Lockstep *group);

}  // namespace

extern "C" {

SKM_API void skm_registration_desc_init(skm_registration_desc *d)
{
    if (!d) return;
    std::memset(d, 0, sizeof(*d));
    d->dim = 3;
    for (int k = 0; k < 3; ++k) {
        d->reference.size[k] = d->target.size[k] = d->weight.size[k] = 1;
        d->reference.spacing[k] = d->target.spacing[k] = d->weight.spacing[k] = 1.0;
    }
    // ConfigUtils.h:29-62; per-scale lists repeat their last entry (CURRENT)
    const double sc4[3] = {64, 32, 16}, sc0[3] = {128, 64, 32}, tw[3] = {400, 200, 100};
    const int sr[3] = {5, 4, 3}, ns[3] = {15, 12, 9};
    const double sv[3] = {2, 1, 0};
    d->num_scales = 3;
    for (int s = 0; s < SKM_MAX_SCALES; ++s) {
        const int k = std::min(s, 2);
        skm_scale_params &p = d->scale[s];
        p.sigma_c4 = sc4[k];
        p.sigma_c0 = sc0[k];
        p.tensor_weight = tw[k];
        p.reg_rd_scaling = p.reg_pg_scaling = 1.0;
        p.reg_bilateral = 0.0;
        p.reg_bilateral_spatial_sigma = 12.0;
        p.reg_bilateral_intensity_sigma = 0.1;
        p.iterations = 500;
        p.initial_step_size = 0.1;
        p.sampling_rate = sr[k];
        p.neighborhood_sampling = ns[k];
        p.smoothing_variance = sv[k];
    }
    d->cp_spacing_scaler = 4.0;
    d->metric = SKM_MEASURE_MSE;
    d->use_intensity_thresholds = 1;
    d->threshold_min = -1500.0;
    d->threshold_max = 3500.0;
    d->restart_at = 0;
}

SKM_API skm_status skm_register(const skm_registration_desc *desc, float *displacement_out,
                                skm_scale_report *reports)
{
    return register_impl(desc, displacement_out, reports, nullptr);
}

}  // extern "C"

namespace {

skm_status register_impl(const skm_registration_desc *desc, float *displacement_out, skm_scale_report *reports,
                         Lockstep *group)
{
    LockstepMember member{group};
    if (!desc || !displacement_out) return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "null argument");
    const int dim = desc->dim;
    if (dim != 2 && dim != 3) return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "dim must be 2 or 3");
    if (!desc->reference_data || !desc->target_data)
        return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "reference and target images are required");
    if (desc->num_scales < 1 || desc->num_scales > SKM_MAX_SCALES)
        return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "num_scales must be 1.." + std::to_string(SKM_MAX_SCALES));
    const long long n_ref = voxels(desc->reference, dim), n_tar = voxels(desc->target, dim);
    if (n_ref < 1 || n_tar < 1) return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "empty image");

    // preprocessing (SKMReg.cxx:84-91) on the device copies of the images; the
    // host copies (result files and warped outputs only) are read back when
    // files are written
    const Output out{desc->temp_directory ? std::string(desc->temp_directory) : std::string(), dim};
    PoolStream setup;
    if (hipGetDevice(&setup.dev) != hipSuccess || skm::take_pool_stream(&setup.s) != hipSuccess)
        return set_last_error(SKM_ERROR_HIP, "registration stream");
    const hipStream_t ss = setup.s;
    DeviceArray<float> d_ref, d_tar, d_ref_s, d_tar_s, d_weight, d_guide, d_tensor, d_prev, d_cp, d_cpw;
    DeviceArray<PreStats> d_pre;
    StreamWait setup_wait{ss};
    SKM_R_TRY(d_ref.upload(desc->reference_data, (size_t)n_ref, ss));
    SKM_R_TRY(d_tar.upload(desc->target_data, (size_t)n_tar, ss));
    SKM_R_TRY(d_pre.alloc(kPreBlocks + 1));
    const bool thresholds = desc->use_intensity_thresholds != 0;
    SKM_R_TRY(preprocess_device(d_ref.p, n_ref, thresholds, desc->threshold_max, desc->threshold_min, d_pre.p, ss));
    SKM_R_TRY(preprocess_device(d_tar.p, n_tar, thresholds, desc->threshold_max, desc->threshold_min, d_pre.p, ss));
    std::vector<float> ref, tar;
    if (out.on()) {
        SKM_R_TRY(d_ref.download(ref, ss));
        SKM_R_TRY(d_tar.download(tar, ss));
    }
    SKM_R_TRY(out.write("reference-preprocessed.vtk", desc->reference, 1, ref.data()));
    SKM_R_TRY(out.write("target-preprocessed.vtk", desc->target, 1, tar.data()));
    const bool landmarks = desc->num_landmarks > 0 && desc->reference_landmarks && desc->target_landmarks;

    // previous displacement (restart, SKMReg.cxx:96-113): given, or read from
    // df-full-<restart_at-2>.vtk; a file that cannot be read (or does not fit
    // the target grid) starts the registration at the first level
    int start = 0;
    std::vector<float> prev;
    if (desc->restart_at > 1 && desc->restart_at <= desc->num_scales) {
        if (desc->restart_displacement) {
            start = desc->restart_at - 1;
            prev.assign(desc->restart_displacement, desc->restart_displacement + n_tar * dim);
        } else if (out.on()) {
            const std::string path = out.dir + "/df-full-" + std::to_string(std::max(0, desc->restart_at - 2)) + ".vtk";
            skm_image_header h;
            std::vector<float> df((size_t)n_tar * dim);
            if (skm_read_image_header(path.c_str(), &h) == SKM_SUCCESS && h.components == dim &&
                voxels(h.grid, dim) == n_tar &&
                skm_read_image(path.c_str(), df.data(), (int64_t)df.size(), nullptr) == SKM_SUCCESS) {
                start = desc->restart_at - 1;
                prev.swap(df);
            }
        }
    }
    const long long n_w = voxels(desc->weight, dim);

    // device copies of the other inputs, made once
    if (desc->weight_image)
        SKM_R_TRY(d_weight.upload(desc->weight_image, (size_t)n_w, ss));
    else if (desc->weight_tensor)
        SKM_R_TRY(d_weight.upload(desc->weight_tensor, (size_t)n_w * dim * dim, ss));
    if (desc->guidance_image) SKM_R_TRY(d_guide.upload(desc->guidance_image, (size_t)n_w, ss));
    bool has_prev = !prev.empty();
    const bool warm_field = !has_prev && desc->initial_displacement != nullptr;
    if (warm_field && desc->initial_displacement_on_device) {
        // device memory of this device, copied on the device
        hipPointerAttribute_t attr{};
        if (hipPointerGetAttributes(&attr, desc->initial_displacement) != hipSuccess ||
            attr.type != hipMemoryTypeDevice || attr.device != setup.dev) {
            (void)hipGetLastError();
            return set_last_error(SKM_ERROR_INVALID_ARGUMENT,
                                  "initial_displacement_on_device: not device memory of the current device");
        }
        SKM_R_TRY(d_prev.alloc((size_t)n_tar * dim));
        hipError_t e = hipMemcpyAsync(d_prev.p, desc->initial_displacement, sizeof(float) * (size_t)n_tar * dim,
                                      hipMemcpyDeviceToDevice, ss);
        if (e == hipSuccess) e = hipStreamSynchronize(ss);
        if (e != hipSuccess)
            return set_last_error(SKM_ERROR_HIP, std::string("initial displacement: ") + hipGetErrorString(e));
        has_prev = true;
    } else {
        if (warm_field) {
            prev.assign(desc->initial_displacement, desc->initial_displacement + n_tar * dim);
            has_prev = true;
        }
        if (has_prev)
            SKM_R_TRY(d_prev.upload(prev.data(), prev.size(), ss));
        else
            SKM_R_TRY(d_prev.alloc((size_t)n_tar * dim));
    }

    std::vector<float> host_tmp;
    for (int scale = start; scale < desc->num_scales; ++scale) {
        const skm_scale_params &P = desc->scale[scale];
        skm_scale_report rep;
        std::memset(&rep, 0, sizeof(rep));
        const auto t_setup = Clock::now();

        // smoothing (SKMReg.cxx:140-150)
        const float *ref_s = d_ref.p, *tar_s = d_tar.p;
        if (P.smoothing_variance > 1e-5) {
            SKM_R_TRY(d_ref_s.alloc(d_ref.n));
            SKM_R_TRY(d_tar_s.alloc(d_tar.n));
            SKM_R_TRY(skm::gaussian_smooth_device(dim, desc->reference, d_ref.p, P.smoothing_variance, d_ref_s.p, ss));
            SKM_R_TRY(skm::gaussian_smooth_device(dim, desc->target, d_tar.p, P.smoothing_variance, d_tar_s.p, ss));
            ref_s = d_ref_s.p;
            tar_s = d_tar_s.p;
        }
        if (out.on() && desc->verbosity >= 2) {
            const DeviceArray<float> &rs = ref_s == d_ref.p ? d_ref : d_ref_s;
            const DeviceArray<float> &ts = tar_s == d_tar.p ? d_tar : d_tar_s;
            SKM_R_TRY(rs.download(host_tmp, ss));
            SKM_R_TRY(out.write("smoothed_reference-" + std::to_string(scale) + ".vtk", desc->reference, 1,
                                host_tmp.data()));
            SKM_R_TRY(ts.download(host_tmp, ss));
            SKM_R_TRY(out.write("smoothed_target-" + std::to_string(scale) + ".vtk", desc->target, 1,
                                host_tmp.data()));
        }

        // kernel: weight image (takes precedence, GpuEvaluator.cu:311-346),
        // else the weight tensor cases of SKMReg.cxx:262-297
        skm_desc ed;
        skm_desc_init(&ed);
        ed.dim = dim;
        ed.fixed = desc->target;
        ed.fixed_data = tar_s;
        ed.moving = desc->reference;
        ed.moving_data = ref_s;
        ed.sigma_c0 = P.sigma_c0;
        ed.sigma_c4 = P.sigma_c4;
        ed.subsample = P.sampling_rate;
        ed.subsample_neighborhood = P.neighborhood_sampling;
        ed.flags = desc->exact_gradient ? SKM_FLAG_EXACT_GRADIENT : 0u;
        const float *host_tensor = nullptr;  // the level's weight tensor in host memory, if any
        if (desc->weight_image) {
            double mx = std::numeric_limits<double>::lowest();
            for (long long i = 0; i < n_w; ++i) mx = std::max(mx, (double)desc->weight_image[i]);
            ed.kernel_mode = SKM_KERNEL_WEIGHT_IMAGE;
            ed.weight = desc->weight;
            ed.weight_data = d_weight.p;
            ed.maximum_weight = mx;
        } else if (desc->weight_tensor && P.tensor_weight == 0) {
            // SetWeightingTensor(tensor, -1): largest |lambda_max| and its tensor
            // (itkTransformAdapter.hxx:341-361)
            const int nn = dim * dim;
            double best = std::numeric_limits<double>::lowest();
            long long at = 0;
            for (long long i = 0; i < n_w; ++i) {
                const double l = std::abs(max_eig_sym(dim, desc->weight_tensor + i * nn));
                if (l > best) {
                    best = l;
                    at = i;
                }
            }
            ed.kernel_mode = SKM_KERNEL_WEIGHT_TENSOR;
            ed.weight = desc->weight;
            ed.weight_data = d_weight.p;
            ed.maximum_weight = best;
            for (int i = 0; i < 9; ++i) ed.maximum_eigenvalue_covariance[i] = 0.0;
            for (int i = 0; i < dim; ++i)
                for (int j = 0; j < dim; ++j)
                    ed.maximum_eigenvalue_covariance[i * 3 + j] = desc->weight_tensor[at * nn + i * dim + j];
            host_tensor = desc->weight_tensor;
        } else if (P.tensor_weight > 0) {
            const bool guided = desc->guidance_image != nullptr;
            const skm_grid &g = guided ? desc->weight : desc->target;
            SKM_R_TRY(d_tensor.alloc((size_t)voxels(g, dim) * dim * dim));
            skm_tensor_options to{P.sigma_c4, P.tensor_weight, 2, 0};
            double tilt = 0, mw = 0, mcov[9];
            SKM_R_TRY(skm::anisotropic_tensor_device(dim, g, guided ? d_guide.p : tar_s, to, d_tensor.p, &tilt, &mw,
                                                     mcov, ss));
            ed.kernel_mode = SKM_KERNEL_WEIGHT_TENSOR;
            ed.weight = g;
            ed.weight_data = d_tensor.p;
            ed.maximum_weight = mw;
            std::memcpy(ed.maximum_eigenvalue_covariance, mcov, sizeof(mcov));
            rep.maximum_tilt = tilt;
            if (out.on()) {
                SKM_R_TRY(d_tensor.download(host_tmp, ss));
                host_tensor = host_tmp.data();
            }
        }
        if (ed.kernel_mode == SKM_KERNEL_WEIGHT_TENSOR)  // SKMReg.cxx:290-297
            SKM_R_TRY(out.write("covariance_tensor-" + std::to_string(scale) + ".vtk", ed.weight, dim * dim,
                                host_tensor));

        // control points (SKMReg.cxx:219-232)
        double sp[3];
        const double smin = std::min(P.sigma_c4, P.sigma_c0);
        for (int d = 0; d < 3; ++d) sp[d] = desc->cp_spacing_scaler < 1 ? smin : smin / desc->cp_spacing_scaler;
        SKM_R_TRY(skm_control_point_grid(dim, &desc->reference, sp, &ed.cp));
        long long n_cp = 1;
        for (int d = 0; d < dim; ++d) {
            n_cp *= ed.cp.size[d];
            rep.cp_size[d] = ed.cp.size[d];
        }

        EvaluatorGuard guard;
        // the evaluator works on the registration's stream too: streams beyond
        // the device's hardware queues share them in order, so each extra
        // stream of a registration can queue behind another's optimizer loop
        SKM_R_TRY(skm::create_evaluator(&ed, true, &guard.ev, ss));
        skm_evaluator *ev = guard.ev;
        if (P.reg_rkhs > 0) SKM_R_TRY(skm_set_regularizer_rkhs(ev, P.reg_rkhs));
        if (P.reg_rd > 0) SKM_R_TRY(skm_set_regularizer_rd(ev, P.reg_rd, P.reg_rd_scaling));
        if (P.reg_pg > 0) SKM_R_TRY(skm_set_regularizer_pg(ev, P.reg_pg, P.reg_pg_scaling));
        if (P.reg_bilateral > 0)
            SKM_R_TRY(skm_set_regularizer_bilateral(ev, P.reg_bilateral, P.reg_bilateral_spatial_sigma,
                                                    P.reg_bilateral_intensity_sigma));
        if (has_prev) SKM_R_TRY(skm::set_displacement_field_device(ev, &desc->target, d_prev.p));
        if (desc->autotune) {
            // launch shapes of this level's shape class, tuned once per process
            // (or loaded with the table)
            if (!skm::tuning_recorded(ev)) {
                std::vector<float> zero((size_t)n_cp * dim, 0.0f), one((size_t)n_cp, 1.0f);
                SKM_R_TRY(d_cp.upload(zero.data(), zero.size(), ss));
                SKM_R_TRY(d_cpw.upload(one.data(), one.size(), ss));
                SKM_R_TRY(skm_autotune(ev, desc->metric, d_cp.p, d_cpw.p, 3, nullptr));
            }
        }
        rep.ms_setup = ms_since(t_setup);

        // optimisation (SKMReg.cxx:375-430)
        const auto t_opt = Clock::now();
        const bool only_l1 = P.reg_l1 > 0 && P.reg_l21 <= 0 && P.reg_l2 <= 0 && P.reg_rkhs <= 0 && P.reg_rd <= 0 &&
                             P.reg_pg <= 0 && P.reg_bilateral <= 0;
        const bool only_l21 = P.reg_l1 <= 0 && P.reg_l21 > 0 && P.reg_l2 <= 0 && P.reg_rkhs <= 0 && P.reg_rd <= 0 &&
                              P.reg_pg <= 0 && P.reg_bilateral <= 0;
        skm_metric_options mo{desc->metric, P.reg_l1, P.reg_l21, P.reg_l2, 1, 1};
        skm_optimizer_options oo{SKM_OPT_ADAPTIVE, P.initial_step_size, 20.0, 1.0, 1e7, -1e6, 1.0,
                                 (only_l1 || only_l21) ? 1 : 0, P.iterations, 1, 0};
        std::vector<float> params((size_t)n_cp * dim, 0.0f), cpw((size_t)n_cp, 1.0f);
        if (scale == start && warm_field) {
            // warm start by a dense field: kept unless it evaluates worse than none
            rep.initial_used = 1;
            if (desc->initial_guard) {
                const std::vector<float> zero(params.size(), 0.0f);
                SKM_R_TRY(level_value(ev, desc->metric, P, zero, cpw, 1, &rep.initial_value));
                SKM_R_TRY(skm_set_displacement_field(ev, nullptr, nullptr));
                SKM_R_TRY(level_value(ev, desc->metric, P, zero, cpw, 0, &rep.zero_value));
                if (rep.initial_value < rep.zero_value) {
                    SKM_R_TRY(skm::set_displacement_field_device(ev, &desc->target, d_prev.p));
                } else {
                    rep.initial_used = 0;
                    has_prev = false;
                }
            }
        }
        if (scale == start && desc->initial_coefficients) {
            // warm start: the given coefficients, unless they evaluate worse than zeros
            std::copy(desc->initial_coefficients, desc->initial_coefficients + params.size(), params.begin());
            rep.initial_used = 1;
            if (desc->initial_guard) {
                SKM_R_TRY(level_value(ev, desc->metric, P, params, cpw, 1, &rep.initial_value));
                const std::vector<float> zero(params.size(), 0.0f);
                SKM_R_TRY(level_value(ev, desc->metric, P, zero, cpw, 0, &rep.zero_value));
                if (!(rep.initial_value < rep.zero_value)) {
                    std::fill(params.begin(), params.end(), 0.0f);
                    rep.initial_used = 0;
                }
            }
        }
        std::vector<skm_iteration_record> rec(std::max(1, P.iterations));
        if (group) {
            oo.graph = 1;  // one graph launch per step for every registration of the group
            Lockstep::Job job{ev, mo, oo, cpw.data(), params.data(), rec.data()};
            SKM_R_TRY(group->optimize(job));
        } else {
            SKM_R_TRY(skm_optimize(ev, &mo, &oo, cpw.data(), params.data(), rec.data()));
        }
        rep.iterations = P.iterations;
        if (P.iterations > 0) {
            int best = 0;
            for (int k = 1; k < P.iterations; ++k)
                if (rec[k].value < rec[best].value) best = k;
            rep.value = rec[best].value;
            rep.measure = rec[best].measure;
            rep.first_value = rec[0].value;
        }
        rep.ms_optimize = ms_since(t_opt);

        // displacement of all levels so far (SKMReg.cxx:436-447), written over
        // the previous field, which the evaluator holds its own copy of
        const auto t_df = Clock::now();
        std::vector<float> cp_aos((size_t)n_cp * dim);
        for (long long j = 0; j < n_cp; ++j)
            for (int d = 0; d < dim; ++d) cp_aos[j * dim + d] = params[d * n_cp + j];
        SKM_R_TRY(d_cp.upload(cp_aos.data(), cp_aos.size(), ss));
        SKM_R_TRY(d_cpw.upload(cpw.data(), cpw.size(), ss));
        SKM_R_TRY(skm_evaluate_displacement_field_device(ev, d_cp.p, d_cpw.p, d_prev.p));
        SKM_R_TRY(skm_synchronize(ev));
        has_prev = true;
        if (out.on() || landmarks) SKM_R_TRY(d_prev.download(prev, ss));
        rep.ms_displacement = ms_since(t_df);

        // results of the level (SKMReg.cxx:459-505)
        const std::string s = std::to_string(scale);
        SKM_R_TRY(out.write("df-full-" + s + ".vtk", desc->target, dim, prev.data()));
        SKM_R_TRY(out.write("cp_image-" + s + ".vtk", ed.cp, dim, cp_aos.data()));
        if (out.on()) {
            const std::vector<float> jac = jacobian_determinant(dim, ed.cp, cp_aos.data());
            SKM_R_TRY(out.write("cp_jacdet-" + s + ".vtk", ed.cp, 1, jac.data()));
        }
        rep.tre = -1.0;
        if (landmarks)
            SKM_R_TRY(skm_landmark_tre(dim, &desc->target, prev.data(), desc->reference_landmarks,
                                       desc->target_landmarks, desc->num_landmarks, &rep.tre));
        if (out.on() && desc->verbosity >= 2) {
            // the level's own transform, without the coarser levels
            std::vector<float> own((size_t)n_tar * dim);
            SKM_R_TRY(skm_set_displacement_field(ev, nullptr, nullptr));
            SKM_R_TRY(skm_evaluate_displacement_field(ev, cp_aos.data(), cpw.data(), own.data()));
            SKM_R_TRY(out.write("df-" + s + ".vtk", desc->target, dim, own.data()));
            std::vector<float> warped((size_t)n_tar);
            SKM_R_TRY(skm_warp_image(dim, &desc->reference, ref.data(), &desc->target, prev.data(), &desc->target,
                                     SKM_INTERPOLATION_BSPLINE3, 0.0, warped.data()));
            SKM_R_TRY(out.write("warped-" + s + ".vtk", desc->target, 1, warped.data()));
        }
        long long zeros = 0;
        for (float v : params) zeros += v == 0.0f;
        rep.zero_parameters = params.empty() ? 0.0 : 100.0 * (double)zeros / (double)params.size();
        if (reports) reports[scale] = rep;
    }
    if (!has_prev)
        prev.assign((size_t)n_tar * dim, 0.0f);
    else if (prev.size() != (size_t)n_tar * dim || !(out.on() || landmarks))
        SKM_R_TRY(d_prev.download(prev, ss));
    std::memcpy(displacement_out, prev.data(), sizeof(float) * prev.size());

    // final results (SKMReg.cxx:507-525): the dense field and the reference
    // resampled onto the target grid through it, linear interpolation, 0 outside
    if (out.on()) {
        SKM_R_TRY(out.write("df.vtk", desc->target, dim, prev.data()));
        std::vector<float> warped((size_t)n_tar);
        SKM_R_TRY(skm_warp_image(dim, &desc->reference, ref.data(), &desc->target, prev.data(), &desc->target,
                                 SKM_INTERPOLATION_LINEAR, 0.0, warped.data()));
        SKM_R_TRY(out.write("warped.vtk", desc->target, 1, warped.data()));
    }
    return SKM_SUCCESS;
}

}  // namespace

// Batched registration: a shared counter hands out the registrations to the
// worker threads in index order; each worker makes the caller's device
// current and runs skm_register, whose evaluators take their own streams from
// the device's pool.  Errors are reported per registration (the error text
// is per thread) and the first by index is returned.
SKM_API skm_status skm_register_batch(const skm_registration_desc *descs, int32_t n, int32_t concurrency,
                                      float *const *displacements_out, skm_scale_report *const *reports,
                                      skm_status *statuses)
{
    if (n < 0) return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "n must be >= 0");
    if (n == 0) return SKM_SUCCESS;
    if (!descs || !displacements_out) return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "null argument");
    if (concurrency < 1) return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "concurrency must be >= 1");
    for (int32_t i = 0; i < n; ++i)
        if (!displacements_out[i])
            return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "displacements_out[" + std::to_string(i) + "] is null");
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) return set_last_error(SKM_ERROR_HIP, "hipGetDevice failed");
    std::vector<skm_status> st((size_t)n, SKM_SUCCESS);
    std::vector<std::string> msg((size_t)n);
    std::atomic<int32_t> next{0};
    auto worker = [&]() {
        if (hipSetDevice(device) != hipSuccess) {
            // no registration runs on this thread; the others take them all
            return;
        }
        for (int32_t i = next.fetch_add(1); i < n; i = next.fetch_add(1)) {
            try {
                st[i] = skm_register(&descs[i], displacements_out[i], reports ? reports[i] : nullptr);
                if (st[i] != SKM_SUCCESS) msg[i] = skm_last_error();
            } catch (const std::bad_alloc &) {
                st[i] = SKM_ERROR_OUT_OF_MEMORY;
                msg[i] = "host allocation failed";
            } catch (const std::exception &e) {
                st[i] = SKM_ERROR_HIP;
                msg[i] = e.what();
            }
        }
    };
    const int32_t workers = std::min(concurrency, n);
    std::vector<std::thread> pool;
    try {
        for (int32_t w = 1; w < workers; ++w) pool.emplace_back(worker);
    } catch (const std::system_error &) {
        // fewer threads than asked: the ones started and this one do the work
    }
    worker();
    for (auto &t : pool) t.join();
    if (next.load() < n) return set_last_error(SKM_ERROR_HIP, "hipSetDevice failed on the worker threads");
    if (statuses)
        for (int32_t i = 0; i < n; ++i) statuses[i] = st[i];
    for (int32_t i = 0; i < n; ++i)
        if (st[i] != SKM_SUCCESS) return set_last_error(st[i], "registration " + std::to_string(i) + ": " + msg[i]);
    return SKM_SUCCESS;
}

// Lockstep batch: one thread per registration, all in flight; every level's
// optimisers run together through the group (one graph launch per optimizer
// step for all registrations with the same options).
SKM_API skm_status skm_register_batch_lockstep(const skm_registration_desc *descs, int32_t n,
                                               float *const *displacements_out, skm_scale_report *const *reports,
                                               skm_status *statuses)
{
    if (n < 0) return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "n must be >= 0");
    if (n == 0) return SKM_SUCCESS;
    if (!descs || !displacements_out) return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "null argument");
    for (int32_t i = 0; i < n; ++i)
        if (!displacements_out[i])
            return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "displacements_out[" + std::to_string(i) + "] is null");
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) return set_last_error(SKM_ERROR_HIP, "hipGetDevice failed");
    std::vector<skm_status> st((size_t)n, SKM_SUCCESS);
    std::vector<std::string> msg((size_t)n);
    Lockstep group;
    group.active = n;
    auto one = [&](int32_t i) {
        if (hipSetDevice(device) != hipSuccess) {
            st[i] = SKM_ERROR_HIP;
            msg[i] = "hipSetDevice failed";
            group.leave();
            return;
        }
        try {
            st[i] = register_impl(&descs[i], displacements_out[i], reports ? reports[i] : nullptr, &group);
            if (st[i] != SKM_SUCCESS) msg[i] = skm_last_error();
        } catch (const std::bad_alloc &) {
            st[i] = SKM_ERROR_OUT_OF_MEMORY;
            msg[i] = "host allocation failed";
        } catch (const std::exception &e) {
            st[i] = SKM_ERROR_HIP;
            msg[i] = e.what();
        }
    };
    std::vector<std::thread> pool;
    for (int32_t i = 1; i < n; ++i) {
        try {
            pool.emplace_back(one, i);
        } catch (const std::system_error &) {
            st[i] = SKM_ERROR_HIP;
            msg[i] = "could not start a thread";
            group.leave();  // the others do not wait for it
        }
    }
    one(0);
    for (auto &t : pool) t.join();
    if (statuses)
        for (int32_t i = 0; i < n; ++i) statuses[i] = st[i];
    for (int32_t i = 0; i < n; ++i)
        if (st[i] != SKM_SUCCESS) return set_last_error(st[i], "registration " + std::to_string(i) + ": " + msg[i]);
    return SKM_SUCCESS;
}
