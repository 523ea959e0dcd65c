// skm_capi.hip -- C ABI of libskm.so: the evaluator object that replaces
// GpuKernelEvaluator (include/cuda/GpuEvaluator.h:69-179), its host-side
// set-up (B-spline prefilter, static windows, per-control-point kernel state)
// and the orchestration of the device passes on one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "skm.h"
#include "skm_internal.hpp"

using namespace skm;

namespace {

thread_local std::string g_error;

skm_status fail(skm_status st, const std::string &msg)
{
    g_error = msg;
    return st;
}

#define SKM_HIP(expr)                                                                                              \
    do {                                                                                                           \
        hipError_t e_ = (expr);                                                                                    \
        if (e_ != hipSuccess)                                                                                      \
            return fail(e_ == hipErrorOutOfMemory ? SKM_ERROR_OUT_OF_MEMORY : SKM_ERROR_HIP,                      \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                                        \
    } while (0)

#define SKM_TRY(expr)                                                                                              \
    do {                                                                                                           \
        skm_status s_ = (expr);                                                                                    \
        if (s_ != SKM_SUCCESS) return s_;                                                                          \
    } while (0)

long long grid_count(const skm_grid &g, int dim)
{
    long long n = 1;
    for (int d = 0; d < dim; ++d) n *= g.size[d];
    return n;
}

Grid to_grid(const skm_grid &g, int dim)
{
    Grid r;
    for (int d = 0; d < 3; ++d) {
        r.n[d] = d < dim ? g.size[d] : 1;
        r.o[d] = d < dim ? (float)g.origin[d] : 0.0f;
        r.h[d] = d < dim ? (float)g.spacing[d] : 1.0f;
        r.ih[d] = (float)(1.0 / (d < dim ? g.spacing[d] : 1.0));
    }
    return r;
}

// ---------------------------------------------------------------- host math

double det_h(int dim, const double *m)
{
    if (dim == 2) return m[0] * m[3] - m[2] * m[1];
    return m[0] * m[4] * m[8] + m[3] * m[7] * m[2] + m[6] * m[1] * m[5] - m[0] * m[7] * m[5] -
           m[6] * m[4] * m[2] - m[3] * m[1] * m[8];
}

double max_eig_h(int dim, const double *m)
{
    if (dim == 2) {
        double tr = m[0] + m[3], det = det_h(2, m);
        return 0.5 * (tr + std::sqrt(std::max(tr * tr - 4.0 * det, 0.0)));
    }
    double tri = m[1] * m[1] + m[2] * m[2] + m[5] * m[5];
    if (tri == 0.0) return std::max(m[0], std::max(m[4], m[8]));
    double q = (1.0 / 3.0) * (m[0] + m[4] + m[8]);
    double psq = 2.0 * tri + (m[0] - q) * (m[0] - q) + (m[4] - q) * (m[4] - q) + (m[8] - q) * (m[8] - q);
    double p = std::sqrt(psq * (1.0 / 6.0));
    double b[9];
    std::memcpy(b, m, sizeof(b));
    b[0] -= q; b[4] -= q; b[8] -= q;
    double r = det_h(3, b) * (0.5 / (p * p * p));
    double phi = (1.0 / 3.0) * std::acos(std::min(1.0, std::max(-1.0, r)));
    return q + 2.0 * p * std::cos(phi);
}

// Symmetric part (m + m^T) / 2 of a row-major dim x dim matrix.
void symmetric_part(int dim, const double *m, double *out)
{
    for (int i = 0; i < dim; ++i)
        for (int j = 0; j < dim; ++j) out[i * dim + j] = 0.5 * (m[i * dim + j] + m[j * dim + i]);
}

// pstate_pack layout of a mode-2 point: {up0, up1, up2, sdet}, {up3, up4, up5, lam}.
void pack_upper(int dim, const double *sym, float sdet, float lam, float *ps)
{
    int k = 0;
    float up[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < dim; ++i)
        for (int j = i; j < dim; ++j) up[k++] = (float)sym[i * dim + j];
    ps[0] = up[0]; ps[1] = up[1]; ps[2] = up[2]; ps[3] = sdet;
    ps[4] = up[3]; ps[5] = up[4]; ps[6] = up[5]; ps[7] = lam;
}

}  // namespace

// ---------------------------------------------------------------- device memory

namespace {

struct BlockCache {
    std::mutex mutex;
    std::unordered_map<void *, std::pair<int, size_t>> live;   // block -> (device, class bytes)
    std::map<std::pair<int, size_t>, std::vector<void *>> idle;  // (device, class bytes) -> blocks
    std::map<int, size_t> idle_bytes;                          // per device
};

BlockCache &block_cache()
{
    // never destroyed: no HIP call may run during the process's exit
    static BlockCache *cache = new BlockCache;
    return *cache;
}

// idle bytes kept per device: SKM_CACHE_IDLE_BYTES in the environment (read
// at the first allocation), skm_set_cache_limit, else 4 GiB
std::atomic<long long> g_idle_cap{-1};

size_t idle_cap()
{
    long long cap = g_idle_cap.load();
    if (cap < 0) {
        cap = (long long)(size_t(4) << 30);
        if (const char *env = std::getenv("SKM_CACHE_IDLE_BYTES")) {
            char *end = nullptr;
            const long long v = std::strtoll(env, &end, 10);
            if (end != env && v >= 0) cap = v;
        }
        long long expected = -1;
        if (!g_idle_cap.compare_exchange_strong(expected, cap)) cap = expected;
    }
    return (size_t)cap;
}

// hipFree of every idle block (of one device, or all with dev < 0)
void drop_idle(int dev)
{
    std::vector<void *> drop;
    {
        BlockCache &c = block_cache();
        std::lock_guard<std::mutex> lock(c.mutex);
        for (auto it = c.idle.begin(); it != c.idle.end();) {
            if (dev < 0 || it->first.first == dev) {
                drop.insert(drop.end(), it->second.begin(), it->second.end());
                it = c.idle.erase(it);
            } else {
                ++it;
            }
        }
        for (auto &kv : c.idle_bytes)
            if (dev < 0 || kv.first == dev) kv.second = 0;
    }
    for (void *q : drop) (void)hipFree(q);
}

// 256 bytes, or bytes rounded up to an eighth of the largest power of two
// below it (at most 12.5% unused)
size_t size_class(size_t bytes)
{
    if (bytes <= 256) return 256;
    size_t pow2 = 1;
    while (pow2 <= (bytes - 1) / 2) pow2 *= 2;
    const size_t step = std::max<size_t>(256, pow2 / 8);
    return (bytes + step - 1) / step * step;
}

}  // namespace

hipError_t skm::device_alloc(void **p, size_t bytes)
{
    *p = nullptr;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    const size_t cls = size_class(bytes);
    BlockCache &c = block_cache();
    {
        std::lock_guard<std::mutex> lock(c.mutex);
        auto it = c.idle.find({dev, cls});
        if (it != c.idle.end() && !it->second.empty()) {
            *p = it->second.back();
            it->second.pop_back();
            c.idle_bytes[dev] -= cls;
            c.live[*p] = {dev, cls};
            return hipSuccess;
        }
    }
    e = hipMalloc(p, cls);
    if (e == hipErrorOutOfMemory) {
        // give this device's idle blocks back and try once more
        drop_idle(dev);
        (void)hipGetLastError();
        e = hipMalloc(p, cls);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        return e;
    }
    std::lock_guard<std::mutex> lock(c.mutex);
    c.live[*p] = {dev, cls};
    return hipSuccess;
}

void skm::device_release(void *p)
{
    if (!p) return;
    BlockCache &c = block_cache();
    {
        std::lock_guard<std::mutex> lock(c.mutex);
        auto it = c.live.find(p);
        if (it != c.live.end()) {
            const std::pair<int, size_t> key = it->second;
            c.live.erase(it);
            if (c.idle_bytes[key.first] + key.second <= idle_cap()) {
                c.idle[key].push_back(p);
                c.idle_bytes[key.first] += key.second;
                return;
            }
        }
    }
    (void)hipFree(p);
}

// ---------------------------------------------------------------- evaluator

namespace {

// Streams of evaluators.  Creating and destroying a HIP stream costs about
// 3 ms on MI355X, more than the rest of an evaluator's set-up, so a
// destroyed evaluator's stream (synchronised, default priority, non-blocking)
// is kept for the next evaluator on the same device.
//
// The HIP runtime backs a process's streams with a few hardware queues
// (GPU_MAX_HW_QUEUES, 4 by default), giving each new stream the queue that
// backs the fewest streams, so streams created one after another land on
// different queues in turn.  Streams that share a queue run their work one
// after the other.  The pool therefore numbers its streams in creation order
// and always hands out the idle stream with the lowest number: streams in use
// at the same time are then consecutive in creation order and spread over
// the queues.  (Handing out the last stream returned instead let, after a
// batch of 16 registrations, four concurrent ones share queues and take 60 %
// longer.)  When too many are idle, the highest numbers are destroyed.
struct StreamPool {
    std::mutex mutex;
    std::map<std::pair<int, unsigned long long>, hipStream_t> idle;  // (device, number) -> stream
    std::unordered_map<hipStream_t, unsigned long long> number;      // every pooled stream's number
    unsigned long long next = 0;
};

StreamPool &stream_pool()
{
    // never destroyed: no HIP call may run during the process's exit
    static StreamPool *pool = new StreamPool;
    return *pool;
}

hipError_t take_stream(hipStream_t *out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    StreamPool &p = stream_pool();
    std::lock_guard<std::mutex> lock(p.mutex);
    auto it = p.idle.lower_bound({dev, 0ull});
    if (it != p.idle.end() && it->first.first == dev) {
        *out = it->second;
        p.idle.erase(it);
        return hipSuccess;
    }
    // created under the lock, so that numbers follow the runtime's creation order
    e = hipStreamCreateWithFlags(out, hipStreamNonBlocking);
    if (e == hipSuccess) p.number[*out] = p.next++;
    return e;
}

void return_stream(int dev, hipStream_t s)
{
    if (!s) return;
    constexpr size_t kKeep = 32;  // idle streams kept per process
    hipStream_t drop = s;
    if (hipStreamSynchronize(s) == hipSuccess) {
        StreamPool &p = stream_pool();
        std::lock_guard<std::mutex> lock(p.mutex);
        auto n = p.number.find(s);
        if (n == p.number.end()) n = p.number.emplace(s, p.next++).first;
        p.idle[{dev, n->second}] = s;
        drop = nullptr;
        if (p.idle.size() > kKeep) {
            // the last entry: the highest number of the highest device
            auto last = std::prev(p.idle.end());
            drop = last->second;
            p.idle.erase(last);
            p.number.erase(drop);
        }
    } else {
        StreamPool &p = stream_pool();
        std::lock_guard<std::mutex> lock(p.mutex);
        p.number.erase(s);
    }
    if (drop) (void)hipStreamDestroy(drop);
}

}  // namespace

hipError_t skm::take_pool_stream(hipStream_t *out) { return take_stream(out); }
void skm::return_pool_stream(int dev, hipStream_t s) { return_stream(dev, s); }

struct skm_evaluator {
    skm_desc desc{};
    int dim = 3;
    Problem pr{};
    hipStream_t stream = nullptr;
    bool own_stream = false;  // taken from the stream pool (returned on destruction)
    int stream_device = 0;
    int max_partial_blocks = 0;

    float *d_fixed = nullptr, *d_coeffs = nullptr, *d_weight = nullptr, *d_df = nullptr;
    float4 *d_cpc = nullptr;
    float *d_cp_state = nullptr;
    float4 *d_cp_pstate = nullptr, *d_smp_pstate = nullptr, *d_vox_pstate = nullptr;
    float *d_cp_row_lam = nullptr, *d_smp_row_lam = nullptr, *d_vox_row_lam = nullptr, *d_cp_reach = nullptr;
    int *d_cp_window = nullptr, *d_reg_window = nullptr, *d_lcc_window = nullptr, *d_dense_window = nullptr;
    int *d_offsets = nullptr;
    uint64_t *d_rng = nullptr;  // PCG32 state of the cube offsets
    float4 *d_sg = nullptr, *d_sv = nullptr;
    float *d_sb = nullptr;
    double *d_partials = nullptr, *d_reg_partials = nullptr, *d_hreg_partials = nullptr;
    double *d_gather_partials = nullptr;
    Scalars *d_sc = nullptr;
    float *d_deriv = nullptr, *d_cp_in = nullptr, *d_cpw_in = nullptr, *d_dst = nullptr;
    float4 *d_dense = nullptr, *d_dense_q = nullptr, *d_ccoef = nullptr, *d_gamma = nullptr;
    double *d_lcc_ref_partials = nullptr;  // chunk partials of the reference-mode LCC derivative
    long long lcc_ref_partials_n = 0;       // doubles held by d_lcc_ref_partials
    skm_tuning tuning = {};                 // launch shapes as set by skm_set_tuning
    double *d_lcc_bound = nullptr;          // exact LCC: per-workgroup bounds of the centre pass
    unsigned long long *d_gamma_fx = nullptr;  // exact LCC: fixed-point scatter sums, 3 per voxel
    unsigned *d_lcc_cpk = nullptr;          // LCC voxel gathers: packed offsets + "contributes" per centre
    double *d_lcc_reft_partials = nullptr;  // reference-mode LCC derivative by voxel tables: block sums
    float *d_params = nullptr, *d_prev_g = nullptr, *d_best = nullptr;
    OptState *d_opt = nullptr;
    skm_iteration_record *d_records = nullptr;
    int records_cap = 0;

    uint64_t rng_inc = 1;
    bool evaluated_once = false;

    // pass timing: events recorded around each pass while enabled
    bool timing = false;
    struct Pending { int pass; hipEvent_t a, b; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    double pass_ms[SKM_NUM_PASSES] = {0};
    uint32_t pass_launches[SKM_NUM_PASSES] = {0};
    double reg_w[4] = {0, 0, 0, 0};        // RKHS, RD, PG, bilateral
    double rd_scaling = 1.0, pg_scaling = 1.0;
    double bl_sigma_s = 1.0, bl_sigma_i = 1.0;
    float *d_cp_intensity = nullptr;       // fixed image at the control points (bilateral term)
    KernelMatrix km;                       // assembled k(x_m, p_j) (subsample 1; see use_kernel_matrix)
    // the regularizers' second stream (MSE / NCC: beside the measure's passes)
    hipStream_t side = nullptr;
    hipEvent_t side_fork = nullptr, side_join = nullptr;
    double *d_reg_add = nullptr;           // their derivative terms (SoA, FP64), added by the gather's finish

    ~skm_evaluator()
    {
        // work queued on the streams may still read these buffers
        if (stream) (void)hipStreamSynchronize(stream);
        if (side) (void)hipStreamSynchronize(side);
        kernel_matrix_release(km);
        void *ptrs[] = {d_fixed, d_coeffs, d_weight, d_df, d_cpc, d_cp_state, d_cp_pstate, d_smp_pstate,
                        d_vox_pstate, d_cp_row_lam, d_smp_row_lam, d_vox_row_lam, d_cp_reach, d_cp_window,
                        d_reg_window,
                        d_lcc_window, d_dense_window, d_offsets, d_sg, d_sv, d_sb, d_partials, d_reg_partials,
                        d_hreg_partials, d_gather_partials, d_sc, d_deriv, d_cp_in, d_cpw_in, d_dst, d_dense,
                        d_dense_q, d_ccoef, d_gamma, d_params, d_prev_g, d_best, d_opt, d_records, d_rng,
                        d_lcc_ref_partials, d_lcc_bound, d_gamma_fx, d_cp_intensity, d_lcc_cpk,
                        d_lcc_reft_partials, d_reg_add};
        for (void *p : ptrs)
            if (p) device_release(p);
        for (auto &p : pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
        for (auto e : event_pool) (void)hipEventDestroy(e);
        if (side_fork) (void)hipEventDestroy(side_fork);
        if (side_join) (void)hipEventDestroy(side_join);
        if (side) return_stream(stream_device, side);
        if (own_stream && stream) return_stream(stream_device, stream);
    }
};

namespace {

hipEvent_t take_event(skm_evaluator *ev)
{
    if (!ev->event_pool.empty()) {
        hipEvent_t e = ev->event_pool.back();
        ev->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Brackets the launches of one pass with events when timing is on.
struct PassTimer {
    skm_evaluator *ev;
    int pass;
    hipEvent_t a = nullptr;
    PassTimer(skm_evaluator *e, int p) : ev(e), pass(p)
    {
        if (!ev->timing) return;
        a = take_event(ev);
        if (a) (void)hipEventRecord(a, ev->stream);
    }
    ~PassTimer()
    {
        if (!ev->timing || !a) return;
        hipEvent_t b = take_event(ev);
        if (!b) return;
        (void)hipEventRecord(b, ev->stream);
        ev->pending.push_back({pass, a, b});
    }
};

template <typename T>
skm_status dev_alloc(T **p, size_t count)
{
    if (count == 0) count = 1;
    SKM_HIP(device_alloc((void **)p, sizeof(T) * count));
    return SKM_SUCCESS;
}

// A device buffer holding a copy of the host array src, complete on return:
// copied on the evaluator's stream s and waited for.  (A plain hipMemcpy runs
// on the null stream, which the evaluator's non-blocking stream does not
// wait for, and a copy from pageable memory may return before its transfer
// has landed.)
template <typename T>
skm_status dev_upload(T **p, const T *src, size_t count, hipStream_t s)
{
    SKM_TRY(dev_alloc(p, count));
    if (src && count) {
        SKM_HIP(hipMemcpyAsync(*p, src, sizeof(T) * count, hipMemcpyHostToDevice, s));
        SKM_HIP(hipStreamSynchronize(s));
    }
    return SKM_SUCCESS;
}

// A device buffer holding a copy of src (a host or, with device, a device
// array); the copy has completed on return (the evaluator's own stream does
// not wait for the null stream).  A device array is copied on the null
// stream, after the work that the caller queued there to produce it.
template <typename T>
skm_status dev_copy_in(T **p, const T *src, size_t count, bool device, hipStream_t s)
{
    if (!device) return dev_upload(p, src, count, s);
    SKM_TRY(dev_alloc(p, count));
    if (src && count) {
        SKM_HIP(hipMemcpy(*p, src, sizeof(T) * count, hipMemcpyDeviceToDevice));
        SKM_HIP(hipStreamSynchronize(nullptr));
    }
    return SKM_SUCCESS;
}

skm_status validate_grid(const skm_grid &g, int dim, const char *name, int min_size)
{
    for (int d = 0; d < dim; ++d) {
        if (g.size[d] < min_size) return fail(SKM_ERROR_INVALID_ARGUMENT, std::string(name) + ": size too small");
        if (!(g.spacing[d] > 0.0)) return fail(SKM_ERROR_INVALID_ARGUMENT, std::string(name) + ": spacing must be > 0");
    }
    return SKM_SUCCESS;
}

// Static per-control-point data: kernel state at p_j (getSigmaAtPoint /
// getCovarianceAtPoint and getRegionSupport, Kernel.h:148-171) and the three
// index windows the derivative passes use, all in FP64 as the reference
// computes them with ScalarType = double.
skm_status build_static(skm_evaluator *ev)
{
    const skm_desc &ds = ev->desc;
    const int dim = ev->dim;
    const long long ncp = grid_count(ds.cp, dim);
    const double smin = std::min(ds.sigma_c4, ds.sigma_c0);
    std::vector<float> state((size_t)ncp * kCpStateStride, 0.0f);
    std::vector<float> pstate((size_t)ncp * 8, 0.0f);  // pstate_pack layout, two float4 per cp
    double lam_cp_max = 0.0;
    std::vector<double> support(ncp, smin);
    std::vector<int> win((size_t)ncp * 6), lwin((size_t)ncp * 6), dwin((size_t)ncp * 6), rwin((size_t)ncp * 6);
    int ns[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d) ns[d] = ds.fixed.size[d] / ds.subsample;
    const int nn = dim * dim;
    // the weight value / tensor at every control point's nearest voxel
    // (ImageNearest, index = int(... + 0.5)), gathered on the device
    std::vector<float> cpw_vals;
    if (ds.kernel_mode != SKM_KERNEL_STATIONARY) {
        const int per = ds.kernel_mode == SKM_KERNEL_WEIGHT_IMAGE ? 1 : nn;
        std::vector<long long> idx((size_t)ncp);
        for (long long j = 0; j < ncp; ++j) {
            int jc[3] = {0, 0, 0};
            long long r = j;
            for (int d = 0; d < dim; ++d) { jc[d] = (int)(r % ds.cp.size[d]); r /= ds.cp.size[d]; }
            long long f = 0, stride = 1;
            bool in = true;
            for (int d = 0; d < dim; ++d) {
                const double p = ds.cp.origin[d] + ds.cp.spacing[d] * jc[d];
                const int i = (int)((1.0 / ds.weight.spacing[d]) * (p - ds.weight.origin[d]) + 0.5);
                in = in && i >= 0 && i < ds.weight.size[d];
                f += stride * i;
                stride *= ds.weight.size[d];
            }
            idx[(size_t)j] = in ? f : -1;
        }
        long long *d_idx = nullptr;
        float *d_vals = nullptr;
        skm_status st = dev_upload(&d_idx, idx.data(), idx.size(), ev->stream);
        if (st == SKM_SUCCESS) st = dev_alloc(&d_vals, (size_t)ncp * per);
        if (st == SKM_SUCCESS) {
            launch_gather_values(ev->d_weight, d_idx, per, ncp, d_vals, ev->stream);
            cpw_vals.resize((size_t)ncp * per);
            hipError_t e = hipMemcpyAsync(cpw_vals.data(), d_vals, sizeof(float) * cpw_vals.size(),
                                          hipMemcpyDeviceToHost, ev->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(ev->stream);
            if (e != hipSuccess) st = fail(SKM_ERROR_HIP, std::string("weight gather: ") + hipGetErrorString(e));
        }
        if (d_idx) device_release(d_idx);
        if (d_vals) device_release(d_vals);
        if (st != SKM_SUCCESS) return st;
    }
    for (long long j = 0; j < ncp; ++j) {
        int jc[3] = {0, 0, 0};
        long long r = j;
        for (int d = 0; d < dim; ++d) { jc[d] = (int)(r % ds.cp.size[d]); r /= ds.cp.size[d]; }
        double p[3] = {0, 0, 0};
        for (int d = 0; d < dim; ++d) p[d] = ds.cp.origin[d] + ds.cp.spacing[d] * jc[d];
        float *st = &state[(size_t)j * kCpStateStride];
        if (ds.kernel_mode != SKM_KERNEL_STATIONARY) {
            float *ps = &pstate[(size_t)j * 8];
            if (ds.kernel_mode == SKM_KERNEL_WEIGHT_IMAGE) {
                double s = cpw_vals[(size_t)j];
                st[0] = (float)s;
                st[1] = (float)std::sqrt(std::fabs(s));
                support[j] = smin / std::sqrt(2.0 / (ds.maximum_weight + s));
                ps[0] = st[0];
                ps[1] = st[1];
                ps[7] = st[0];
                lam_cp_max = std::max(lam_cp_max, (double)ps[7]);
            } else {
                double cov[9] = {0}, sum[9] = {0};
                for (int i = 0; i < nn; ++i) cov[i] = cpw_vals[(size_t)j * nn + i];
                for (int i = 0; i < nn; ++i) st[2 + i] = (float)cov[i];
                st[1] = (float)std::sqrt(det_h(dim, cov));
                for (int i = 0; i < nn; ++i) sum[i] = ds.maximum_eigenvalue_covariance[i] + cov[i];
                support[j] = smin / std::sqrt(2.0 / max_eig_h(dim, sum));
                double sym[9];
                symmetric_part(dim, cov, sym);
                const double lam = max_eig_h(dim, sym);
                st[0] = (float)lam;
                pack_upper(dim, sym, st[1], (float)lam, ps);
                lam_cp_max = std::max(lam_cp_max, lam);
            }
        }
        st[11] = (float)support[j];
        const bool exact = (ds.flags & SKM_FLAG_EXACT_GRADIENT) != 0;
        for (int d = 0; d < 3; ++d) {
            int *w = &win[(size_t)j * 6], *lw = &lwin[(size_t)j * 6], *dw = &dwin[(size_t)j * 6], *rw = &rwin[(size_t)j * 6];
            if (d >= dim) { w[d] = w[d + 3] = lw[d] = lw[d + 3] = dw[d] = dw[d + 3] = rw[d] = rw[d + 3] = 0; continue; }
            double idx = (1.0 / ds.fixed.spacing[d]) * (p[d] - ds.fixed.origin[d]);
            double sp = support[j] / ds.fixed.spacing[d];
            double sub = (double)ds.subsample;
            int lim = (ds.subsample > 1 ? ns[d] : ds.fixed.size[d]) - 1;
            int lo, hi;
            if (dim == 3 || exact) {  // GpuEvaluatorMSE.cu:126-132 (3D)
                lo = (int)((idx - sp) / sub);
                hi = (int)((idx + sp) / sub + 1.0);
            } else {                  // GpuEvaluatorMSE.cu:267-273 (2D)
                lo = (int)((idx - sp) / sub + 1.0);
                hi = (int)((idx + sp) / sub);
            }
            if (exact) { lo -= 1; hi += 1; }
            w[d] = std::min(std::max(lo, 0), lim);
            w[d + 3] = std::min(std::max(hi, 0), lim);
            // LCC centre window, tight in 2D and 3D (GpuEvaluatorLCC.cu:239-245, 724-730)
            lw[d] = std::min(std::max((int)((idx - sp) / sub + 1.0), 0), lim);
            lw[d + 3] = std::min(std::max((int)((idx + sp) / sub), 0), lim);
            // dense voxel window for the exact LCC gather
            dw[d] = std::min(std::max((int)(idx - sp) - 1, 0), ds.fixed.size[d] - 1);
            dw[d + 3] = std::min(std::max((int)(idx + sp + 1.0) + 1, 0), ds.fixed.size[d] - 1);
            // regularizer window on the cp grid (GpuEvaluatorRegularizers.cu:52-54)
            double csp = support[j] / ds.cp.spacing[d];
            rw[d] = std::min(std::max((int)((double)jc[d] - csp + 1.0), 0), ds.cp.size[d] - 1);
            rw[d + 3] = std::min(std::max((int)((double)jc[d] + csp), 0), ds.cp.size[d] - 1);
        }
    }
    SKM_TRY(dev_upload(&ev->d_cp_state, state.data(), state.size(), ev->stream));
    if (ds.kernel_mode != SKM_KERNEL_STATIONARY) {
        SKM_TRY(dev_upload(&ev->d_cp_pstate, reinterpret_cast<const float4 *>(pstate.data()), (size_t)ncp * 2,
                             ev->stream));
        // largest lam a sample can look up: every weight voxel, and the
        // outside value (sigma 0 / identity covariance)
        const long long nw = grid_count(ds.weight, dim);
        const double lam_out = ds.kernel_mode == SKM_KERNEL_WEIGHT_IMAGE ? 0.0 : 1.0;
        double lam_w = lam_out;
        // device scan of the weight voxels: validation (itkTransformAdapter.hxx:321-324,
        // 346-349) and lam per voxel
        float *wlam = nullptr;
        double *bmax = nullptr;
        unsigned *bbad = nullptr;
        const unsigned nb = weight_lam_blocks(nw);
        auto release = [&]() {
            (void)hipStreamSynchronize(ev->stream);  // error paths: the scan may still run
            if (wlam) device_release(wlam);
            if (bmax) device_release(bmax);
            if (bbad) device_release(bbad);
        };
        skm_status st;
        if ((st = dev_alloc(&wlam, (size_t)nw)) != SKM_SUCCESS || (st = dev_alloc(&bmax, nb)) != SKM_SUCCESS ||
            (st = dev_alloc(&bbad, nb)) != SKM_SUCCESS) {
            release();
            return st;
        }
        launch_weight_lam(dim, ds.kernel_mode, nw, ev->d_weight, wlam, bmax, bbad, ev->stream);
        std::vector<double> hmax(nb);
        std::vector<unsigned> hbad(nb);
        hipError_t e = hipMemcpyAsync(hmax.data(), bmax, sizeof(double) * nb, hipMemcpyDeviceToHost, ev->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(hbad.data(), bbad, sizeof(unsigned) * nb, hipMemcpyDeviceToHost, ev->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ev->stream);
        if (e != hipSuccess) {
            release();
            return fail(SKM_ERROR_HIP, std::string("weight scan: ") + hipGetErrorString(e));
        }
        unsigned long long bad = 0;
        for (unsigned b = 0; b < nb; ++b) {
            lam_w = std::max(lam_w, hmax[b]);
            bad += hbad[b];
        }
        if (bad) {
            release();
            return fail(SKM_ERROR_INVALID_ARGUMENT, ds.kernel_mode == SKM_KERNEL_WEIGHT_IMAGE
                                                        ? "Weighting image has to be strictly positive everywhere."
                                                        : "Weighting tensors must be positive definite.");
        }
        // rounded up so that the FP32 values stay upper bounds
        ev->pr.lam_cp_max = (float)(lam_cp_max * (1.0 + 1e-6));
        ev->pr.lam_w_max = (float)(lam_w * (1.0 + 1e-6));
        // lam bound per row of control points (cy, cz): the value pass's reach
        std::vector<float> cp_row((size_t)ds.cp.size[1] * (dim == 3 ? ds.cp.size[2] : 1), 0.0f);
        for (long long j = 0; j < ncp; ++j) {
            const long long row = j / ds.cp.size[0];
            cp_row[row] = std::max(cp_row[row], (float)(pstate[(size_t)j * 8 + 7] * (1.0 + 1e-6)));
        }
        SKM_TRY(dev_upload(&ev->d_cp_row_lam, cp_row.data(), cp_row.size(), ev->stream));
        // each control point's share of the pair reach, pair_reach2(smin, lam, 0) in float
        const float sminf = (float)std::min(ds.sigma_c4, ds.sigma_c0);
        std::vector<float> cp_reach((size_t)ncp);
        for (long long j = 0; j < ncp; ++j)
            cp_reach[(size_t)j] = sminf * sminf * 0.5f * pstate[(size_t)j * 8 + 7] * 1.0001f;
        SKM_TRY(dev_upload(&ev->d_cp_reach, cp_reach.data(), cp_reach.size(), ev->stream));
        // lam bound per row of sample cells (my, mz), jitter included, for the
        // gather's reach; once for the evaluator's subsample and once for the
        // dense voxel grid
        int cells[3] = {1, 1, 1}, vox[3] = {1, 1, 1};
        for (int d = 0; d < dim; ++d) {
            cells[d] = ds.fixed.size[d] / ds.subsample;
            vox[d] = ds.fixed.size[d];
        }
        st = dev_alloc(&ev->d_smp_row_lam, (size_t)cells[1] * cells[2]);
        if (st == SKM_SUCCESS && ds.subsample > 1) st = dev_alloc(&ev->d_vox_row_lam, (size_t)vox[1] * vox[2]);
        if (st != SKM_SUCCESS) {
            release();
            return st;
        }
        launch_sample_row_lams(dim, ev->pr.fixed, ev->pr.weight, ds.subsample, cells, wlam, (float)lam_out,
                               ev->d_smp_row_lam, ev->stream);
        // the dense voxel grid (exact LCC gather); the same table at subsample 1
        if (ds.subsample > 1)
            launch_sample_row_lams(dim, ev->pr.fixed, ev->pr.weight, 1, vox, wlam, (float)lam_out, ev->d_vox_row_lam,
                                   ev->stream);
        e = hipStreamSynchronize(ev->stream);
        release();
        if (e != hipSuccess) return fail(SKM_ERROR_HIP, std::string("row bounds: ") + hipGetErrorString(e));
    }
    SKM_TRY(dev_upload(&ev->d_cp_window, win.data(), win.size(), ev->stream));
    SKM_TRY(dev_upload(&ev->d_lcc_window, lwin.data(), lwin.size(), ev->stream));
    SKM_TRY(dev_upload(&ev->d_dense_window, dwin.data(), dwin.size(), ev->stream));
    SKM_TRY(dev_upload(&ev->d_reg_window, rwin.data(), rwin.size(), ev->stream));
    return SKM_SUCCESS;
}

skm_status ensure_lcc_buffers(skm_evaluator *ev)
{
    if (!ev->d_dense) SKM_TRY(dev_alloc(&ev->d_dense, ev->pr.n_vox));
    if (!ev->d_ccoef) SKM_TRY(dev_alloc(&ev->d_ccoef, ev->pr.n_sub));
    const long long np = lcc_reference_partials(ev->pr);
    if (np > 0 && !ev->d_lcc_ref_partials) {
        SKM_TRY(dev_alloc(&ev->d_lcc_ref_partials, (size_t)np));
        ev->lcc_ref_partials_n = np;
    }
    return SKM_SUCCESS;
}

// Largest kernel matrix an evaluator may assemble: SKM_KERNEL_MATRIX_MAX_BYTES,
// else a quarter of the device's memory or half of its free memory.
long long kernel_matrix_limit()
{
    if (const char *env = std::getenv("SKM_KERNEL_MATRIX_MAX_BYTES")) {
        char *end = nullptr;
        const long long v = std::strtoll(env, &end, 10);
        if (end && end != env && v >= 0) return v;
    }
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return 0;
    return (long long)std::min(total_b / 4, free_b / 2);
}

// Assembles the kernel matrix now (on the evaluator's stream, timed as pass 10).
skm_status assemble_kernel_matrix(skm_evaluator *ev)
{
    PassTimer t(ev, 10);
    return kernel_matrix_build(ev->pr, ev->km, ev->stream, kernel_matrix_limit());
}

// Whether this MSE / NCC evaluation reads the kernel matrix, assembling it
// first.  skm_tuning.kernel_matrix: 2 always, 1 never, 0 for the tensor kernel
// only (the stationary kernels cost less to evaluate than their stored values
// to stream: 0.69 against 0.87 ms per evaluation at 128x128x96).  None is
// assembled when it was refused, or while the stream is being captured into a
// graph (the assembly waits for its own counts; the captured evaluation then
// runs on the fly).  A matrix whose gather slab no longer matches the tuning
// is assembled again.
skm_status use_kernel_matrix(skm_evaluator *ev, bool *use)
{
    *use = false;
    const int policy = ev->tuning.kernel_matrix;
    if (policy == 1 || (policy == 0 && ev->pr.kc.mode != 2) || !kernel_matrix_applies(ev->pr) ||
        ev->km.state == 2)
        return SKM_SUCCESS;
    if (ev->km.state == 1 && ev->km.slab == kernel_matrix_slab(ev->pr)) {
        *use = true;
        return SKM_SUCCESS;
    }
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    SKM_HIP(hipStreamIsCapturing(ev->stream, &cap));
    if (cap != hipStreamCaptureStatusNone) return SKM_SUCCESS;
    SKM_TRY(assemble_kernel_matrix(ev));
    *use = ev->km.state == 1;
    return SKM_SUCCESS;
}

// The regularizers' second stream and its fork / join events (and, with a
// derivative, the buffer of their terms), made at the first use; false (and
// the regularizers run in sequence) while the evaluator's stream is being
// captured into a graph, or where they cannot be made.  (A capture holding
// the fork ended the process in the runtime during skm_optimize_batch's
// capture of three evaluators; captures keep the sequential order.)
bool side_stream(skm_evaluator *ev, bool with_derivative)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(ev->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return false;
    if (ev->side && ev->side_fork && ev->side_join && (!with_derivative || ev->d_reg_add)) return true;
    // made on the evaluator's device only (the pool hands out the current device's streams)
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != ev->stream_device) return false;
    if (with_derivative && !ev->d_reg_add &&
        dev_alloc(&ev->d_reg_add, (size_t)ev->pr.n_cp * ev->pr.dim) != SKM_SUCCESS) {
        ev->d_reg_add = nullptr;
        return false;
    }
    if (!ev->side && take_stream(&ev->side) != hipSuccess) ev->side = nullptr;
    if (!ev->side_fork && hipEventCreateWithFlags(&ev->side_fork, hipEventDisableTiming) != hipSuccess)
        ev->side_fork = nullptr;
    if (!ev->side_join && hipEventCreateWithFlags(&ev->side_join, hipEventDisableTiming) != hipSuccess)
        ev->side_join = nullptr;
    return ev->side && ev->side_fork && ev->side_join;
}

// One evaluation on the device: value (+ derivative) of the measure and the
// regularizers, everything left in d_sc / deriv (GpuEvaluator.cu:94-357, 361-714).
skm_status run_eval(skm_evaluator *ev, int measure, int do_resampling, float *deriv)
{
    Problem &pr = ev->pr;
    hipStream_t s = ev->stream;
    const bool regs = ev->reg_w[0] > 0 || ev->reg_w[1] > 0 || ev->reg_w[2] > 0 || ev->reg_w[3] > 0;
    bool regs_done = false;
    uint64_t drawn = 0;  // random numbers drawn for the offsets of this evaluation
    if (pr.sub > 1 && (!ev->evaluated_once || do_resampling)) {
        PassTimer t(ev, 0);
        launch_offsets(pr, ev->d_offsets, ev->d_rng, ev->rng_inc, s);
        drawn = (uint64_t)pr.n_sub * pr.dim;
        ev->evaluated_once = true;
    }
    int blocks = 0;
    if (measure == SKM_MEASURE_MSE || measure == SKM_MEASURE_NCC) {
        bool km = false;
        SKM_TRY(use_kernel_matrix(ev, &km));
        // The regularizers read the control points only: on the second stream,
        // beside the value pass, joined before the gather (whose finish adds
        // their derivative terms, in the same FP64 sum the regularizers would
        // make) or, without a derivative, at the end.  For evaluations of at
        // least kBesideSamples samples: below, the fork and join cost more than
        // they hide (12k-25k samples: 7-10 % slower per iteration); in
        // sequence while the passes are timed.
        constexpr long long kBesideSamples = 1ll << 18;
        Problem gp = pr;
        const bool beside = regs && !ev->timing && pr.n_sub >= kBesideSamples && side_stream(ev, deriv != nullptr);
        if (beside) {
            gp.reg_add = deriv ? ev->d_reg_add : nullptr;
            SKM_HIP(hipEventRecord(ev->side_fork, s));
            SKM_HIP(hipStreamWaitEvent(ev->side, ev->side_fork, 0));
            const BilateralArgs bl = make_bilateral_args(pr, ev->reg_w[3], ev->bl_sigma_s, ev->bl_sigma_i);
            launch_regularizers(gp, ev->reg_w, ev->pg_scaling, bl, deriv != nullptr, deriv, ev->d_reg_partials,
                                ev->d_sc, ev->side);
            SKM_HIP(hipEventRecord(ev->side_join, ev->side));
            regs_done = true;
        }
        {
            PassTimer t(ev, 1);
            if (km) launch_value_km(pr, ev->km, measure, ev->d_sg, ev->d_sb, ev->d_partials, &blocks, s);
            else launch_value(pr, measure, ev->d_sg, ev->d_sb, ev->d_partials, &blocks, s);
        }
        {
            PassTimer t(ev, 2);
            launch_reduce(pr, measure, ev->d_partials, blocks, ev->d_sc, s, ev->d_rng, drawn, ev->rng_inc, beside);
        }
        if (deriv) {
            { PassTimer t(ev, 3); launch_prepare(pr, measure, ev->d_sg, ev->d_sb, ev->d_sc, ev->d_sv, s); }
            if (beside) SKM_HIP(hipStreamWaitEvent(s, ev->side_join, 0));
            {
                PassTimer t(ev, 4);
                if (km) launch_gather_km(gp, ev->km, ev->d_sv, deriv, s);
                else launch_gather(gp, ev->d_sv, deriv, s);
            }
        } else if (beside) {
            SKM_HIP(hipStreamWaitEvent(s, ev->side_join, 0));
        }
    } else if (measure == SKM_MEASURE_LCC) {
        SKM_TRY(ensure_lcc_buffers(ev));
        { PassTimer t(ev, 6); launch_dense_warp(pr, 0, ev->d_dense, s); }
        const bool exact_deriv = deriv && pr.exact;
        // stationary exact mode gathers gamma per voxel; the scatter (weighted
        // modes) needs the centres' bounds for its fixed-point step
        const bool gamma_gather = exact_deriv && lcc_gamma_gather_applies(pr);
        const bool need_bound = exact_deriv && !gamma_gather;
        // reference mode: the voxel-table form of the windowed derivative
        const bool ref_table = deriv && !pr.exact && lcc_ref_table_applies(pr);
        const bool need_cpk = gamma_gather || ref_table;
        if (need_bound && !ev->d_lcc_bound)
            SKM_TRY(dev_alloc(&ev->d_lcc_bound, (size_t)((pr.n_sub + kThreads - 1) / kThreads)));
        if (need_cpk && !ev->d_lcc_cpk) SKM_TRY(dev_alloc(&ev->d_lcc_cpk, (size_t)pr.n_sub));
        if (ref_table && !ev->d_lcc_reft_partials)
            SKM_TRY(dev_alloc(&ev->d_lcc_reft_partials, (size_t)lcc_ref_table_partials(pr)));
        {
            PassTimer t(ev, 7);
            launch_lcc_centres(pr, ev->d_dense, ev->d_ccoef, ev->d_partials, &blocks, s,
                               need_bound ? ev->d_lcc_bound : nullptr, need_cpk ? ev->d_lcc_cpk : nullptr);
        }
        {
            PassTimer t(ev, 2);
            launch_reduce(pr, measure, ev->d_partials, blocks, ev->d_sc, s, ev->d_rng, drawn, ev->rng_inc);
        }
        if (deriv) {
            PassTimer t(ev, 8);
            if (pr.exact) {
                if (!ev->d_gamma) SKM_TRY(dev_alloc(&ev->d_gamma, pr.n_vox));
                if (gamma_gather) {
                    launch_lcc_gamma(pr, ev->d_dense, ev->d_ccoef, ev->d_lcc_cpk, ev->d_sc, ev->d_gamma, s);
                } else {
                    if (!ev->d_gamma_fx) SKM_TRY(dev_alloc(&ev->d_gamma_fx, (size_t)pr.n_vox * 3));
                    launch_zero_words(ev->d_gamma_fx, 6LL * pr.n_vox, s);
                    SKM_HIP(hipGetLastError());
                    launch_lcc_scatter(pr, ev->d_dense, ev->d_ccoef, ev->d_sc, ev->d_lcc_bound, blocks,
                                       ev->d_gamma_fx, ev->d_gamma, s);
                }
                Problem dp = pr;  // gather over the dense voxel grid
                if (!is_stationary(pr.kc.mode)) {
                    if (!ev->d_vox_pstate) {
                        SKM_TRY(dev_alloc(&ev->d_vox_pstate, (size_t)pr.n_vox * 2));
                        launch_voxel_states(pr, ev->d_vox_pstate, s);
                    }
                    dp.smp_pstate = ev->d_vox_pstate;
                    if (ev->d_vox_row_lam) dp.smp_row_lam = ev->d_vox_row_lam;
                }
                dp.sub = 1;
                for (int d = 0; d < 3; ++d) dp.nsub[d] = pr.fixed.n[d];
                dp.n_sub = pr.n_vox;
                dp.cp_window = ev->d_dense_window;
                launch_gather(dp, ev->d_gamma, deriv, s);
            } else {
                const float4 *dense = ev->d_dense;
                Problem lp = pr;
                lp.cp_window = ev->d_lcc_window;
                if (pr.dim == 3 && !is_stationary(pr.kc.mode)) {
                    // the 3D derivative kernel warps neighbours with sigma(y) twice
                    if (!ev->d_dense_q) SKM_TRY(dev_alloc(&ev->d_dense_q, pr.n_vox));
                    launch_dense_warp(pr, 1, ev->d_dense_q, s);
                    int b2 = 0;
                    launch_lcc_centres(pr, ev->d_dense_q, ev->d_ccoef, ev->d_partials + (size_t)blocks * kPartialsK,
                                       &b2, s);
                    dense = ev->d_dense_q;
                }
                if (ref_table)
                    launch_lcc_ref_table(lp, dense, ev->d_ccoef, ev->d_sc, deriv, ev->d_lcc_reft_partials, s);
                else
                    launch_lcc_gather_reference(lp, dense, ev->d_ccoef, ev->d_sc, deriv, ev->d_lcc_ref_partials, s);
            }
        }
    } else {
        return fail(SKM_ERROR_INVALID_ARGUMENT, "unknown measure");
    }
    if (regs && !regs_done) {
        PassTimer t(ev, 5);
        const BilateralArgs bl = make_bilateral_args(pr, ev->reg_w[3], ev->bl_sigma_s, ev->bl_sigma_i);
        launch_regularizers(pr, ev->reg_w, ev->pg_scaling, bl, deriv != nullptr, deriv, ev->d_reg_partials, ev->d_sc,
                            s);
    }
    SKM_HIP(hipGetLastError());
    return SKM_SUCCESS;
}

skm_status upload_cp(skm_evaluator *ev, const float *cp_aos, const float *cpw)
{
    if (!cp_aos || !cpw) return fail(SKM_ERROR_INVALID_ARGUMENT, "null control-point data");
    size_t n = (size_t)ev->pr.n_cp;
    SKM_HIP(hipMemcpyAsync(ev->d_cp_in, cp_aos, sizeof(float) * n * ev->dim, hipMemcpyHostToDevice, ev->stream));
    SKM_HIP(hipMemcpyAsync(ev->d_cpw_in, cpw, sizeof(float) * n, hipMemcpyHostToDevice, ev->stream));
    launch_pack_cp(ev->pr, ev->d_cp_in, ev->d_cpw_in, 0, ev->d_cpc, ev->stream);
    return SKM_SUCCESS;
}

skm_status read_result(skm_evaluator *ev, skm_result *out)
{
    Scalars sc;
    SKM_HIP(hipMemcpyAsync(&sc, ev->d_sc, sizeof(Scalars), hipMemcpyDeviceToHost, ev->stream));
    SKM_HIP(hipStreamSynchronize(ev->stream));
    if (out) {
        out->measure = sc.measure;
        out->reg_value_rkhs = sc.reg[0];
        out->reg_value_rd = sc.reg[1];
        out->reg_value_pg = sc.reg[2];
        out->num_pixels_counted = (uint32_t)sc.count;
        out->reserved = 0;
        out->reg_value_bilateral = sc.reg[3];
    }
    return SKM_SUCCESS;
}

// ---- least-squares fit of a dense displacement onto the kernel basis

__global__ __launch_bounds__(kThreads) void k_aos_to_float4(const float *__restrict__ u, int dim, long long n,
                                                            float4 *__restrict__ out)
{
    const long long i = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (i >= n) return;
    out[i] = make_float4(u[i * dim], u[i * dim + 1], dim == 3 ? u[i * dim + 2] : 0.0f, 0.0f);
}

// FP64 partial sums of a.b, one per workgroup
__global__ __launch_bounds__(kThreads) void k_dot(const float *__restrict__ a, const float *__restrict__ b,
                                                  long long n, double *__restrict__ part)
{
    __shared__ double lds[kThreads / 64];
    double s = 0.0;
    for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads)
        s += (double)a[i] * (double)b[i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kThreads / 64; ++w) t += lds[w];
        part[blockIdx.x] = t;
    }
}

// out = a x + b y
__global__ __launch_bounds__(kThreads) void k_lincomb(float *__restrict__ out, double a, const float *__restrict__ x,
                                                      double b, const float *__restrict__ y, long long n)
{
    const long long i = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (i < n) out[i] = (float)(a * (double)x[i] + b * (double)y[i]);
}

constexpr int kDotBlocks = 1024;

skm_status dot(const float *a, const float *b, long long n, double *part, hipStream_t s, double *out)
{
    k_dot<<<kDotBlocks, kThreads, 0, s>>>(a, b, n, part);
    std::vector<double> h(kDotBlocks);
    SKM_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * kDotBlocks, hipMemcpyDeviceToHost, s));
    SKM_HIP(hipStreamSynchronize(s));
    double t = 0.0;
    for (double v : h) t += v;
    *out = t;
    return SKM_SUCCESS;
}

unsigned grid_of(long long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

}  // namespace

// CGLS for min |K c - u|^2 + ridge |c|^2, with K the evaluator's kernel sum
// (control-point weights 1, no previous displacement) from SoA coefficients
// to the dense displacement on the fixed grid.  K^T is the derivative gather
// over the dense voxel grid.  Deterministic: fixed-order reductions.
skm_status skm::fit_coefficients_device(skm_evaluator *ev, const float *d_u_aos, int iterations, double ridge,
                                        float *d_c_soa)
{
    if (!ev || !d_u_aos || !d_c_soa) return fail(SKM_ERROR_INVALID_ARGUMENT, "null argument");
    if (iterations < 0 || !(ridge >= 0.0)) return fail(SKM_ERROR_INVALID_ARGUMENT, "invalid fit settings");
    if (!is_stationary(ev->pr.kc.mode))
        return fail(SKM_ERROR_NOT_SUPPORTED, "the coefficient fit needs a stationary kernel (exact adjoint)");
    const int dim = ev->dim;
    hipStream_t s = ev->stream;
    Problem pk = ev->pr;     // K: kernel sum only
    pk.has_df = 0;
    Problem pt = ev->pr;     // K^T: gather over the dense voxel grid
    if (!is_stationary(pt.kc.mode)) {
        if (!ev->d_vox_pstate) {
            SKM_TRY(dev_alloc(&ev->d_vox_pstate, (size_t)pt.n_vox * 2));
            launch_voxel_states(pt, ev->d_vox_pstate, s);
        }
        pt.smp_pstate = ev->d_vox_pstate;
        if (ev->d_vox_row_lam) pt.smp_row_lam = ev->d_vox_row_lam;
    }
    pt.sub = 1;
    for (int d = 0; d < 3; ++d) pt.nsub[d] = pt.fixed.n[d];
    pt.n_sub = pt.n_vox;
    pt.cp_window = ev->d_dense_window;
    const long long nc = pt.n_cp * dim, nv = pt.n_vox * dim;
    struct Tmp {
        float *r = nullptr, *q = nullptr, *p = nullptr, *sv = nullptr, *ones = nullptr;
        float4 *r4 = nullptr;
        double *part = nullptr;
        ~Tmp()
        {
            void *ptrs[] = {r, q, p, sv, ones, r4, part};
            for (void *x : ptrs)
                if (x) device_release(x);
        }
    } t;
    SKM_TRY(dev_alloc(&t.r, (size_t)nv));
    SKM_TRY(dev_alloc(&t.q, (size_t)nv));
    SKM_TRY(dev_alloc(&t.p, (size_t)nc));
    SKM_TRY(dev_alloc(&t.sv, (size_t)nc));
    SKM_TRY(dev_alloc(&t.ones, (size_t)pt.n_cp));
    SKM_TRY(dev_alloc(&t.r4, (size_t)pt.n_vox));
    SKM_TRY(dev_alloc(&t.part, (size_t)kDotBlocks));
    // declared after the temporaries: waits for the stream before they are released, on every return
    struct SyncOnExit {
        hipStream_t s;
        ~SyncOnExit() { (void)hipStreamSynchronize(s); }
    } sync_before_release{s};
    std::vector<float> one((size_t)pt.n_cp, 1.0f);
    SKM_HIP(hipMemcpyAsync(t.ones, one.data(), sizeof(float) * one.size(), hipMemcpyHostToDevice, s));
    auto apply_kt = [&](const float *r, float *out) {
        k_aos_to_float4<<<grid_of(pt.n_vox), kThreads, 0, s>>>(r, dim, pt.n_vox, t.r4);
        launch_gather(pt, t.r4, out, s);
    };
    auto apply_k = [&](const float *c, float *out) {
        launch_soa_to_cpc(pk, c, t.ones, ev->d_cpc, s);
        launch_displacement(pk, out, s);
    };
    // c = 0, r = u, s = K^T r, p = s; the packed control points carry the
    // weights (1) that the transpose multiplies by
    SKM_HIP(hipMemsetAsync(d_c_soa, 0, sizeof(float) * nc, s));
    launch_soa_to_cpc(pk, d_c_soa, t.ones, ev->d_cpc, s);
    SKM_HIP(hipMemcpyAsync(t.r, d_u_aos, sizeof(float) * nv, hipMemcpyDeviceToDevice, s));
    apply_kt(t.r, t.sv);
    SKM_HIP(hipMemcpyAsync(t.p, t.sv, sizeof(float) * nc, hipMemcpyDeviceToDevice, s));
    double gamma = 0.0;
    SKM_TRY(dot(t.sv, t.sv, nc, t.part, s, &gamma));
    for (int it = 0; it < iterations && gamma > 0.0; ++it) {
        apply_k(t.p, t.q);
        double qq = 0.0, pp = 0.0;
        SKM_TRY(dot(t.q, t.q, nv, t.part, s, &qq));
        SKM_TRY(dot(t.p, t.p, nc, t.part, s, &pp));
        const double delta = qq + ridge * pp;
        if (!(delta > 0.0)) break;
        const double alpha = gamma / delta;
        k_lincomb<<<grid_of(nc), kThreads, 0, s>>>(d_c_soa, 1.0, d_c_soa, alpha, t.p, nc);
        k_lincomb<<<grid_of(nv), kThreads, 0, s>>>(t.r, 1.0, t.r, -alpha, t.q, nv);
        apply_kt(t.r, t.sv);
        k_lincomb<<<grid_of(nc), kThreads, 0, s>>>(t.sv, 1.0, t.sv, -ridge, d_c_soa, nc);
        double g2 = 0.0;
        SKM_TRY(dot(t.sv, t.sv, nc, t.part, s, &g2));
        const double beta = g2 / gamma;
        gamma = g2;
        k_lincomb<<<grid_of(nc), kThreads, 0, s>>>(t.p, 1.0, t.sv, beta, t.p, nc);
    }
    SKM_HIP(hipGetLastError());
    SKM_HIP(hipStreamSynchronize(s));
    return SKM_SUCCESS;
}

skm_status skm::set_last_error(skm_status st, const std::string &msg) { return fail(st, msg); }

// ---------------------------------------------------------------- C ABI

extern "C" {

SKM_API void skm_desc_init(skm_desc *desc)
{
    if (!desc) return;
    std::memset(desc, 0, sizeof(*desc));
    desc->dim = 3;
    for (int d = 0; d < 3; ++d) {
        desc->fixed.spacing[d] = desc->moving.spacing[d] = desc->cp.spacing[d] = desc->weight.spacing[d] = 1.0;
        desc->fixed.size[d] = desc->moving.size[d] = desc->cp.size[d] = desc->weight.size[d] = 1;
    }
    desc->sigma_c0 = desc->sigma_c4 = 1.0;
    desc->kernel_mode = SKM_KERNEL_STATIONARY;
    desc->maximum_weight = 1.0;
    for (int i = 0; i < 3; ++i) desc->maximum_eigenvalue_covariance[i * 3 + i] = 1.0;
    desc->subsample = 1;
    desc->subsample_neighborhood = 1;
    desc->rng_state = 0xBA5EBA11DEADBEEFULL;
    desc->rng_sequence = 0;
}

SKM_API const char *skm_last_error(void) { return g_error.c_str(); }

SKM_API skm_status skm_create(const skm_desc *desc, skm_evaluator **out)
{
    return skm::create_evaluator(desc, false, out);
}

}  // extern "C"

// skm_create; with device_data the image, weight and tensor pointers of the
// descriptor are device pointers (copied on the device, no PCIe transfer).
// ---- launch-shape table (skm_autotune results), per shape class

namespace {

std::mutex g_tuning_mutex;
std::map<std::string, skm_tuning> g_tuning;

// what the launch shapes depend on: dimension, kernel mode, windows, sampling,
// image and control-point grids
std::string shape_class(const skm_evaluator *ev)
{
    const Problem &pr = ev->pr;
    std::string k = std::to_string(pr.dim) + ":" + std::to_string(pr.kc.mode) + ":" + std::to_string(pr.exact) + ":" +
                    std::to_string(pr.sub) + ":" + std::to_string(pr.subn);
    for (int d = 0; d < pr.dim; ++d) k += ":" + std::to_string(pr.fixed.n[d]);
    for (int d = 0; d < pr.dim; ++d) k += ":" + std::to_string(pr.cp.n[d]);
    return k;
}

bool tuning_lookup(const std::string &key, skm_tuning *t)
{
    std::lock_guard<std::mutex> lock(g_tuning_mutex);
    auto it = g_tuning.find(key);
    if (it == g_tuning.end()) return false;
    *t = it->second;
    return true;
}

void tuning_store(const std::string &key, const skm_tuning &t)
{
    std::lock_guard<std::mutex> lock(g_tuning_mutex);
    g_tuning[key] = t;
}

}  // namespace

bool skm::tuning_recorded(const skm_evaluator *ev)
{
    skm_tuning t;
    return tuning_lookup(shape_class(ev), &t);
}

skm_status skm::create_evaluator(const skm_desc *desc, bool device_data, skm_evaluator **out, hipStream_t stream)
{
    if (!desc || !out) return fail(SKM_ERROR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    const int dim = desc->dim;
    if (dim != 2 && dim != 3) return fail(SKM_ERROR_INVALID_ARGUMENT, "dim must be 2 or 3");
    SKM_TRY(validate_grid(desc->fixed, dim, "fixed", 1));
    SKM_TRY(validate_grid(desc->moving, dim, "moving", 3));
    SKM_TRY(validate_grid(desc->cp, dim, "cp", 1));
    if (!desc->fixed_data || !desc->moving_data) return fail(SKM_ERROR_INVALID_ARGUMENT, "image data missing");
    if (desc->flags & ~(SKM_FLAG_EXACT_GRADIENT | SKM_FLAG_MOVING_COEFFICIENTS))
        return fail(SKM_ERROR_INVALID_ARGUMENT, "unknown flags");
    if (!(desc->sigma_c0 > 0) || !(desc->sigma_c4 > 0))
        return fail(SKM_ERROR_INVALID_ARGUMENT, "Kernel parameters have to be positive.");
    if (desc->subsample < 1 || desc->subsample > 100)
        return fail(SKM_ERROR_INVALID_ARGUMENT, "Sampling rate inproper.");
    if (desc->subsample_neighborhood < 1 || desc->subsample_neighborhood > 100)
        return fail(SKM_ERROR_INVALID_ARGUMENT, "Neighborhood sampling rate inproper.");
    if (desc->kernel_mode < 0 || desc->kernel_mode > 2) return fail(SKM_ERROR_INVALID_ARGUMENT, "bad kernel mode");
    for (int d = 0; d < dim; ++d)
        if (desc->fixed.size[d] / desc->subsample < 1)
            return fail(SKM_ERROR_INVALID_ARGUMENT, "subsample larger than the fixed image");
    const long long nw = grid_count(desc->weight, dim);
    if (desc->kernel_mode != SKM_KERNEL_STATIONARY) {
        SKM_TRY(validate_grid(desc->weight, dim, "weight", 1));
        if (!desc->weight_data) return fail(SKM_ERROR_INVALID_ARGUMENT, "Weight image not provided.");
        // the values are validated on the device (build_static)
    }
    if (desc->moving.size[0] > 0xffff || desc->subsample > 255)
        return fail(SKM_ERROR_NOT_SUPPORTED, "image or subsample too large");

    skm_evaluator *ev = new skm_evaluator();
    ev->desc = *desc;
    ev->desc.fixed_data = ev->desc.moving_data = ev->desc.weight_data = nullptr;
    // The descriptor holds the maximum-eigenvalue covariance as 3x3 row-major
    // (a 2D kernel uses the upper-left block); internally it is packed D*D.
    double mec[9] = {0};
    for (int i = 0; i < dim; ++i)
        for (int k = 0; k < dim; ++k) mec[i * dim + k] = desc->maximum_eigenvalue_covariance[i * 3 + k];
    std::memcpy(ev->desc.maximum_eigenvalue_covariance, mec, sizeof(mec));
    ev->dim = dim;
    ev->rng_inc = desc->rng_sequence | 1ULL;
    auto cleanup = [&](skm_status st) {
        delete ev;
        return st;
    };
    skm_status st;
    if (hipGetDevice(&ev->stream_device) != hipSuccess) return cleanup(fail(SKM_ERROR_HIP, "hipGetDevice failed"));
    if (stream) {
        ev->stream = stream;  // the caller's (not returned to the pool)
    } else {
        if (take_stream(&ev->stream) != hipSuccess) return cleanup(fail(SKM_ERROR_HIP, "hipStreamCreate failed"));
        ev->own_stream = true;
    }

    Problem &pr = ev->pr;
    pr.dim = dim;
    pr.fixed = to_grid(desc->fixed, dim);
    pr.moving = to_grid(desc->moving, dim);
    pr.cp = to_grid(desc->cp, dim);
    pr.weight = to_grid(desc->weight, dim);
    pr.df = to_grid(desc->fixed, dim);
    pr.kc.mode = desc->kernel_mode;
    if (desc->kernel_mode == SKM_KERNEL_STATIONARY && desc->sigma_c0 == desc->sigma_c4)
        pr.kc.mode = kModeStationaryEqual;  // C4 * C0 share one (1 - t) factor
    pr.kc.inv_s4 = (float)(1.0 / desc->sigma_c4);
    pr.kc.inv_s0 = (float)(1.0 / desc->sigma_c0);
    pr.kc.smin = (float)std::min(desc->sigma_c4, desc->sigma_c0);
    pr.kc.max_weight = (float)desc->maximum_weight;
    for (int i = 0; i < 9; ++i) pr.kc.max_eig_cov[i] = (float)ev->desc.maximum_eigenvalue_covariance[i];
    pr.sub = desc->subsample;
    pr.subn = desc->subsample_neighborhood;
    for (int d = 0; d < 3; ++d) pr.nsub[d] = d < dim ? desc->fixed.size[d] / desc->subsample : 1;
    pr.n_sub = (long long)pr.nsub[0] * pr.nsub[1] * pr.nsub[2];
    pr.n_cp = grid_count(desc->cp, dim);
    pr.n_vox = grid_count(desc->fixed, dim);
    pr.exact = (desc->flags & SKM_FLAG_EXACT_GRADIENT) ? 1 : 0;

    // images
    int msize[3] = {desc->moving.size[0], desc->moving.size[1], dim == 3 ? desc->moving.size[2] : 1};
    if (desc->flags & SKM_FLAG_MOVING_COEFFICIENTS) {
        // the caller's prefiltered, padded coefficients (BsplineImage's layout)
        const long long np = (long long)(msize[0] + 3) * (msize[1] + 3) * (dim == 3 ? msize[2] + 3 : 1);
        if ((st = dev_copy_in(&ev->d_coeffs, desc->moving_data, (size_t)np, device_data, ev->stream)) != SKM_SUCCESS)
            return cleanup(st);
    } else {
        // cubic B-spline coefficients of the moving image, prefiltered on the device in FP64
        const long long nm = (long long)msize[0] * msize[1] * msize[2];
        const long long np = (long long)(msize[0] + 3) * (msize[1] + 3) * (dim == 3 ? msize[2] + 3 : 1);
        float *raw = nullptr;
        double *work = nullptr;
        if ((!device_data && (st = dev_upload(&raw, desc->moving_data, (size_t)nm, ev->stream)) != SKM_SUCCESS) ||
            (st = dev_alloc(&work, (size_t)nm)) != SKM_SUCCESS || (st = dev_alloc(&ev->d_coeffs, (size_t)np)) != SKM_SUCCESS) {
            if (raw) device_release(raw);
            if (work) device_release(work);
            return cleanup(st);
        }
        launch_bspline_coefficients(dim, msize, device_data ? desc->moving_data : raw, work, ev->d_coeffs, ev->stream);
        const hipError_t e = hipStreamSynchronize(ev->stream);
        device_release(raw);
        device_release(work);
        if (e != hipSuccess) return cleanup(fail(SKM_ERROR_HIP, std::string("prefilter: ") + hipGetErrorString(e)));
    }
    if ((st = dev_copy_in(&ev->d_fixed, desc->fixed_data, (size_t)pr.n_vox, device_data, ev->stream)) != SKM_SUCCESS)
        return cleanup(st);
    if (desc->kernel_mode != SKM_KERNEL_STATIONARY) {
        size_t per = desc->kernel_mode == SKM_KERNEL_WEIGHT_IMAGE ? 1 : (size_t)dim * dim;
        if ((st = dev_copy_in(&ev->d_weight, desc->weight_data, (size_t)nw * per, device_data, ev->stream)) != SKM_SUCCESS)
            return cleanup(st);
    }
    if ((st = build_static(ev)) != SKM_SUCCESS) return cleanup(st);

    // per-sample and per-cp buffers
    int tiles = 1;
    {
        int tx = dim == 3 ? 8 : 16, ty = dim == 3 ? 8 : 16, tz = dim == 3 ? 4 : 1;
        long long a = (long long)((pr.nsub[0] + tx - 1) / tx) * ((pr.nsub[1] + ty - 1) / ty) * ((pr.nsub[2] + tz - 1) / tz);
        long long b = (pr.n_sub + kThreads - 1) / kThreads;
        // the value pass writes one partial per block tile, or per wave brick
        // (four per tile) when it splits control-point rows over the waves
        tiles = (int)std::max(4 * a, b) * 2 + 8;
    }
    ev->max_partial_blocks = tiles;
    if ((st = dev_alloc(&ev->d_cpc, pr.n_cp)) || (st = dev_alloc(&ev->d_cp_in, pr.n_cp * dim)) ||
        (st = dev_alloc(&ev->d_cpw_in, pr.n_cp)) || (st = dev_alloc(&ev->d_deriv, pr.n_cp * dim)) ||
        (st = dev_alloc(&ev->d_sg, pr.n_sub)) || (st = dev_alloc(&ev->d_sb, pr.n_sub)) ||
        (st = dev_alloc(&ev->d_sv, pr.n_sub)) || (st = dev_alloc(&ev->d_partials, (size_t)tiles * kPartialsK)) ||
        (st = dev_alloc(&ev->d_reg_partials, (size_t)reg_blocks(pr.n_cp) * 4 + 4)) ||
        (st = dev_alloc(&ev->d_hreg_partials, (size_t)((pr.n_cp + kThreads - 1) / kThreads) * 4 + 4)) ||
        (st = dev_alloc(&ev->d_sc, 1)) ||
        (st = dev_alloc(&ev->d_offsets, pr.sub > 1 ? pr.n_sub * dim : 1)) ||
        (st = dev_upload(&ev->d_rng, &desc->rng_state, 1, ev->stream)))
        return cleanup(st);
    pr.fixed_data = ev->d_fixed;
    pr.coeffs = ev->d_coeffs;
    pr.weight_data = ev->d_weight;
    pr.df_data = nullptr;
    pr.has_df = 0;
    pr.cpc = ev->d_cpc;
    pr.cp_state = ev->d_cp_state;
    pr.cp_pstate = ev->d_cp_pstate;
    if (desc->kernel_mode != SKM_KERNEL_STATIONARY)
        if ((st = dev_alloc(&ev->d_smp_pstate, (size_t)pr.n_sub * 2)) != SKM_SUCCESS) return cleanup(st);
    {
        // sized for the dense voxel grid, which the exact LCC gather walks, and
        // for the thinnest slab skm_set_tuning can select
        Problem thin = pr;
        thin.gather_split = 4;
        const long long np = gather_partials_size(thin, pr.fixed.n[dim - 1]);
        if ((st = dev_alloc(&ev->d_gather_partials, (size_t)np)) != SKM_SUCCESS) return cleanup(st);
    }
    pr.gather_partials = ev->d_gather_partials;
    pr.smp_pstate = ev->d_smp_pstate;
    pr.cp_row_lam = ev->d_cp_row_lam;
    pr.cp_reach = ev->d_cp_reach;
    pr.smp_row_lam = ev->d_smp_row_lam;
    pr.cp_window = ev->d_cp_window;
    pr.reg_window = ev->d_reg_window;
    pr.offsets = ev->d_offsets;
    // on the evaluator's stream, complete on return: a memset on the null
    // stream is not ordered before the stream's kernels, and could land
    // after the first draw of the offsets
    if (hipMemsetAsync(ev->d_offsets, 0, sizeof(int) * (pr.sub > 1 ? pr.n_sub * dim : 1), ev->stream) != hipSuccess ||
        hipStreamSynchronize(ev->stream) != hipSuccess)
        return cleanup(fail(SKM_ERROR_HIP, "clearing the offsets failed"));
    // launch shapes tuned for this shape class earlier in the process (or loaded)
    skm_tuning t;
    if (tuning_lookup(shape_class(ev), &t)) {
        ev->tuning = t;
        pr.lcc_chunks = t.lcc_ref_chunks;
        pr.gather_split = t.gather_split;
        pr.value_split = t.value_split;
        pr.gather_regions = t.gather_regions;
        pr.lcc_ref_walk = t.lcc_ref_walk;
    }
    *out = ev;
    return SKM_SUCCESS;
}

extern "C" {

SKM_API void skm_destroy(skm_evaluator *ev)
{
    if (!ev) return;
    if (ev->stream) (void)hipStreamSynchronize(ev->stream);
    delete ev;
}

// SetRegularizerRKHS / RD / PG (GpuEvaluator.cu:786-831).  The scaling of RD
// is stored but, as in the reference, not used by the RD term (:813).
SKM_API skm_status skm_set_regularizer_rkhs(skm_evaluator *ev, double weight)
{
    if (!ev) return fail(SKM_ERROR_INVALID_ARGUMENT, "null evaluator");
    ev->reg_w[0] = weight;
    return SKM_SUCCESS;
}

SKM_API skm_status skm_set_regularizer_rd(skm_evaluator *ev, double weight, double scaling)
{
    if (!ev) return fail(SKM_ERROR_INVALID_ARGUMENT, "null evaluator");
    if (!(scaling > 0.0)) return fail(SKM_ERROR_INVALID_ARGUMENT, "scaling of regularizer must be strictly positive");
    ev->reg_w[1] = weight;
    ev->rd_scaling = scaling;
    return SKM_SUCCESS;
}

SKM_API skm_status skm_set_regularizer_pg(skm_evaluator *ev, double weight, double scaling)
{
    if (!ev) return fail(SKM_ERROR_INVALID_ARGUMENT, "null evaluator");
    if (!(scaling > 0.0)) return fail(SKM_ERROR_INVALID_ARGUMENT, "scaling of regularizer must be strictly positive");
    ev->reg_w[2] = weight;
    ev->pg_scaling = scaling;
    return SKM_SUCCESS;
}

SKM_API skm_status skm_set_regularizer_bilateral(skm_evaluator *ev, double weight, double spatial_sigma,
                                                double intensity_sigma)
{
    if (!ev) return fail(SKM_ERROR_INVALID_ARGUMENT, "null evaluator");
    if (!(spatial_sigma > 0.0) || !(intensity_sigma > 0.0))
        return fail(SKM_ERROR_INVALID_ARGUMENT, "bilateral sigmas must be strictly positive");
    if (weight > 0 && !ev->d_cp_intensity) {
        // the fixed image at every control point's nearest voxel (ImageNearest, 0 outside)
        const int dim = ev->dim;
        const skm_grid &cg = ev->desc.cp, &fg = ev->desc.fixed;
        const long long ncp = ev->pr.n_cp;
        std::vector<long long> idx((size_t)ncp);
        for (long long j = 0; j < ncp; ++j) {
            long long r = j, f = 0, stride = 1;
            bool in = true;
            for (int d = 0; d < dim; ++d) {
                const int jc = (int)(r % cg.size[d]);
                r /= cg.size[d];
                const double p = cg.origin[d] + cg.spacing[d] * jc;
                const int i = (int)((1.0 / fg.spacing[d]) * (p - fg.origin[d]) + 0.5);
                in = in && i >= 0 && i < fg.size[d];
                f += stride * i;
                stride *= fg.size[d];
            }
            idx[(size_t)j] = in ? f : -1;
        }
        long long *d_idx = nullptr;
        skm_status st = dev_upload(&d_idx, idx.data(), idx.size(), ev->stream);
        if (st == SKM_SUCCESS) st = dev_alloc(&ev->d_cp_intensity, (size_t)ncp);
        if (st == SKM_SUCCESS) {
            launch_gather_values(ev->d_fixed, d_idx, 1, ncp, ev->d_cp_intensity, ev->stream);
            const hipError_t e = hipStreamSynchronize(ev->stream);
            if (e != hipSuccess) st = fail(SKM_ERROR_HIP, std::string("intensity gather: ") + hipGetErrorString(e));
        }
        if (d_idx) device_release(d_idx);
        if (st != SKM_SUCCESS) return st;
        ev->pr.cp_intensity = ev->d_cp_intensity;
    }
    ev->reg_w[3] = weight;
    ev->bl_sigma_s = spatial_sigma;
    ev->bl_sigma_i = intensity_sigma;
    return SKM_SUCCESS;
}

SKM_API skm_status skm_set_displacement_field(skm_evaluator *ev, const skm_grid *grid, const float *df_aos)
{
    return skm::set_displacement_field_impl(ev, grid, df_aos, false);
}

}  // extern "C"

skm_status skm::set_displacement_field_device(skm_evaluator *ev, const skm_grid *grid, const float *d_df_aos)
{
    return set_displacement_field_impl(ev, grid, d_df_aos, true);
}

skm_status skm::set_displacement_field_impl(skm_evaluator *ev, const skm_grid *grid, const float *df_aos, bool device)
{
    if (!ev) return fail(SKM_ERROR_INVALID_ARGUMENT, "null evaluator");
    SKM_HIP(hipStreamSynchronize(ev->stream));
    if (ev->d_df) {
        device_release(ev->d_df);
        ev->d_df = nullptr;
    }
    ev->pr.has_df = 0;
    ev->pr.df_data = nullptr;
    if (!grid || !df_aos) return SKM_SUCCESS;
    SKM_TRY(validate_grid(*grid, ev->dim, "displacement field", 1));
    size_t n = (size_t)grid_count(*grid, ev->dim) * ev->dim;
    SKM_TRY(dev_copy_in(&ev->d_df, df_aos, n, device, ev->stream));
    ev->pr.df = to_grid(*grid, ev->dim);
    ev->pr.df_data = ev->d_df;
    ev->pr.has_df = 1;
    return SKM_SUCCESS;
}

extern "C" {

SKM_API skm_status skm_set_stream(skm_evaluator *ev, void *hip_stream)
{
    if (!ev) return fail(SKM_ERROR_INVALID_ARGUMENT, "null evaluator");
    SKM_HIP(hipStreamSynchronize(ev->stream));
    if (ev->own_stream) return_stream(ev->stream_device, ev->stream);
    ev->own_stream = false;
    ev->stream = (hipStream_t)hip_stream;
    return SKM_SUCCESS;
}

SKM_API skm_status skm_synchronize(skm_evaluator *ev)
{
    if (!ev) return fail(SKM_ERROR_INVALID_ARGUMENT, "null evaluator");
    SKM_HIP(hipStreamSynchronize(ev->stream));
    return SKM_SUCCESS;
}

SKM_API skm_status skm_get_value(skm_evaluator *ev, int32_t measure, const float *cp_aos, const float *cpw,
                                 int32_t do_resampling, skm_result *out)
{
    if (!ev) return fail(SKM_ERROR_INVALID_ARGUMENT, "null evaluator");
    SKM_TRY(upload_cp(ev, cp_aos, cpw));
    SKM_TRY(run_eval(ev, measure, do_resampling, nullptr));
    return read_result(ev, out);
}

SKM_API skm_status skm_get_value_and_derivative(skm_evaluator *ev, int32_t measure, const float *cp_aos,
                                                const float *cpw, float *derivative_soa, int32_t do_resampling,
                                                skm_result *out)
{
    if (!ev || !derivative_soa) return fail(SKM_ERROR_INVALID_ARGUMENT, "null argument");
    SKM_TRY(upload_cp(ev, cp_aos, cpw));
    SKM_TRY(run_eval(ev, measure, do_resampling, ev->d_deriv));
    SKM_HIP(hipMemcpyAsync(derivative_soa, ev->d_deriv, sizeof(float) * ev->pr.n_cp * ev->dim,
                           hipMemcpyDeviceToHost, ev->stream));
    return read_result(ev, out);
}

SKM_API skm_status skm_get_value_and_derivative_device(skm_evaluator *ev, int32_t measure, const float *d_cp_aos,
                                                       const float *d_cpw, float *d_derivative_soa,
                                                       int32_t do_resampling, skm_result *d_result)
{
    if (!ev || !d_cp_aos || !d_cpw || !d_derivative_soa) return fail(SKM_ERROR_INVALID_ARGUMENT, "null argument");
    launch_pack_cp(ev->pr, d_cp_aos, d_cpw, 0, ev->d_cpc, ev->stream);
    SKM_TRY(run_eval(ev, measure, do_resampling, d_derivative_soa));
    if (d_result) launch_result(ev->d_sc, d_result, ev->stream);
    SKM_HIP(hipGetLastError());
    return SKM_SUCCESS;
}

SKM_API skm_status skm_evaluate_displacement_field(skm_evaluator *ev, const float *cp_aos, const float *cpw,
                                                   float *dst_aos)
{
    if (!ev || !dst_aos) return fail(SKM_ERROR_INVALID_ARGUMENT, "null argument");
    SKM_TRY(upload_cp(ev, cp_aos, cpw));
    if (!ev->d_dst) SKM_TRY(dev_alloc(&ev->d_dst, ev->pr.n_vox * ev->dim));
    {
        PassTimer t(ev, 9);
        launch_displacement(ev->pr, ev->d_dst, ev->stream);
    }
    SKM_HIP(hipGetLastError());
    SKM_HIP(hipMemcpyAsync(dst_aos, ev->d_dst, sizeof(float) * ev->pr.n_vox * ev->dim, hipMemcpyDeviceToHost,
                           ev->stream));
    SKM_HIP(hipStreamSynchronize(ev->stream));
    return SKM_SUCCESS;
}

SKM_API skm_status skm_evaluate_displacement_field_device(skm_evaluator *ev, const float *d_cp_aos,
                                                          const float *d_cpw, float *d_dst_aos)
{
    if (!ev || !d_cp_aos || !d_cpw || !d_dst_aos) return fail(SKM_ERROR_INVALID_ARGUMENT, "null argument");
    launch_pack_cp(ev->pr, d_cp_aos, d_cpw, 0, ev->d_cpc, ev->stream);
    {
        PassTimer t(ev, 9);
        launch_displacement(ev->pr, d_dst_aos, ev->stream);
    }
    SKM_HIP(hipGetLastError());
    return SKM_SUCCESS;
}

SKM_API skm_status skm_enable_timing(skm_evaluator *ev, int32_t enable)
{
    if (!ev) return fail(SKM_ERROR_INVALID_ARGUMENT, "null evaluator");
    ev->timing = enable != 0;
    return SKM_SUCCESS;
}

SKM_API skm_status skm_read_timing(skm_evaluator *ev, skm_pass_times *out)
{
    if (!ev || !out) return fail(SKM_ERROR_INVALID_ARGUMENT, "null argument");
    SKM_HIP(hipStreamSynchronize(ev->stream));
    for (auto &p : ev->pending) {
        float ms = 0.0f;
        SKM_HIP(hipEventElapsedTime(&ms, p.a, p.b));
        ev->pass_ms[p.pass] += ms;
        ev->pass_launches[p.pass] += 1;
        ev->event_pool.push_back(p.a);
        ev->event_pool.push_back(p.b);
    }
    ev->pending.clear();
    for (int i = 0; i < SKM_NUM_PASSES; ++i) {
        out->ms[i] = ev->pass_ms[i];
        out->launches[i] = ev->pass_launches[i];
        ev->pass_ms[i] = 0.0;
        ev->pass_launches[i] = 0;
    }
    return SKM_SUCCESS;
}

SKM_API skm_status skm_fit_coefficients_device(skm_evaluator *ev, const float *d_displacement_aos, int32_t iterations,
                                               double ridge, float *d_coefficients_soa)
{
    return skm::fit_coefficients_device(ev, d_displacement_aos, iterations, ridge, d_coefficients_soa);
}

namespace {

skm_status validate_tuning(const skm_tuning &t)
{
    if (t.lcc_ref_chunks < 0 || t.lcc_ref_chunks > 64)
        return fail(SKM_ERROR_INVALID_ARGUMENT, "lcc_ref_chunks must be 0 (automatic) or 1..64");
    if (t.gather_split != 0 && t.gather_split != 4 && t.gather_split != 8)
        return fail(SKM_ERROR_INVALID_ARGUMENT, "gather_split must be 0 (automatic), 4 or 8");
    if (t.value_split != 0 && t.value_split != 1 && t.value_split != 4 && t.value_split != 8 && t.value_split != 16)
        return fail(SKM_ERROR_INVALID_ARGUMENT, "value_split must be 0 (automatic), 1, 4, 8 or 16");
    if (t.gather_regions != 0 && t.gather_regions != 1 && t.gather_regions != 8)
        return fail(SKM_ERROR_INVALID_ARGUMENT, "gather_regions must be 0 (automatic), 1 or 8");
    if (t.lcc_ref_walk < 0 || t.lcc_ref_walk > 2)
        return fail(SKM_ERROR_INVALID_ARGUMENT, "lcc_ref_walk must be 0 (automatic), 1 or 2");
    if (t.kernel_matrix < 0 || t.kernel_matrix > 2)
        return fail(SKM_ERROR_INVALID_ARGUMENT, "kernel_matrix must be 0 (automatic), 1 (never) or 2 (always)");
    return SKM_SUCCESS;
}

}  // namespace

SKM_API skm_status skm_set_tuning(skm_evaluator *ev, const skm_tuning *tuning)
{
    if (!ev || !tuning) return fail(SKM_ERROR_INVALID_ARGUMENT, "null argument");
    SKM_TRY(validate_tuning(*tuning));
    SKM_HIP(hipStreamSynchronize(ev->stream));  // queued work may use the buffers resized below
    // a new policy, or a new gather slab (the gather layout follows the slabs), drops the kernel matrix
    if (tuning->kernel_matrix != ev->tuning.kernel_matrix || tuning->gather_split != ev->tuning.gather_split)
        kernel_matrix_release(ev->km);
    ev->tuning = *tuning;
    ev->pr.lcc_chunks = tuning->lcc_ref_chunks;
    ev->pr.gather_split = tuning->gather_split;
    ev->pr.value_split = tuning->value_split;
    ev->pr.gather_regions = tuning->gather_regions;
    ev->pr.lcc_ref_walk = tuning->lcc_ref_walk;
    const long long np = lcc_reference_partials(ev->pr);
    if (ev->d_lcc_ref_partials && np > ev->lcc_ref_partials_n) {
        device_release(ev->d_lcc_ref_partials);
        ev->d_lcc_ref_partials = nullptr;
        ev->lcc_ref_partials_n = 0;
    }
    return SKM_SUCCESS;
}

SKM_API skm_status skm_autotune(skm_evaluator *ev, int32_t measure, const float *d_cp_aos, const float *d_cpw,
                                int32_t repeats, skm_autotune_report *report)
{
    if (!ev || !d_cp_aos || !d_cpw) return fail(SKM_ERROR_INVALID_ARGUMENT, "null argument");
    if (measure < SKM_MEASURE_MSE || measure > SKM_MEASURE_LCC) return fail(SKM_ERROR_INVALID_ARGUMENT, "bad measure");
    if (ev->timing) return fail(SKM_ERROR_INVALID_ARGUMENT, "disable pass timing before tuning");
    repeats = std::max(1, repeats);
    hipStream_t s = ev->stream;
    // axes of the search, each tuned with the others at their best so far
    // (the value pass and the gather are separate kernels: their shapes do
    // not interact beyond the order of the evaluation's launches)
    using Field = int32_t skm_tuning::*;
    struct Axis {
        Field field;
        std::vector<int> values;  // the first is the automatic choice (0)
    };
    std::vector<Axis> axes;
    const bool d3 = ev->pr.dim == 3;
    if (measure == SKM_MEASURE_LCC && !ev->pr.exact) {
        Problem tp = ev->pr;
        tp.lcc_ref_walk = 2;
        if (lcc_ref_table_applies(tp)) axes.push_back({&skm_tuning::lcc_ref_walk, {0, 1, 2}});
        axes.push_back({&skm_tuning::lcc_ref_chunks, {0, 1, 2, 4, 8, 16, 32, 64}});
    } else {
        if (measure != SKM_MEASURE_LCC && kernel_matrix_applies(ev->pr))
            axes.push_back({&skm_tuning::kernel_matrix, {0, 1, 2}});
        if (measure != SKM_MEASURE_LCC) axes.push_back({&skm_tuning::value_split, {0, 1, 4, 8, 16}});
        axes.push_back({&skm_tuning::gather_split, {0, 4}});
        if (d3) axes.push_back({&skm_tuning::gather_regions, {0, 8}});
    }
    hipEvent_t a = nullptr, b = nullptr;
    SKM_HIP(hipEventCreate(&a));
    SKM_HIP(hipEventCreate(&b));
    struct Events {
        hipEvent_t &a, &b;
        ~Events()
        {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } events{a, b};
    launch_pack_cp(ev->pr, d_cp_aos, d_cpw, 0, ev->d_cpc, s);
    auto time_of = [&](const skm_tuning &t, double *out) -> skm_status {
        SKM_TRY(skm_set_tuning(ev, &t));
        // first use of the shape (allocations, the kernel matrix), untimed
        SKM_TRY(run_eval(ev, measure, 0, ev->d_deriv));
        std::vector<double> ms;
        for (int r = 0; r < repeats; ++r) {
            SKM_HIP(hipEventRecord(a, s));
            SKM_TRY(run_eval(ev, measure, 0, ev->d_deriv));
            SKM_HIP(hipEventRecord(b, s));
            SKM_HIP(hipEventSynchronize(b));
            float t_ms = 0.0f;
            SKM_HIP(hipEventElapsedTime(&t_ms, a, b));
            ms.push_back(t_ms);
        }
        std::sort(ms.begin(), ms.end());
        *out = ms[ms.size() / 2];
        return SKM_SUCCESS;
    };
    skm_tuning best = ev->tuning;
    for (const Axis &ax : axes) best.*(ax.field) = 0;
    double best_ms = 0.0;
    SKM_TRY(time_of(best, &best_ms));
    const double default_ms = best_ms;
    int candidates = 1;
    for (const Axis &ax : axes) {
        for (size_t k = 1; k < ax.values.size(); ++k) {
            skm_tuning t = best;
            t.*(ax.field) = ax.values[k];
            double ms = 0.0;
            SKM_TRY(time_of(t, &ms));
            ++candidates;
            if (ms < best_ms) {
                best_ms = ms;
                best = t;
            }
        }
    }
    SKM_TRY(skm_set_tuning(ev, &best));
    tuning_store(shape_class(ev), best);
    if (report) {
        report->best = best;
        report->best_ms = best_ms;
        report->first_ms = default_ms;
        report->candidates = candidates;
    }
    return SKM_SUCCESS;
}

SKM_API skm_status skm_tuning_table_save(const char *path)
{
    if (!path) return fail(SKM_ERROR_INVALID_ARGUMENT, "null path");
    std::ofstream f(path);
    if (!f) return fail(SKM_ERROR_IO, std::string("cannot write ") + path);
    std::lock_guard<std::mutex> lock(g_tuning_mutex);
    f << "# skm launch shapes: shape class, lcc_ref_chunks, gather_split, value_split, gather_regions, "
         "lcc_ref_walk, kernel_matrix\n";
    for (const auto &kv : g_tuning)
        f << kv.first << " " << kv.second.lcc_ref_chunks << " " << kv.second.gather_split << " "
          << kv.second.value_split << " " << kv.second.gather_regions << " " << kv.second.lcc_ref_walk << " "
          << kv.second.kernel_matrix << "\n";
    return f ? SKM_SUCCESS : fail(SKM_ERROR_IO, std::string("cannot write ") + path);
}

SKM_API skm_status skm_tuning_table_load(const char *path)
{
    if (!path) return fail(SKM_ERROR_INVALID_ARGUMENT, "null path");
    std::ifstream f(path);
    if (!f) return fail(SKM_ERROR_IO, std::string("cannot read ") + path);
    std::string line;
    std::map<std::string, skm_tuning> read;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream in(line);
        std::string key;
        skm_tuning t{};
        if (!(in >> key >> t.lcc_ref_chunks >> t.gather_split >> t.value_split >> t.gather_regions >>
              t.lcc_ref_walk))
            return fail(SKM_ERROR_IO, "malformed tuning line: " + line);
        // kernel_matrix: an optional last column (tables written before it hold five)
        if (!(in >> t.kernel_matrix)) t.kernel_matrix = 0;
        if (validate_tuning(t) != SKM_SUCCESS) return fail(SKM_ERROR_IO, "launch shape out of range: " + line);
        read[key] = t;
    }
    std::lock_guard<std::mutex> lock(g_tuning_mutex);
    for (const auto &kv : read) g_tuning[kv.first] = kv.second;
    return SKM_SUCCESS;
}

SKM_API skm_status skm_set_cache_limit(uint64_t idle_bytes_per_device)
{
    if (idle_bytes_per_device > (uint64_t)INT64_MAX) return fail(SKM_ERROR_INVALID_ARGUMENT, "limit out of range");
    g_idle_cap.store((long long)idle_bytes_per_device);
    // devices over the new limit give their idle blocks back
    std::vector<int> over;
    {
        BlockCache &c = block_cache();
        std::lock_guard<std::mutex> lock(c.mutex);
        for (const auto &kv : c.idle_bytes)
            if (kv.second > idle_bytes_per_device) over.push_back(kv.first);
    }
    for (int dev : over) drop_idle(dev);
    return SKM_SUCCESS;
}

SKM_API skm_status skm_get_cache_info(uint64_t *idle_bytes, uint64_t *live_bytes, uint64_t *limit)
{
    BlockCache &c = block_cache();
    std::lock_guard<std::mutex> lock(c.mutex);
    uint64_t idle = 0, live = 0;
    for (const auto &kv : c.idle_bytes) idle += kv.second;
    for (const auto &kv : c.live) live += kv.second.second;
    if (idle_bytes) *idle_bytes = idle;
    if (live_bytes) *live_bytes = live;
    if (limit) *limit = idle_cap();
    return SKM_SUCCESS;
}

SKM_API skm_status skm_release_cached_memory(void)
{
    drop_idle(-1);
    return SKM_SUCCESS;
}

SKM_API skm_status skm_get_kernel_matrix_info(skm_evaluator *ev, skm_kernel_matrix_info *info)
{
    if (!ev || !info) return fail(SKM_ERROR_INVALID_ARGUMENT, "null argument");
    *info = skm_kernel_matrix_info{};
    info->state = ev->km.state;
    info->bytes = (uint64_t)ev->km.bytes;
    info->pairs_value = (uint64_t)ev->km.nnz_v;
    info->pairs_gather = (uint64_t)ev->km.nnz_g;
    info->read_value = (uint64_t)ev->km.read_v;
    info->read_gather = (uint64_t)ev->km.read_g;
    info->build_ms = ev->km.build_ms;
    return SKM_SUCCESS;
}

SKM_API skm_status skm_release_kernel_matrix(skm_evaluator *ev)
{
    if (!ev) return fail(SKM_ERROR_INVALID_ARGUMENT, "null argument");
    SKM_HIP(hipStreamSynchronize(ev->stream));  // queued evaluations may still read it
    kernel_matrix_release(ev->km);
    return SKM_SUCCESS;
}

SKM_API skm_status skm_get_tuning(skm_evaluator *ev, skm_tuning *tuning)
{
    if (!ev || !tuning) return fail(SKM_ERROR_INVALID_ARGUMENT, "null argument");
    *tuning = ev->tuning;
    return SKM_SUCCESS;
}

SKM_API skm_status skm_get_cube_offsets(skm_evaluator *ev, int32_t *out, int64_t *n_cells)
{
    if (!ev) return fail(SKM_ERROR_INVALID_ARGUMENT, "null evaluator");
    if (n_cells) *n_cells = ev->pr.n_sub;
    if (!out) return SKM_SUCCESS;
    if (ev->pr.sub <= 1) {
        std::memset(out, 0, sizeof(int32_t) * ev->pr.n_sub * ev->dim);
        return SKM_SUCCESS;
    }
    SKM_HIP(hipMemcpyAsync(out, ev->d_offsets, sizeof(int32_t) * ev->pr.n_sub * ev->dim, hipMemcpyDeviceToHost,
                           ev->stream));
    SKM_HIP(hipStreamSynchronize(ev->stream));
    return SKM_SUCCESS;
}

SKM_API skm_status skm_control_point_grid(int32_t dim, const skm_grid *image, const double *cp_spacing,
                                          skm_grid *cp_out)
{
    if (!image || !cp_spacing || !cp_out || (dim != 2 && dim != 3))
        return fail(SKM_ERROR_INVALID_ARGUMENT, "bad argument");
    skm_grid g;
    std::memset(&g, 0, sizeof(g));
    for (int d = 0; d < 3; ++d) { g.size[d] = 1; g.spacing[d] = 1.0; }
    for (int d = 0; d < dim; ++d) {
        if (!(cp_spacing[d] > 0)) return fail(SKM_ERROR_INVALID_ARGUMENT, "control-point spacing must be > 0");
        double lower = image->origin[d];
        double upper = image->origin[d] + image->spacing[d] * (image->size[d] - 1);
        g.spacing[d] = cp_spacing[d];
        g.size[d] = (int)std::ceil(std::fabs((upper - lower) / cp_spacing[d]));
        double cp_upper = image->origin[d] + cp_spacing[d] * (g.size[d] - 1);
        g.origin[d] = image->origin[d] + std::fabs(upper - cp_upper) / 2.0;
    }
    *cp_out = g;
    return SKM_SUCCESS;
}

SKM_API skm_status skm_host_regularizers(int32_t dim, int64_t n_cp, const double *params_soa, double reg_l1,
                                         double reg_l21, double reg_l2, int32_t use_negative_gradient,
                                         double *derivative_soa, double *reg_value, double *l1, double *l21,
                                         double *l2)
{
    if (!params_soa || (dim != 2 && dim != 3)) return fail(SKM_ERROR_INVALID_ARGUMENT, "bad argument");
    double v1 = 0, v21 = 0, v2 = 0;
    const long long np = n_cp * dim;
    if (reg_l1 > 0) {
        for (long long p = 0; p < np; ++p) v1 += std::fabs(params_soa[p]);
        if (derivative_soa)
            for (long long p = 0; p < np; ++p) {
                double par = params_soa[p];
                double &g = derivative_soa[p];
                if (std::fabs(par) > 0) g += reg_l1 * (par < 0 ? -1.0 : 1.0);
                else if (g < -reg_l1) g += reg_l1;
                else if (g > reg_l1) g -= reg_l1;
                else if (-reg_l1 <= g && g <= reg_l1) g = 0.0;
                else return fail(SKM_ERROR_INVALID_ARGUMENT, "Error in calculating subgradient");
            }
    }
    if (reg_l2 > 0)
        for (long long j = 0; j < n_cp; ++j) {
            double n2 = 0;
            for (int d = 0; d < dim; ++d) n2 += params_soa[j + d * n_cp] * params_soa[j + d * n_cp];
            v2 += n2;
            if (derivative_soa)
                for (int d = 0; d < dim; ++d) derivative_soa[j + d * n_cp] += reg_l2 * 2.0 * params_soa[j + d * n_cp];
        }
    if (reg_l21 > 0)
        for (long long j = 0; j < n_cp; ++j) {
            double n2 = 0;
            for (int d = 0; d < dim; ++d) n2 += params_soa[j + d * n_cp] * params_soa[j + d * n_cp];
            double nrm = std::sqrt(n2);
            v21 += nrm;
            if (!derivative_soa) continue;
            for (int d = 0; d < dim; ++d) {
                double &g = derivative_soa[j + d * n_cp];
                if (nrm > 0) g += reg_l21 * params_soa[j + d * n_cp] / nrm;
                else if (g < -reg_l21) g += reg_l21;
                else if (g > reg_l21) g -= reg_l21;
                else if (-reg_l21 <= g && g <= reg_l21) g = 0.0;
                else return fail(SKM_ERROR_INVALID_ARGUMENT, "Error in calculating subgradient");
            }
        }
    if (use_negative_gradient && derivative_soa)
        for (long long p = 0; p < np; ++p) derivative_soa[p] = -derivative_soa[p];
    double reg = 0;
    if (reg_l1 > 0) reg += reg_l1 * v1;
    if (reg_l21 > 0) reg += reg_l21 * v21;
    if (reg_l2 > 0) reg += reg_l2 * v2;
    if (reg_value) *reg_value = reg;
    if (l1) *l1 = v1;
    if (l21) *l21 = v21;
    if (l2) *l2 = v2;
    return SKM_SUCCESS;
}

namespace {

skm_status opt_check(const skm_evaluator *ev, const skm_metric_options *metric, const skm_optimizer_options *opt,
                     const float *cpw, const float *params_soa)
{
    if (!ev || !metric || !opt || !cpw || !params_soa) return fail(SKM_ERROR_INVALID_ARGUMENT, "null argument");
    if (opt->iterations < 0) return fail(SKM_ERROR_INVALID_ARGUMENT, "iterations must be >= 0");
    if (opt->kind < 0 || opt->kind > 2) return fail(SKM_ERROR_INVALID_ARGUMENT, "bad optimizer kind");
    if (metric->reg_l1 < 0 || metric->reg_l21 < 0 || metric->reg_l2 < 0)
        return fail(SKM_ERROR_INVALID_ARGUMENT, "Regularization weights must be positive.");
    return SKM_SUCCESS;
}

// buffers, start point and state of a device optimisation, on the evaluator's stream
skm_status opt_prepare(skm_evaluator *ev, const skm_optimizer_options *opt, const float *cpw, const float *params_soa)
{
    Problem &pr = ev->pr;
    hipStream_t s = ev->stream;
    const size_t np = (size_t)pr.n_cp * pr.dim;
    if (!ev->d_params) SKM_TRY(dev_alloc(&ev->d_params, np));
    if (!ev->d_prev_g) SKM_TRY(dev_alloc(&ev->d_prev_g, np));
    if (!ev->d_opt) SKM_TRY(dev_alloc(&ev->d_opt, 1));
    if (opt->return_best && !ev->d_best) SKM_TRY(dev_alloc(&ev->d_best, np));
    if (ev->records_cap < opt->iterations) {
        SKM_HIP(hipStreamSynchronize(s));  // the previous records may still be read
        device_release(ev->d_records);
        ev->d_records = nullptr;
        SKM_TRY(dev_alloc(&ev->d_records, (size_t)opt->iterations));
        ev->records_cap = opt->iterations;
    }
    SKM_HIP(hipMemcpyAsync(ev->d_params, params_soa, sizeof(float) * np, hipMemcpyHostToDevice, s));
    SKM_HIP(hipMemcpyAsync(ev->d_cpw_in, cpw, sizeof(float) * pr.n_cp, hipMemcpyHostToDevice, s));
    SKM_HIP(hipMemsetAsync(ev->d_prev_g, 0, sizeof(float) * np, s));
    SKM_HIP(hipMemsetAsync(ev->d_opt, 0, sizeof(OptState), s));
    launch_soa_to_cpc(pr, ev->d_params, ev->d_cpw_in, ev->d_cpc, s);
    return SKM_SUCCESS;
}

// One iteration: every value it needs lives on the device (parameters and
// their packed control points, optimizer state and iteration count, random
// stream), so the launches of one iteration serve every other one; the
// update packs the control points of the next evaluation.
skm_status opt_iteration(skm_evaluator *ev, const skm_metric_options *metric, const skm_optimizer_options *opt)
{
    Problem &pr = ev->pr;
    hipStream_t s = ev->stream;
    SKM_TRY(run_eval(ev, metric->measure, metric->do_resampling, ev->d_deriv));
    launch_host_regs_device(pr, ev->d_params, metric->reg_l1, metric->reg_l21, metric->reg_l2,
                            metric->use_negative_gradient, ev->d_deriv, ev->d_prev_g, ev->d_hreg_partials, s);
    launch_opt_step(pr, opt->kind, opt->a, opt->A, opt->alpha, opt->fmax, opt->fmin, opt->omega,
                    opt->orthant_projection, ev->d_params, ev->d_deriv, ev->d_prev_g, ev->d_opt, ev->d_sc,
                    ev->d_hreg_partials, metric->reg_l1, metric->reg_l21, metric->reg_l2, ev->reg_w, ev->d_records,
                    opt->return_best ? ev->d_best : nullptr, ev->d_cpw_in, ev->d_cpc, s);
    return SKM_SUCCESS;
}

// the result (best or last position) and the records, back to the host
skm_status opt_finish(skm_evaluator *ev, const skm_optimizer_options *opt, float *params_soa,
                      skm_iteration_record *records)
{
    hipStream_t s = ev->stream;
    const size_t np = (size_t)ev->pr.n_cp * ev->pr.dim;
    SKM_HIP(hipGetLastError());
    // with return_best (and at least one evaluation) the lowest-value position
    const float *result = (opt->return_best && opt->iterations > 0) ? ev->d_best : ev->d_params;
    SKM_HIP(hipMemcpyAsync(params_soa, result, sizeof(float) * np, hipMemcpyDeviceToHost, s));
    if (records && opt->iterations > 0)
        SKM_HIP(hipMemcpyAsync(records, ev->d_records, sizeof(skm_iteration_record) * opt->iterations,
                               hipMemcpyDeviceToHost, s));
    SKM_HIP(hipStreamSynchronize(s));
    return SKM_SUCCESS;
}

// Iterations first..iterations of every evaluator as one captured HIP graph
// per step: the capture forks from the first evaluator's stream to every
// other one and joins back, so the graph holds one branch per evaluator; one
// hipGraphLaunch then runs one optimizer step of all of them.
skm_status opt_replay(skm_evaluator *const *evs, int n, const skm_metric_options *metric,
                      const skm_optimizer_options *opt, int first)
{
    hipStream_t s0 = evs[0]->stream;
    SKM_HIP(hipGetLastError());
    std::vector<hipEvent_t> events;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    auto release = [&]() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
    };
    for (int i = 0; i < 2 * n; ++i) {
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            release();
            return fail(SKM_ERROR_HIP, "graph events");
        }
        events.push_back(e);
    }
    // the evaluators' streams must be idle and after s0's work before the capture joins them
    for (int i = 1; i < n; ++i)
        if (evs[i]->stream != s0 && hipEventRecord(events[n + i], evs[i]->stream) == hipSuccess)
            (void)hipStreamWaitEvent(s0, events[n + i], 0);
    SKM_HIP(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
    skm_status captured = SKM_SUCCESS;
    hipError_t e = hipEventRecord(events[0], s0);
    for (int i = 0; i < n && e == hipSuccess && captured == SKM_SUCCESS; ++i) {
        hipStream_t si = evs[i]->stream;
        if (si != s0) e = hipStreamWaitEvent(si, events[0], 0);
        if (e == hipSuccess) captured = opt_iteration(evs[i], metric, opt);
        if (e == hipSuccess && captured == SKM_SUCCESS && si != s0) {
            e = hipEventRecord(events[n + i], si);
            if (e == hipSuccess) e = hipStreamWaitEvent(s0, events[n + i], 0);
        }
    }
    const hipError_t ended = hipStreamEndCapture(s0, &graph);
    if (captured != SKM_SUCCESS) {
        release();
        return captured;
    }
    if (e != hipSuccess || ended != hipSuccess) {
        release();
        return fail(SKM_ERROR_HIP, std::string("optimizer capture: ") + hipGetErrorString(e != hipSuccess ? e : ended));
    }
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    for (int it = first; e == hipSuccess && it < opt->iterations; ++it) e = hipGraphLaunch(exec, s0);
    if (e == hipSuccess) e = hipStreamSynchronize(s0);
    release();
    if (e != hipSuccess) return fail(SKM_ERROR_HIP, std::string("optimizer replay: ") + hipGetErrorString(e));
    return SKM_SUCCESS;
}

}  // namespace

SKM_API skm_status skm_optimize(skm_evaluator *ev, const skm_metric_options *metric,
                                const skm_optimizer_options *opt, const float *cpw, float *params_soa,
                                skm_iteration_record *records)
{
    SKM_TRY(opt_check(ev, metric, opt, cpw, params_soa));
    SKM_TRY(opt_prepare(ev, opt, cpw, params_soa));
    // with graph replay the first iteration runs kernel by kernel (it
    // allocates what a measure needs on first use and draws the first
    // offsets) and the rest replay a capture of the second one; pass timing
    // needs the kernel-by-kernel launches
    const bool replay = opt->graph && !ev->timing && opt->iterations > 2;
    int it = 0;
    for (; it < (replay ? 1 : opt->iterations); ++it) SKM_TRY(opt_iteration(ev, metric, opt));
    if (replay) SKM_TRY(opt_replay(&ev, 1, metric, opt, 1));
    return opt_finish(ev, opt, params_soa, records);
}

SKM_API skm_status skm_optimize_batch(skm_evaluator *const *evaluators, int32_t n, const skm_metric_options *metric,
                                      const skm_optimizer_options *opt, const float *const *cpw,
                                      float *const *params_soa, skm_iteration_record *const *records)
{
    if (n < 1 || !evaluators || !cpw || !params_soa) return fail(SKM_ERROR_INVALID_ARGUMENT, "null argument");
    for (int32_t i = 0; i < n; ++i) {
        SKM_TRY(opt_check(evaluators[i], metric, opt, cpw[i], params_soa[i]));
        for (int32_t k = 0; k < i; ++k)
            if (evaluators[k] == evaluators[i]) return fail(SKM_ERROR_INVALID_ARGUMENT, "an evaluator appears twice");
    }
    for (int32_t i = 0; i < n; ++i) SKM_TRY(opt_prepare(evaluators[i], opt, cpw[i], params_soa[i]));
    bool timing = false;
    for (int32_t i = 0; i < n; ++i) timing = timing || evaluators[i]->timing;
    const bool replay = opt->graph && !timing && opt->iterations > 1;
    // the first iteration kernel by kernel (first-use allocations, first offsets)
    int it = 0;
    for (; it < (replay ? 1 : opt->iterations); ++it)
        for (int32_t i = 0; i < n; ++i) SKM_TRY(opt_iteration(evaluators[i], metric, opt));
    if (replay) SKM_TRY(opt_replay(evaluators, n, metric, opt, 1));
    for (int32_t i = 0; i < n; ++i)
        SKM_TRY(opt_finish(evaluators[i], opt, params_soa[i], records ? records[i] : nullptr));
    return SKM_SUCCESS;
}

}  // extern "C"
