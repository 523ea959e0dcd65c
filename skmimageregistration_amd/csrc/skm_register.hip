// skm_register.hip -- the scale-level loop of SKMReg (src/SKMReg.cxx:84-525)
// as a library call over this library's own C ABI: preprocessing, per-level
// smoothing and weight tensor, a new evaluator per level whose previous
// displacement is the sum of the coarser levels, the device-resident
// adaptive-step optimizer, and the dense displacement of all levels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
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

// ThresholdImage (include/itkUtils.h:45-100): values at or above thresh_max
// become the largest value below it, values at or below thresh_min the
// smallest value above it (two ThresholdOutside passes bounded by the image's
// own minimum and maximum).
void threshold_image(std::vector<float> &img, double thresh_max, double thresh_min)
{
    double below_max = std::numeric_limits<double>::lowest(), above_min = std::numeric_limits<double>::max();
    double lo = std::numeric_limits<double>::max(), hi = std::numeric_limits<double>::lowest();
    for (float v : img) {
        const double x = v;
        if (below_max < x && x < thresh_max) below_max = x;
        if (above_min > x && x > thresh_min) above_min = x;
        lo = std::min(lo, x);
        hi = std::max(hi, x);
    }
    for (float &v : img) {
        if (v < lo || v > below_max) v = (float)below_max;  // ThresholdOutside(min, max), outside value max
    }
    for (float &v : img) {
        if (v < above_min || v > hi) v = (float)above_min;  // ThresholdOutside(min, image max), value min
    }
}

// RescaleImage(image, 0, 1) (itkUtils.h:103-112): RescaleIntensityImageFilter,
// out = in * scale + shift with scale = (1 - 0) / (max - min).
void rescale_image(std::vector<float> &img, double out_min, double out_max)
{
    if (img.empty()) return;
    const auto mm = std::minmax_element(img.begin(), img.end());
    const double lo = *mm.first, hi = *mm.second;
    double scale = 0.0;
    if (hi != lo) scale = (out_max - out_min) / (hi - lo);
    else if (hi != 0.0) scale = (out_max - out_min) / hi;
    const double shift = out_min - lo * scale;
    for (float &v : img) v = (float)std::min(out_max, std::max(out_min, (double)v * scale + shift));
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
    if (!desc || !displacement_out) return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "null argument");
    const int dim = desc->dim;
    if (dim != 2 && dim != 3) return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "dim must be 2 or 3");
    if (!desc->reference_data || !desc->target_data)
        return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "reference and target images are required");
    if (desc->num_scales < 1 || desc->num_scales > SKM_MAX_SCALES)
        return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "num_scales must be 1.." + std::to_string(SKM_MAX_SCALES));
    const long long n_ref = voxels(desc->reference, dim), n_tar = voxels(desc->target, dim);
    if (n_ref < 1 || n_tar < 1) return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "empty image");

    // preprocessing (SKMReg.cxx:84-91)
    std::vector<float> ref(desc->reference_data, desc->reference_data + n_ref);
    std::vector<float> tar(desc->target_data, desc->target_data + n_tar);
    if (desc->use_intensity_thresholds) {
        threshold_image(ref, desc->threshold_max, desc->threshold_min);
        threshold_image(tar, desc->threshold_max, desc->threshold_min);
    }
    rescale_image(ref, 0.0, 1.0);
    rescale_image(tar, 0.0, 1.0);

    // previous displacement (restart, SKMReg.cxx:96-113)
    int start = 0;
    std::vector<float> prev;
    if (desc->restart_at > 1 && desc->restart_at <= desc->num_scales && desc->restart_displacement) {
        start = desc->restart_at - 1;
        prev.assign(desc->restart_displacement, desc->restart_displacement + n_tar * dim);
    }
    const long long n_w = voxels(desc->weight, dim);

    std::vector<float> ref_s(n_ref), tar_s(n_tar), tensor;
    for (int scale = start; scale < desc->num_scales; ++scale) {
        const skm_scale_params &P = desc->scale[scale];
        skm_scale_report rep;
        std::memset(&rep, 0, sizeof(rep));
        const auto t_setup = Clock::now();

        // smoothing (SKMReg.cxx:140-150)
        if (P.smoothing_variance > 1e-5) {
            SKM_R_TRY(skm_gaussian_smooth(dim, &desc->reference, ref.data(), P.smoothing_variance, ref_s.data()));
            SKM_R_TRY(skm_gaussian_smooth(dim, &desc->target, tar.data(), P.smoothing_variance, tar_s.data()));
        } else {
            ref_s = ref;
            tar_s = tar;
        }

        // kernel: weight image (takes precedence, GpuEvaluator.cu:311-346),
        // else the weight tensor cases of SKMReg.cxx:262-297
        skm_desc ed;
        skm_desc_init(&ed);
        ed.dim = dim;
        ed.fixed = desc->target;
        ed.fixed_data = tar_s.data();
        ed.moving = desc->reference;
        ed.moving_data = ref_s.data();
        ed.sigma_c0 = P.sigma_c0;
        ed.sigma_c4 = P.sigma_c4;
        ed.subsample = P.sampling_rate;
        ed.subsample_neighborhood = P.neighborhood_sampling;
        if (desc->weight_image) {
            double mx = std::numeric_limits<double>::lowest();
            for (long long i = 0; i < n_w; ++i) mx = std::max(mx, (double)desc->weight_image[i]);
            ed.kernel_mode = SKM_KERNEL_WEIGHT_IMAGE;
            ed.weight = desc->weight;
            ed.weight_data = desc->weight_image;
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
            ed.weight_data = desc->weight_tensor;
            ed.maximum_weight = best;
            for (int i = 0; i < 9; ++i) ed.maximum_eigenvalue_covariance[i] = 0.0;
            for (int i = 0; i < dim; ++i)
                for (int j = 0; j < dim; ++j)
                    ed.maximum_eigenvalue_covariance[i * 3 + j] = desc->weight_tensor[at * nn + i * dim + j];
        } else if (P.tensor_weight > 0) {
            const bool guided = desc->guidance_image != nullptr;
            const skm_grid &g = guided ? desc->weight : desc->target;
            tensor.resize((size_t)voxels(g, dim) * dim * dim);
            skm_tensor_options to{P.sigma_c4, P.tensor_weight, 2, 0};
            double tilt = 0, mw = 0, mcov[9];
            SKM_R_TRY(skm_anisotropic_tensor(dim, &g, guided ? desc->guidance_image : tar_s.data(), &to,
                                             tensor.data(), &tilt, &mw, mcov));
            ed.kernel_mode = SKM_KERNEL_WEIGHT_TENSOR;
            ed.weight = g;
            ed.weight_data = tensor.data();
            ed.maximum_weight = mw;
            std::memcpy(ed.maximum_eigenvalue_covariance, mcov, sizeof(mcov));
            rep.maximum_tilt = tilt;
        }

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
        SKM_R_TRY(skm_create(&ed, &guard.ev));
        skm_evaluator *ev = guard.ev;
        if (P.reg_rkhs > 0) SKM_R_TRY(skm_set_regularizer_rkhs(ev, P.reg_rkhs));
        if (P.reg_rd > 0) SKM_R_TRY(skm_set_regularizer_rd(ev, P.reg_rd, P.reg_rd_scaling));
        if (P.reg_pg > 0) SKM_R_TRY(skm_set_regularizer_pg(ev, P.reg_pg, P.reg_pg_scaling));
        if (!prev.empty()) SKM_R_TRY(skm_set_displacement_field(ev, &desc->target, prev.data()));
        rep.ms_setup = ms_since(t_setup);

        // optimisation (SKMReg.cxx:375-430)
        const auto t_opt = Clock::now();
        const bool only_l1 = P.reg_l1 > 0 && P.reg_l21 <= 0 && P.reg_l2 <= 0 && P.reg_rkhs <= 0 && P.reg_rd <= 0 &&
                             P.reg_pg <= 0;
        const bool only_l21 = P.reg_l1 <= 0 && P.reg_l21 > 0 && P.reg_l2 <= 0 && P.reg_rkhs <= 0 && P.reg_rd <= 0 &&
                              P.reg_pg <= 0;
        skm_metric_options mo{desc->metric, P.reg_l1, P.reg_l21, P.reg_l2, 1, 1};
        skm_optimizer_options oo{SKM_OPT_ADAPTIVE, P.initial_step_size, 20.0, 1.0, 1e7, -1e6, 1.0,
                                 (only_l1 || only_l21) ? 1 : 0, P.iterations, 1};
        std::vector<float> params((size_t)n_cp * dim, 0.0f), cpw((size_t)n_cp, 1.0f);
        std::vector<skm_iteration_record> rec(std::max(1, P.iterations));
        SKM_R_TRY(skm_optimize(ev, &mo, &oo, cpw.data(), params.data(), rec.data()));
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

        // displacement of all levels so far (SKMReg.cxx:436-447)
        const auto t_df = Clock::now();
        std::vector<float> cp_aos((size_t)n_cp * dim);
        for (long long j = 0; j < n_cp; ++j)
            for (int d = 0; d < dim; ++d) cp_aos[j * dim + d] = params[d * n_cp + j];
        std::vector<float> df((size_t)n_tar * dim);
        SKM_R_TRY(skm_evaluate_displacement_field(ev, cp_aos.data(), cpw.data(), df.data()));
        prev.swap(df);
        rep.ms_displacement = ms_since(t_df);
        if (reports) reports[scale] = rep;
    }
    if (prev.empty()) prev.assign((size_t)n_tar * dim, 0.0f);
    std::memcpy(displacement_out, prev.data(), sizeof(float) * prev.size());
    return SKM_SUCCESS;
}

}  // extern "C"
