// skm/GpuEvaluator.h -- C++ class over the C ABI with the member functions of
// the reference's GpuKernelEvaluator (include/cuda/GpuEvaluator.h:69-179).
//
// The reference's default build uses ScalarType = double for parameters and
// derivatives (CommonConfig.h:29-35); this class keeps that interface and
// converts to the library's float arrays.  Errors throw skm::ExceptionObject.
#pragma once

#include <cstring>
#include <string>
#include <vector>

#include "../skm.h"
#include "Image.h"

namespace skm {

inline void check_status(skm_status st, const char *what)
{
    if (st != SKM_SUCCESS) throw ExceptionObject(std::string(what) + ": " + skm_last_error());
}

class GpuKernelEvaluator {
public:
    enum class MEASURE { MSE = SKM_MEASURE_MSE, NCC = SKM_MEASURE_NCC, LCC = SKM_MEASURE_LCC };

    struct EvaluationResult {
        double measure;
        double reg_value_rkhs;
        double reg_value_rd;
        double reg_value_pg;
        unsigned numPixelsCountedGpu;
        double reg_value_bilateral;  // no reference counterpart (SetRegularizerBilateral)
    };

    explicit GpuKernelEvaluator(const skm_desc &desc) : dim_(desc.dim)
    {
        check_status(skm_create(&desc, &ev_), "GpuKernelEvaluator");
        n_cp_ = 1;
        for (int d = 0; d < dim_; ++d) n_cp_ *= desc.cp.size[d];
        n_vox_ = 1;
        for (int d = 0; d < dim_; ++d) n_vox_ *= desc.fixed.size[d];
    }
    ~GpuKernelEvaluator()
    {
        if (ev_) skm_destroy(ev_);
    }
    GpuKernelEvaluator(const GpuKernelEvaluator &) = delete;
    GpuKernelEvaluator &operator=(const GpuKernelEvaluator &) = delete;

    void SetRegularizerRKHS(double weight) { check_status(skm_set_regularizer_rkhs(ev_, weight), "SetRegularizerRKHS"); }
    void SetRegularizerRD(double weight, double scaling = 1.0)
    {
        check_status(skm_set_regularizer_rd(ev_, weight, scaling), "SetRegularizerRD");
    }
    void SetRegularizerPG(double weight, double scaling = 1.0)
    {
        check_status(skm_set_regularizer_pg(ev_, weight, scaling), "SetRegularizerPG");
    }
    // the bilateral regularizer of skm_set_regularizer_bilateral (no reference counterpart)
    void SetRegularizerBilateral(double weight, double spatial_sigma = 12.0, double intensity_sigma = 0.1)
    {
        check_status(skm_set_regularizer_bilateral(ev_, weight, spatial_sigma, intensity_sigma),
                     "SetRegularizerBilateral");
    }

    // Previous-scale field (AoS doubles on `grid`); nullptr clears it.  The
    // reference caches the device copy by host pointer (GpuEvaluator.cu:120-126).
    void SetDisplacementField(const skm_grid *grid, const double *df_aos)
    {
        if (df_aos == df_key_) return;
        df_key_ = df_aos;
        if (!grid || !df_aos) {
            check_status(skm_set_displacement_field(ev_, nullptr, nullptr), "displacement field");
            return;
        }
        long long n = 1;
        for (int d = 0; d < dim_; ++d) n *= grid->size[d];
        std::vector<float> f(df_aos, df_aos + n * dim_);
        check_status(skm_set_displacement_field(ev_, grid, f.data()), "displacement field");
    }

    EvaluationResult getValue(MEASURE metric, const double *cpData, const double *cpwData, bool do_resampling = true)
    {
        to_float(cpData, cpwData);
        skm_result r;
        check_status(skm_get_value(ev_, (int)metric, cp_.data(), cpw_.data(), do_resampling, &r), "getValue");
        return convert(r);
    }

    EvaluationResult getValueAndDerivative(MEASURE metric, const double *cpData, const double *cpwData,
                                           double *derivatives, bool do_resampling = true)
    {
        to_float(cpData, cpwData);
        der_.resize((size_t)n_cp_ * dim_);
        skm_result r;
        check_status(skm_get_value_and_derivative(ev_, (int)metric, cp_.data(), cpw_.data(), der_.data(),
                                                  do_resampling, &r),
                     "getValueAndDerivative");
        for (size_t i = 0; i < der_.size(); ++i) derivatives[i] = der_[i];
        return convert(r);
    }

    void evaluateDisplacementField(const double *cpData, const double *cpwData, double *dst)
    {
        to_float(cpData, cpwData);
        std::vector<float> out((size_t)n_vox_ * dim_);
        check_status(skm_evaluate_displacement_field(ev_, cp_.data(), cpw_.data(), out.data()),
                     "evaluateDisplacementField");
        for (size_t i = 0; i < out.size(); ++i) dst[i] = out[i];
    }

    skm_evaluator *handle() { return ev_; }
    long long numberOfControlPoints() const { return n_cp_; }

private:
    void to_float(const double *cp, const double *cpw)
    {
        cp_.assign(cp, cp + (size_t)n_cp_ * dim_);
        cpw_.assign(cpw, cpw + (size_t)n_cp_);
    }
    static EvaluationResult convert(const skm_result &r)
    {
        return {r.measure, r.reg_value_rkhs, r.reg_value_rd, r.reg_value_pg, r.num_pixels_counted,
                r.reg_value_bilateral};
    }

    skm_evaluator *ev_ = nullptr;
    int dim_;
    long long n_cp_ = 0, n_vox_ = 0;
    std::vector<float> cp_, cpw_, der_;
    const double *df_key_ = nullptr;
};

}  // namespace skm
