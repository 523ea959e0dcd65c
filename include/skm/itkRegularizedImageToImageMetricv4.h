// skm/itkRegularizedImageToImageMetricv4.h -- the SKM metric, the summation
// transform and the two gradient-descent optimizers of the reference's ITK
// layer, over skm::Image and the C ABI.
//
//   RegularizedImageToImageMetricv4  itkRegularizedImageToImageMetricv4.{h,hxx}
//   SummationTransform               itkSummationTransform.{h,hxx}
//   StandardGradientDescentOptimizerv4, AdaptiveStepGradientDescentOptimizerv4
//                                    itk{Standard,AdaptiveStep}GradientDescentOptimizerv4.{h,hxx}
#pragma once

#include <cmath>
#include <fstream>
#include <functional>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "itkTransformAdapter.h"

namespace skm {

// A previous-scale displacement field (front) followed by the SKM transform
// (back); TransformPoint applies the queue in reverse (itkSummationTransform.hxx:43-64).
// The field is looked up nearest-neighbour, as the GPU evaluator does
// (ImageNearest, GpuEvaluatorMSE.cu:45-46).
template <typename TScalar, unsigned N>
class SummationTransform {
public:
    using Adapter = TransformAdapter<TScalar, N>;
    using FieldImage = Image<std::array<TScalar, N>, N>;
    using Pointer = std::shared_ptr<SummationTransform>;
    using PointType = std::array<double, N>;

    static Pointer New() { return Pointer(new SummationTransform()); }
    void AddDisplacementField(typename FieldImage::Pointer df) { field_ = df; }
    void AddTransform(typename Adapter::Pointer t) { back_ = t; }
    Adapter *GetBackTransform() const { return back_.get(); }
    FieldImage *GetDisplacementField() const { return field_.get(); }
    unsigned GetNumberOfTransforms() const { return (field_ ? 1u : 0u) + (back_ ? 1u : 0u); }

    PointType TransformPoint(const PointType &p) const
    {
        PointType out = back_ ? back_->TransformPoint(p) : p;
        if (field_) {
            typename FieldImage::IndexType idx;
            bool in = true;
            for (unsigned d = 0; d < N; ++d) {
                idx[d] = (long)((out[d] - field_->GetOrigin()[d]) / field_->GetSpacing()[d] + 0.5);
                in = in && idx[d] >= 0 && idx[d] < field_->GetSize()[d];
            }
            if (in)
                for (unsigned d = 0; d < N; ++d) out[d] += field_->GetPixel(idx)[d];
        }
        return out;
    }

private:
    typename FieldImage::Pointer field_;
    typename Adapter::Pointer back_;
};

template <typename TScalar, unsigned N>
class RegularizedImageToImageMetricv4 {
public:
    using Adapter = TransformAdapter<TScalar, N>;
    using Summation = SummationTransform<TScalar, N>;
    using ImageType = Image<TScalar, N>;
    using MeasureType = double;
    using DerivativeType = std::vector<double>;
    using ParametersType = std::vector<TScalar>;
    using TemporaryValuesType = std::vector<double>;
    using PointType = std::array<double, N>;
    using PointListType = std::vector<PointType>;
    using Pointer = std::shared_ptr<RegularizedImageToImageMetricv4>;

    static Pointer New() { return Pointer(new RegularizedImageToImageMetricv4()); }

    void SetFixedImage(typename ImageType::Pointer i) { fixed_ = i; }
    void SetMovingImage(typename ImageType::Pointer i) { moving_ = i; }
    void SetTransform(typename Adapter::Pointer t)
    {
        adapter_ = t;
        sum_.reset();
    }
    void SetTransform(typename Summation::Pointer s)
    {
        sum_ = s;
        adapter_.reset();
    }

#define SKM_ACCESSOR(Name, type, member)                                                                          \
    void Set##Name(type v) { member = v; }                                                                         \
    type Get##Name() const { return member; }
    SKM_ACCESSOR(RegularizerL1, double, l1_)
    SKM_ACCESSOR(RegularizerL21, double, l21_)
    SKM_ACCESSOR(RegularizerL2, double, l2_)
    SKM_ACCESSOR(RegularizerRKHS, double, rkhs_)
    SKM_ACCESSOR(RegularizerRD, double, rd_)
    SKM_ACCESSOR(RegularizerPG, double, pg_)
    SKM_ACCESSOR(RegularizerRDScaling, double, rd_scaling_)
    SKM_ACCESSOR(RegularizerPGScaling, double, pg_scaling_)
    // bilateral regularizer (skm_set_regularizer_bilateral; no reference counterpart)
    SKM_ACCESSOR(RegularizerBilateral, double, bl_)
    SKM_ACCESSOR(RegularizerBilateralSpatialSigma, double, bl_sigma_s_)
    SKM_ACCESSOR(RegularizerBilateralIntensitySigma, double, bl_sigma_i_)
    SKM_ACCESSOR(UseNccMetric, bool, use_ncc_)
    SKM_ACCESSOR(UseLccMetric, bool, use_lcc_)
    SKM_ACCESSOR(UseNegativeGradient, bool, negative_)
    SKM_ACCESSOR(DoResampling, bool, resampling_)
    SKM_ACCESSOR(SubSample, unsigned, subsample_)
    SKM_ACCESSOR(SubSampleNeighborhood, unsigned, subsample_neighborhood_)
    SKM_ACCESSOR(OutputFilename, std::string, output_filename_)
    SKM_ACCESSOR(Verbosity, unsigned, verbosity_)
#undef SKM_ACCESSOR
    void UseNccMetricOn() { use_ncc_ = true; }
    void UseNccMetricOff() { use_ncc_ = false; }
    void UseLccMetricOn() { use_lcc_ = true; }
    void UseLccMetricOff() { use_lcc_ = false; }
    void UseNegativeGradientOn() { negative_ = true; }
    void UseNegativeGradientOff() { negative_ = false; }
    void DoResamplingOn() { resampling_ = true; }
    void DoResamplingOff() { resampling_ = false; }

    void SetFixedLandmarks(const PointListType &l) { fixed_landmarks_ = l; }
    void SetMovingLandmarks(const PointListType &l) { moving_landmarks_ = l; }

    template <typename Printer>
    void SetPrintFunction(Printer &&p)
    {
        print_ = p;
    }

    // Initialize (itkRegularizedImageToImageMetricv4.hxx:71-106)
    void Initialize()
    {
        if (l1_ < 0 || l21_ < 0 || l2_ < 0 || rkhs_ < 0 || rd_ < 0 || pg_ < 0 || bl_ < 0)
            throw ExceptionObject("Regularization weights must be positive.");
        if (rd_scaling_ <= 0 || pg_scaling_ <= 0)
            throw ExceptionObject("Robust regularization scales must be strictly positive.");
        if (subsample_ == 0 || subsample_ > 100) throw ExceptionObject("Sampling rate inproper.");
        if (subsample_neighborhood_ == 0 || subsample_neighborhood_ > 100)
            throw ExceptionObject("Neighborhood sampling rate inproper.");
        if (!fixed_) throw ExceptionObject("Fixed image not set.");
        if (!moving_) throw ExceptionObject("Moving image not set.");
    }

    Adapter *GetAdapter() const
    {
        Adapter *t = sum_ ? sum_->GetBackTransform() : adapter_.get();
        if (!t) throw ExceptionObject("Transform type not supported.");
        return t;
    }
    std::size_t GetNumberOfParameters() const { return GetAdapter()->GetNumberOfParameters(); }
    ParametersType GetParameters() const { return GetAdapter()->GetParameters(); }
    void SetParameters(const ParametersType &p) { GetAdapter()->SetParameters(p); }

    MeasureType GetValue() const
    {
        MeasureType v;
        evaluate(v, nullptr);
        return v;
    }
    void GetDerivative(DerivativeType &d) const
    {
        MeasureType v;
        GetValueAndDerivative(v, d);
    }
    // GetValueAndDerivative (itkRegularizedImageToImageMetricv4.hxx:316-573)
    void GetValueAndDerivative(MeasureType &value, DerivativeType &derivative) const
    {
        if (derivative.size() != GetNumberOfParameters()) derivative.assign(GetNumberOfParameters(), 0.0);
        evaluate(value, &derivative);
    }

    // GetTRE (itkRegularizedImageToImageMetricv4.hxx:590-610)
    MeasureType GetTRE() const
    {
        if (fixed_landmarks_.empty() || moving_landmarks_.empty() || fixed_landmarks_.size() != moving_landmarks_.size())
            return -1;
        double tre = 0.0;
        for (std::size_t i = 0; i < fixed_landmarks_.size(); ++i) {
            PointType p = sum_ ? sum_->TransformPoint(fixed_landmarks_[i]) : adapter_->TransformPoint(fixed_landmarks_[i]);
            double n2 = 0.0;
            for (unsigned d = 0; d < N; ++d) n2 += (p[d] - moving_landmarks_[i][d]) * (p[d] - moving_landmarks_[i][d]);
            tre += std::sqrt(n2);
        }
        return tre / fixed_landmarks_.size();
    }

    GpuKernelEvaluator::MEASURE measure() const
    {
        if (!use_ncc_ && !use_lcc_) return GpuKernelEvaluator::MEASURE::MSE;
        if (use_ncc_ && !use_lcc_) return GpuKernelEvaluator::MEASURE::NCC;
        if (!use_ncc_ && use_lcc_) return GpuKernelEvaluator::MEASURE::LCC;
        throw ExceptionObject("Do not know which metric to execute");
    }

private:
    RegularizedImageToImageMetricv4() = default;

    void evaluate(MeasureType &value, DerivativeType *derivative) const
    {
        Adapter *t = GetAdapter();
        GpuKernelEvaluator *ev = t->GetGpuKernelEvaluator();
        if (!ev) throw ExceptionObject("Transform is not initialized.");
        const ParametersType &params = t->GetParameters();
        if (sum_ && sum_->GetDisplacementField()) {
            auto *df = sum_->GetDisplacementField();
            skm_grid g{};
            for (int d = 0; d < 3; ++d) {
                g.size[d] = d < (int)N ? (int)df->GetSize()[d] : 1;
                g.origin[d] = d < (int)N ? df->GetOrigin()[d] : 0.0;
                g.spacing[d] = d < (int)N ? df->GetSpacing()[d] : 1.0;
            }
            ev->SetDisplacementField(&g, reinterpret_cast<const double *>(df->GetBufferPointer()));
        } else {
            ev->SetDisplacementField(nullptr, nullptr);
        }
        if (rkhs_ > 0) ev->SetRegularizerRKHS(rkhs_);
        if (rd_ > 0) ev->SetRegularizerRD(rd_, rd_scaling_);
        if (pg_ > 0) ev->SetRegularizerPG(pg_, pg_scaling_);
        if (bl_ > 0) ev->SetRegularizerBilateral(bl_, bl_sigma_s_, bl_sigma_i_);
        const auto cp = t->GetControlPointImage();
        const auto cpw = t->GetControlPointWeightImage();
        std::vector<double> cpd(cp->GetNumberOfPixels() * N), cpwd(cpw->GetBufferPointer(),
                                                                    cpw->GetBufferPointer() + cpw->GetNumberOfPixels());
        for (std::size_t j = 0; j < cp->GetNumberOfPixels(); ++j)
            for (unsigned d = 0; d < N; ++d) cpd[j * N + d] = cp->GetBufferPointer()[j][d];
        GpuKernelEvaluator::EvaluationResult r;
        if (derivative)
            r = ev->getValueAndDerivative(measure(), cpd.data(), cpwd.data(), derivative->data(), resampling_);
        else
            r = ev->getValue(measure(), cpd.data(), cpwd.data(), resampling_);
        std::vector<double> pd(params.begin(), params.end());
        double reg_l = 0.0, v1 = 0.0, v21 = 0.0, v2 = 0.0;
        check_status(skm_host_regularizers((int)N, (int64_t)cp->GetNumberOfPixels(), pd.data(), l1_, l21_, l2_,
                                           derivative && negative_, derivative ? derivative->data() : nullptr,
                                           &reg_l, &v1, &v21, &v2),
                     "host regularizers");
        double reg = reg_l;
        if (rkhs_ > 0) reg += rkhs_ * r.reg_value_rkhs;
        if (rd_ > 0) reg += rd_ * r.reg_value_rd;
        if (pg_ > 0) reg += pg_ * r.reg_value_pg;
        if (bl_ > 0) reg += bl_ * r.reg_value_bilateral;
        if (verbosity_ >= 1) {
            std::stringstream s;
            s << "metric: " << r.measure + reg << ",\t value: " << r.measure;
            if (l1_ > 0) s << ",\t l1: " << l1_ * v1;
            if (l21_ > 0) s << ",\t l21: " << l21_ * v21;
            if (l2_ > 0) s << ",\t l2: " << l2_ * v2;
            if (rkhs_ > 0) s << ",\t rkhs: " << rkhs_ * r.reg_value_rkhs;
            if (rd_ > 0) s << ",\t rd: " << rd_ * r.reg_value_rd;
            if (pg_ > 0) s << ",\t pg: " << pg_ * r.reg_value_pg;
            if (bl_ > 0) s << ",\t bilateral: " << bl_ * r.reg_value_bilateral;
            if (verbosity_ >= 2) std::cout << s.str() << std::endl;
            std::ofstream fs(output_filename_, std::ios::app);
            fs << s.str() << std::endl;
        }
        if (print_)
            print_(TemporaryValuesType{(double)r.numPixelsCountedGpu, r.measure, reg, v1, v21, v2, r.reg_value_rkhs,
                                       r.reg_value_rd, r.reg_value_pg});
        value = r.measure + reg;
    }

    typename ImageType::Pointer fixed_, moving_;
    typename Adapter::Pointer adapter_;
    typename Summation::Pointer sum_;
    double l1_ = 0, l21_ = 0, l2_ = 0, rkhs_ = 0, rd_ = 0, pg_ = 0, rd_scaling_ = 1.0, pg_scaling_ = 1.0;
    double bl_ = 0, bl_sigma_s_ = 12.0, bl_sigma_i_ = 0.1;
    bool use_ncc_ = false, use_lcc_ = false, negative_ = true, resampling_ = true;
    unsigned subsample_ = N == 3 ? 4 : 1, subsample_neighborhood_ = N == 3 ? 4 : 1;
    std::string output_filename_ = "/tmp/output.txt";
    unsigned verbosity_ = 0;
    PointListType fixed_landmarks_, moving_landmarks_;
    std::function<void(const TemporaryValuesType &)> print_;
};

// Gradient descent with the step (a / (A + k + 1))^alpha and optional orthant
// projection (itkStandardGradientDescentOptimizerv4.hxx:83-166).  The update is
// theta_{k+1} = theta_k + eta_k d_k with d the metric derivative (negated by
// the metric by default).
template <typename TMetric>
class StandardGradientDescentOptimizerv4 {
public:
    using ParametersType = typename TMetric::ParametersType;
    using DerivativeType = typename TMetric::DerivativeType;

    virtual ~StandardGradientDescentOptimizerv4() = default;
    void SetMetric(std::shared_ptr<TMetric> m) { metric_ = m; }
    void Seta(double v) { a_ = v; }
    void SetA(double v) { A_ = v; }
    void Setalpha(double v) { alpha_ = v; }
    void SetOrthantProjection(bool v) { orthant_ = v; }
    void SetNumberOfIterations(unsigned n) { iterations_ = n; }
    void CalculateMagnitudeOn() { magnitude_ = true; }
    void CountZeroParamsOn() { count_zero_ = true; }
    // ITK's ReturnBestParametersAndValue: the result is the evaluated position
    // with the lowest value (SKMReg.cxx:383 switches it on)
    void ReturnBestParametersAndValueOn() { return_best_ = true; }
    void ReturnBestParametersAndValueOff() { return_best_ = false; }
    void SetIterationCallback(std::function<void(unsigned, double)> cb) { callback_ = cb; }
    unsigned GetCurrentIteration() const { return iteration_; }
    double GetValue() const { return value_; }
    double GetLearningRate() const { return rate_; }
    const ParametersType &GetCurrentPosition() const { return position_; }

    void StartOptimization()
    {
        if (!metric_) throw ExceptionObject("m_Metric must be set.");
        position_ = metric_->GetParameters();
        previous_gradient_.assign(position_.size(), 0.0);
        start();
        ParametersType best;
        double best_value = 0.0;
        for (iteration_ = 0; iteration_ < iterations_; ++iteration_) {
            metric_->GetValueAndDerivative(value_, gradient_);
            if (return_best_ && (iteration_ == 0 || value_ < best_value)) {
                best_value = value_;
                best = position_;
            }
            EstimateLearningRate();
            ModifyGradientByLearningRate();
            metric_->SetParameters(position_);
            if (callback_) callback_(iteration_, value_);
        }
        if (return_best_ && iterations_ > 0) {
            position_ = best;
            value_ = best_value;
            metric_->SetParameters(position_);
        }
    }

    // The same loop on the device in one call (skm_optimize); no host round trip per step.
    void StartOptimizationOnDevice(int kind)
    {
        if (!metric_) throw ExceptionObject("m_Metric must be set.");
        auto *t = metric_->GetAdapter();
        std::vector<float> p(t->GetParameters().begin(), t->GetParameters().end());
        const auto cpw = t->GetControlPointWeightImage();
        std::vector<float> w(cpw->GetBufferPointer(), cpw->GetBufferPointer() + cpw->GetNumberOfPixels());
        skm_metric_options mo{(int)metric_->measure(), metric_->GetRegularizerL1(), metric_->GetRegularizerL21(),
                              metric_->GetRegularizerL2(), metric_->GetUseNegativeGradient(),
                              metric_->GetDoResampling()};
        skm_optimizer_options oo{kind, a_, A_, alpha_, fmax_, fmin_, omega_, orthant_, (int)iterations_, return_best_,
                                 0};
        std::vector<skm_iteration_record> rec(iterations_ ? iterations_ : 1);
        auto *ev = t->GetGpuKernelEvaluator();
        if (metric_->GetRegularizerRKHS() > 0) ev->SetRegularizerRKHS(metric_->GetRegularizerRKHS());
        if (metric_->GetRegularizerRD() > 0) ev->SetRegularizerRD(metric_->GetRegularizerRD(), metric_->GetRegularizerRDScaling());
        if (metric_->GetRegularizerPG() > 0) ev->SetRegularizerPG(metric_->GetRegularizerPG(), metric_->GetRegularizerPGScaling());
        if (metric_->GetRegularizerBilateral() > 0)
            ev->SetRegularizerBilateral(metric_->GetRegularizerBilateral(), metric_->GetRegularizerBilateralSpatialSigma(),
                                        metric_->GetRegularizerBilateralIntensitySigma());
        check_status(skm_optimize(ev->handle(), &mo, &oo, w.data(), p.data(), rec.data()), "skm_optimize");
        position_.assign(p.begin(), p.end());
        metric_->SetParameters(position_);
        iteration_ = iterations_;
        value_ = iterations_ ? rec[iterations_ - 1].value : 0.0;
        if (return_best_)
            for (unsigned k = 0; k < iterations_; ++k) value_ = k == 0 ? rec[0].value : std::min(value_, rec[k].value);
        if (callback_)
            for (unsigned k = 0; k < iterations_; ++k) callback_(k, rec[k].value);
    }

protected:
    virtual void start() {}

    // EstimateLearningRate (Standard...hxx:159-166)
    virtual void EstimateLearningRate() { rate_ = std::pow(a_ / (A_ + iteration_ + 1), alpha_); }

    // ModifyGradientByLearningRateOverSubRange (Standard...hxx:116-157)
    virtual void ModifyGradientByLearningRate()
    {
        if (magnitude_) {
            double m = 0.0;
            for (double g : gradient_) m += g * g;
            std::cout << "\t magnitude: " << std::sqrt(m) << std::flush;
        }
        for (std::size_t j = 0; j < position_.size(); ++j) {
            double prev = position_[j];
            double v = prev + gradient_[j] * rate_;
            if (orthant_ && v * prev < 0) v = 0;
            position_[j] = v;
        }
        if (count_zero_) {
            std::size_t z = 0;
            for (auto v : position_) z += v == 0;
            std::cout << "\t zero params: " << 100.0 * z / position_.size() << "%" << std::flush;
        }
    }

    std::shared_ptr<TMetric> metric_;
    double a_ = 1, A_ = 5, alpha_ = 1, rate_ = 0, value_ = 0;
    double fmax_ = 1, fmin_ = -0.5, omega_ = 1;
    bool orthant_ = false, magnitude_ = false, count_zero_ = false, return_best_ = false;
    unsigned iterations_ = 100, iteration_ = 0;
    ParametersType position_;
    DerivativeType gradient_, previous_gradient_;
    std::function<void(unsigned, double)> callback_;
};

// Adaptive step a / (A + t + 1)^alpha with t <- max(0, t + f(-<g_k, g_{k-1}>))
// (itkAdaptiveStepGradientDescentOptimizerv4.hxx:71-128).
template <typename TMetric>
class AdaptiveStepGradientDescentOptimizerv4 : public StandardGradientDescentOptimizerv4<TMetric> {
    using Base = StandardGradientDescentOptimizerv4<TMetric>;

public:
    void SetFmax(double v) { this->fmax_ = v; }
    void SetFmin(double v) { this->fmin_ = v; }
    void Setomega(double v) { this->omega_ = v; }
    double GetCurrentTime() const { return time_; }

protected:
    void start() override
    {
        time_ = 0.0;
        temporary_gradient_.assign(this->position_.size(), 0.0);
    }
    void EstimateLearningRate() override { this->rate_ = this->a_ / std::pow(this->A_ + time_ + 1, this->alpha_); }
    void ModifyGradientByLearningRate() override
    {
        double sp = 0.0;
        for (std::size_t j = 0; j < this->gradient_.size(); ++j) sp += this->gradient_[j] * temporary_gradient_[j];
        temporary_gradient_ = this->gradient_;
        Base::ModifyGradientByLearningRate();
        double x = -sp;
        double f = this->fmin_ + (this->fmax_ - this->fmin_) / (1 - (this->fmax_ / this->fmin_) * std::exp(-x / this->omega_));
        time_ = std::max(0.0, time_ + f);
    }

private:
    double time_ = 0.0;
    std::vector<double> temporary_gradient_;
};

}  // namespace skm
