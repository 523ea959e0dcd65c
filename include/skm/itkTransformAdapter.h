// skm/itkTransformAdapter.h -- the SKM transform (itkTransformAdapter.h / .hxx of
// the reference) over skm::Image and the C ABI.
//
// Parameters are the SoA vector i + N(j + M(k + O d)) (GetFlatIndex,
// itkTransformAdapter.hxx:202-236); the control-point image holds them as
// D-vectors (AoS), which is what the evaluator receives.
#pragma once

#include <algorithm>
#include <cmath>
#include <limits>
#include <memory>
#include <vector>

#include "GpuEvaluator.h"
#include "Image.h"

namespace skm {

namespace detail {

inline double det_n(unsigned n, const double *m)
{
    if (n == 2) return m[0] * m[3] - m[2] * m[1];
    return m[0] * m[4] * m[8] + m[3] * m[7] * m[2] + m[6] * m[1] * m[5] - m[0] * m[7] * m[5] - m[6] * m[4] * m[2] -
           m[3] * m[1] * m[8];
}

// Largest eigenvalue of a symmetric matrix (closed forms; vnl's
// symmetric eigensystem in the reference, itkTransformAdapter.hxx:353-358).
inline double max_eig_sym(unsigned n, const double *m)
{
    if (n == 2) {
        double tr = m[0] + m[3], det = det_n(2, m);
        return 0.5 * (tr + std::sqrt(std::max(tr * tr - 4.0 * det, 0.0)));
    }
    double tri = m[1] * m[1] + m[2] * m[2] + m[5] * m[5];
    if (tri == 0.0) return std::max(m[0], std::max(m[4], m[8]));
    double q = (m[0] + m[4] + m[8]) / 3.0;
    double psq = 2.0 * tri + (m[0] - q) * (m[0] - q) + (m[4] - q) * (m[4] - q) + (m[8] - q) * (m[8] - q);
    double p = std::sqrt(psq / 6.0);
    double b[9] = {m[0] - q, m[1], m[2], m[3], m[4] - q, m[5], m[6], m[7], m[8] - q};
    double r = det_n(3, b) * (0.5 / (p * p * p));
    double phi = std::acos(std::min(1.0, std::max(-1.0, r))) / 3.0;
    return q + 2.0 * p * std::cos(phi);
}

inline double wendland(double t, bool c4)
{
    double f = std::max(0.0, 1.0 - t);
    if (!c4) return f * f;
    return f * f * f * f * f * f * (3.0 + 18.0 * t + 35.0 * t * t) / 3.0;
}

}  // namespace detail

template <typename TScalar, unsigned NDimensions>
class TransformAdapter {
public:
    static constexpr unsigned SpaceDimension = NDimensions;
    using Self = TransformAdapter;
    using Pointer = std::shared_ptr<Self>;
    using ImageType = Image<TScalar, NDimensions>;
    using VectorType = std::array<TScalar, NDimensions>;
    using ControlPointImageType = Image<VectorType, NDimensions>;
    using ControlPointWeightImageType = Image<TScalar, NDimensions>;
    using CovarianceMatrixType = std::array<TScalar, NDimensions * NDimensions>;
    using TensorImageType = Image<CovarianceMatrixType, NDimensions>;
    using ParametersType = std::vector<TScalar>;
    using PointType = std::array<double, NDimensions>;

    static Pointer New() { return Pointer(new Self()); }

    void SetSigmaC0(double v) { sigma_c0_ = v; }
    double GetSigmaC0() const { return sigma_c0_; }
    void SetSigmaC4(double v) { sigma_c4_ = v; }
    double GetSigmaC4() const { return sigma_c4_; }
    void SetSubsample(unsigned v) { subsample_ = v; }
    unsigned GetSubsample() const { return subsample_; }
    void SetSubsampleNeighborhood(unsigned v) { subsample_neighborhood_ = v; }
    unsigned GetSubsampleNeighborhood() const { return subsample_neighborhood_; }
    bool GetUseWeightImage() const { return use_weight_image_; }
    bool GetUseWeightTensor() const { return use_weight_tensor_; }
    double GetMaximumWeight() const { return maximum_weight_; }
    const CovarianceMatrixType &GetMaximumEigenvalueCovariance() const { return max_eig_cov_; }
    void SetExactGradient(bool v) { exact_gradient_ = v; }

    std::size_t GetNumberOfParameters() const { return cp_image_ ? cp_image_->GetNumberOfPixels() * NDimensions : 0; }

    // SetParameters (itkTransformAdapter.hxx:56-98)
    void SetParameters(const ParametersType &parameters)
    {
        if (!cp_image_ || cp_image_->GetNumberOfPixels() == 0) {
            parameters_.clear();
            return;
        }
        if (cp_image_->GetNumberOfPixels() * NDimensions != parameters.size())
            throw ExceptionObject("Number of provided parameters (" + std::to_string(parameters.size()) +
                                  ") does not correspond to #Dimensions times the number of control points (" +
                                  std::to_string(cp_image_->GetNumberOfPixels() * NDimensions) + ").");
        parameters_ = parameters;
        const std::size_t n = cp_image_->GetNumberOfPixels();
        VectorType *cp = cp_image_->GetBufferPointer();
        for (std::size_t j = 0; j < n; ++j)
            for (unsigned d = 0; d < NDimensions; ++d) cp[j][d] = parameters_[j + n * d];
    }

    // UpdateParameters / GetParameters (itkTransformAdapter.hxx:100-131)
    const ParametersType &GetParameters() const
    {
        const std::size_t n = cp_image_ ? cp_image_->GetNumberOfPixels() : 0;
        parameters_.assign(n * NDimensions, TScalar(0));
        const VectorType *cp = n ? cp_image_->GetBufferPointer() : nullptr;
        for (std::size_t j = 0; j < n; ++j)
            for (unsigned d = 0; d < NDimensions; ++d) parameters_[j + n * d] = cp[j][d];
        return parameters_;
    }

    // GetFlatIndex (itkTransformAdapter.hxx:202-236)
    unsigned long GetFlatIndex(const typename ControlPointImageType::IndexType &idx, unsigned dimension) const
    {
        const auto &s = cp_image_->GetSize();
        unsigned long r = (unsigned long)dimension;
        for (int d = (int)NDimensions - 1; d >= 0; --d) r = r * (unsigned long)s[d] + (unsigned long)idx[d];
        return r;
    }

    void SetParameter(const typename ControlPointImageType::IndexType &idx, const VectorType &v)
    {
        if (!cp_image_->IsInside(idx)) throw ExceptionObject("Index is outside the control points region");
        cp_image_->SetPixel(idx, v);
    }

    typename ControlPointImageType::Pointer GetControlPointImage() const { return cp_image_; }
    void SetControlPointImage(typename ControlPointImageType::Pointer image) { cp_image_ = image; }
    typename ControlPointWeightImageType::Pointer GetControlPointWeightImage() const { return cpw_image_; }

    // SetControlPointWeightImage (itkTransformAdapter.hxx:178-192)
    void SetControlPointWeightImage(typename ControlPointWeightImageType::Pointer w)
    {
        if (cp_image_->GetNumberOfPixels() != w->GetNumberOfPixels())
            throw ExceptionObject("Control point weight image must have the same amount of pixels than the control point image.");
        if (cp_image_->GetSize() != w->GetSize())
            throw ExceptionObject("The largest possible region of the control point weight and the control point image must be equal.");
        cpw_image_ = w;
    }

    // ComputeControlPointImage (itkTransformAdapter.hxx:246-308)
    void ComputeControlPointImage(typename ImageType::Pointer reference_image, const std::array<double, NDimensions> &sigma)
    {
        skm_grid img{}, cp{};
        fill_grid(*reference_image, img);
        double sp[3] = {1.0, 1.0, 1.0};
        for (unsigned d = 0; d < NDimensions; ++d) sp[d] = sigma[d];
        check_status(skm_control_point_grid((int)NDimensions, &img, sp, &cp), "ComputeControlPointImage");
        cp_image_ = ControlPointImageType::New();
        typename ControlPointImageType::SizeType size;
        typename ControlPointImageType::PointType origin;
        typename ControlPointImageType::SpacingType spacing;
        for (unsigned d = 0; d < NDimensions; ++d) {
            size[d] = cp.size[d];
            origin[d] = cp.origin[d];
            spacing[d] = cp.spacing[d];
        }
        cp_image_->SetRegions(size);
        cp_image_->SetOrigin(origin);
        cp_image_->SetSpacing(spacing);
        cp_image_->Allocate();
        cp_image_->FillBuffer(VectorType{});
        cpw_image_ = ControlPointWeightImageType::New();
        cpw_image_->SetRegions(size);
        cpw_image_->SetOrigin(origin);
        cpw_image_->SetSpacing(spacing);
        cpw_image_->Allocate();
        cpw_image_->FillBuffer(TScalar(1));
        GetParameters();
    }

    // SetWeightingImage (itkTransformAdapter.hxx:310-332)
    void SetWeightingImage(typename ImageType::Pointer weight_image)
    {
        use_weight_image_ = true;
        weight_image_ = weight_image;
        double maximum = std::numeric_limits<double>::lowest();
        for (const auto &v : weight_image->Buffer()) {
            if (maximum < v) maximum = v;
            if (v <= 1e-8) throw ExceptionObject("Weighting image has to be strictly positive everywhere.");
        }
        maximum_weight_ = maximum;
    }

    // SetWeightingTensor (itkTransformAdapter.hxx:334-366)
    void SetWeightingTensor(typename TensorImageType::Pointer tensor, TScalar maximum_tilt = -1)
    {
        use_weight_tensor_ = true;
        weight_tensor_ = tensor;
        if (maximum_tilt < 0) {
            double maximum = std::numeric_limits<double>::lowest();
            for (const auto &M : tensor->Buffer()) {
                double m[9];
                for (unsigned i = 0; i < NDimensions * NDimensions; ++i) m[i] = M[i];
                if (detail::det_n(NDimensions, m) <= 0)
                    throw ExceptionObject("Weighting tensors must be positive definite.");
                double lmax = std::abs(detail::max_eig_sym(NDimensions, m));
                if (lmax > maximum) {
                    maximum = lmax;
                    max_eig_cov_ = M;
                }
            }
            maximum_weight_ = maximum;
        } else {
            maximum_weight_ = maximum_tilt;
        }
    }

    void SetReferenceImage(typename ImageType::Pointer image) { reference_ = image; }
    void SetTargetImage(typename ImageType::Pointer image) { target_ = image; }

    // Initialize (itkTransformAdapter.hxx:386-462): validates and creates the evaluator.
    void Initialize()
    {
        if (!reference_) throw ExceptionObject("Reference image not set.");
        if (!target_) throw ExceptionObject("Target image not set.");
        if (!cp_image_) throw ExceptionObject("Control point image has not been computed jet.");
        if (!cpw_image_) throw ExceptionObject("Control point weight image has not been computed jet.");
        if (!weight_image_ && use_weight_image_) throw ExceptionObject("Weight image not provided.");
        if (!weight_tensor_ && use_weight_tensor_) throw ExceptionObject("Weight tensor not set or not constructed.");
        if (sigma_c0_ <= 0 || sigma_c4_ <= 0) throw ExceptionObject("Kernel parameters have to be positive.");
        skm_desc desc;
        skm_desc_init(&desc);
        desc.dim = (int)NDimensions;
        fill_grid(*target_, desc.fixed);
        fill_grid(*reference_, desc.moving);
        fill_grid(*cp_image_, desc.cp);
        fixed_f_.assign(target_->Buffer().begin(), target_->Buffer().end());
        moving_f_.assign(reference_->Buffer().begin(), reference_->Buffer().end());
        desc.fixed_data = fixed_f_.data();
        desc.moving_data = moving_f_.data();
        desc.sigma_c0 = sigma_c0_;
        desc.sigma_c4 = sigma_c4_;
        desc.maximum_weight = maximum_weight_;
        for (int i = 0; i < 9; ++i) desc.maximum_eigenvalue_covariance[i] = 0.0;
        for (unsigned i = 0; i < NDimensions; ++i)
            for (unsigned k = 0; k < NDimensions; ++k)
                desc.maximum_eigenvalue_covariance[i * 3 + k] = max_eig_cov_[i * NDimensions + k];
        if (use_weight_image_) {
            desc.kernel_mode = SKM_KERNEL_WEIGHT_IMAGE;
            fill_grid(*weight_image_, desc.weight);
            weight_f_.assign(weight_image_->Buffer().begin(), weight_image_->Buffer().end());
            desc.weight_data = weight_f_.data();
        } else if (use_weight_tensor_) {
            desc.kernel_mode = SKM_KERNEL_WEIGHT_TENSOR;
            fill_grid(*weight_tensor_, desc.weight);
            weight_f_.clear();
            for (const auto &M : weight_tensor_->Buffer()) weight_f_.insert(weight_f_.end(), M.begin(), M.end());
            desc.weight_data = weight_f_.data();
        }
        desc.subsample = (int)subsample_;
        desc.subsample_neighborhood = (int)subsample_neighborhood_;
        desc.flags = exact_gradient_ ? SKM_FLAG_EXACT_GRADIENT : 0u;
        evaluator_.reset();
        evaluator_.reset(new GpuKernelEvaluator(desc));
    }

    GpuKernelEvaluator *GetGpuKernelEvaluator() const { return evaluator_.get(); }

    // TransformPoint (itkTransformAdapter.hxx:472-587): host kernel sum at one point.
    PointType TransformPoint(const PointType &x) const
    {
        const std::size_t n = cp_image_->GetNumberOfPixels();
        const auto &cs = cp_image_->GetSize();
        const auto &co = cp_image_->GetOrigin();
        const auto &ch = cp_image_->GetSpacing();
        double sigmaP = 0.0, covP[9] = {0};
        double smin = std::min(sigma_c4_, sigma_c0_), support = smin;
        if (use_weight_image_) {
            sigmaP = lookup_weight(x);
            support = smin / std::sqrt(2.0 / (maximum_weight_ + sigmaP));
        } else if (use_weight_tensor_) {
            lookup_tensor(x, covP);
            double sum[9];
            for (unsigned i = 0; i < NDimensions * NDimensions; ++i) sum[i] = max_eig_cov_[i] + covP[i];
            support = smin / std::sqrt(2.0 / detail::max_eig_sym(NDimensions, sum));
        }
        long lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        for (unsigned d = 0; d < NDimensions; ++d) {
            double idx = (x[d] - co[d]) / ch[d];
            lo[d] = std::min(std::max((long)(idx - support / ch[d] + 1.0), 0L), cs[d] - 1);
            hi[d] = std::min(std::max((long)(idx + support / ch[d]), 0L), cs[d] - 1);
        }
        PointType out = x;
        typename ControlPointImageType::IndexType i{};
        const long z1 = NDimensions == 3 ? hi[NDimensions - 1] : 0;
        for (long iz = NDimensions == 3 ? lo[2] : 0; iz <= z1; ++iz)
            for (long iy = lo[1]; iy <= hi[1]; ++iy)
                for (long ix = lo[0]; ix <= hi[0]; ++ix) {
                    i[0] = ix;
                    i[1] = iy;
                    if (NDimensions == 3) i[NDimensions - 1] = iz;
                    PointType p = cp_image_->TransformIndexToPhysicalPoint(i);
                    double k = kernel(x, p, sigmaP, covP);
                    const VectorType &c = cp_image_->GetPixel(i);
                    double w = cpw_image_->GetPixel(i);
                    for (unsigned d = 0; d < NDimensions; ++d) out[d] += w * k * c[d];
                }
        (void)n;
        return out;
    }

private:
    TransformAdapter() { max_eig_cov_.fill(0); for (unsigned i = 0; i < NDimensions; ++i) max_eig_cov_[i * NDimensions + i] = 1; }

    template <typename Img>
    static void fill_grid(const Img &im, skm_grid &g)
    {
        for (int d = 0; d < 3; ++d) {
            g.size[d] = d < (int)NDimensions ? (int)im.GetSize()[d] : 1;
            g.origin[d] = d < (int)NDimensions ? im.GetOrigin()[d] : 0.0;
            g.spacing[d] = d < (int)NDimensions ? im.GetSpacing()[d] : 1.0;
        }
    }

    template <typename Img>
    static bool nearest(const Img &im, const PointType &p, typename Img::IndexType &idx)
    {
        for (unsigned d = 0; d < NDimensions; ++d) {
            idx[d] = (long)((p[d] - im.GetOrigin()[d]) / im.GetSpacing()[d] + 0.5);
            if (idx[d] < 0 || idx[d] >= im.GetSize()[d]) return false;
        }
        return true;
    }

    double lookup_weight(const PointType &p) const
    {
        typename ImageType::IndexType idx;
        return nearest(*weight_image_, p, idx) ? (double)weight_image_->GetPixel(idx) : 0.0;
    }

    void lookup_tensor(const PointType &p, double *cov) const
    {
        typename TensorImageType::IndexType idx;
        if (nearest(*weight_tensor_, p, idx)) {
            const auto &M = weight_tensor_->GetPixel(idx);
            for (unsigned i = 0; i < NDimensions * NDimensions; ++i) cov[i] = M[i];
        } else {
            for (unsigned i = 0; i < NDimensions * NDimensions; ++i) cov[i] = (i % (NDimensions + 1)) == 0 ? 1.0 : 0.0;
        }
    }

    // GKernel::evaluate in the three modes (Kernel.h:96-131, 173-186), FP64.
    double kernel(const PointType &p, const PointType &q, double sigmaP, const double *covP) const
    {
        double r2 = 0.0;
        for (unsigned d = 0; d < NDimensions; ++d) r2 += (p[d] - q[d]) * (p[d] - q[d]);
        if (use_weight_image_) {
            double sigmaQ = lookup_weight(q);
            double F = std::pow(std::abs(sigmaP), 0.25) * std::pow(std::abs(sigmaQ), 0.25) *
                       std::pow(std::abs((sigmaP + sigmaQ) / 2.0), -0.5);
            double r = std::sqrt(2.0 / (sigmaP + sigmaQ)) * std::sqrt(r2);
            return F * detail::wendland(r / sigma_c4_, true) * F * detail::wendland(r / sigma_c0_, false);
        }
        if (use_weight_tensor_) {
            double covQ[9], m[9], inv[9];
            lookup_tensor(q, covQ);
            const unsigned n = NDimensions;
            for (unsigned i = 0; i < n * n; ++i) m[i] = 0.5 * (covP[i] + covQ[i]);
            double det = detail::det_n(n, m);
            if (n == 2) {
                inv[0] = m[3] / det; inv[1] = -m[1] / det; inv[2] = -m[2] / det; inv[3] = m[0] / det;
            } else {
                inv[0] = (m[4] * m[8] - m[5] * m[7]) / det; inv[1] = (m[2] * m[7] - m[1] * m[8]) / det;
                inv[2] = (m[1] * m[5] - m[2] * m[4]) / det; inv[3] = (m[5] * m[6] - m[3] * m[8]) / det;
                inv[4] = (m[0] * m[8] - m[2] * m[6]) / det; inv[5] = (m[2] * m[3] - m[0] * m[5]) / det;
                inv[6] = (m[3] * m[7] - m[4] * m[6]) / det; inv[7] = (m[1] * m[6] - m[0] * m[7]) / det;
                inv[8] = (m[0] * m[4] - m[1] * m[3]) / det;
            }
            double v[3] = {0, 0, 0}, quad = 0.0;
            for (unsigned d = 0; d < n; ++d) v[d] = p[d] - q[d];
            for (unsigned i = 0; i < n; ++i)
                for (unsigned k = 0; k < n; ++k) quad += v[k] * inv[k * n + i] * v[i];
            double F = std::pow(detail::det_n(n, covP), 0.25) * std::pow(detail::det_n(n, covQ), 0.25) *
                       std::pow(det, -0.5);
            double r = std::sqrt(quad);
            return F * detail::wendland(r / sigma_c4_, true) * F * detail::wendland(r / sigma_c0_, false);
        }
        double r = std::sqrt(r2);
        return detail::wendland(r / sigma_c4_, true) * detail::wendland(r / sigma_c0_, false);
    }

    double sigma_c0_ = 1.0, sigma_c4_ = 1.0;
    bool use_weight_image_ = false, use_weight_tensor_ = false, exact_gradient_ = false;
    double maximum_weight_ = 1.0;
    CovarianceMatrixType max_eig_cov_{};
    unsigned subsample_ = 1, subsample_neighborhood_ = 1;
    typename ControlPointImageType::Pointer cp_image_;
    typename ControlPointWeightImageType::Pointer cpw_image_;
    typename ImageType::Pointer weight_image_, reference_, target_;
    typename TensorImageType::Pointer weight_tensor_;
    mutable ParametersType parameters_;
    std::vector<float> fixed_f_, moving_f_, weight_f_;
    std::unique_ptr<GpuKernelEvaluator> evaluator_;
};

}  // namespace skm
