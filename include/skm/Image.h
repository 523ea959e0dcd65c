// skm/Image.h -- minimal image container with ITK-style geometry accessors.
//
// ITK 4.10 is not a dependency of this library; the classes of the SKM ITK
// layer (TransformAdapter, RegularizedImageToImageMetricv4, the optimizers)
// are mirrored in this directory over this type.  Geometry is origin +
// spacing * index (direction cosines are not used by the path, itkGlue.h:59-63).
#pragma once

#include <array>
#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace skm {

// itk::ExceptionObject stand-in: every validation error of the mirrored
// classes is thrown as this type with the reference's message.
class ExceptionObject : public std::runtime_error {
public:
    explicit ExceptionObject(const std::string &what) : std::runtime_error(what) {}
};

template <typename TPixel, unsigned D>
class Image {
public:
    using PixelType = TPixel;
    using SizeType = std::array<long, D>;
    using PointType = std::array<double, D>;
    using SpacingType = std::array<double, D>;
    using IndexType = std::array<long, D>;
    using Pointer = std::shared_ptr<Image>;
    static constexpr unsigned ImageDimension = D;

    static Pointer New() { return std::make_shared<Image>(); }

    void SetRegions(const SizeType &size)
    {
        size_ = size;
    }
    const SizeType &GetSize() const { return size_; }
    void SetOrigin(const PointType &o) { origin_ = o; }
    const PointType &GetOrigin() const { return origin_; }
    void SetSpacing(const SpacingType &s) { spacing_ = s; }
    const SpacingType &GetSpacing() const { return spacing_; }
    void Allocate() { data_.assign(GetNumberOfPixels(), TPixel()); }
    void FillBuffer(const TPixel &v) { std::fill(data_.begin(), data_.end(), v); }
    std::size_t GetNumberOfPixels() const
    {
        std::size_t n = 1;
        for (unsigned d = 0; d < D; ++d) n *= (std::size_t)size_[d];
        return n;
    }
    TPixel *GetBufferPointer() { return data_.data(); }
    const TPixel *GetBufferPointer() const { return data_.data(); }
    std::size_t ComputeOffset(const IndexType &idx) const
    {
        std::size_t off = 0, stride = 1;
        for (unsigned d = 0; d < D; ++d) {
            off += stride * (std::size_t)idx[d];
            stride *= (std::size_t)size_[d];
        }
        return off;
    }
    TPixel &GetPixel(const IndexType &idx) { return data_[ComputeOffset(idx)]; }
    const TPixel &GetPixel(const IndexType &idx) const { return data_[ComputeOffset(idx)]; }
    void SetPixel(const IndexType &idx, const TPixel &v) { data_[ComputeOffset(idx)] = v; }
    PointType TransformIndexToPhysicalPoint(const IndexType &idx) const
    {
        PointType p;
        for (unsigned d = 0; d < D; ++d) p[d] = origin_[d] + spacing_[d] * (double)idx[d];
        return p;
    }
    bool IsInside(const IndexType &idx) const
    {
        for (unsigned d = 0; d < D; ++d)
            if (idx[d] < 0 || idx[d] >= size_[d]) return false;
        return true;
    }
    void CopyInformation(const Image &o)
    {
        size_ = o.size_;
        origin_ = o.origin_;
        spacing_ = o.spacing_;
    }
    std::vector<TPixel> &Buffer() { return data_; }
    const std::vector<TPixel> &Buffer() const { return data_; }

private:
    SizeType size_{};
    PointType origin_{};
    SpacingType spacing_ = [] {
        SpacingType s;
        s.fill(1.0);
        return s;
    }();
    std::vector<TPixel> data_;
};

}  // namespace skm
