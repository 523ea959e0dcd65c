// C++ test of the ITK-style layer (include/skm/*.h) over libskm.so, modelled
// on the reference's test/itkFiniteDifferenceTest.cpp: identity and random
// parameters, GetValue == GetValueAndDerivative, finite differences of the
// value against the analytic derivative, plus TransformPoint against the
// dense GPU field and the host / device optimizer loops.  Exit code 0 = pass.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "skm/ConfigUtils.h"
#include "skm/itkRegularizedImageToImageMetricv4.h"
#include "skm/itkTransformAdapter.h"

using namespace skm;

constexpr unsigned D = 3;
using ImageType = Image<double, D>;
using TransformType = TransformAdapter<double, D>;
using MetricType = RegularizedImageToImageMetricv4<double, D>;

static int failures = 0;
#define EXPECT(cond, ...)                                                                                          \
    do {                                                                                                           \
        if (!(cond)) {                                                                                             \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__);                                                       \
            std::printf(__VA_ARGS__);                                                                              \
            std::printf("\n");                                                                                     \
            ++failures;                                                                                            \
        }                                                                                                          \
    } while (0)

static ImageType::Pointer blobs(const ImageType::SizeType &size, unsigned seed, double shift)
{
    auto img = ImageType::New();
    img->SetRegions(size);
    img->Allocate();
    std::mt19937 gen(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    struct Blob { double c[3], s, a; };
    std::vector<Blob> b(10);
    for (auto &x : b) {
        for (unsigned d = 0; d < D; ++d) x.c[d] = u(gen) * (size[d] - 1);
        x.s = 3.0 + 4.0 * u(gen);
        x.a = 0.3 + 0.7 * u(gen);
    }
    ImageType::IndexType i;
    for (i[2] = 0; i[2] < size[2]; ++i[2])
        for (i[1] = 0; i[1] < size[1]; ++i[1])
            for (i[0] = 0; i[0] < size[0]; ++i[0]) {
                double v = 0.0;
                for (auto &x : b) {
                    double r2 = 0.0;
                    for (unsigned d = 0; d < D; ++d) {
                        double q = i[d] + (d == 0 ? shift : 0.0) - x.c[d];
                        r2 += q * q;
                    }
                    v += x.a * std::exp(-r2 / (2 * x.s * x.s));
                }
                img->SetPixel(i, v);
            }
    return img;
}

static TransformType::Pointer make_transform(ImageType::Pointer fixed, ImageType::Pointer moving, unsigned sub,
                                             unsigned subn)
{
    auto t = TransformType::New();
    const double sc0 = 16.0, sc4 = 12.0;
    t->SetSigmaC0(sc0);
    t->SetSigmaC4(sc4);
    std::array<double, D> spacing;
    spacing.fill(std::min(sc0, sc4) / 3.0);
    t->ComputeControlPointImage(fixed, spacing);
    t->SetReferenceImage(moving);
    t->SetTargetImage(fixed);
    t->SetSubsample(sub);
    t->SetSubsampleNeighborhood(subn);
    t->Initialize();
    return t;
}

static void fd_test(bool ncc, bool regularized = false)
{
    ImageType::SizeType size{{40, 36, 32}};
    auto fimage = blobs(size, 7, 0.0);
    auto mimage = fimage;  // itkFiniteDifferenceTest.cpp:36-42 registers an image to itself
    auto transform = make_transform(fimage, mimage, 2, 3);
    auto metric = MetricType::New();
    metric->UseNegativeGradientOff();
    metric->SetFixedImage(fimage);
    metric->SetMovingImage(mimage);
    metric->SetTransform(transform);
    if (ncc) metric->UseNccMetricOn();
    if (regularized) {
        // RD and the bilateral term (spatial sigma 2 control-point spacings,
        // intensity sigma 0.3) add their values and derivatives to the metric's
        metric->SetRegularizerRD(2e-3);
        metric->SetRegularizerBilateral(3e-3);
        metric->SetRegularizerBilateralSpatialSigma(8.0);
        metric->SetRegularizerBilateralIntensitySigma(0.3);
    }
    metric->SetSubSample(2);
    metric->SetSubSampleNeighborhood(3);
    metric->DoResamplingOff();
    metric->Initialize();

    std::mt19937 gen(1234);
    std::uniform_real_distribution<double> dis_real(-1.0, 1.0);
    MetricType::ParametersType r(transform->GetNumberOfParameters(), 0.0);
    std::uniform_int_distribution<std::size_t> dis_int(0, r.size() - 1);
    double val;
    MetricType::DerivativeType deriv;
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1)
            for (auto &v : r) v = dis_real(gen);
        transform->SetParameters(r);
        metric->GetValueAndDerivative(val, deriv);
        double val2 = metric->GetValue();
        EXPECT(val == val2, "GetValue %.12g != GetValueAndDerivative %.12g", val2, val);
    }
    if (ncc) EXPECT(std::abs(val) <= 1.0, "NCC out of range %g", val);
    double maxd = 0.0;
    for (double d : deriv) maxd = std::max(maxd, std::abs(d));
    double err_sum = 0.0;
    const double delta = 0.1;  // float builds use 0.1 (itkFiniteDifferenceTest.cpp:126-130)
    for (int i = 0; i < 20; ++i) {
        std::size_t idx = dis_int(gen);
        auto pp = r, pm = r;
        pp[idx] += delta;
        pm[idx] -= delta;
        transform->SetParameters(pp);
        double vp = metric->GetValue();
        transform->SetParameters(pm);
        double vm = metric->GetValue();
        double fd = (vp - vm) / (2 * delta);
        double err = std::abs(fd - deriv[idx]);
        err_sum += err;
        EXPECT(err <= 2e-2 * maxd + 1e-7, "FD %zu: fd %.6g derivative %.6g", idx, fd, deriv[idx]);
    }
    std::printf("%s%s: value %.8f, average |FD - derivative| %.3e (max |derivative| %.3e)\n", ncc ? "NCC" : "MSE",
                regularized ? " + RD + bilateral" : "", val, err_sum / 20, maxd);
    transform->SetParameters(r);
}

static void transform_point_test()
{
    ImageType::SizeType size{{30, 28, 26}};
    auto fimage = blobs(size, 3, 0.0);
    auto transform = make_transform(fimage, fimage, 1, 1);
    std::mt19937 gen(5);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    MetricType::ParametersType p(transform->GetNumberOfParameters());
    for (auto &v : p) v = u(gen);
    transform->SetParameters(p);
    auto cp = transform->GetControlPointImage();
    auto cpw = transform->GetControlPointWeightImage();
    std::vector<double> cpd(cp->GetNumberOfPixels() * D), df(fimage->GetNumberOfPixels() * D);
    for (std::size_t j = 0; j < cp->GetNumberOfPixels(); ++j)
        for (unsigned d = 0; d < D; ++d) cpd[j * D + d] = cp->GetBufferPointer()[j][d];
    transform->GetGpuKernelEvaluator()->evaluateDisplacementField(cpd.data(), cpw->GetBufferPointer(), df.data());
    double worst = 0.0;
    for (std::size_t v = 0; v < fimage->GetNumberOfPixels(); v += 97) {
        ImageType::IndexType i{{(long)(v % 30), (long)((v / 30) % 28), (long)(v / (30 * 28))}};
        auto x = fimage->TransformIndexToPhysicalPoint(i);
        auto y = transform->TransformPoint(x);
        for (unsigned d = 0; d < D; ++d) worst = std::max(worst, std::abs((y[d] - x[d]) - df[v * D + d]));
    }
    EXPECT(worst < 1e-5, "TransformPoint vs GPU displacement: %g", worst);
}

static void optimizer_test()
{
    ImageType::SizeType size{{36, 32, 28}};
    auto fimage = blobs(size, 11, 0.0);
    auto mimage = blobs(size, 11, 1.5);  // reference shifted by 1.5 voxels in x
    std::vector<double> host_final, device_final;
    double first = 0, last_host = 0, last_dev = 0;
    for (int device = 0; device < 2; ++device) {
        auto transform = make_transform(fimage, mimage, 1, 1);
        auto metric = MetricType::New();
        metric->SetFixedImage(fimage);
        metric->SetMovingImage(mimage);
        metric->SetTransform(transform);
        metric->SetSubSample(1);
        metric->SetRegularizerL2(1e-6);
        metric->Initialize();
        AdaptiveStepGradientDescentOptimizerv4<MetricType> opt;
        opt.SetMetric(metric);
        opt.SetNumberOfIterations(15);
        opt.Seta(2e4);  // MSE derivatives are O(1e-5) after the 1/N normalisation
        opt.SetA(20);
        opt.Setalpha(1);
        opt.SetFmax(1e7);
        opt.SetFmin(-1e6);
        opt.Setomega(1);
        std::vector<double> values;
        opt.SetIterationCallback([&](unsigned, double v) { values.push_back(v); });
        if (device) opt.StartOptimizationOnDevice(SKM_OPT_ADAPTIVE);
        else opt.StartOptimization();
        first = values.front();
        (device ? last_dev : last_host) = values.back();
        (device ? device_final : host_final).assign(opt.GetCurrentPosition().begin(), opt.GetCurrentPosition().end());
    }
    EXPECT(last_host < 0.5 * first, "host loop did not reduce the MSE: %g -> %g", first, last_host);
    double diff = 0.0, scale = 0.0;
    for (std::size_t i = 0; i < host_final.size(); ++i) {
        diff = std::max(diff, std::abs(host_final[i] - device_final[i]));
        scale = std::max(scale, std::abs(host_final[i]));
    }
    EXPECT(diff <= 1e-3 * scale, "device loop differs from host loop: %g (scale %g)", diff, scale);
    std::printf("optimizer: MSE %.6g -> %.6g (host), %.6g (device), max param diff %.3g\n", first, last_host, last_dev,
                diff);
}

static void validation_test()
{
    auto t = TransformType::New();
    t->SetSigmaC0(-1);
    ImageType::SizeType size{{8, 8, 8}};
    auto img = blobs(size, 1, 0);
    std::array<double, D> sp{{2, 2, 2}};
    t->ComputeControlPointImage(img, sp);
    t->SetReferenceImage(img);
    t->SetTargetImage(img);
    bool thrown = false;
    try {
        t->Initialize();
    } catch (const ExceptionObject &e) {
        thrown = std::string(e.what()).find("Kernel parameters") != std::string::npos;
    }
    EXPECT(thrown, "negative sigma accepted");
    auto m = MetricType::New();
    m->SetRegularizerL1(-1);
    thrown = false;
    try {
        m->Initialize();
    } catch (const ExceptionObject &e) {
        thrown = true;
    }
    EXPECT(thrown, "negative regularizer accepted");
    Config c;
    const char *argv[] = {"SKMReg", "sigma_C4", "[24, 12]", "metric", "\"ncc\""};
    c.apply_argv(5, argv);
    EXPECT(c.CURRENT("sigma_C4", 0) == 24 && c.CURRENT("sigma_C4", 5) == 12, "CURRENT clamp");
    EXPECT(c.text("metric") == "ncc", "string option");
    EXPECT(c.CURRENT("sampling_rate", 1) == 4, "default");
}

int main()
{
    validation_test();
    fd_test(false);
    fd_test(true);
    fd_test(false, true);
    transform_point_test();
    optimizer_test();
    std::printf(failures ? "FAILED (%d)\n" : "OK\n", failures);
    return failures ? 1 : 0;
}
