/*
 * skm.h -- C ABI of the MI355X (gfx950) SKM kernel-sum / metric / regularizer
 * library (libskm.so).
 *
 * The boundary replaces the C++ class GpuKernelEvaluator of the reference
 * (include/cuda/GpuEvaluator.h:69-179).  Each entry point names the member it
 * replaces.  Arrays are plain pointers; host-pointer entry points copy in and
 * out, *_device entry points take device pointers and enqueue work on the
 * evaluator's HIP stream without synchronising.
 *
 * Layouts (as in the reference):
 *   images           float, x fastest, geometry = origin + spacing*index,
 *                    direction cosines are not part of the geometry
 *                    (include/itkGlue.h:59-63);
 *   control points   AoS float, D values per control point (GpuEvaluator.cu:104);
 *   derivative       SoA float, index i + N(j + M(k + O d))
 *                    (itkTransformAdapter.hxx:202-236, GpuEvaluatorMSE.cu:165-169);
 *   weight tensor    D*D floats per voxel, row-major.
 *
 * Errors: every call returns an skm_status; skm_last_error() describes the
 * last failure of the calling thread.  An evaluator is not re-entrant: calls
 * on one evaluator must not overlap (GpuKernelEvaluator is not either).
 */
#ifndef SKM_H
#define SKM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SKM_API __attribute__((visibility("default")))

typedef enum skm_status {
    SKM_SUCCESS = 0,
    SKM_ERROR_INVALID_ARGUMENT = 1,
    SKM_ERROR_HIP = 2,
    SKM_ERROR_OUT_OF_MEMORY = 3,
    SKM_ERROR_NOT_SUPPORTED = 4,
    SKM_ERROR_IO = 5
} skm_status;

/* GpuKernelEvaluator::MEASURE (GpuEvaluator.h:160) */
typedef enum skm_measure { SKM_MEASURE_MSE = 0, SKM_MEASURE_NCC = 1, SKM_MEASURE_LCC = 2 } skm_measure;

/* Kernel modes selected by GKernel::useWeightImage / useWeightTensor (Kernel.h:188-196). */
typedef enum skm_kernel_mode {
    SKM_KERNEL_STATIONARY = 0,
    SKM_KERNEL_WEIGHT_IMAGE = 1,
    SKM_KERNEL_WEIGHT_TENSOR = 2
} skm_kernel_mode;

/* SKM_FLAG_EXACT_GRADIENT: derivative windows that cover every sample in the
 * kernel support, so the derivative is the exact gradient of the value.  Off
 * by default: the reference windows are used, including the 2D subsampled
 * lower bound (GpuEvaluatorMSE.cu:267-273) and the LCC centre truncation
 * (GpuEvaluatorLCC.cu:239-245).  For LCC, centres that fail the reference's
 * derivative guard |u| > 1e-16 (GpuEvaluatorLCC.cu:517, a flat fixed or
 * moving neighbourhood) add cc to the value but nothing to the derivative in
 * both modes, so the exact gradient covers the other centres. */
#define SKM_FLAG_EXACT_GRADIENT 0x1u
/* moving_data holds the moving image's cubic B-spline coefficients as the
 * reference's BsplineImage takes them (BsplineImage::computeCoeffs,
 * BsplineImage.h:209-221, called by TransformAdapter::Initialize,
 * itkTransformAdapter.hxx:435-439): prefiltered, shifted by one and padded to
 * (n+3)^D with mirrored borders, x fastest; moving.size stays the image size
 * n.  skm_create then copies them as they are instead of prefiltering. */
#define SKM_FLAG_MOVING_COEFFICIENTS 0x2u

/* Origin, spacing and size of an image (ImageBase, include/cuda/ImageBase.h:25-50). */
typedef struct skm_grid {
    int32_t size[3];      /* size[2] is ignored when dim == 2 */
    double origin[3];
    double spacing[3];
} skm_grid;

/* Construction arguments of GpuKernelEvaluator (GpuEvaluator.h:147-154) and
 * of the GKernel it holds (Kernel.h:134-146). */
typedef struct skm_desc {
    int32_t dim;                              /* 2 or 3 (SpaceDimensions)              */
    skm_grid fixed;  const float *fixed_data; /* target image T                        */
    skm_grid moving; const float *moving_data;/* reference image R, raw intensities;   */
                                              /* the cubic B-spline prefilter runs     */
                                              /* inside skm_create (with               */
                                              /* SKM_FLAG_MOVING_COEFFICIENTS: the     */
                                              /* (n+3)^D coefficients instead)          */
    skm_grid cp;                              /* control-point grid                    */
    double sigma_c0, sigma_c4;                /* Wendland C0 / C4 supports             */
    int32_t kernel_mode;                      /* skm_kernel_mode                       */
    double maximum_weight;                    /* max of the weight image               */
    double maximum_eigenvalue_covariance[9];  /* 3x3 row-major; 2D uses the upper-left
                                                 2x2 block                              */
    skm_grid weight; const float *weight_data;/* weight image / tensor field, or NULL  */
    int32_t subsample;                        /* sampling rate (>= 1)                  */
    int32_t subsample_neighborhood;           /* LCC neighbourhood stride (>= 1)       */
    uint64_t rng_state, rng_sequence;         /* PCG32 state (UniformRng.h:32)         */
    uint32_t flags;                           /* SKM_FLAG_*                            */
} skm_desc;

/* GpuKernelEvaluator::EvaluationResult (GpuEvaluator.h:138-145). */
typedef struct skm_result {
    double measure;
    double reg_value_rkhs;
    double reg_value_rd;
    double reg_value_pg;
    uint32_t num_pixels_counted;
    uint32_t reserved;
    double reg_value_bilateral;   /* no reference counterpart (skm_set_regularizer_bilateral) */
} skm_result;

typedef struct skm_evaluator skm_evaluator;

/* Fills the reference defaults: subsample 1, neighbourhood 1, stationary kernel,
 * identity maximum covariance, PCG32 state 0xBA5EBA11DEADBEEF / sequence 0. */
SKM_API void skm_desc_init(skm_desc *desc);

/* GpuKernelEvaluator::GpuKernelEvaluator (GpuEvaluator.h:147-154, .cu:30-91),
 * including the host prefilter done by TransformAdapter::Initialize
 * (itkTransformAdapter.hxx:435-439). */
SKM_API skm_status skm_create(const skm_desc *desc, skm_evaluator **out);
SKM_API void skm_destroy(skm_evaluator *ev);
SKM_API const char *skm_last_error(void);

/* SetRegularizerRKHS / RD / PG (GpuEvaluator.h:156-158). */
SKM_API skm_status skm_set_regularizer_rkhs(skm_evaluator *ev, double weight);
SKM_API skm_status skm_set_regularizer_rd(skm_evaluator *ev, double weight, double scaling);
SKM_API skm_status skm_set_regularizer_pg(skm_evaluator *ev, double weight, double scaling);

/* Bilateral regularizer, no counterpart in the reference (whose closest term
 * is RD, GpuEvaluatorRegularizers.cu:96-166; a radiometric term appears only
 * as a comment, itkRegularizedImageToImageMetricv4.h:204).  Per control point
 * j over the points i within 3 spatial_sigma:
 *   B_ij    = exp(-|p_j - p_i|^2 / (2 spatial_sigma^2))
 *             * exp(-(I(p_j) - I(p_i))^2 / (2 intensity_sigma^2)),
 *   value   = sum_j sum_i w_j B_ij |c_j - c_i|^2   (skm_result.reg_value_bilateral),
 *   d/dc_j  = weight * sum_i 2 B_ij (w_j + w_i)(c_j - c_i), added to the derivative,
 * with I the fixed image at the control point's nearest voxel (0 outside) and
 * w the control-point weights.  weight 0 switches it off. */
SKM_API skm_status skm_set_regularizer_bilateral(skm_evaluator *ev, double weight, double spatial_sigma,
                                                double intensity_sigma);

/* Previous-scale displacement field (the ImageNearest<VecF> argument of
 * getValue / getValueAndDerivative / evaluateDisplacementField, cached by the
 * reference on pointer identity, GpuEvaluator.cu:120-126).  Host AoS data;
 * grid == NULL or data == NULL clears it. */
SKM_API skm_status skm_set_displacement_field(skm_evaluator *ev, const skm_grid *grid, const float *df_aos);

/* GpuKernelEvaluator::getValue (GpuEvaluator.h:162-166). */
SKM_API skm_status skm_get_value(skm_evaluator *ev, int32_t measure, const float *cp_aos,
                                 const float *cpw, int32_t do_resampling, skm_result *out);

/* GpuKernelEvaluator::getValueAndDerivative (GpuEvaluator.h:168-173). */
SKM_API skm_status skm_get_value_and_derivative(skm_evaluator *ev, int32_t measure, const float *cp_aos,
                                                const float *cpw, float *derivative_soa,
                                                int32_t do_resampling, skm_result *out);

/* GpuKernelEvaluator::evaluateDisplacementField (GpuEvaluator.h:175-178):
 * displacement at every fixed voxel, AoS. */
SKM_API skm_status skm_evaluate_displacement_field(skm_evaluator *ev, const float *cp_aos,
                                                   const float *cpw, float *dst_aos);

/* Device-pointer forms of the two calls above (no counterpart in the
 * reference, which copies through the host every call).  Work is enqueued on
 * the evaluator's stream; d_result (device memory, may be NULL) receives the
 * skm_result when the work completes. */
SKM_API skm_status skm_set_stream(skm_evaluator *ev, void *hip_stream);
SKM_API skm_status skm_get_value_and_derivative_device(skm_evaluator *ev, int32_t measure,
                                                       const float *d_cp_aos, const float *d_cpw,
                                                       float *d_derivative_soa, int32_t do_resampling,
                                                       skm_result *d_result);
SKM_API skm_status skm_evaluate_displacement_field_device(skm_evaluator *ev, const float *d_cp_aos,
                                                          const float *d_cpw, float *d_dst_aos);
SKM_API skm_status skm_synchronize(skm_evaluator *ev);

/* Device time per pass, accumulated with hipEvents on the evaluator's stream
 * while timing is enabled (no counterpart in the reference, which prints host
 * wall-clock only, itkIterationObserver.h:59-63).  Passes: 0 jitter offsets,
 * 1 value, 2 reduction, 3 derivative preparation, 4 derivative gather,
 * 5 regularizers, 6 LCC dense warp, 7 LCC centres, 8 LCC derivative,
 * 9 displacement field, 10 kernel-matrix assembly. */
#define SKM_NUM_PASSES 11
typedef struct skm_pass_times {
    double ms[SKM_NUM_PASSES];
    uint32_t launches[SKM_NUM_PASSES];
} skm_pass_times;
SKM_API skm_status skm_enable_timing(skm_evaluator *ev, int32_t enable);
/* Synchronises the stream, returns the totals since the last read and resets them. */
SKM_API skm_status skm_read_timing(skm_evaluator *ev, skm_pass_times *out);

/* Least-squares coefficients of a dense displacement in the evaluator's
 * kernel basis (no counterpart in the reference): CGLS with `iterations`
 * steps for min_c |K c - u|^2 + ridge |c|^2, where K c is the displacement of
 * skm_evaluate_displacement_field with control-point weights 1 and no
 * previous field.  d_displacement_aos: D floats per fixed voxel; d_coefficients_soa:
 * the control points' coefficients in the derivative's SoA layout.  Device
 * arrays; returns when done.  Stationary kernels only (SKM_ERROR_NOT_SUPPORTED
 * otherwise): there the transpose, a gather over every voxel in a control
 * point's support, is exactly the adjoint of K, whose per-voxel windows hold
 * the whole support; the weighted modes' windows (the reference's support
 * estimate of the voxel) can truncate a pair that the transpose would keep.
 * Used to turn a registration's final field into the coefficients the warm
 * start learns. */
SKM_API skm_status skm_fit_coefficients_device(skm_evaluator *ev, const float *d_displacement_aos, int32_t iterations,
                                               double ridge, float *d_coefficients_soa);

/* Launch shapes of an evaluator's kernels (no counterpart in the reference,
 * whose launch shapes are fixed, GpuEvaluator.cu:157-171, 424-438, 578-586).
 * 0 in a field selects the library's own choice; results do not depend on
 * the choice beyond FP32 rounding of the per-workgroup partial sums (the
 * sums across workgroups are FP64, in a fixed order).
 *   lcc_ref_chunks  reference-window LCC derivative: workgroups sharing one
 *                   tile of control points (1..64; auto: ~8192 workgroups)
 *   gather_split    derivative gather: sample planes per workgroup slab
 *                   (4 or 8; auto 8)
 *   value_split     value pass: waves sharing one brick of samples
 *                   (1, 4, 8 or 16; auto from the sample count)
 *   gather_regions  3D derivative gather: 8 splits the control-point tiles
 *                   into eight regions of the (x, y) plane, one per XCD
 *                   (workgroups reach the XCDs in turn), so that each XCD's
 *                   L2 holds only its region's samples; 1 keeps the tiles in
 *                   grid order (1 or 8; auto 1: at 128x128x96 the regions
 *                   measured 1.20 ms against 1.05-1.10 ms per gather)
 *   lcc_ref_walk    reference-window LCC derivative: 1 walks the centres of
 *                   each control point's window (the reference's order); 2
 *                   regroups the same sum per neighbour voxel where that
 *                   applies (stationary kernels in 3D, at most 8 control
 *                   points per axis within the support of a voxel), which
 *                   evaluates the kernel pair once per (control point,
 *                   voxel) instead of once per (control point, centre,
 *                   neighbour); auto (0) takes the form with the smaller
 *                   expected time (the walk's triples against the voxels);
 *                   results agree to FP rounding
 *   kernel_matrix   MSE / NCC with subsample 1: 2 assembles the evaluator's
 *                   kernel matrix at its first evaluation and reads it from
 *                   then on, 1 evaluates the kernels on the fly, 0 (auto) the
 *                   former for the tensor kernel and the latter for the
 *                   stationary and weight-image kernels, which cost less to
 *                   evaluate than their stored values to stream (see
 *                   skm_get_kernel_matrix_info)
 * skm_get_tuning returns the fields as set (0 where automatic). */
typedef struct skm_tuning {
    int32_t lcc_ref_chunks;
    int32_t gather_split;
    int32_t value_split;
    int32_t gather_regions;
    int32_t lcc_ref_walk;
    int32_t kernel_matrix;
    int32_t reserved[2];
} skm_tuning;
SKM_API skm_status skm_set_tuning(skm_evaluator *ev, const skm_tuning *tuning);
SKM_API skm_status skm_get_tuning(skm_evaluator *ev, skm_tuning *tuning);

/* The kernel matrix (no counterpart in the reference, which evaluates every
 * (voxel, control point) kernel pair in the value pass and again in the
 * derivative pass of every GetValueAndDerivative, GpuEvaluatorMSE.cu:61-80,
 * 134-162; GpuEvaluatorNCC.cu:61-82, 145-176).  With subsample 1 the samples
 * are the fixed voxels and never move, and the kernel depends only on the
 * positions and the weight image / tensor field, all fixed for an evaluator's
 * life: its non-zero k(x_m, p_j) are assembled once in device memory, in the
 * order each pass reads them, and every later MSE / NCC evaluation streams
 * them instead of evaluating the kernel.  It is assembled at the first MSE /
 * NCC evaluation when skm_tuning.kernel_matrix asks for it (automatically for
 * the tensor kernel), unless that evaluation is being captured into a graph;
 * it is refused (the kernels then run on the fly)
 * when it would hold more than SKM_KERNEL_MATRIX_MAX_BYTES (environment) or,
 * by default, a quarter of the device's memory or half of its free memory.
 * Results equal the on-the-fly passes' within FP32 rounding of the sums'
 * order, and are the same bits on every call.  skm_release_kernel_matrix
 * frees it (the next evaluation assembles it again). */
typedef struct skm_kernel_matrix_info {
    int32_t state;          /* 0 not assembled, 1 assembled, 2 refused (too large) */
    int32_t reserved;
    uint64_t bytes;         /* device memory it holds (or would hold, when refused) */
    uint64_t pairs_value;   /* non-zero pairs stored for the value pass */
    uint64_t pairs_gather;  /* non-zero pairs stored for the derivative gather */
    uint64_t read_value;    /* bytes of the value layout one evaluation reads */
    uint64_t read_gather;   /* bytes of the gather layout one evaluation reads */
    double build_ms;        /* wall time of the assembly */
} skm_kernel_matrix_info;
SKM_API skm_status skm_get_kernel_matrix_info(skm_evaluator *ev, skm_kernel_matrix_info *info);
SKM_API skm_status skm_release_kernel_matrix(skm_evaluator *ev);

/* The project's autotuner (no counterpart in the reference): times every
 * launch shape skm_set_tuning offers for `measure` on this evaluator (value
 * split x gather slab for MSE / NCC, workgroups per tile for the reference
 * LCC derivative, gather slab for the exact LCC derivative) with HIP events,
 * `repeats` evaluations each (median) at the device control points given,
 * sets the fastest and records it for the evaluator's shape class (dimension,
 * kernel mode, windows, sampling, image and control-point grids); evaluators
 * created later in the process with the same class start with it.  The
 * table can be saved and loaded (one line per class). */
typedef struct skm_autotune_report {
    skm_tuning best;
    double best_ms;          /* median time of one evaluation with it        */
    double first_ms;         /* the same for the automatic shapes            */
    int32_t candidates;
    int32_t reserved;
} skm_autotune_report;
SKM_API skm_status skm_autotune(skm_evaluator *ev, int32_t measure, const float *d_cp_aos, const float *d_cpw,
                                int32_t repeats, skm_autotune_report *report);
SKM_API skm_status skm_tuning_table_save(const char *path);
SKM_API skm_status skm_tuning_table_load(const char *path);

/* The library's device-memory cache (no counterpart in the reference, which
 * calls cudaMalloc / cudaFree per evaluator, GpuEvaluator.cu:30-91, 93-113):
 * blocks that skm_destroy and the registration driver give back are kept,
 * per device, up to a limit of idle bytes, for the next evaluator (hipFree
 * waits for the whole device, which stalls registrations running on other
 * streams).  The limit is 4 GiB per device, or SKM_CACHE_IDLE_BYTES from the
 * environment at the first allocation; skm_set_cache_limit changes it (idle
 * blocks beyond a lower limit are freed at once; 0 disables the cache).
 * skm_release_cached_memory frees every idle block of every device (a host
 * application that needs the memory back); live blocks stay.  Sizes are
 * rounded up to a size class (256 bytes, else an eighth of the largest power
 * of two below the size). */
SKM_API skm_status skm_set_cache_limit(uint64_t idle_bytes_per_device);
SKM_API skm_status skm_get_cache_info(uint64_t *idle_bytes, uint64_t *live_bytes, uint64_t *limit);
SKM_API skm_status skm_release_cached_memory(void);

/* Jitter offsets of the last evaluation, D ints per subsampled cell
 * (GpuEvaluator.cu:107-118); n_cells receives prod(fixed.size / subsample). */
SKM_API skm_status skm_get_cube_offsets(skm_evaluator *ev, int32_t *out, int64_t *n_cells);

/* TransformAdapter::ComputeControlPointImage geometry (itkTransformAdapter.hxx:246-308). */
SKM_API skm_status skm_control_point_grid(int32_t dim, const skm_grid *image, const double *cp_spacing,
                                          skm_grid *cp_out);

/* ---- image preprocessing ----------------------------------------------- */

/* GaussianSmoothing(image, variance) (include/itkUtils.h:140-151):
 * itk::DiscreteGaussianImageFilter with the variance in physical units
 * (UseImageSpacing), maximum error 0.01, maximum kernel width 64 and edge
 * replication at the border.  Host pointers; in and out may be the same
 * array.  A per-axis variance beyond ~709 pixels^2, where the discrete
 * kernel's exp(t) overflows, is rejected. */
SKM_API skm_status skm_gaussian_smooth(int32_t dim, const skm_grid *grid, const float *in, double variance,
                                       float *out);

/* skm_gaussian_smooth on device arrays without any allocation: no
 * counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): its temporaries
 * (the Gaussian taps and two intermediate planes) live in a fixed device
 * array of the library (64 MiB + 32 KiB per GPU, held until the process
 * ends).  Limit: 8,388,608 voxels (e.g. 256 x 256 x 128); larger images are
 * rejected with SKM_ERROR_INVALID_ARGUMENT.  d_in and d_out may be the same
 * array.  The work is enqueued on hip_stream (NULL: the null stream) and the
 * call returns without waiting; the result equals skm_gaussian_smooth's bit
 * for bit.  Calls that overlap on one GPU share the array and race: order
 * them on one stream or synchronise between them. */
SKM_API skm_status skm_gaussian_smooth_device_noalloc(int32_t dim, const skm_grid *grid, const float *d_in,
                                                      double variance, float *d_out, void *hip_stream);

/* Arguments of ComputeAnisotropicTensor (include/AnisotropicTensorUtils.h:37-39). */
typedef struct skm_tensor_options {
    double sigma;                   /* sigma_C4 of the scale: image smoothing variance; sigma/4
                                       smooths the structure tensor                            */
    double weight;                  /* tensor weight (gradient scale), tensor_weights[scale]   */
    int32_t alignment_direction;    /* 0 x, 1 y, 2 z, negative none; SKMReg.cxx:271 passes 2   */
    int32_t anisotropic_filtering;  /* must be 0; SKMReg.cxx:272 passes false                  */
} skm_tensor_options;

/* ComputeAnisotropicTensor(image, sigma, weight, alignment, maximum_tilt,
 * false) (AnisotropicTensorUtils.h:37-355) followed by
 * TransformAdapter::SetWeightingTensor(tensor, maximum_tilt)
 * (itkTransformAdapter.hxx:335-366), on the GPU.  tensor_out receives D*D
 * floats per voxel on the image grid (the layout of skm_desc.weight_data);
 * maximum_tilt the value the reference returns (the lowest double when no
 * voxel is re-weighted); maximum_weight and maximum_eigenvalue_covariance
 * (9 doubles, 3x3 row-major like skm_desc) the values skm_desc takes.  Output pointers other than
 * tensor_out may be NULL. */
SKM_API skm_status skm_anisotropic_tensor(int32_t dim, const skm_grid *grid, const float *image,
                                          const skm_tensor_options *opt, float *tensor_out, double *maximum_tilt,
                                          double *maximum_weight, double *maximum_eigenvalue_covariance);

/* ---- warped images ------------------------------------------------------ */

typedef enum skm_interpolation {
    SKM_INTERPOLATION_LINEAR = 1,    /* itk::LinearInterpolateImageFunction       */
    SKM_INTERPOLATION_BSPLINE3 = 3   /* itk::BSplineInterpolateImageFunction, order 3 */
} skm_interpolation;

/* Image warped by a dense displacement field, on the GPU: out(x) = I(x + u(x))
 * at every voxel x of out_grid, default_value where x + u(x) leaves the
 * image's buffer ([-0.5, n - 0.5) in index space).
 *   LINEAR:   the final registered image of SKMReg (src/SKMReg.cxx:507-525):
 *             ResampleImageFilter with a DisplacementFieldTransform -- u
 *             interpolated linearly in df (zero outside its buffer), the image
 *             interpolated linearly;
 *   BSPLINE3: WarpImage (include/itkUtils.h:163-179): WarpImageFilter on the
 *             field's own grid (out_grid must equal df_grid), the image
 *             interpolated by cubic B-splines with the mirror boundary.
 * Host pointers; df_aos holds D floats per df voxel. */
SKM_API skm_status skm_warp_image(int32_t dim, const skm_grid *image_grid, const float *image,
                                  const skm_grid *df_grid, const float *df_aos, const skm_grid *out_grid,
                                  int32_t interpolation, double default_value, float *out);

/* Synthetic 3D registration pair on the GPU (no counterpart in the
 * reference; inputs for tests, benchmarks and the warm start's training):
 * reference = n_ellipsoids random ellipsoids (centres U[0.15, 0.85] of the
 * extent, semi-axes U[0.05, 0.2] of it, intensities U[0, 1], later ones over
 * earlier ones) smoothed by a Gaussian of standard deviation smoothing_sigma;
 * d = sum of n_rbf Gaussian RBFs (standard deviation rbf_sigma, centres
 * uniform in the volume, weights U[-1, 1]^3) scaled so that max |d| =
 * max_displacement, times displacement_scale; target(x) = reference(x + d(x))
 * (cubic B-spline, 0 outside) + N(0, noise_sigma^2).  The random stream is
 * fixed by seed.  Host outputs: reference and target (x fastest), d AoS (may
 * be NULL). */
typedef struct skm_synthetic_desc {
    int32_t size[3];
    int32_t n_ellipsoids;        /* <= 64 */
    double smoothing_sigma;
    int32_t n_rbf;               /* 1..32 */
    double rbf_sigma;
    double max_displacement;
    double displacement_scale;
    double noise_sigma;
    uint64_t seed;
} skm_synthetic_desc;
SKM_API skm_status skm_synthetic_pair(const skm_synthetic_desc *desc, float *reference, float *target,
                                      float *displacement_aos);

/* ---- image files and landmarks ------------------------------------------ */

/* What an image file holds: dimension, geometry and values per voxel
 * (1: scalar image, D: displacement field). */
typedef struct skm_image_header {
    int32_t dim;
    skm_grid grid;
    int32_t components;
} skm_image_header;

/* The files of itk::ImageFileReader / ImageFileWriter as SKMReg uses them
 * (include/IOUtils.h:27-52): VTK legacy STRUCTURED_POINTS (.vtk; ASCII or
 * big-endian BINARY; SCALARS, VECTORS or TENSORS point data) and MetaImage
 * (.mhd with its data file, .mha; uncompressed, identity direction).  A VTK
 * file with DIMENSIONS nz = 1 reads as 2D.  Stored types are converted to
 * float; writing stores float (VTK: big-endian BINARY, 3 components as
 * VECTORS; MetaImage: little-endian, .mhd writes <name>.raw beside it). */
SKM_API skm_status skm_read_image_header(const char *path, skm_image_header *header);
/* capacity: floats the buffer holds (voxels * components needed); header may be NULL. */
SKM_API skm_status skm_read_image(const char *path, float *data, int64_t capacity, skm_image_header *header);
SKM_API skm_status skm_write_image(const char *path, const skm_image_header *header, const float *data);

/* ReadLandmarks (include/IOUtils.h:65-107): lines starting with '#' are
 * skipped, "POINTS n ..." declares n, every complete triple of numbers on any
 * other line is a point of which the first dim coordinates are kept.  points
 * may be NULL to count; *declared (may be NULL) receives the POINTS count,
 * which the reference only warns about when it differs. */
SKM_API skm_status skm_read_landmarks(const char *path, int32_t dim, double *points, int64_t capacity,
                                      int64_t *count, int64_t *declared);

/* CalculateTRE (include/itkUtils.h:268-293): the mean over i of
 * |m_i + u(m_i) - f_i| with u interpolated linearly in the dense field
 * (the point is left unchanged outside the field's buffer); -1 when count is 0. */
SKM_API skm_status skm_landmark_tre(int32_t dim, const skm_grid *df_grid, const float *df_aos,
                                    const double *fixed_points, const double *moving_points, int64_t count,
                                    double *tre);

/* ---- metric assembly and optimisation, device resident ---------------- */

/* RegularizedImageToImageMetricv4 switches (itkRegularizedImageToImageMetricv4.h:124-160). */
typedef struct skm_metric_options {
    int32_t measure;
    double reg_l1, reg_l21, reg_l2;
    int32_t use_negative_gradient;   /* default 1 (itkRegularizedImageToImageMetricv4.hxx:47) */
    int32_t do_resampling;           /* default 1 */
} skm_metric_options;

/* Optimizer kinds: Standard (itkStandardGradientDescentOptimizerv4.hxx:159-166),
 * AdaptiveStep (itkAdaptiveStepGradientDescentOptimizerv4.hxx:86-128) and a
 * fixed step of size a. */
typedef enum skm_optimizer_kind { SKM_OPT_STANDARD = 0, SKM_OPT_ADAPTIVE = 1, SKM_OPT_FIXED = 2 } skm_optimizer_kind;

typedef struct skm_optimizer_options {
    int32_t kind;
    double a, A, alpha;              /* step schedule                                  */
    double fmax, fmin, omega;        /* adaptive time function                          */
    int32_t orthant_projection;
    int32_t iterations;
    int32_t return_best;             /* return the evaluated position with the lowest value
                                        (ITK's ReturnBestParametersAndValue, set by
                                        SKMReg.cxx:383) instead of the last step's     */
    int32_t graph;                   /* 1: the first iteration is launched kernel by
                                        kernel, the others replay it as one captured HIP
                                        graph (one host call per iteration; same results);
                                        0: every iteration kernel by kernel, as also while
                                        pass timing is on                               */
} skm_optimizer_options;

/* Metric value per iteration as reported by
 * RegularizedImageToImageMetricv4::GetValueAndDerivative (measure + weighted
 * regularizers, .hxx:531-572). */
typedef struct skm_iteration_record {
    double value;          /* measure + regularizers */
    double measure;
    double learning_rate;
    uint32_t num_pixels_counted;
    uint32_t reserved;
} skm_iteration_record;

/* Runs `iterations` steps of gradient descent entirely on the device:
 * GetValueAndDerivative, host-side regularizers L1/L21/L2 and sign flip
 * (itkRegularizedImageToImageMetricv4.hxx:430-529), then the optimizer
 * update theta_{k+1} = theta_k + eta_k d_k with optional orthant projection
 * (itkStandardGradientDescentOptimizerv4.hxx:116-157).  cpw (host) holds the
 * control-point weights; params_soa (host)
 * holds the start point and receives the result; records (host, may be NULL)
 * receives one entry per iteration. */
SKM_API skm_status skm_optimize(skm_evaluator *ev, const skm_metric_options *metric,
                                const skm_optimizer_options *opt, const float *cpw, float *params_soa,
                                skm_iteration_record *records);

/* skm_optimize of n evaluators at once (no counterpart in the reference,
 * which optimises one pair per process): the same metric and optimizer
 * options, each evaluator's own cpw, start point (params_soa[i], host, SoA)
 * and records (records[i], may be NULL; records itself may be NULL).  The
 * first iteration runs kernel by kernel; with opt->graph the others replay
 * one captured HIP graph holding one iteration of every evaluator, each on
 * its own stream, so that one graph launch performs one optimizer step of
 * all n.  Every result equals skm_optimize's on its evaluator. */
SKM_API skm_status skm_optimize_batch(skm_evaluator *const *evaluators, int32_t n, const skm_metric_options *metric,
                                      const skm_optimizer_options *opt, const float *const *cpw,
                                      float *const *params_soa, skm_iteration_record *const *records);

/* Adds L1 / L21 / L2 values and subgradients to a host derivative and applies
 * the sign flip exactly as the metric does on the host
 * (itkRegularizedImageToImageMetricv4.hxx:430-529).  Returns the weighted sum
 * of the three regularizers in *reg_value. */
SKM_API skm_status skm_host_regularizers(int32_t dim, int64_t n_cp, const double *params_soa,
                                         double reg_l1, double reg_l21, double reg_l2,
                                         int32_t use_negative_gradient, double *derivative_soa,
                                         double *reg_value, double *l1, double *l21, double *l2);

/* ---- multi-scale registration -------------------------------------------- */

/* The scale-level loop of SKMReg (src/SKMReg.cxx:123-500) over this library:
 * per level Gaussian smoothing of both images, the weight tensor (SKMReg's
 * cases 1-4, SKMReg.cxx:262-297), a control-point grid with spacing
 * min(sigma_C4, sigma_C0) / cp_spacing_scaler, a new evaluator whose previous
 * displacement is the sum of the coarser levels (SummationTransform with a
 * nearest-neighbour displacement field), the adaptive-step optimizer of
 * SKMReg.cxx:375-430 run on the device (a = initial_step_size, A = 20,
 * alpha = 1, Fmax = 1e7, Fmin = -1e6, omega = 1, best parameters returned,
 * orthant projection when L1 or L2,1 is the only regularizer), and the dense
 * displacement of all levels so far.  Images are preprocessed as
 * SKMReg.cxx:84-91 does: intensity thresholds, then rescaling to [0, 1]. */
#define SKM_MAX_SCALES 8

/* Per-level options, the per-scale entries of ConfigUtils.h:29-62. */
typedef struct skm_scale_params {
    double sigma_c4, sigma_c0;
    double tensor_weight;            /* > 0: anisotropic tensor from the target
                                        (or the guidance image); 0: none, or the
                                        given weight tensor                        */
    double reg_l1, reg_l21, reg_l2, reg_rkhs, reg_rd, reg_rd_scaling, reg_pg, reg_pg_scaling;
    int32_t iterations;              /* num_function_evaluations                   */
    double initial_step_size;
    int32_t sampling_rate, neighborhood_sampling;
    double smoothing_variance;       /* > 1e-5: GaussianSmoothing of both images   */
    double reg_bilateral;            /* skm_set_regularizer_bilateral (no SKMReg key) */
    double reg_bilateral_spatial_sigma, reg_bilateral_intensity_sigma;
} skm_scale_params;

typedef struct skm_registration_desc {
    int32_t dim;
    skm_grid reference; const float *reference_data;   /* moving image R            */
    skm_grid target;    const float *target_data;      /* fixed image T, the output
                                                          grid of the displacement  */
    int32_t num_scales;                                /* <= SKM_MAX_SCALES        */
    skm_scale_params scale[SKM_MAX_SCALES];
    double cp_spacing_scaler;                          /* < 1: spacing = min sigma  */
    int32_t metric;                                    /* skm_measure               */
    int32_t use_intensity_thresholds;                  /* intensity_thresholds set  */
    double threshold_min, threshold_max;
    skm_grid weight; const float *weight_image;        /* weight_filename, or NULL  */
    const float *weight_tensor;                        /* weight_tensor_filename as a
                                                          D*D tensor field on the
                                                          weight grid (tensor_weight
                                                          0), or NULL               */
    const float *guidance_image;                       /* weight_tensor_filename as a
                                                          guidance image on the weight
                                                          grid (tensor_weight > 0)  */
    int32_t restart_at;                                /* restart_at (1-based level) */
    const float *restart_displacement;                 /* df-full of level restart_at-1,
                                                          AoS on the target grid    */
    const char *temp_directory;                        /* temp_directory, or NULL: SKMReg's
                                                          result files are written there
                                                          (SKMReg.cxx:92-94, 150-153,
                                                          459-525) and restart_at without
                                                          restart_displacement reads
                                                          df-full-<restart_at-2>.vtk
                                                          (SKMReg.cxx:96-113)        */
    int32_t verbosity;                                 /* >= 2: also the smoothed images,
                                                          df-<s> and warped-<s>     */
    int64_t num_landmarks;                             /* landmark pairs for the TRE */
    const double *reference_landmarks;                 /* D doubles per point       */
    const double *target_landmarks;
    int32_t exact_gradient;                            /* 1: every level's evaluator with
                                                          SKM_FLAG_EXACT_GRADIENT (no SKMReg
                                                          counterpart); 0: the reference's
                                                          derivative windows        */
    /* warm start (no SKMReg counterpart) */
    const float *initial_coefficients;                 /* start point of the first level run,
                                                          SoA on its control-point grid, or
                                                          NULL for zeros (SKMReg's)  */
    int32_t initial_guard;                             /* 1: start from zeros instead when the
                                                          given coefficients (or displacement)
                                                          evaluate worse than zeros (none) */
    int32_t autotune;                                  /* 1: each level's evaluator runs
                                                          skm_autotune first unless its shape
                                                          class is already tuned (no SKMReg
                                                          counterpart)               */
    const float *initial_displacement;                 /* previous displacement of the first
                                                          level run, AoS on the target grid
                                                          (as restart_displacement), or NULL;
                                                          with initial_guard dropped when it
                                                          evaluates worse than none  */
    int32_t initial_displacement_on_device;            /* 1: initial_displacement is device
                                                          memory of the current device, its
                                                          writes complete before the call
                                                          (the caller synchronises); copied on
                                                          the device, no host round trip;
                                                          any other pointer is an
                                                          SKM_ERROR_INVALID_ARGUMENT */
} skm_registration_desc;

/* What one level did. */
typedef struct skm_scale_report {
    double value;                    /* lowest metric value (measure + regularizers) */
    double measure;                  /* the measure at that iteration               */
    double first_value;              /* value at the first iteration                */
    int32_t iterations;
    int32_t cp_size[3];
    double maximum_tilt;             /* ComputeAnisotropicTensor's, or 0           */
    double ms_setup, ms_optimize, ms_displacement;
    double tre;                      /* CalculateTRE of df-full, -1 without landmarks */
    double zero_parameters;          /* percentage of parameters equal to 0          */
    int32_t initial_used;            /* 1: the level started from initial_coefficients
                                        or initial_displacement                      */
    int32_t reserved;
    double initial_value, zero_value;/* with initial_guard: value at the given start
                                        and at zeros                                 */
} skm_scale_report;

/* ConfigUtils.h:29-62 defaults (3 levels, metric MSE, sigma_C4 64/32/16, ...). */
SKM_API void skm_registration_desc_init(skm_registration_desc *desc);

/* Runs the registration.  displacement_out (host, D floats per target voxel)
 * receives the displacement of all levels; reports (host, num_scales entries,
 * may be NULL) the per-level summaries (levels before restart_at untouched).
 * With temp_directory set, the files written there are SKMReg's:
 * reference-/target-preprocessed.vtk, per level s (0-based)
 * covariance_tensor-<s>.vtk (tensor kernels), df-full-<s>.vtk,
 * cp_image-<s>.vtk, cp_jacdet-<s>.vtk (and at verbosity >= 2
 * smoothed_reference-<s>.vtk, smoothed_target-<s>.vtk, df-<s>.vtk,
 * warped-<s>.vtk), then df.vtk and warped.vtk. */
SKM_API skm_status skm_register(const skm_registration_desc *desc, float *displacement_out,
                                skm_scale_report *reports);

/* Batched registration, no counterpart in the reference (whose SKMReg
 * registers one pair per process): n independent registrations, e.g. the
 * frames of a time series each registered to frame 0, run by up to
 * `concurrency` host threads at once on the calling thread's device, each
 * registration on its own streams, so that the small per-iteration launches
 * of one frame overlap those of the others.  Every registration gives the
 * same result as skm_register on its descriptor.  displacements_out[i] and
 * reports[i] are as skm_register's for descs[i] (reports and reports[i] may be
 * NULL); statuses (n entries, may be NULL) receives each registration's
 * status.  Returns SKM_SUCCESS when all succeed, else the status of the first
 * failing registration by index, with skm_last_error() naming it. */
SKM_API skm_status skm_register_batch(const skm_registration_desc *descs, int32_t n, int32_t concurrency,
                                      float *const *displacements_out, skm_scale_report *const *reports,
                                      skm_status *statuses);

/* The same batch in lockstep (no counterpart in the reference): all n
 * registrations in flight, one host thread each, their set-up concurrent;
 * at every level the optimisers of all registrations still running run
 * together through skm_optimize_batch, so that each optimizer step of the
 * whole batch is one launch of one captured HIP graph (one branch per
 * registration; registrations whose levels have different optimizer options
 * form separate graphs).  Results, reports and statuses as skm_register_batch;
 * each result equals skm_register's on its descriptor. */
SKM_API skm_status skm_register_batch_lockstep(const skm_registration_desc *descs, int32_t n,
                                               float *const *displacements_out, skm_scale_report *const *reports,
                                               skm_status *statuses);

/* ---- learned warm start ------------------------------------------------- */

/* A 3D convolutional encoder-decoder that predicts a registration's initial
 * control-point coefficients from the downsampled image pair, trained and
 * run by the library's own kernels (no counterpart in the reference, which
 * starts every registration from zero coefficients, src/SKMReg.cxx:375-430).
 * Levels of 16 / 32 / 48 / 64 channels, 3x3x3 convolutions with BF16 weights
 * and FP32 accumulation on the matrix cores, a decoder back to half the input
 * resolution, trilinear sampling at the control points and a linear head to
 * D = 3 coefficients per control point.
 *   input   [batch][z][y][x][2] floats (channel 0 moving, 1 fixed), device
 *   output  [batch][cp z][cp y][cp x][3] floats, device
 * Control points are given in input voxel units (index i of the input is
 * coordinate i).  Parameters: FP32 master weights, host layout as
 * skm_ws_get_params returns them. */
typedef struct skm_ws_desc {
    int32_t in_size[3];      /* x, y, z of the network input; multiples of 8 */
    int32_t cp_size[3];      /* control-point grid, x, y, z                   */
    double cp_origin[3];     /* control point 0 in input voxel coordinates    */
    double cp_spacing[3];    /* control-point spacing in input voxels (>= 2)  */
    int32_t max_batch;
} skm_ws_desc;
typedef struct skm_ws_net skm_ws_net;

SKM_API skm_status skm_ws_create(const skm_ws_desc *desc, skm_ws_net **out);
SKM_API void skm_ws_destroy(skm_ws_net *net);
SKM_API int64_t skm_ws_num_params(const skm_ws_net *net);
SKM_API skm_status skm_ws_set_stream(skm_ws_net *net, void *hip_stream);
/* host FP32 parameters in and out; set also resets the Adam moments */
SKM_API skm_status skm_ws_set_params(skm_ws_net *net, const float *params);
SKM_API skm_status skm_ws_get_params(skm_ws_net *net, float *params);
/* inference on device arrays, enqueued on the network's stream */
SKM_API skm_status skm_ws_forward(skm_ws_net *net, const float *d_input, int32_t batch, float *d_out);
/* gradients of every parameter for the output gradient d_dout of the last
 * forward; d_grad (device, may be NULL) receives them in the parameter layout */
SKM_API skm_status skm_ws_backward(skm_ws_net *net, const float *d_dout, float *d_grad);
/* sample `index` of a network input from a moving and a fixed image of
 * twice the input size (device, x fastest): their 2 x 2 x 2 block means */
SKM_API skm_status skm_ws_prepare_input(skm_ws_net *net, const float *d_moving, const float *d_fixed, int32_t index,
                                        float *d_input);
/* output of conv `layer` (0..9, after its ReLU) of the last forward,
 * [batch][z][y][x][channels] at that layer's resolution, copied to d_dst */
SKM_API skm_status skm_ws_get_activation(skm_ws_net *net, int32_t layer, float *d_dst);
/* one Adam step (beta 0.9 / 0.999, eps 1e-8) on the mean squared error
 * between the output and d_target; *loss (may be NULL) receives the loss of
 * the step's forward pass (the call then waits for the stream) */
SKM_API skm_status skm_ws_train_step(skm_ws_net *net, const float *d_input, const float *d_target, int32_t batch,
                                     double learning_rate, double *loss);

#ifdef __cplusplus
}
#endif
#endif
