// skm_internal.hpp -- argument blocks shared by the kernels and the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "skm.h"
#include "skm_device.hpp"

namespace skm {

// Stores msg as the calling thread's last error (skm_last_error) and returns st.
skm_status set_last_error(skm_status st, const std::string &msg);

// Device memory of the library.  A released block is kept for later requests
// of its size class on its device, so that repeated set-ups (levels, frames of
// a batch) call neither hipMalloc nor hipFree: hipFree waits for all work on
// the device, which would stall every other registration in flight.  A block
// may be released only once the work that uses it has completed.
hipError_t device_alloc(void **p, size_t bytes);
void device_release(void *p);

// Per-control-point kernel state (static for an evaluator): sig, sdet, cov[9], support.
constexpr int kCpStateStride = 12;

// Everything a kernel needs to know about one registration problem.
struct Problem {
    int dim;
    Grid fixed, moving, cp, weight, df;
    KernelConst kc;
    int sub, subn;
    int nsub[3];
    long long n_sub, n_cp, n_vox;
    int has_df;
    int exact;
    const float *fixed_data;
    const float *coeffs;       // moving B-spline coefficients, (n+3)^D
    const float *weight_data;  // weight image / tensor field
    const float *df_data;      // previous displacement, AoS
    const float4 *cpc;         // (c, w) per control point
    const float *cp_state;     // kCpStateStride floats per control point
    const int *cp_window;      // per cp: lo[3], hi[3] in sample space (derivative gather)
    const int *reg_window;     // per cp: lo[3], hi[3] in cp space (regularizers)
    const int *offsets;        // D ints per subsampled cell
    // weighted modes (1, 2): packed point states, two float4 per point (pstate_pack)
    const float4 *cp_pstate;   // per control point, static
    float4 *smp_pstate;        // per sample of the current evaluation, written by the value pass
    const float *cp_reach;     // per control point: pair_reach2(smin, lam, 0), its share of the pair reach
    const float *cp_row_lam;   // largest lam per row of control points (cy + n[1] cz)
    const float *cp_intensity; // fixed image at each control point (nearest neighbour), bilateral term
    const float *smp_row_lam;  // largest lam per row of sample cells (my + nsub[1] mz)
    double *gather_partials;   // weighted-mode gather: FP64 sums per (slab, control-point tile)
    double *reg_add;           // regularizers run beside the measure's passes: their derivative terms
                               // (SoA, FP64), added to deriv by the gather's finish; nullptr: the
                               // regularizers add them to deriv themselves
    int lcc_chunks;            // reference-window LCC derivative: workgroups per cp tile (0: automatic)
    int gather_split;          // derivative gather: sample planes per slab, 4 or 8 (0: automatic, 8)
    int value_split;           // value pass: waves per sample brick, 1/4/8/16 (0: automatic)
    int gather_regions;        // 3D derivative gather: 8 XCD regions of cp tiles, or 1 (0: automatic, 1)
    int lcc_ref_walk;          // reference-window LCC derivative: 1 the centre walk, 2 the voxel tables (0: automatic)
    float lam_cp_max;          // largest lam over the control points
    float lam_w_max;           // largest lam over the points a sample can look up (weight image, outside)
};

// Device-resident scalars of one evaluation (all FP64).
struct Scalars {
    double count;
    double sums[6];        // MSE: sum d^2 | NCC: sf, sm, sff, smm, sfm | LCC: sum cc
    double measure;
    double mean_f, mean_m, rho, inv_denom;  // NCC derivative coefficients
    double coef_scale;     // MSE: 2/N, LCC: 1/N
    double reg[4];         // RKHS, RD, PG, bilateral values
    double fx_unit;        // exact LCC scatter: value of one fixed-point step
    double fx_inv_unit;    // 1 / fx_unit
};

enum { kPartialsK = 8 };

// ---- launchers (skm_kernels.hip) ----
// Draws the cube offsets from the PCG32 state at *state (device); the
// evaluation's launch_reduce advances the state past them.
void launch_offsets(const Problem &pr, int *offsets, uint64_t *state, uint64_t inc, hipStream_t s);
void launch_pack_cp(const Problem &pr, const float *cp, const float *cpw, int soa, float4 *cpc, hipStream_t s);
void launch_value(const Problem &pr, int measure, float4 *sg, float *sb, double *partials, int *blocks_out,
                  hipStream_t s);
// Also moves the PCG32 state at *rng rng_delta draws ahead (rng_delta 0: untouched),
// and sets the regularizer values to 0 unless keep_reg (the regularizers'
// reduction writes them in this evaluation).
void launch_reduce(const Problem &pr, int measure, const double *partials, int blocks, Scalars *sc,
                   hipStream_t s, uint64_t *rng = nullptr, uint64_t rng_delta = 0, uint64_t rng_inc = 1,
                   bool keep_reg = false);
void launch_prepare(const Problem &pr, int measure, const float4 *sg, const float *sb, const Scalars *sc,
                    float4 *sv, hipStream_t s);
void launch_gather(const Problem &pr, const float4 *sv, float *deriv_soa, hipStream_t s);
// deriv_j = w_j * the sum of j's (slab, tile) partials in slab order (plus
// pr.reg_add's term, where set, as the regularizers would add it)
void launch_gather_finish(const Problem &pr, float *deriv_soa, hipStream_t s);
// doubles of Problem::gather_partials for a sample grid with `planes` cells along the last axis
long long gather_partials_size(const Problem &pr, long long planes);
// The bilateral term (no reference counterpart): spatial Gaussian of standard
// deviation sigma_s cut at 3 sigma_s, intensity Gaussian of standard deviation
// sigma_i of the fixed image's values at the two control points.
struct BilateralArgs {
    double weight;
    float cut2;        // (3 sigma_s)^2
    float inv2ss;      // 1 / (2 sigma_s^2)
    float inv2si;      // 1 / (2 sigma_i^2)
    float rad[3];      // 3 sigma_s / spacing per axis (window in control-point indices)
};
BilateralArgs make_bilateral_args(const Problem &pr, double weight, double sigma_s, double sigma_i);
// weights: RKHS, RD, PG; bl: the bilateral term (weight 0: off)
void launch_regularizers(const Problem &pr, const double *weights, double pg_scaling, const BilateralArgs &bl,
                         int with_derivative, float *deriv_soa, double *partials, Scalars *sc, hipStream_t s);
unsigned reg_blocks(long long n_cp);  // workgroups (and 4-double partial slots) of launch_regularizers
void launch_displacement(const Problem &pr, float *dst, hipStream_t s);
// packed states of the weighted modes at every voxel of the dense fixed grid
void launch_voxel_states(const Problem &pr, float4 *out, hipStream_t s);

// The kernel matrix of an evaluator whose samples do not move between
// evaluations (subsample 1: sample m is voxel m): every non-zero k(x_m, p_j)
// that the value pass and the derivative gather use, assembled once in device
// memory (skm_kmatrix.hip).  k depends on the sample and control-point
// positions and on the weight image / tensor field only, none of which changes
// over an evaluator's life; the coefficients c_j do.  Each evaluation then reads
// the stored values in place of evaluating the kernel twice per pair.
//   value layout: per sample brick (one wave of the value pass), the control
//     points of the walk (in its order) with a pair passing the window and
//     reach tests, in groups of four: one record {lane mask lo, lane mask hi,
//     point 0, point 1}, {point 2, point 3, 0, 0} per group, followed in v_val
//     by one float4 per lane of the mask: the lane's k with the four points (0
//     where a pair is zero or outside the lane's window);
//   gather layout: per wave of a (slab, control-point tile) workgroup of the
//     derivative gather, one record {lane mask lo, lane mask hi, first sample, 0}
//     per sample block with a non-zero pair, followed in g_val by one float4 per
//     lane of the mask: k of the four tile points (0 where a pair is zero or
//     outside the point's window).
struct KernelMatrix {
    int state = 0;  // 0 not built, 1 built, 2 refused (larger than the limit); see kernel_matrix_build
    int slab = 0;   // gather slab of the gather layout
    long long units_v = 0, units_g = 0;
    long long cap_ent_v = 0, cap_val_v = 0, cap_ent_g = 0, cap_val_g = 0;  // allocated records and float4 slots
    long long nnz_v = 0, nnz_g = 0;    // non-zero pairs stored
    long long read_v = 0, read_g = 0;  // bytes of each layout an evaluation reads (records and slots)
    long long bytes = 0;
    double build_ms = 0.0;
    uint4 *v_ent = nullptr;  // two per record
    float4 *v_val = nullptr;
    long long *v_ent_off = nullptr, *v_val_off = nullptr;
    int *v_cnt = nullptr;
    uint4 *g_ent = nullptr;
    float4 *g_val = nullptr;
    long long *g_ent_off = nullptr, *g_val_off = nullptr;
    int *g_cnt = nullptr;
};
// whether the evaluator's problem has a kernel matrix (subsample 1, sizes within 32-bit indices)
bool kernel_matrix_applies(const Problem &pr);
// sample planes per slab of the kernel-matrix gather for this problem's tuning
int kernel_matrix_slab(const Problem &pr);
// Assembles km on stream s (returns when done).  Refuses (state 2, SKM_SUCCESS)
// when its device memory would exceed max_bytes; errors leave state 0.
skm_status kernel_matrix_build(const Problem &pr, KernelMatrix &km, hipStream_t s, long long max_bytes);
// frees the buffers (the caller has synchronised the work that reads them)
void kernel_matrix_release(KernelMatrix &km);
// value pass and derivative gather from the kernel matrix (same outputs as
// launch_value / launch_gather)
void launch_value_km(const Problem &pr, const KernelMatrix &km, int measure, float4 *sg, float *sb,
                     double *partials, int *blocks_out, hipStream_t s);
void launch_gather_km(const Problem &pr, const KernelMatrix &km, const float4 *sv, float *deriv_soa,
                      hipStream_t s);

// skm_create with, when device_data, device pointers in the descriptor (skm_capi.hip)
skm_status create_evaluator(const skm_desc *desc, bool device_data, skm_evaluator **out,
                            hipStream_t stream = nullptr);
// least-squares fit of a dense displacement (AoS, fixed grid, device) onto the
// evaluator's kernel basis: CGLS for min |K c - u|^2 + ridge |c|^2, control-point
// weights 1; c (SoA, device) starts at zero
skm_status fit_coefficients_device(skm_evaluator *ev, const float *d_u_aos, int iterations, double ridge,
                                   float *d_c_soa);
// skm_set_displacement_field from a device array
skm_status set_displacement_field_device(skm_evaluator *ev, const skm_grid *grid, const float *d_df_aos);
// whether the process's launch-shape table holds this evaluator's shape class
bool tuning_recorded(const skm_evaluator *ev);
skm_status set_displacement_field_impl(skm_evaluator *ev, const skm_grid *grid, const float *df_aos, bool device);
// GaussianSmoothing / ComputeAnisotropicTensor on device arrays (skm_tensor.hip),
// on `stream`, complete on return
skm_status gaussian_smooth_device(int dim, const skm_grid &grid, const float *d_in, double variance, float *d_out,
                                  hipStream_t stream = nullptr);
skm_status anisotropic_tensor_device(int dim, const skm_grid &grid, const float *d_image,
                                     const skm_tensor_options &opt, float *d_tensor_out, double *maximum_tilt,
                                     double *maximum_weight, double *maximum_eigenvalue_covariance,
                                     hipStream_t stream = nullptr);
// A synchronised non-blocking stream of the current device's pool, and its
// return (synchronised, then kept for the next user on device `dev`).
hipError_t take_pool_stream(hipStream_t *out);
void return_pool_stream(int dev, hipStream_t s);

// set-up (skm_setup.hip)
// cubic B-spline coefficients (mirror boundary) in FP64, unpadded, in d_work
void launch_bspline_prefilter(int dim, const int *size, const float *d_src, double *d_work, hipStream_t s);
// the same, then shifted by one and padded to (n+3)^D in float
void launch_bspline_coefficients(int dim, const int *size, const float *d_src, double *d_work, float *d_coeffs,
                                 hipStream_t s);
// dst[i * per + k] = src[idx[i] * per + k], or (idx < 0) 0 for per == 1 and
// the identity's entries for per == D * D
void launch_zero_words(void *d_ptr, long long n_words, hipStream_t s);
void launch_gather_values(const float *d_src, const long long *d_idx, int per, long long n, float *d_dst,
                          hipStream_t s);
unsigned weight_lam_blocks(long long nw);
void launch_weight_lam(int dim, int mode, long long nw, const float *d_w, float *d_lam, double *d_blk_max,
                       unsigned *d_blk_bad, hipStream_t s);
void launch_sample_row_lams(int dim, const Grid &fixed, const Grid &weight, int sub, const int *ncell,
                            const float *d_wlam, float lam_out, float *d_out, hipStream_t s);
void launch_result(const Scalars *sc, void *result, hipStream_t s);

// LCC (skm_lcc.hip)
void launch_dense_warp(const Problem &pr, int quirk3d, float4 *dense_g, hipStream_t s);
// bound_blk (optional, one double per workgroup): the largest bound per
// workgroup of |beta kw gradM| over a centre's neighbours, before the 1/N scale;
// cpk (optional, one word per centre): its jitter offsets, one byte per axis,
// and bit 31 set when its derivative coefficients are not all zero; the word
// is also stored in the centre's centre_coef .w (0 without cpk)
void launch_lcc_centres(const Problem &pr, const float4 *dense_g, float4 *centre_coef, double *partials,
                        int *blocks_out, hipStream_t s, double *bound_blk = nullptr, unsigned *cpk = nullptr);
// partials: lcc_reference_partials(pr) doubles (none needed when it is 0)
void launch_lcc_gather_reference(const Problem &pr, const float4 *dense_g, const float4 *centre_coef,
                                 const Scalars *sc, float *deriv_soa, double *partials, hipStream_t s);
long long lcc_reference_partials(const Problem &pr);
// Exact-mode scatter in 64-bit fixed point (order-independent, so the sums
// are the same bits on every run): the step fx_unit is chosen from the
// centres' bounds so that no voxel's sum can overflow; then gamma = fx * unit.
// fx: 3 long longs per fixed voxel, zeroed by the caller.
void launch_lcc_scatter(const Problem &pr, const float4 *dense_g, const float4 *centre_coef, Scalars *sc,
                        const double *bound_blk, int bound_blocks, unsigned long long *fx, float4 *gamma,
                        hipStream_t s);
// Exact mode, stationary kernels: the same gamma gathered per voxel (no
// atomics, FP64 sums in a fixed order); applies when every axis' reach of a
// voxel fits the gather's 64-position masks (support / spacing <= ~30).
bool lcc_gamma_gather_applies(const Problem &pr);
void launch_lcc_gamma(const Problem &pr, const float4 *dense_g, const float4 *centre_coef, const unsigned *cpk,
                      const Scalars *sc, float4 *gamma, hipStream_t s);
// Reference-window derivative regrouped per neighbour voxel (stationary
// kernels, 3D, a control-point box of at most 8 per axis around each voxel;
// chosen by lcc_ref_walk or, automatically, where its expected time is the
// smaller): the same sum as launch_lcc_gather_reference, with
// the kernel pair evaluated once per (control point, voxel).  centre_coef as
// from launch_lcc_centres given a cpk buffer (the packed word then rides in
// each centre's .w); partials: lcc_ref_table_partials(pr) doubles.
bool lcc_ref_table_applies(const Problem &pr);
long long lcc_ref_table_partials(const Problem &pr);
void launch_lcc_ref_table(const Problem &pr, const float4 *dense_g, const float4 *centre_coef, const Scalars *sc,
                          float *deriv_soa, double *partials, hipStream_t s);

// optimizer (skm_kernels.hip)
struct OptState {
    double time;           // adaptive time
    double rate;
    double prev_dot;
    long long iteration;
    double best_value;     // lowest metric value evaluated so far
    int improved;          // this iteration's value is the lowest so far
};
// L1 / L2,1 / L2 subgradients and the sign flip on the device derivative,
// with block partials of the three values and of <g, prev_grad>.
void launch_host_regs_device(const Problem &pr, const float *params_soa, double l1, double l21, double l2,
                             int negate, float *deriv_soa, const float *prev_grad, double *partials, hipStream_t s);
// Step scalars and record, then the update of params_soa (best position kept
// when best_params is set) and the packed control points of the next step.
void launch_opt_step(const Problem &pr, int kind, double a, double A, double alpha, double fmax, double fmin,
                     double omega, int orthant, float *params_soa, const float *grad_soa, float *prev_grad,
                     OptState *st, const Scalars *sc, const double *reg_partials, double l1w, double l21w,
                     double l2w, const double *gpu_reg_weights, void *record, float *best_params, const float *cpw,
                     float4 *cpc, hipStream_t s);
void launch_soa_to_cpc(const Problem &pr, const float *params_soa, const float *cpw, float4 *cpc,
                       hipStream_t s);

}  // namespace skm
