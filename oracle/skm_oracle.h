/*
 * skm_oracle.h -- FP64 CPU restatement of the SKM kernel-sum / metric /
 * regularizer path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * product path.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product library never links it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root).  Arithmetic is double precision throughout;
 * where the reference evaluates in float (the kernel, GKernel<float>,
 * include/cuda/Kernel.h:209) the oracle keeps double, so it is the exact
 * mathematical value the reference approximates.
 */
#ifndef SKM_ORACLE_H
#define SKM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct skmo_grid {
    int size[3];          /* voxels per axis; size[2] ignored in 2D        */
    double origin[3];     /* physical position of index 0                 */
    double spacing[3];    /* physical spacing                              */
} skmo_grid;

enum { SKMO_KERNEL_STATIONARY = 0, SKMO_KERNEL_WEIGHT_IMAGE = 1, SKMO_KERNEL_WEIGHT_TENSOR = 2 };
enum { SKMO_MSE = 0, SKMO_NCC = 1, SKMO_LCC = 2 };

typedef struct skmo_problem {
    int dim;
    skmo_grid fixed;   const double *fixed_data;     /* target image T, x fastest            */
    skmo_grid moving;  const double *moving_coeffs;  /* B-spline coefficients, (size+3)^D     */
    skmo_grid cp;      const double *cp_aos;         /* control-point coefficients, AoS       */
    const double *cpw;                               /* control-point weights                 */
    double sigma_c0, sigma_c4;
    int kernel_mode;
    double max_weight;                               /* maximum of the weight image           */
    double max_eig_cov[9];                           /* covariance with largest eigenvalue    */
    skmo_grid weight;  const double *weight_data;    /* scalar image or DxD tensors per voxel */
    int has_df;        skmo_grid df; const double *df_data;  /* previous-scale field, AoS   */
    int subsample, subsample_neighborhood;
    const int *offsets;                              /* per subsampled cell, D ints           */
    int exact;                                       /* 0: reference windows, 1: exact        */
} skmo_problem;

typedef struct skmo_result {
    double measure;
    double reg_rkhs, reg_rd, reg_pg;
    long long pixels_counted;
} skmo_result;

/* a1 kernel */
double skmo_wendland_c4(double t);
double skmo_wendland_c0(double t);
double skmo_kernel_eval(const skmo_problem *pr, const double *p, const double *q,
                        double sigmaP, const double *covP);
double skmo_region_support(const skmo_problem *pr, const double *p);
double skmo_max_eigenvalue(int dim, const double *m);
double skmo_determinant(int dim, const double *m);

/* a2 geometry */
int skmo_nearest_index(const skmo_grid *g, int dim, const double *p, int *idx);

/* a3 B-spline */
void skmo_bspline_coefficients(int dim, const int *size, const double *src, double *coeffs);
int  skmo_bspline_sample(int dim, const skmo_grid *g, const double *coeffs, const double *p,
                         double *value, double *grad);

/* a5 PCG32 and cube offsets */
uint32_t skmo_pcg32_next(uint64_t *state, uint64_t sequence);
void skmo_cube_offsets(uint64_t *state, uint64_t sequence, long long n_cells, int dim,
                       int subsample, int *out);

/* a6 dense displacement */
void skmo_displacement(const skmo_problem *pr, double *out_aos);
void skmo_displacement_at(const skmo_problem *pr, const double *x, double *u);

/* a7-a10, a12 evaluator */
int skmo_evaluate(const skmo_problem *pr, int measure, int with_derivative,
                  double reg_rkhs, double reg_rd, double reg_pg, double pg_scaling,
                  double *deriv_soa, double *per_sample, skmo_result *res);

/* a10 regularizers alone: per-cp values (optional) and derivative increment */
void skmo_regularizer(const skmo_problem *pr, int which, double weight, double pg_scaling,
                      double *values, double *deriv_soa);
/* bilateral regularizer (no reference counterpart; definition in skm_oracle.c) */
void skmo_regularizer_bilateral(const skmo_problem *pr, double weight, double sigma_s, double sigma_i,
                                double *values, double *deriv_soa);

/* a11 host regularizers and assembly */
double skmo_host_regularizers(int dim, long long n_cp, const double *params_soa,
                              double reg_l1, double reg_l21, double reg_l2,
                              int use_negative_gradient, double *deriv_soa,
                              double *l1, double *l21, double *l2);

/* a13 optimizer steps */
double skmo_standard_rate(double a, double A, double alpha, long long iteration);
double skmo_adaptive_rate(double a, double A, double alpha, double time);
double skmo_adaptive_time(double time, double scalar_product, double fmax, double fmin,
                          double omega);
void skmo_gd_update(long long n, const double *prev, const double *grad, double rate,
                    int orthant, double *out);

/* a14 control-point grid */
void skmo_control_point_grid(int dim, const skmo_grid *image, const double *sigma, skmo_grid *cp);

/* Threads of the OpenMP build of this file (1 in the serial build); n > 0 sets them. */
int skmo_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
