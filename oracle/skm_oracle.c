/*
 * skm_oracle.c -- FP64 CPU restatement of the SKM hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP product library.
 * Loaded only by tests/, __graft_entry__.smoke() and bench.py (cpu_baseline).
 *
 * Citations are `file:line` relative to the reference repository root.
 * The restatement follows the reference's loop structure literally (one
 * "thread" of the reference = one loop iteration here), single threaded,
 * in double precision.
 */
#include "skm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* small helpers                                                       */
/* ------------------------------------------------------------------ */

static int imin(int a, int b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }
static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

/* VecI(VecF) is a truncating C++ cast (include/cuda/Vec.h:44-49). */
static int trunc_int(double v) { return (int)v; }

static long long grid_count(const skmo_grid *g, int dim)
{
    long long n = 1;
    for (int d = 0; d < dim; ++d) n *= g->size[d];
    return n;
}

/* ImageBase::toGlobal / toLocal / toIndex (include/cuda/ImageBase.h:54-72). */
static void to_global(const skmo_grid *g, int dim, const double *idx, double *p)
{
    for (int d = 0; d < dim; ++d) p[d] = g->origin[d] + g->spacing[d] * idx[d];
}
static void to_local(const skmo_grid *g, int dim, const double *p, double *idx)
{
    for (int d = 0; d < dim; ++d) idx[d] = (1.0 / g->spacing[d]) * (p[d] - g->origin[d]);
}
static long long to_index(const skmo_grid *g, int dim, const int *i)
{
    long long r = i[0];
    long long stride = g->size[0];
    for (int d = 1; d < dim; ++d) { r += stride * i[d]; stride *= g->size[d]; }
    return r;
}

/* ImageNearest::atGlobal with toNearest = int(idx + 0.5) and inside()
 * (include/cuda/ImageNearest.h:80-84, ImageBase.h:64-67, 89-95). */
int skmo_nearest_index(const skmo_grid *g, int dim, const double *p, int *idx)
{
    double loc[3];
    to_local(g, dim, p, loc);
    int inside = 1;
    for (int d = 0; d < dim; ++d) {
        idx[d] = trunc_int(loc[d] + 0.5);
        if (idx[d] < 0 || idx[d] >= g->size[d]) inside = 0;
    }
    return inside;
}

/* ------------------------------------------------------------------ */
/* a4: determinant and maximum eigenvalue (include/cuda/Mat.h:227-281) */
/* ------------------------------------------------------------------ */

double skmo_determinant(int dim, const double *m)
{
    if (dim == 2) return m[0] * m[3] - m[2] * m[1];
    return m[0] * m[4] * m[8] + m[3] * m[7] * m[2] + m[6] * m[1] * m[5]
         - m[0] * m[7] * m[5] - m[6] * m[4] * m[2] - m[3] * m[1] * m[8];
}

static void invert(int dim, const double *m, double *out)
{
    double det = skmo_determinant(dim, m);
    if (dim == 2) {
        out[0] = m[3] / det;  out[1] = -m[1] / det;
        out[2] = -m[2] / det; out[3] = m[0] / det;
        return;
    }
    out[0] = (m[4] * m[8] - m[5] * m[7]) / det;
    out[1] = (m[2] * m[7] - m[1] * m[8]) / det;
    out[2] = (m[1] * m[5] - m[2] * m[4]) / det;
    out[3] = (m[5] * m[6] - m[3] * m[8]) / det;
    out[4] = (m[0] * m[8] - m[2] * m[6]) / det;
    out[5] = (m[2] * m[3] - m[0] * m[5]) / det;
    out[6] = (m[3] * m[7] - m[4] * m[6]) / det;
    out[7] = (m[1] * m[6] - m[0] * m[7]) / det;
    out[8] = (m[0] * m[4] - m[1] * m[3]) / det;
}

double skmo_max_eigenvalue(int dim, const double *m)
{
    if (dim == 2) {
        double trace = m[0] + m[3];
        double det = skmo_determinant(2, m);
        return 0.5 * (trace + sqrt(dmax(trace * trace - 4.0 * det, 0.0)));
    }
    double tri = m[1] * m[1] + m[2] * m[2] + m[5] * m[5];
    if (tri == 0.0) return dmax(m[0], dmax(m[4], m[8]));
    double q = (1.0 / 3.0) * (m[0] + m[4] + m[8]);
    double psq = 2.0 * tri + (m[0] - q) * (m[0] - q) + (m[4] - q) * (m[4] - q) + (m[8] - q) * (m[8] - q);
    double p = sqrt(psq * (1.0 / 6.0));
    double b[9];
    memcpy(b, m, sizeof(b));
    b[0] -= q; b[4] -= q; b[8] -= q;
    double r = skmo_determinant(3, b) * (0.5 / (p * p * p));
    double phi = (1.0 / 3.0) * acos(dmin(1.0, dmax(-1.0, r)));
    return q + 2.0 * p * cos(phi);
}

/* ------------------------------------------------------------------ */
/* a1: Wendland kernels (include/cuda/Kernel.h:48-186)                 */
/* ------------------------------------------------------------------ */

/* evaluateWendlandStationaryC4 with r already divided by sigmaC4 (Kernel.h:72-77). */
double skmo_wendland_c4(double t)
{
    double f = dmax(0.0, 1.0 - t);
    return f * f * f * f * f * f * (3.0 + 18.0 * t + 35.0 * t * t) * (560.0 / 1680.0);
}

/* evaluateWendlandStationaryC0 (Kernel.h:48-53). */
double skmo_wendland_c0(double t)
{
    double f = dmax(0.0, 1.0 - t);
    return f * f;
}

static const double *weight_tensor_at(const skmo_problem *pr, const double *q, double *ident)
{
    int idx[3];
    int n = pr->dim;
    if (skmo_nearest_index(&pr->weight, n, q, idx))
        return pr->weight_data + (size_t)to_index(&pr->weight, n, idx) * n * n;
    /* ImageNearest<Mat> default is the identity (ImageNearest.h:44-48). */
    for (int i = 0; i < n * n; ++i) ident[i] = 0.0;
    for (int i = 0; i < n; ++i) ident[i * n + i] = 1.0;
    return ident;
}

static double weight_scalar_at(const skmo_problem *pr, const double *q)
{
    int idx[3];
    if (skmo_nearest_index(&pr->weight, pr->dim, q, idx))
        return pr->weight_data[to_index(&pr->weight, pr->dim, idx)];
    return 0.0;  /* ImageNearest default for scalars (ImageNearest.h:38-42) */
}

/* getSigmaAtPoint / getCovarianceAtPoint (Kernel.h:148-156). */
static double sigma_at(const skmo_problem *pr, const double *p) { return weight_scalar_at(pr, p); }

/* evaluateNonStationary with explicit sigmaQ (Kernel.h:105-117 with :55-61, :79-85). */
static double kernel_nonstationary(const skmo_problem *pr, const double *p, const double *q,
                                   double sigmaP, double sigmaQ)
{
    double F = pow(fabs(sigmaP), 0.25) * pow(fabs(sigmaQ), 0.25) * pow(fabs((sigmaP + sigmaQ) / 2.0), -0.5);
    double r2 = 0.0;
    for (int d = 0; d < pr->dim; ++d) r2 += (p[d] - q[d]) * (p[d] - q[d]);
    double sq = sqrt(2.0 / (sigmaP + sigmaQ)) * sqrt(r2);
    double c4 = F * skmo_wendland_c4(sq / pr->sigma_c4);
    double c0 = F * skmo_wendland_c0(sq / pr->sigma_c0);
    return c4 * c0;  /* F appears in both factors (Kernel.h:112-113) */
}

/* evaluateAnisotropic with explicit covQ (Kernel.h:119-131 with :63-70, :87-94). */
static double kernel_anisotropic(const skmo_problem *pr, const double *p, const double *q,
                                 const double *covP, const double *covQ)
{
    int n = pr->dim;
    double m[9], inv[9], v[3];
    for (int i = 0; i < n * n; ++i) m[i] = (covP[i] + covQ[i]) * 0.5;
    double F = pow(skmo_determinant(n, covP), 0.25) * pow(skmo_determinant(n, covQ), 0.25)
             * pow(skmo_determinant(n, m), -0.5);
    invert(n, m, inv);
    for (int d = 0; d < n; ++d) v[d] = p[d] - q[d];
    double quad = 0.0;
    for (int i = 0; i < n; ++i) {
        double vi = 0.0;                    /* (v*invert(M))[i] (Mat.h:206-216) */
        for (int k = 0; k < n; ++k) vi += v[k] * inv[k * n + i];
        quad += vi * v[i];
    }
    double sq = sqrt(quad);
    double c4 = F * skmo_wendland_c4(sq / pr->sigma_c4);
    double c0 = F * skmo_wendland_c0(sq / pr->sigma_c0);
    return c4 * c0;
}

/* GKernel::evaluate(p, q[, sigmaP | covP]) (Kernel.h:173-186): the second
 * point's sigma / covariance is looked up in the weight image. */
double skmo_kernel_eval(const skmo_problem *pr, const double *p, const double *q,
                        double sigmaP, const double *covP)
{
    if (pr->kernel_mode == SKMO_KERNEL_WEIGHT_IMAGE)
        return kernel_nonstationary(pr, p, q, sigmaP, sigma_at(pr, q));
    if (pr->kernel_mode == SKMO_KERNEL_WEIGHT_TENSOR) {
        double ident[9];
        return kernel_anisotropic(pr, p, q, covP, weight_tensor_at(pr, q, ident));
    }
    double r2 = 0.0;
    for (int d = 0; d < pr->dim; ++d) r2 += (p[d] - q[d]) * (p[d] - q[d]);
    double r = sqrt(r2);
    return skmo_wendland_c4(r / pr->sigma_c4) * skmo_wendland_c0(r / pr->sigma_c0);
}

/* Per-point kernel state used by the reference kernels: sigmaP / covP at the
 * point and the region support (Kernel.h:158-171; e.g. GpuEvaluatorMSE.cu:48-59). */
typedef struct point_state {
    double sigma;
    double cov[9];
    double support;
} point_state;

static void point_state_at(const skmo_problem *pr, const double *p, point_state *st)
{
    double smin = dmin(pr->sigma_c4, pr->sigma_c0);
    st->sigma = 0.0;
    if (pr->kernel_mode == SKMO_KERNEL_WEIGHT_IMAGE) {
        st->sigma = sigma_at(pr, p);
        st->support = smin / sqrt(2.0 / (pr->max_weight + st->sigma));
    } else if (pr->kernel_mode == SKMO_KERNEL_WEIGHT_TENSOR) {
        double ident[9], sum[9];
        const double *c = weight_tensor_at(pr, p, ident);
        int nn = pr->dim * pr->dim;
        memcpy(st->cov, c, sizeof(double) * nn);
        for (int i = 0; i < nn; ++i) sum[i] = pr->max_eig_cov[i] + c[i];
        st->support = smin / sqrt(2.0 / skmo_max_eigenvalue(pr->dim, sum));
    } else {
        st->support = smin;
    }
}

double skmo_region_support(const skmo_problem *pr, const double *p)
{
    point_state st;
    point_state_at(pr, p, &st);
    return st.support;
}

static double kernel_with_state(const skmo_problem *pr, const double *p, const double *q,
                                const point_state *st)
{
    return skmo_kernel_eval(pr, p, q, st->sigma, st->cov);
}

/* The reference cp window around a physical point (e.g. GpuEvaluatorMSE.cu:61-63):
 * lower = int(idx - s/h + 1), upper = int(idx + s/h), clamped to [0, size-1]. */
static void tight_window(const skmo_grid *g, int dim, const double *p, double support,
                         int *lo, int *hi)
{
    double idx[3];
    to_local(g, dim, p, idx);
    for (int d = 0; d < dim; ++d) {
        lo[d] = imin(imax(trunc_int(idx[d] - support / g->spacing[d] + 1.0), 0), g->size[d] - 1);
        hi[d] = imin(imax(trunc_int(idx[d] + support / g->spacing[d]), 0), g->size[d] - 1);
    }
}

/* ------------------------------------------------------------------ */
/* a3: cubic B-spline (include/cuda/BsplineImage.h)                    */
/* ------------------------------------------------------------------ */

/* initialCausalCoefficient (BsplineImage.h:42-55). */
static double causal_init(const double *c, int length, long long stride, double z)
{
    double zn = z;
    double iz = 1.0 / z;
    double z2n = pow(z, (double)(length - 1));
    double sum = c[0] + z2n * c[(long long)(length - 1) * stride];
    z2n *= z2n * iz;
    for (int n = 1; n < length - 1; n++) {
        sum += (zn + z2n) * c[(long long)n * stride];
        zn *= z;
        z2n *= iz;
    }
    return sum / (1.0 - zn * zn);
}

/* initialAntiCausalCoefficient (BsplineImage.h:37-40). */
static double anticausal_init(const double *c, int length, long long stride, double z)
{
    return (z / (z * z - 1.0)) * (z * c[(long long)(length - 2) * stride] + c[(long long)(length - 1) * stride]);
}

/* convertToInterpolationCoefficients (BsplineImage.h:57-70). */
static void convert_line(double *c, int length, long long stride, double z)
{
    double lambda = 2.0 - z - 1.0 / z;
    for (int n = 0; n < length; n++) c[(long long)n * stride] *= lambda;
    c[0] = causal_init(c, length, stride, z);
    for (int n = 1; n < length; n++) c[(long long)n * stride] += z * c[(long long)(n - 1) * stride];
    c[(long long)(length - 1) * stride] = anticausal_init(c, length, stride, z);
    for (int n = length - 2; n >= 0; n--)
        c[(long long)n * stride] = z * (c[(long long)(n + 1) * stride] - c[(long long)n * stride]);
}

/* computeCoeffs (BsplineImage.h:209-221) with computeCoefficients (:174-195)
 * and mirrorCoeffBorders (:72-108).  `coeffs` holds prod(size+3) values. */
void skmo_bspline_coefficients(int dim, const int *size, const double *src, double *coeffs)
{
    long long n = 1;
    for (int d = 0; d < dim; ++d) n *= size[d];
    double *c = (double *)malloc(sizeof(double) * n);
    memcpy(c, src, sizeof(double) * n);
    const double pole = sqrt(3.0) - 2.0;
    if (dim == 2) {
        for (int y = 0; y < size[1]; y++) convert_line(c + (long long)y * size[0], size[0], 1, pole);
        for (int x = 0; x < size[0]; x++) convert_line(c + x, size[1], size[0], pole);
    } else {
        long long sxy = (long long)size[0] * size[1];
        for (int y = 0; y < size[1]; ++y)
            for (int x = 0; x < size[0]; x++) convert_line(c + x + (long long)y * size[0], size[2], sxy, pole);
        for (int z = 0; z < size[2]; ++z)
            for (int x = 0; x < size[0]; x++) convert_line(c + x + z * sxy, size[1], size[0], pole);
        for (int z = 0; z < size[2]; ++z)
            for (int y = 0; y < size[1]; ++y) convert_line(c + (long long)size[0] * (y + (long long)z * size[1]), size[0], 1, pole);
    }
    int ps[3] = {size[0] + 3, size[1] + 3, dim == 3 ? size[2] + 3 : 1};
    int mz = dim == 3 ? ps[2] : 1;
    for (int z = 0; z < mz; ++z)
        for (int y = 0; y < ps[1]; ++y)
            for (int x = 0; x < ps[0]; ++x) {
                int xi = (size[0] - 1) - abs((size[0] - 1) - abs(x - 1));
                int yi = (size[1] - 1) - abs((size[1] - 1) - abs(y - 1));
                int zi = dim == 3 ? (size[2] - 1) - abs((size[2] - 1) - abs(z - 1)) : 0;
                coeffs[x + (long long)ps[0] * (y + (long long)ps[1] * z)] =
                    c[xi + (long long)size[0] * (yi + (long long)size[1] * zi)];
            }
    free(c);
}

/* insideGlobal + derivsGlobal (BsplineImage.h:247-296): `g` is the raw image
 * geometry; the coefficient grid is (size+3) per axis, shifted by one. */
int skmo_bspline_sample(int dim, const skmo_grid *g, const double *coeffs, const double *p,
                        double *value, double *grad)
{
    double idx[3];
    to_local(g, dim, p, idx);
    for (int d = 0; d < dim; ++d)
        if (idx[d] < -0.5 || idx[d] + 3.5 >= g->size[d] + 3) return 0;
    int base[3] = {0, 0, 0};
    double w[3][4] = {{0}}, dw[3][4] = {{0}};
    for (int d = 0; d < dim; ++d) {
        base[d] = trunc_int(idx[d]);
        double t = idx[d] - base[d];
        w[d][3] = (1.0 / 6.0) * t * t * t;
        w[d][0] = (1.0 / 6.0) + 0.5 * t * (t - 1.0) - w[d][3];
        w[d][2] = t + w[d][0] - 2.0 * w[d][3];
        w[d][1] = 1.0 - w[d][0] - w[d][2] - w[d][3];
        dw[d][3] = (3.0 / 6.0) * t * t;
        dw[d][0] = t - 0.5 - dw[d][3];
        dw[d][2] = 1.0 + dw[d][0] - 2.0 * dw[d][3];
        dw[d][1] = -dw[d][0] - dw[d][2] - dw[d][3];
    }
    long long sx = g->size[0] + 3, sy = g->size[1] + 3;
    double val = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
    if (dim == 2) {
        for (int y = 0; y < 4; ++y)
            for (int x = 0; x < 4; ++x) {
                double c = coeffs[(base[0] + x) + (base[1] + y) * sx];
                val += w[0][x] * w[1][y] * c;
                gx += dw[0][x] * w[1][y] * c;
                gy += w[0][x] * dw[1][y] * c;
            }
    } else {
        for (int z = 0; z < 4; ++z)
            for (int y = 0; y < 4; ++y)
                for (int x = 0; x < 4; ++x) {
                    double c = coeffs[(base[0] + x) + (base[1] + y) * sx + (base[2] + z) * sx * sy];
                    val += w[0][x] * w[1][y] * w[2][z] * c;
                    gx += dw[0][x] * w[1][y] * w[2][z] * c;
                    gy += w[0][x] * dw[1][y] * w[2][z] * c;
                    gz += w[0][x] * w[1][y] * dw[2][z] * c;
                }
    }
    *value = val;
    grad[0] = gx / g->spacing[0];
    grad[1] = gy / g->spacing[1];
    if (dim == 3) grad[2] = gz / g->spacing[2];
    return 1;
}

/* ------------------------------------------------------------------ */
/* a5: PCG32 XSH-RR (include/cuda/UniformRng.h:34-55)                  */
/* ------------------------------------------------------------------ */

uint32_t skmo_pcg32_next(uint64_t *state, uint64_t sequence)
{
    uint64_t old = *state;
    *state = old * 6364136223846793005ULL + (sequence | 1);
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((uint32_t)(-(int32_t)rot) & 31));
}

/* Cube offsets: offset_d = int(nextF()*subsample), nextF = nextI/UINT_MAX,
 * drawn cell by cell, axis by axis (GpuEvaluator.cu:107-113; UniformRng.h:43-55). */
void skmo_cube_offsets(uint64_t *state, uint64_t sequence, long long n_cells, int dim,
                       int subsample, int *out)
{
    for (long long i = 0; i < n_cells; ++i)
        for (int d = 0; d < dim; ++d) {
            double f = skmo_pcg32_next(state, sequence) / (double)4294967295u;
            out[i * dim + d] = trunc_int(f * (double)subsample);
        }
}

/* ------------------------------------------------------------------ */
/* a6: kernel-sum displacement                                          */
/* ------------------------------------------------------------------ */

static void df_lookup(const skmo_problem *pr, const double *p, double *u)
{
    for (int d = 0; d < pr->dim; ++d) u[d] = 0.0;
    if (!pr->has_df) return;
    int idx[3];
    if (skmo_nearest_index(&pr->df, pr->dim, p, idx)) {
        long long k = to_index(&pr->df, pr->dim, idx);
        for (int d = 0; d < pr->dim; ++d) u[d] = pr->df_data[k * pr->dim + d];
    }
}

/* Kernel sum at x with the state at x (GpuEvaluatorDisplacementField.cu:35-72):
 * out = sum_j w_j k(x, p_j) c_j, window with the support at x. */
static void kernel_sum(const skmo_problem *pr, const double *x, const point_state *st, double *out)
{
    int n = pr->dim, lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, i[3] = {0, 0, 0};
    tight_window(&pr->cp, n, x, st->support, lo, hi);
    for (int d = 0; d < n; ++d) out[d] = 0.0;
    int z0 = n == 3 ? lo[2] : 0, z1 = n == 3 ? hi[2] : 0;
    for (i[2] = z0; i[2] <= z1; ++i[2])
        for (i[1] = lo[1]; i[1] <= hi[1]; ++i[1])
            for (i[0] = lo[0]; i[0] <= hi[0]; ++i[0]) {
                double fi[3] = {i[0], i[1], i[2]}, pj[3];
                to_global(&pr->cp, n, fi, pj);
                double k = kernel_with_state(pr, x, pj, st);
                long long j = to_index(&pr->cp, n, i);
                for (int d = 0; d < n; ++d) out[d] += pr->cpw[j] * k * pr->cp_aos[j * n + d];
            }
}

void skmo_displacement_at(const skmo_problem *pr, const double *x, double *u)
{
    point_state st;
    double s[3];
    point_state_at(pr, x, &st);
    df_lookup(pr, x, u);
    kernel_sum(pr, x, &st, s);
    for (int d = 0; d < pr->dim; ++d) u[d] += s[d];
}

/* evaluateDisplacement over the full fixed grid (GpuEvaluatorDisplacementField.cu:24-75). */
void skmo_displacement(const skmo_problem *pr, double *out_aos)
{
    int n = pr->dim;
    long long nv = grid_count(&pr->fixed, n);
    for (long long v = 0; v < nv; ++v) {
        int i[3] = {0, 0, 0};
        long long r = v;
        for (int d = 0; d < n; ++d) { i[d] = (int)(r % pr->fixed.size[d]); r /= pr->fixed.size[d]; }
        double fi[3] = {i[0], i[1], n == 3 ? i[2] : 0}, x[3];
        to_global(&pr->fixed, n, fi, x);
        skmo_displacement_at(pr, x, out_aos + v * n);
    }
}

/* ------------------------------------------------------------------ */
/* sampling lattice                                                     */
/* ------------------------------------------------------------------ */

static void subsampled_size(const skmo_problem *pr, int *ns)
{
    for (int d = 0; d < 3; ++d) ns[d] = 1;
    for (int d = 0; d < pr->dim; ++d) ns[d] = pr->fixed.size[d] / pr->subsample;
}

/* samplePos = m*subsample + gridShift[m] (GpuEvaluatorMSE.cu:40-42). */
static void sample_position(const skmo_problem *pr, const int *m, long long flat, int *pos)
{
    for (int d = 0; d < pr->dim; ++d) {
        pos[d] = m[d];
        if (pr->subsample > 1) pos[d] = m[d] * pr->subsample + pr->offsets[flat * pr->dim + d];
    }
}

static double fixed_at(const skmo_problem *pr, const int *pos)
{
    /* A shift equal to `subsample` (nextF()==1) can leave the image; the
     * reference reads out of bounds there.  The oracle clamps, as the
     * product library does. */
    int c[3] = {0, 0, 0};
    for (int d = 0; d < pr->dim; ++d) c[d] = imin(imax(pos[d], 0), pr->fixed.size[d] - 1);
    return pr->fixed_data[to_index(&pr->fixed, pr->dim, c)];
}

/* ------------------------------------------------------------------ */
/* a10: regularizers (include/cuda/GpuEvaluatorRegularizers.cu)        */
/* ------------------------------------------------------------------ */

void skmo_regularizer(const skmo_problem *pr, int which, double weight, double pg_scaling,
                      double *values, double *deriv_soa)
{
    int n = pr->dim;
    long long ncp = grid_count(&pr->cp, n);
    for (long long jj = 0; jj < ncp; ++jj) {
        int j[3] = {0, 0, 0};
        long long r = jj;
        for (int d = 0; d < n; ++d) { j[d] = (int)(r % pr->cp.size[d]); r /= pr->cp.size[d]; }
        double fj[3] = {j[0], j[1], j[2]}, pj[3];
        to_global(&pr->cp, n, fj, pj);
        double wj = pr->cpw[jj];
        const double *cj = pr->cp_aos + jj * n;
        point_state st;
        point_state_at(pr, pj, &st);
        /* the window is built from the integer cp index itself (:52-54) */
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        for (int d = 0; d < n; ++d) {
            lo[d] = imin(imax(trunc_int(fj[d] - st.support / pr->cp.spacing[d] + 1.0), 0), pr->cp.size[d] - 1);
            hi[d] = imin(imax(trunc_int(fj[d] + st.support / pr->cp.spacing[d]), 0), pr->cp.size[d] - 1);
        }
        double value = 0.0, der[3] = {0, 0, 0};
        double s = pg_scaling;
        int i[3] = {0, 0, 0};
        int z0 = n == 3 ? lo[2] : 0, z1 = n == 3 ? hi[2] : 0;
        for (i[2] = z0; i[2] <= z1; ++i[2])
            for (i[1] = lo[1]; i[1] <= hi[1]; ++i[1])
                for (i[0] = lo[0]; i[0] <= hi[0]; ++i[0]) {
                    double fi[3] = {i[0], i[1], i[2]}, pi[3];
                    to_global(&pr->cp, n, fi, pi);
                    const double *ci = pr->cp_aos + to_index(&pr->cp, n, i) * n;
                    double k = kernel_with_state(pr, pj, pi, &st);
                    if (which == 0) {            /* RKHS (:59-79) */
                        double dot = 0.0;
                        for (int d = 0; d < n; ++d) dot += k * wj * cj[d] * ci[d];
                        value += dot;
                        for (int d = 0; d < n; ++d) der[d] += 2.0 * ci[d] * k * wj;
                    } else if (which == 1) {     /* RD (:131-151) */
                        double l2 = 0.0;
                        for (int d = 0; d < n; ++d) l2 += (cj[d] - ci[d]) * (cj[d] - ci[d]);
                        value += l2 * k * wj;
                        for (int d = 0; d < n; ++d) der[d] += 4.0 * (cj[d] - ci[d]) * k * wj;
                    } else {                     /* PG (:203-240; 2D :431-456) */
                        if (n == 3) {
                            double cr0 = ci[1] * cj[2] - ci[2] * cj[1];
                            double cr1 = ci[2] * cj[0] - ci[0] * cj[2];
                            double cr2 = ci[0] * cj[1] - ci[1] * cj[0];
                            double nrm2 = cr0 * cr0 + cr1 * cr1 + cr2 * cr2;
                            value += k * wj * s * s / 2.0 * log(1.0 + nrm2 / s);
                            double dp0 = 2.0 * ci[2] * (ci[2] * cj[0] - ci[0] * cj[2]) - 2.0 * ci[1] * (ci[0] * cj[1] - ci[1] * cj[0]);
                            double dp1 = 2.0 * ci[0] * (ci[0] * cj[1] - ci[1] * cj[0]) - 2.0 * ci[2] * (ci[1] * cj[2] - ci[2] * cj[1]);
                            double dp2 = 2.0 * ci[1] * (ci[1] * cj[2] - ci[2] * cj[1]) - 2.0 * ci[0] * (ci[2] * cj[0] - ci[0] * cj[2]);
                            double den = 2.0 * (nrm2 / s + 1.0);
                            der[0] += 2.0 * k * wj * (s * dp0 / den);
                            der[1] += 2.0 * k * wj * (s * dp1 / den);
                            der[2] += 2.0 * k * wj * (s * dp2 / den);
                        } else {
                            double val = ci[0] * cj[1] - ci[1] * cj[0];
                            double val2 = val * val;
                            value += k * wj * s * s / 2.0 * log(1.0 + val2 / s);
                            double factor = val2 / s + 1.0;
                            der[0] += 2.0 * k * wj * (-s * ci[1] * val / factor);
                            der[1] += 2.0 * k * wj * (s * ci[0] * val / factor);
                        }
                    }
                }
        if (values) values[jj] = value;
        if (deriv_soa)
            for (int d = 0; d < n; ++d) deriv_soa[jj + d * ncp] += weight * der[d];
    }
}

/* Bilateral regularizer: no counterpart in the reference, whose closest term
 * is RD (GpuEvaluatorRegularizers.cu:96-166, sum_i k(p_j, p_i) w_j |c_j - c_i|^2
 * with the transform kernel) and which names a radiometric term only in a
 * comment (itkRegularizedImageToImageMetricv4.h:204).  Here the spatial
 * factor is a Gaussian of standard deviation sigma_s cut at 3 sigma_s and
 * the intensity factor a Gaussian of standard deviation sigma_i of the
 * difference of the fixed image at the two control points (nearest
 * neighbour, 0 outside):
 *   B_ij = exp(-r_ij^2 / (2 sigma_s^2)) exp(-(I_j - I_i)^2 / (2 sigma_i^2)),  r_ij < 3 sigma_s
 *   value_j = sum_i w_j B_ij |c_j - c_i|^2        (unweighted by `weight`, like RD)
 *   d/dc_j (sum_j value_j) = sum_i 2 B_ij (w_j + w_i)(c_j - c_i)   (B symmetric)
 * The derivative, times `weight`, is added to deriv_soa.  Offsets between
 * control points are (index difference) x spacing. */
void skmo_regularizer_bilateral(const skmo_problem *pr, double weight, double sigma_s, double sigma_i,
                                double *values, double *deriv_soa)
{
    int n = pr->dim;
    long long ncp = grid_count(&pr->cp, n);
    double *inten = (double *)calloc((size_t)ncp, sizeof(double));
    for (long long jj = 0; jj < ncp; ++jj) {
        int j[3] = {0, 0, 0}, f[3] = {0, 0, 0};
        long long r = jj;
        for (int d = 0; d < n; ++d) { j[d] = (int)(r % pr->cp.size[d]); r /= pr->cp.size[d]; }
        double fj[3] = {j[0], j[1], j[2]}, pj[3];
        to_global(&pr->cp, n, fj, pj);
        inten[jj] = skmo_nearest_index(&pr->fixed, n, pj, f) ? pr->fixed_data[to_index(&pr->fixed, n, f)] : 0.0;
    }
    const double cut = 3.0 * sigma_s;
    for (long long jj = 0; jj < ncp; ++jj) {
        int j[3] = {0, 0, 0};
        long long r = jj;
        for (int d = 0; d < n; ++d) { j[d] = (int)(r % pr->cp.size[d]); r /= pr->cp.size[d]; }
        const double wj = pr->cpw[jj];
        const double *cj = pr->cp_aos + jj * n;
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        for (int d = 0; d < n; ++d) {
            lo[d] = imin(imax(trunc_int((double)j[d] - cut / pr->cp.spacing[d] + 1.0), 0), pr->cp.size[d] - 1);
            hi[d] = imin(imax(trunc_int((double)j[d] + cut / pr->cp.spacing[d]), 0), pr->cp.size[d] - 1);
        }
        double value = 0.0, der[3] = {0, 0, 0};
        int i[3] = {0, 0, 0};
        int z0 = n == 3 ? lo[2] : 0, z1 = n == 3 ? hi[2] : 0;
        for (i[2] = z0; i[2] <= z1; ++i[2])
            for (i[1] = lo[1]; i[1] <= hi[1]; ++i[1])
                for (i[0] = lo[0]; i[0] <= hi[0]; ++i[0]) {
                    double r2 = 0.0;
                    for (int d = 0; d < n; ++d) {
                        double v = (double)(j[d] - i[d]) * pr->cp.spacing[d];
                        r2 += v * v;
                    }
                    if (!(r2 < cut * cut)) continue;
                    long long ii = to_index(&pr->cp, n, i);
                    double di = inten[jj] - inten[ii];
                    double b = exp(-r2 / (2.0 * sigma_s * sigma_s)) * exp(-di * di / (2.0 * sigma_i * sigma_i));
                    const double *ci = pr->cp_aos + ii * n;
                    double l2 = 0.0;
                    for (int d = 0; d < n; ++d) l2 += (cj[d] - ci[d]) * (cj[d] - ci[d]);
                    value += wj * b * l2;
                    for (int d = 0; d < n; ++d) der[d] += 2.0 * b * (wj + pr->cpw[ii]) * (cj[d] - ci[d]);
                }
        if (values) values[jj] = value;
        if (deriv_soa)
            for (int d = 0; d < n; ++d) deriv_soa[jj + d * ncp] += weight * der[d];
    }
    free(inten);
}

/* ------------------------------------------------------------------ */
/* a9: LCC helpers (include/cuda/GpuEvaluatorLCC.cu)                    */
/* ------------------------------------------------------------------ */

/* neighborhoodWeight (GpuEvaluatorLCC.cu:28-35). */
static double neighborhood_weight(int dim, const double *p, const double *q, double support, double beta)
{
    double r2 = 0.0;
    for (int d = 0; d < dim; ++d) r2 += (p[d] - q[d]) * (p[d] - q[d]);
    double r = pow(sqrt(r2) / support, beta);
    double f = dmax(0.0, 1.0 - r);
    return f * f;
}

/* Warped neighbour: y + df(y) + kernel sum with the state at y; `quirk3d`
 * reproduces the 3D derivative kernel's evaluate(cp_i, y, sigma(y)) call, which
 * looks the second sigma up at y as well (GpuEvaluatorLCC.cu:428-443). */
static void warp_neighbor(const skmo_problem *pr, const double *y, const point_state *sty,
                          int quirk3d, double *out)
{
    int n = pr->dim;
    double u[3];
    df_lookup(pr, y, u);
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, i[3] = {0, 0, 0};
    tight_window(&pr->cp, n, y, sty->support, lo, hi);
    int z0 = n == 3 ? lo[2] : 0, z1 = n == 3 ? hi[2] : 0;
    for (i[2] = z0; i[2] <= z1; ++i[2])
        for (i[1] = lo[1]; i[1] <= hi[1]; ++i[1])
            for (i[0] = lo[0]; i[0] <= hi[0]; ++i[0]) {
                double fi[3] = {i[0], i[1], i[2]}, pj[3];
                to_global(&pr->cp, n, fi, pj);
                double k;
                if (quirk3d && pr->kernel_mode == SKMO_KERNEL_WEIGHT_IMAGE)
                    k = kernel_nonstationary(pr, pj, y, sty->sigma, sty->sigma);
                else if (quirk3d && pr->kernel_mode == SKMO_KERNEL_WEIGHT_TENSOR)
                    k = kernel_anisotropic(pr, pj, y, sty->cov, sty->cov);
                else
                    k = kernel_with_state(pr, y, pj, sty);
                long long j = to_index(&pr->cp, n, i);
                for (int d = 0; d < n; ++d) u[d] += pr->cpw[j] * k * pr->cp_aos[j * n + d];
            }
    for (int d = 0; d < n; ++d) out[d] = y[d] + u[d];
}

/* ka = evaluate(globalPoint, y, sigmaP) inside the derivative: sigmaP is the
 * neighbour's (shadowed, :411-422) and the second lookup is at y as well. */
static double kernel_ka(const skmo_problem *pr, const double *pj, const double *y, const point_state *sty)
{
    if (pr->kernel_mode == SKMO_KERNEL_WEIGHT_IMAGE) return kernel_nonstationary(pr, pj, y, sty->sigma, sty->sigma);
    if (pr->kernel_mode == SKMO_KERNEL_WEIGHT_TENSOR) return kernel_anisotropic(pr, pj, y, sty->cov, sty->cov);
    return skmo_kernel_eval(pr, pj, y, 0.0, NULL);
}

/* Neighbour box of a centre on the fixed grid: tight window with the centre's
 * support, stepped by subsampleNeighborhood from the lower bound (:95-107). */
static void neighbor_box(const skmo_problem *pr, const double *x, double support, int *lo, int *hi)
{
    tight_window(&pr->fixed, pr->dim, x, support, lo, hi);
}

typedef struct lcc_sums {
    double sf, sm, sff, smm, sfm;
    int count;
} lcc_sums;

/* The warped neighbour y + u(y) and its B-spline sample depend only on the
 * fixed voxel y (and, in the 3D derivative, on the sigma(y)-twice quirk), so
 * they are computed once per voxel and looked up; the reference recomputes
 * them for every (centre, neighbour) pair and gets the same numbers. */
typedef struct lcc_cache {
    int *in[2];
    double *val[2];
    double *grad[2];
} lcc_cache;

static int lcc_cache_build(const skmo_problem *pr, int with_quirk, lcc_cache *c)
{
    int n = pr->dim;
    long long nv = grid_count(&pr->fixed, n);
    memset(c, 0, sizeof(*c));
    for (int q = 0; q <= with_quirk; ++q) {
        c->in[q] = (int *)calloc((size_t)nv, sizeof(int));
        c->val[q] = (double *)calloc((size_t)nv, sizeof(double));
        c->grad[q] = (double *)calloc((size_t)nv * 3, sizeof(double));
        if (!c->in[q] || !c->val[q] || !c->grad[q]) return -1;
    }
    if (!with_quirk) { c->in[1] = c->in[0]; c->val[1] = c->val[0]; c->grad[1] = c->grad[0]; }
#pragma omp parallel for schedule(dynamic, 256)
    for (long long v = 0; v < nv; ++v) {
        int f[3] = {0, 0, 0};
        long long r = v;
        for (int d = 0; d < n; ++d) { f[d] = (int)(r % pr->fixed.size[d]); r /= pr->fixed.size[d]; }
        double ff[3] = {f[0], f[1], f[2]}, y[3];
        to_global(&pr->fixed, n, ff, y);
        point_state sty;
        point_state_at(pr, y, &sty);
        for (int q = 0; q <= with_quirk; ++q) {
            double my[3], mv = 0.0, grad[3] = {0, 0, 0};
            warp_neighbor(pr, y, &sty, q, my);
            c->in[q][v] = skmo_bspline_sample(n, &pr->moving, pr->moving_coeffs, my, &mv, grad);
            c->val[q][v] = mv;
            for (int d = 0; d < 3; ++d) c->grad[q][v * 3 + d] = grad[d];
        }
    }
    return 0;
}

static void lcc_cache_free(lcc_cache *c)
{
    int distinct = c->in[1] != c->in[0];
    for (int q = 0; q < 1 + distinct; ++q) { free(c->in[q]); free(c->val[q]); free(c->grad[q]); }
    memset(c, 0, sizeof(*c));
}

/* Sample of the warped neighbour at fixed index f (quirk: 3D derivative form). */
static int lcc_sample(const skmo_problem *pr, const lcc_cache *c, int quirk, const int *f, double *mv, double *grad)
{
    long long v = to_index(&pr->fixed, pr->dim, f);
    *mv = c->val[quirk][v];
    for (int d = 0; d < 3; ++d) grad[d] = c->grad[quirk][v * 3 + d];
    return c->in[quirk][v];
}

/* LCC value at one centre (GpuEvaluatorLCC.cu:94-191). */
static double lcc_centre_value(const skmo_problem *pr, const lcc_cache *cache, const double *x,
                               const point_state *stx)
{
    int n = pr->dim, lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, f[3] = {0, 0, 0};
    neighbor_box(pr, x, stx->support, lo, hi);
    int sn = pr->subsample_neighborhood;
    lcc_sums s = {0, 0, 0, 0, 0, 0};
    int z0 = n == 3 ? lo[2] : 0, z1 = n == 3 ? hi[2] : 0;
    for (f[2] = z0; f[2] <= z1; f[2] += sn)
        for (f[1] = lo[1]; f[1] <= hi[1]; f[1] += sn)
            for (f[0] = lo[0]; f[0] <= hi[0]; f[0] += sn) {
                double fi[3] = {f[0], f[1], f[2]}, y[3], mv, grad[3];
                to_global(&pr->fixed, n, fi, y);
                point_state sty;
                point_state_at(pr, y, &sty);
                double kw = neighborhood_weight(n, x, y, sty.support, 10.0);
                if (kw > 0.0 && lcc_sample(pr, cache, 0, f, &mv, grad)) {
                    mv *= kw;
                    double fv = kw * pr->fixed_data[to_index(&pr->fixed, n, f)];
                    s.sf += fv; s.sm += mv; s.sff += fv * fv; s.smm += mv * mv; s.sfm += fv * mv;
                    s.count++;
                }
            }
    double cc = 0.0;
    if (s.count > 0) {
        double msf = s.sf / s.count, msm = s.sm / s.count;
        double d1 = s.sff - 2.0 * msf * s.sf + s.count * msf * msf;
        double d2 = s.smm - 2.0 * msm * s.sm + s.count * msm * msm;
        if (d1 > 0 && d2 > 0) {
            cc = s.sfm - msm * s.sf - msf * s.sm + s.count * msf * msm;
            cc = -cc * cc / (d1 * d2);
        }
    }
    return cc;
}

/* Reference-mode LCC derivative of one cp (GpuEvaluatorLCC.cu:206-530; 2D 696-961):
 * centres restricted to the tight cp window, w_j not applied, |u|>1e-16 guard. */
static void lcc_derivative_reference(const skmo_problem *pr, const lcc_cache *cache, const int *j, double *der)
{
    int n = pr->dim, ns[3];
    subsampled_size(pr, ns);
    double fj[3] = {j[0], j[1], j[2]}, pj[3], idx[3];
    to_global(&pr->cp, n, fj, pj);
    point_state stj;
    point_state_at(pr, pj, &stj);
    to_local(&pr->fixed, n, pj, idx);
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int d = 0; d < n; ++d) {
        double sp = stj.support / pr->fixed.spacing[d];
        if (pr->subsample > 1) {
            lo[d] = imin(imax(trunc_int((idx[d] - sp) / (double)pr->subsample + 1.0), 0), ns[d] - 1);
            hi[d] = imin(imax(trunc_int((idx[d] + sp) / (double)pr->subsample), 0), ns[d] - 1);
        } else {
            lo[d] = imin(imax(trunc_int(idx[d] - sp + 1.0), 0), pr->fixed.size[d] - 1);
            hi[d] = imin(imax(trunc_int(idx[d] + sp), 0), pr->fixed.size[d] - 1);
        }
    }
    for (int d = 0; d < n; ++d) der[d] = 0.0;
    int m[3] = {0, 0, 0};
    int z0 = n == 3 ? lo[2] : 0, z1 = n == 3 ? hi[2] : 0;
    for (m[2] = z0; m[2] <= z1; ++m[2])
        for (m[1] = lo[1]; m[1] <= hi[1]; ++m[1])
            for (m[0] = lo[0]; m[0] <= hi[0]; ++m[0]) {
                long long flat = m[0] + (long long)ns[0] * (m[1] + (long long)ns[1] * m[2]);
                int pos[3];
                sample_position(pr, m, flat, pos);
                double fp[3] = {pos[0], pos[1], n == 3 ? pos[2] : 0}, xi[3];
                to_global(&pr->fixed, n, fp, xi);
                point_state sti;
                point_state_at(pr, xi, &sti);
                int flo[3] = {0, 0, 0}, fhi[3] = {0, 0, 0}, f[3] = {0, 0, 0};
                neighbor_box(pr, xi, sti.support, flo, fhi);
                double sf = 0, sm = 0, sff = 0, smm = 0, sfm = 0;
                double mkcka[3] = {0, 0, 0}, mkcmkcka[3] = {0, 0, 0}, fkcmka[3] = {0, 0, 0};
                int count = 0;
                int sn = pr->subsample_neighborhood;
                int fz0 = n == 3 ? flo[2] : 0, fz1 = n == 3 ? fhi[2] : 0;
                for (f[2] = fz0; f[2] <= fz1; f[2] += sn)
                    for (f[1] = flo[1]; f[1] <= fhi[1]; f[1] += sn)
                        for (f[0] = flo[0]; f[0] <= fhi[0]; f[0] += sn) {
                            double ff[3] = {f[0], f[1], f[2]}, y[3], mv, grad[3];
                            to_global(&pr->fixed, n, ff, y);
                            point_state sty;
                            point_state_at(pr, y, &sty);
                            double kw = neighborhood_weight(n, xi, y, sty.support, 10.0);
                            double ka = kernel_ka(pr, pj, y, &sty);
                            if (kw > 0.0 && lcc_sample(pr, cache, n == 3, f, &mv, grad)) {
                                mv *= kw;
                                double fv = kw * pr->fixed_data[to_index(&pr->fixed, n, f)];
                                sf += fv; sm += mv; sff += fv * fv; smm += mv * mv; sfm += fv * mv;
                                for (int d = 0; d < n; ++d) {
                                    double gk = grad[d] * kw;
                                    mkcmkcka[d] += mv * gk * ka;
                                    mkcka[d] += gk * ka;
                                    fkcmka[d] += fv * gk * ka;
                                }
                                count++;
                            }
                        }
                if (count > 0) {
                    double msf = sf / count, msm = sm / count;
                    double d1 = sff - 2.0 * msf * sf + count * msf * msf;
                    double u = d1 * (smm - 2.0 * msm * sm + count * msm * msm);
                    double v = sfm - msm * sf - msf * sm + count * msf * msm;
                    double vdev[3], udev[3];
                    for (int d = 0; d < n; ++d) {
                        vdev[d] = 2.0 * v * (fkcmka[d] - msf * mkcka[d]);
                        udev[d] = 2.0 * d1 * (mkcmkcka[d] - msm * mkcka[d]);
                    }
                    v = v * v;
                    if (fabs(u) > 1e-16)
                        for (int d = 0; d < n; ++d) der[d] += -(vdev[d] * u - udev[d] * v) / (u * u);
                }
            }
}

/* ------------------------------------------------------------------ */
/* a7-a9, a12: evaluator (include/cuda/GpuEvaluator.cu:94-714)          */
/* ------------------------------------------------------------------ */

int skmo_evaluate(const skmo_problem *pr, int measure, int with_derivative,
                  double reg_rkhs, double reg_rd, double reg_pg, double pg_scaling,
                  double *deriv_soa, double *per_sample, skmo_result *res)
{
    int n = pr->dim, ns[3];
    subsampled_size(pr, ns);
    long long nsub = (long long)ns[0] * ns[1] * (n == 3 ? ns[2] : 1);
    long long ncp = grid_count(&pr->cp, n);
    double *a = (double *)calloc((size_t)nsub, sizeof(double));   /* diff | f | cc */
    double *b = (double *)calloc((size_t)nsub, sizeof(double));   /* m              */
    double *g = (double *)calloc((size_t)nsub * n, sizeof(double));
    double *xs = (double *)calloc((size_t)nsub * n, sizeof(double));
    int *inside = (int *)calloc((size_t)nsub, sizeof(int));
    if (!a || !b || !g || !xs || !inside) return -1;
    lcc_cache cache;
    memset(&cache, 0, sizeof(cache));
    if (measure == SKMO_LCC &&
        lcc_cache_build(pr, with_derivative && !pr->exact && n == 3 && pr->kernel_mode != SKMO_KERNEL_STATIONARY,
                        &cache) != 0)
        return -1;

    /* value pass: one iteration per subsampled sample (e.g. GpuEvaluatorMSE.cu:29-95);
       iterations are independent (the OpenMP build runs them in parallel) */
#pragma omp parallel for schedule(dynamic, 256)
    for (long long s = 0; s < nsub; ++s) {
        int m[3] = {0, 0, 0}, pos[3] = {0, 0, 0};
        long long r = s;
        for (int d = 0; d < n; ++d) { m[d] = (int)(r % ns[d]); r /= ns[d]; }
        sample_position(pr, m, s, pos);
        double fp[3] = {pos[0], pos[1], n == 3 ? pos[2] : 0}, x[3], w[3], val, grad[3] = {0, 0, 0};
        to_global(&pr->fixed, n, fp, x);
        for (int d = 0; d < n; ++d) xs[s * n + d] = x[d];
        skmo_displacement_at(pr, x, w);
        for (int d = 0; d < n; ++d) w[d] += x[d];
        int in = skmo_bspline_sample(n, &pr->moving, pr->moving_coeffs, w, &val, grad);
        inside[s] = in;
        if (measure == SKMO_LCC) {
            point_state st;
            point_state_at(pr, x, &st);
            a[s] = lcc_centre_value(pr, &cache, x, &st);
            continue;
        }
        if (in) {
            double fv = fixed_at(pr, pos);
            if (measure == SKMO_MSE) a[s] = val - fv;
            else { a[s] = fv; b[s] = val; }
            for (int d = 0; d < n; ++d) g[s * n + d] = grad[d];
        }
    }

    /* host reductions (GpuEvaluator.cu:464-471, 501-534, 560-570) */
    long long count = 0;
    double measure_value = 0.0, sf = 0, sm = 0, sff = 0, smm = 0, sfm = 0, denom = 0;
    for (long long s = 0; s < nsub; ++s) count += inside[s];
    if (measure == SKMO_MSE) {
        for (long long s = 0; s < nsub; ++s) measure_value += a[s] * a[s];
        measure_value /= (double)count;
    } else if (measure == SKMO_NCC) {
        for (long long s = 0; s < nsub; ++s) {
            sff += a[s] * a[s]; smm += b[s] * b[s]; sfm += a[s] * b[s]; sf += a[s]; sm += b[s];
        }
        sff -= sf * sf / count;
        smm -= sm * sm / count;
        sfm -= sf * sm / count;
        denom = -1.0 * sqrt(sff * smm);
        measure_value = denom != 0 ? sfm / denom : 0.0;
    } else {
        for (long long s = 0; s < nsub; ++s) measure_value += a[s];
        measure_value /= (double)count;
    }
    if (per_sample) for (long long s = 0; s < nsub; ++s) per_sample[s] = a[s];

    res->measure = measure_value;
    res->pixels_counted = count;
    res->reg_rkhs = res->reg_rd = res->reg_pg = 0.0;

    if (with_derivative) {
        double mean_f = sf / count, mean_m = sm / count;
        for (long long jj = 0; jj < ncp * n; ++jj) deriv_soa[jj] = 0.0;
        /* one control point per iteration, independent (parallel in the OpenMP build) */
#pragma omp parallel for schedule(dynamic, 4)
        for (long long jj = 0; jj < ncp; ++jj) {
            int j[3] = {0, 0, 0};
            long long r = jj;
            for (int d = 0; d < n; ++d) { j[d] = (int)(r % pr->cp.size[d]); r /= pr->cp.size[d]; }
            double der[3] = {0, 0, 0};
            if (measure == SKMO_LCC) {
                if (pr->exact) continue;  /* exact LCC gradient handled below */
                lcc_derivative_reference(pr, &cache, j, der);
                for (int d = 0; d < n; ++d) deriv_soa[jj + d * ncp] = der[d] / (double)count;
                continue;
            }
            double fj[3] = {j[0], j[1], j[2]}, pj[3], idx[3];
            to_global(&pr->cp, n, fj, pj);
            double wj = pr->cpw[jj];
            point_state st;
            point_state_at(pr, pj, &st);
            to_local(&pr->fixed, n, pj, idx);
            int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
            for (int d = 0; d < n; ++d) {
                double sp = st.support / pr->fixed.spacing[d];
                int sub = pr->subsample;
                int lim = (sub > 1 ? ns[d] : pr->fixed.size[d]) - 1;
                double dsub = (double)sub;
                if (n == 3 || pr->exact) {
                    /* 3D superset window (GpuEvaluatorMSE.cu:126-132) */
                    lo[d] = trunc_int((idx[d] - sp) / dsub);
                    hi[d] = trunc_int((idx[d] + sp) / dsub + 1.0);
                } else {
                    /* 2D tight window (GpuEvaluatorMSE.cu:267-273) */
                    lo[d] = trunc_int((idx[d] - sp) / dsub + 1.0);
                    hi[d] = trunc_int((idx[d] + sp) / dsub);
                }
                if (pr->exact) { lo[d] -= 1; hi[d] += 1; }
                lo[d] = imin(imax(lo[d], 0), lim);
                hi[d] = imin(imax(hi[d], 0), lim);
            }
            double derF[3] = {0, 0, 0}, derM[3] = {0, 0, 0};
            int m[3] = {0, 0, 0};
            int z0 = n == 3 ? lo[2] : 0, z1 = n == 3 ? hi[2] : 0;
            for (m[2] = z0; m[2] <= z1; ++m[2])
                for (m[1] = lo[1]; m[1] <= hi[1]; ++m[1])
                    for (m[0] = lo[0]; m[0] <= hi[0]; ++m[0]) {
                        long long s = m[0] + (long long)ns[0] * (m[1] + (long long)ns[1] * m[2]);
                        double k = kernel_with_state(pr, pj, xs + s * n, &st);
                        for (int d = 0; d < n; ++d) {
                            double gd = g[s * n + d];
                            if (measure == SKMO_MSE) {
                                /* 2*diff*w_i*k*gradient (GpuEvaluatorMSE.cu:160) */
                                der[d] += 2.0 * a[s] * wj * k * gd;
                            } else {
                                /* derivativeF / derivativeM (GpuEvaluatorNCC.cu:177-179) */
                                double diff = wj * k * gd;
                                derF[d] += a[s] * diff - diff * mean_f;
                                derM[d] += b[s] * diff - diff * mean_m;
                            }
                        }
                    }
            for (int d = 0; d < n; ++d) {
                double v;
                if (measure == SKMO_MSE) v = der[d] / (double)count;          /* GpuEvaluator.cu:602-603 */
                else v = denom != 0 ? (derF[d] - (sfm / smm) * derM[d]) / denom : 0.0;  /* :628-633 */
                deriv_soa[jj + d * ncp] = v;
            }
        }
        if (measure == SKMO_LCC && pr->exact) {
            /* Exact gradient of sum_x cc(x)/N: per-centre coefficients are
             * scattered onto their neighbours, then gathered per cp with the
             * true kernel k(p_j, y) and the weight w_j. */
            long long nv = grid_count(&pr->fixed, n);
            double *gam = (double *)calloc((size_t)nv * n, sizeof(double));
            for (long long s = 0; s < nsub; ++s) {
                double *x = xs + s * n;
                point_state stx;
                point_state_at(pr, x, &stx);
                int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, f[3] = {0, 0, 0};
                neighbor_box(pr, x, stx.support, lo, hi);
                int sn = pr->subsample_neighborhood;
                double sfx = 0, smx = 0, sffx = 0, smmx = 0, sfmx = 0;
                int cnt = 0;
                int z0 = n == 3 ? lo[2] : 0, z1 = n == 3 ? hi[2] : 0;
                for (int pass = 0; pass < 2; ++pass) {
                    double msf = cnt ? sfx / cnt : 0, msm = cnt ? smx / cnt : 0;
                    double d1 = sffx - 2.0 * msf * sfx + cnt * msf * msf;
                    double d2 = smmx - 2.0 * msm * smx + cnt * msm * msm;
                    double v = sfmx - msm * sfx - msf * smx + cnt * msf * msm;
                    /* cc's condition (GpuEvaluatorLCC.cu:186) and the
                     * derivative guard |u| > 1e-16 (:517) */
                    if (pass == 1 && !(cnt > 0 && d1 > 0 && d2 > 0 && fabs(d1 * d2) > 1e-16)) break;
                    for (f[2] = z0; f[2] <= z1; f[2] += sn)
                        for (f[1] = lo[1]; f[1] <= hi[1]; f[1] += sn)
                            for (f[0] = lo[0]; f[0] <= hi[0]; f[0] += sn) {
                                double ff[3] = {f[0], f[1], f[2]}, y[3], mv, grad[3];
                                to_global(&pr->fixed, n, ff, y);
                                point_state sty;
                                point_state_at(pr, y, &sty);
                                double kw = neighborhood_weight(n, x, y, sty.support, 10.0);
                                if (!(kw > 0.0 && lcc_sample(pr, &cache, 0, f, &mv, grad)))
                                    continue;
                                double fv = kw * pr->fixed_data[to_index(&pr->fixed, n, f)];
                                double mvk = kw * mv;
                                if (pass == 0) {
                                    sfx += fv; smx += mvk; sffx += fv * fv; smmx += mvk * mvk; sfmx += fv * mvk;
                                    cnt++;
                                } else {
                                    /* d(-v^2/(d1 d2))/dM~(y), M~ = kw*M(y+u) */
                                    double dv = fv - msf;
                                    double dd2 = 2.0 * (mvk - msm);
                                    double beta = -(2.0 * v * dv * d2 - v * v * dd2) / (d1 * d2 * d2);
                                    long long yi = to_index(&pr->fixed, n, f);
                                    for (int d = 0; d < n; ++d) gam[yi * n + d] += beta * kw * grad[d];
                                }
                            }
                }
            }
            for (long long jj = 0; jj < ncp; ++jj) {
                int j[3] = {0, 0, 0};
                long long r = jj;
                for (int d = 0; d < n; ++d) { j[d] = (int)(r % pr->cp.size[d]); r /= pr->cp.size[d]; }
                double fj[3] = {j[0], j[1], j[2]}, pj[3];
                to_global(&pr->cp, n, fj, pj);
                point_state stj;
                point_state_at(pr, pj, &stj);
                double der[3] = {0, 0, 0};
                for (long long v = 0; v < nv; ++v) {
                    int dz = 0;
                    for (int d = 0; d < n; ++d) if (gam[v * n + d] != 0.0) dz = 1;
                    if (!dz) continue;
                    int yi[3] = {0, 0, 0};
                    long long rr = v;
                    for (int d = 0; d < n; ++d) { yi[d] = (int)(rr % pr->fixed.size[d]); rr /= pr->fixed.size[d]; }
                    double fy[3] = {yi[0], yi[1], yi[2]}, y[3];
                    to_global(&pr->fixed, n, fy, y);
                    double k = kernel_with_state(pr, pj, y, &stj);
                    for (int d = 0; d < n; ++d) der[d] += pr->cpw[jj] * k * gam[v * n + d];
                }
                for (int d = 0; d < n; ++d) deriv_soa[jj + d * ncp] = der[d] / (double)count;
            }
            free(gam);
        }
    }

    /* regularizers: value per cp summed, derivative added after normalisation
     * (GpuEvaluator.cu:652-711) */
    double *vals = (double *)calloc((size_t)ncp, sizeof(double));
    double w[3] = {reg_rkhs, reg_rd, reg_pg};
    double *out[3] = {&res->reg_rkhs, &res->reg_rd, &res->reg_pg};
    for (int which = 0; which < 3; ++which) {
        if (!(w[which] > 0)) continue;
        skmo_regularizer(pr, which, w[which], pg_scaling, vals, with_derivative ? deriv_soa : NULL);
        double sum = 0.0;
        for (long long jj = 0; jj < ncp; ++jj) sum += vals[jj];
        *out[which] = sum;
    }
    free(vals);
    free(a); free(b); free(g); free(xs); free(inside);
    if (measure == SKMO_LCC) lcc_cache_free(&cache);
    return 0;
}

/* ------------------------------------------------------------------ */
/* a11: host regularizers (itkRegularizedImageToImageMetricv4.hxx:430-538) */
/* ------------------------------------------------------------------ */

double skmo_host_regularizers(int dim, long long n_cp, const double *params_soa,
                              double reg_l1, double reg_l21, double reg_l2,
                              int use_negative_gradient, double *deriv_soa,
                              double *l1, double *l21, double *l2)
{
    long long np = n_cp * dim;
    *l1 = *l21 = *l2 = 0.0;
    if (reg_l1 > 0) {
        for (long long p = 0; p < np; ++p) *l1 += fabs(params_soa[p]);
        for (long long p = 0; p < np; ++p) {
            double par = params_soa[p];
            if (deriv_soa == NULL) break;
            if (fabs(par) > 0) deriv_soa[p] += reg_l1 * (par < 0 ? -1.0 : 1.0);
            else if (deriv_soa[p] < -reg_l1) deriv_soa[p] += reg_l1;
            else if (deriv_soa[p] > reg_l1) deriv_soa[p] -= reg_l1;
            else deriv_soa[p] = 0.0;
        }
    }
    if (reg_l2 > 0) {
        for (long long j = 0; j < n_cp; ++j) {
            double nrm = 0.0;
            for (int d = 0; d < dim; ++d) nrm += params_soa[j + d * n_cp] * params_soa[j + d * n_cp];
            *l2 += nrm;
            if (deriv_soa)
                for (int d = 0; d < dim; ++d) deriv_soa[j + d * n_cp] += reg_l2 * 2.0 * params_soa[j + d * n_cp];
        }
    }
    if (reg_l21 > 0) {
        for (long long j = 0; j < n_cp; ++j) {
            double nrm = 0.0;
            for (int d = 0; d < dim; ++d) nrm += params_soa[j + d * n_cp] * params_soa[j + d * n_cp];
            nrm = sqrt(nrm);
            *l21 += nrm;
            if (!deriv_soa) continue;
            for (int d = 0; d < dim; ++d) {
                long long p = j + d * n_cp;
                if (nrm > 0) deriv_soa[p] += reg_l21 * params_soa[p] / nrm;
                else if (deriv_soa[p] < -reg_l21) deriv_soa[p] += reg_l21;
                else if (deriv_soa[p] > reg_l21) deriv_soa[p] -= reg_l21;
                else deriv_soa[p] = 0.0;
            }
        }
    }
    if (use_negative_gradient && deriv_soa)
        for (long long p = 0; p < np; ++p) deriv_soa[p] = -deriv_soa[p];
    double reg = 0.0;
    if (reg_l1 > 0) reg += reg_l1 * *l1;
    if (reg_l21 > 0) reg += reg_l21 * *l21;
    if (reg_l2 > 0) reg += reg_l2 * *l2;
    return reg;
}

/* ------------------------------------------------------------------ */
/* a13: optimizers                                                      */
/* ------------------------------------------------------------------ */

/* EstimateLearningRate (itkStandardGradientDescentOptimizerv4.hxx:159-166). */
double skmo_standard_rate(double a, double A, double alpha, long long iteration)
{
    return pow(a / (A + (double)iteration + 1.0), alpha);
}

/* gamma(t) (itkAdaptiveStepGradientDescentOptimizerv4.hxx:86-97). */
double skmo_adaptive_rate(double a, double A, double alpha, double time)
{
    return a / pow(A + time + 1.0, alpha);
}

/* ModifyCurrentTime (itkAdaptiveStepGradientDescentOptimizerv4.hxx:118-128). */
double skmo_adaptive_time(double time, double scalar_product, double fmax, double fmin, double omega)
{
    double x = -scalar_product;
    double f = fmin + (fmax - fmin) / (1.0 - (fmax / fmin) * exp(-x / omega));
    return dmax(0.0, time + f);
}

/* ModifyGradientByLearningRateOverSubRange (Standard...hxx:133-143). */
void skmo_gd_update(long long n, const double *prev, const double *grad, double rate,
                    int orthant, double *out)
{
    for (long long j = 0; j < n; ++j) {
        double v = prev[j] + grad[j] * rate;
        if (orthant && v * prev[j] < 0) v = 0.0;
        out[j] = v;
    }
}

/* ------------------------------------------------------------------ */
/* a14: control-point grid (itkTransformAdapter.hxx:246-308)            */
/* ------------------------------------------------------------------ */

void skmo_control_point_grid(int dim, const skmo_grid *image, const double *sigma, skmo_grid *cp)
{
    memset(cp, 0, sizeof(*cp));
    for (int d = 0; d < 3; ++d) { cp->size[d] = 1; cp->spacing[d] = 1.0; }
    for (int d = 0; d < dim; ++d) {
        double lower = image->origin[d];
        double upper = image->origin[d] + image->spacing[d] * (image->size[d] - 1);
        cp->spacing[d] = sigma[d];
        cp->size[d] = (int)ceil(fabs((upper - lower) / sigma[d]));
        double cp_upper = image->origin[d] + sigma[d] * (cp->size[d] - 1);
        cp->origin[d] = image->origin[d] + fabs(upper - cp_upper) / 2.0;
    }
}

/* Threads of the OpenMP build (1 in the serial build); n > 0 sets them. */
int skmo_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}
