/*
 * cpu_fp32.c -- FP32 multi-core CPU restatement of the reference's
 * GetValueAndDerivative for the 3D NCC workload (stationary or anisotropic
 * Wendland kernel, RD regularizer, subsample 1), the CPU baseline timed
 * beside the GPU.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by bench.py (cpu_baseline) and tests/.
 *
 * It follows the reference's loops (per-thread work of the CUDA kernels run
 * as OpenMP iterations) in single precision, as the reference's kernels run
 * (GKernel<float>, include/cuda/Kernel.h:209):
 *   value pass, one iteration per voxel: the kernel sum over the voxel's
 *     control-point window (evaluateValueNCC, GpuEvaluatorNCC.cu:37-109),
 *     the cubic B-spline moving value and gradient (BsplineImage.h:229-296);
 *   host sums of the NCC terms (GpuEvaluator.cu:202-262, here FP64);
 *   derivative pass, one iteration per control point over the voxels of its
 *     3D window (evaluateDerivativeNCC, GpuEvaluatorNCC.cu:111-190);
 *   RD regularizer per control point over its control-point window
 *     (regularizerRD, GpuEvaluatorRegularizers.cu:96-166);
 *   the bilateral regularizer (no reference counterpart; the library's
 *   definition, RD's structure with a spatial and an intensity Gaussian).
 * Images: spacing 1, origin 0, x fastest.  Build: -O3 -march=native -fopenmp.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct skmf_problem {
    int size[3];                 /* fixed = moving grid, x y z                */
    const float *fixed;
    const float *coeffs;         /* (size + 3)^3 padded cubic B-spline coefficients */
    int cp_size[3];
    float cp_origin[3], cp_spacing[3];
    const float *cp_aos;         /* 3 per control point                        */
    const float *cpw;
    float sigma;                 /* sigma_C4 = sigma_C0                        */
    int tensor;                  /* 0 stationary, 1 weight tensor              */
    const float *cov;            /* 9 per voxel (tensor mode)                  */
    float max_eig_cov[9];
} skmf_problem;

static int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }

static float det3(const float *m)
{
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
}

/* largest eigenvalue of a symmetric 3x3 matrix (trigonometric form, Mat.h:259-281) */
static float max_eig3(const float *m)
{
    float tri = m[1] * m[1] + m[2] * m[2] + m[5] * m[5];
    if (tri == 0.0f) return fmaxf(m[0], fmaxf(m[4], m[8]));
    float q = (m[0] + m[4] + m[8]) / 3.0f;
    float p2 = (m[0] - q) * (m[0] - q) + (m[4] - q) * (m[4] - q) + (m[8] - q) * (m[8] - q) + 2.0f * tri;
    float p = sqrtf(p2 / 6.0f);
    float b[9];
    memcpy(b, m, sizeof(b));
    b[0] -= q; b[4] -= q; b[8] -= q;
    float r = det3(b) / (2.0f * p * p * p);
    r = fminf(1.0f, fmaxf(-1.0f, r));
    return q + 2.0f * p * cosf(acosf(r) / 3.0f);
}

static float wendland(float t)
{
    float f = fmaxf(0.0f, 1.0f - t);
    float f2 = f * f;
    return f2 * f2 * f2 * (3.0f + 18.0f * t + 35.0f * t * t) * (1.0f / 3.0f) * f2;
}

/* k(p, q) with the states of both points (Kernel.h:96-131) */
static float kernel(const skmf_problem *pr, const float *v, const float *cp, const float *cq)
{
    if (!pr->tensor) return wendland(sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / pr->sigma);
    float m[9], inv[9];
    for (int i = 0; i < 9; ++i) m[i] = 0.5f * (cp[i] + cq[i]);
    float dm = det3(m);
    inv[0] = (m[4] * m[8] - m[5] * m[7]) / dm; inv[1] = (m[2] * m[7] - m[1] * m[8]) / dm;
    inv[2] = (m[1] * m[5] - m[2] * m[4]) / dm; inv[3] = (m[5] * m[6] - m[3] * m[8]) / dm;
    inv[4] = (m[0] * m[8] - m[2] * m[6]) / dm; inv[5] = (m[2] * m[3] - m[0] * m[5]) / dm;
    inv[6] = (m[3] * m[7] - m[4] * m[6]) / dm; inv[7] = (m[1] * m[6] - m[0] * m[7]) / dm;
    inv[8] = (m[0] * m[4] - m[1] * m[3]) / dm;
    float quad = 0.0f;
    for (int i = 0; i < 3; ++i) {
        float vi = v[0] * inv[0 * 3 + i] + v[1] * inv[1 * 3 + i] + v[2] * inv[2 * 3 + i];
        quad += vi * v[i];
    }
    float F = powf(det3(cp), 0.25f) * powf(det3(cq), 0.25f) * powf(dm, -0.5f);
    return F * F * wendland(sqrtf(quad) / pr->sigma);
}

static const float ident[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};

static const float *cov_at(const skmf_problem *pr, const float *p)
{
    int i[3];
    for (int d = 0; d < 3; ++d) {
        i[d] = (int)(p[d] + 0.5f);
        if (i[d] < 0 || i[d] >= pr->size[d]) return ident;
    }
    return pr->cov + 9 * ((long long)i[0] + pr->size[0] * ((long long)i[1] + (long long)pr->size[1] * i[2]));
}

static float support_at(const skmf_problem *pr, const float *c)
{
    if (!pr->tensor) return pr->sigma;
    float s[9];
    for (int i = 0; i < 9; ++i) s[i] = pr->max_eig_cov[i] + c[i];
    return pr->sigma / sqrtf(2.0f / max_eig3(s));
}

/* value and gradient of the cubic B-spline at index position p (BsplineImage.h:229-296) */
static int bspline(const skmf_problem *pr, const float *p, float *val, float *g)
{
    int base[3];
    float w[3][4], dw[3][4];
    for (int d = 0; d < 3; ++d) {
        if (p[d] < -0.5f || p[d] >= pr->size[d] - 0.5f) return 0;
        base[d] = (int)p[d];
        float t = p[d] - base[d];
        w[d][3] = t * t * t / 6.0f;
        w[d][0] = 1.0f / 6.0f + 0.5f * t * (t - 1.0f) - w[d][3];
        w[d][2] = t + w[d][0] - 2.0f * w[d][3];
        w[d][1] = 1.0f - w[d][0] - w[d][2] - w[d][3];
        dw[d][3] = 0.5f * t * t;
        dw[d][0] = t - 0.5f - dw[d][3];
        dw[d][2] = 1.0f + dw[d][0] - 2.0f * dw[d][3];
        dw[d][1] = -dw[d][0] - dw[d][2] - dw[d][3];
    }
    const long long sx = pr->size[0] + 3, sy = pr->size[1] + 3;
    float v = 0, gx = 0, gy = 0, gz = 0;
    for (int z = 0; z < 4; ++z)
        for (int y = 0; y < 4; ++y) {
            const float *row = pr->coeffs + base[0] + (base[1] + y) * sx + (base[2] + z) * sx * sy;
            float rv = 0, rd = 0;
            for (int x = 0; x < 4; ++x) {
                rv += w[0][x] * row[x];
                rd += dw[0][x] * row[x];
            }
            v += w[1][y] * w[2][z] * rv;
            gx += w[1][y] * w[2][z] * rd;
            gy += dw[1][y] * w[2][z] * rv;
            gz += w[1][y] * dw[2][z] * rv;
        }
    *val = v;
    g[0] = gx; g[1] = gy; g[2] = gz;
    return 1;
}

int skmf_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* NCC value (returned) and derivative (SoA, before the host regularizers);
 * rd_weight > 0 adds RD.  Returns 0 on success. */
int skmf_value_and_derivative_ncc(const skmf_problem *pr, double rd_weight, double *measure, double *rd_value,
                                  float *deriv_soa)
{
    const long long nv = (long long)pr->size[0] * pr->size[1] * pr->size[2];
    const long long ncp = (long long)pr->cp_size[0] * pr->cp_size[1] * pr->cp_size[2];
    float *f = malloc(sizeof(float) * nv), *m = malloc(sizeof(float) * nv), *g = malloc(sizeof(float) * nv * 3);
    unsigned char *in = malloc(nv);
    float *cpcov = malloc(sizeof(float) * 9 * ncp), *cpsup = malloc(sizeof(float) * ncp);
    if (!f || !m || !g || !in || !cpcov || !cpsup) return -1;
#pragma omp parallel for schedule(static)
    for (long long j = 0; j < ncp; ++j) {
        int jc[3] = {(int)(j % pr->cp_size[0]), (int)((j / pr->cp_size[0]) % pr->cp_size[1]),
                     (int)(j / ((long long)pr->cp_size[0] * pr->cp_size[1]))};
        float p[3];
        for (int d = 0; d < 3; ++d) p[d] = pr->cp_origin[d] + pr->cp_spacing[d] * jc[d];
        const float *c = pr->tensor ? cov_at(pr, p) : ident;
        memcpy(cpcov + 9 * j, c, sizeof(float) * 9);
        cpsup[j] = support_at(pr, c);
    }
    /* value pass */
#pragma omp parallel for schedule(dynamic, 64)
    for (long long v = 0; v < nv; ++v) {
        int x[3] = {(int)(v % pr->size[0]), (int)((v / pr->size[0]) % pr->size[1]),
                    (int)(v / ((long long)pr->size[0] * pr->size[1]))};
        float px[3] = {(float)x[0], (float)x[1], (float)x[2]};
        const float *cx = pr->tensor ? cov_at(pr, px) : ident;
        float sup = support_at(pr, cx);
        int lo[3], hi[3];
        for (int d = 0; d < 3; ++d) {
            float idx = (px[d] - pr->cp_origin[d]) / pr->cp_spacing[d];
            lo[d] = clampi((int)(idx - sup / pr->cp_spacing[d] + 1.0f), 0, pr->cp_size[d] - 1);
            hi[d] = clampi((int)(idx + sup / pr->cp_spacing[d]), 0, pr->cp_size[d] - 1);
        }
        float u[3] = {0, 0, 0};
        for (int k = lo[2]; k <= hi[2]; ++k)
            for (int jy = lo[1]; jy <= hi[1]; ++jy)
                for (int jx = lo[0]; jx <= hi[0]; ++jx) {
                    long long j = jx + (long long)pr->cp_size[0] * (jy + (long long)pr->cp_size[1] * k);
                    float vv[3] = {px[0] - (pr->cp_origin[0] + pr->cp_spacing[0] * jx),
                                   px[1] - (pr->cp_origin[1] + pr->cp_spacing[1] * jy),
                                   px[2] - (pr->cp_origin[2] + pr->cp_spacing[2] * k)};
                    float kk = kernel(pr, vv, cx, cpcov + 9 * j);
                    float wk = pr->cpw[j] * kk;
                    for (int d = 0; d < 3; ++d) u[d] += wk * pr->cp_aos[3 * j + d];
                }
        float q[3] = {px[0] + u[0], px[1] + u[1], px[2] + u[2]}, val, gr[3];
        in[v] = (unsigned char)bspline(pr, q, &val, gr);
        f[v] = in[v] ? pr->fixed[v] : 0.0f;
        m[v] = in[v] ? val : 0.0f;
        for (int d = 0; d < 3; ++d) g[v * 3 + d] = in[v] ? gr[d] : 0.0f;
    }
    /* host sums (GpuEvaluator.cu:229-261) */
    double sf = 0, sm = 0, sff = 0, smm = 0, sfm = 0;
    long long count = 0;
#pragma omp parallel for reduction(+ : sf, sm, sff, smm, sfm, count)
    for (long long v = 0; v < nv; ++v) {
        sf += f[v]; sm += m[v]; sff += (double)f[v] * f[v]; smm += (double)m[v] * m[v]; sfm += (double)f[v] * m[v];
        count += in[v];
    }
    sff -= sf * sf / count;
    smm -= sm * sm / count;
    sfm -= sf * sm / count;
    const double denom = -sqrt(sff * smm);
    *measure = denom != 0 ? sfm / denom : 0.0;
    const float mean_f = (float)(sf / count), mean_m = (float)(sm / count);
    const float rho = (float)(sfm / smm), inv_denom = denom != 0 ? (float)(1.0 / denom) : 0.0f;
    /* derivative pass, one control point per iteration (3D superset window) */
#pragma omp parallel for schedule(dynamic, 4)
    for (long long j = 0; j < ncp; ++j) {
        int jc[3] = {(int)(j % pr->cp_size[0]), (int)((j / pr->cp_size[0]) % pr->cp_size[1]),
                     (int)(j / ((long long)pr->cp_size[0] * pr->cp_size[1]))};
        float p[3], lo[3], hi[3];
        int ilo[3], ihi[3];
        for (int d = 0; d < 3; ++d) {
            p[d] = pr->cp_origin[d] + pr->cp_spacing[d] * jc[d];
            lo[d] = p[d] - cpsup[j];
            hi[d] = p[d] + cpsup[j] + 1.0f;
            ilo[d] = clampi((int)lo[d], 0, pr->size[d] - 1);
            ihi[d] = clampi((int)hi[d], 0, pr->size[d] - 1);
        }
        float dF[3] = {0, 0, 0}, dM[3] = {0, 0, 0};
        const float wj = pr->cpw[j];
        for (int z = ilo[2]; z <= ihi[2]; ++z)
            for (int y = ilo[1]; y <= ihi[1]; ++y)
                for (int x = ilo[0]; x <= ihi[0]; ++x) {
                    long long v = x + (long long)pr->size[0] * (y + (long long)pr->size[1] * z);
                    float vv[3] = {p[0] - x, p[1] - y, p[2] - z};
                    float px[3] = {(float)x, (float)y, (float)z};
                    float kk = kernel(pr, vv, cpcov + 9 * j, pr->tensor ? cov_at(pr, px) : ident);
                    if (kk == 0.0f) continue;
                    for (int d = 0; d < 3; ++d) {
                        float diff = wj * kk * g[v * 3 + d];
                        dF[d] += f[v] * diff - diff * mean_f;
                        dM[d] += m[v] * diff - diff * mean_m;
                    }
                }
        for (int d = 0; d < 3; ++d) deriv_soa[j + d * ncp] = (dF[d] - rho * dM[d]) * inv_denom;
    }
    /* RD: values and weighted derivatives (GpuEvaluatorRegularizers.cu:96-166) */
    double rd = 0.0;
    if (rd_weight > 0) {
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : rd)
        for (long long j = 0; j < ncp; ++j) {
            int jc[3] = {(int)(j % pr->cp_size[0]), (int)((j / pr->cp_size[0]) % pr->cp_size[1]),
                         (int)(j / ((long long)pr->cp_size[0] * pr->cp_size[1]))};
            int lo[3], hi[3];
            for (int d = 0; d < 3; ++d) {
                lo[d] = clampi((int)(jc[d] - cpsup[j] / pr->cp_spacing[d] + 1.0f), 0, pr->cp_size[d] - 1);
                hi[d] = clampi((int)(jc[d] + cpsup[j] / pr->cp_spacing[d]), 0, pr->cp_size[d] - 1);
            }
            float val = 0.0f, der[3] = {0, 0, 0};
            const float *cj = pr->cp_aos + 3 * j;
            for (int k = lo[2]; k <= hi[2]; ++k)
                for (int iy = lo[1]; iy <= hi[1]; ++iy)
                    for (int ix = lo[0]; ix <= hi[0]; ++ix) {
                        long long i = ix + (long long)pr->cp_size[0] * (iy + (long long)pr->cp_size[1] * k);
                        float vv[3] = {pr->cp_spacing[0] * (jc[0] - ix), pr->cp_spacing[1] * (jc[1] - iy),
                                       pr->cp_spacing[2] * (jc[2] - k)};
                        float kk = kernel(pr, vv, cpcov + 9 * j, cpcov + 9 * i) * pr->cpw[j];
                        const float *ci = pr->cp_aos + 3 * i;
                        float e[3] = {cj[0] - ci[0], cj[1] - ci[1], cj[2] - ci[2]};
                        val += kk * (e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
                        for (int d = 0; d < 3; ++d) der[d] += 4.0f * kk * e[d];
                    }
            rd += val;
            for (int d = 0; d < 3; ++d) deriv_soa[j + d * ncp] += (float)rd_weight * der[d];
        }
    }
    *rd_value = rd;
    free(f); free(m); free(g); free(in); free(cpcov); free(cpsup);
    return 0;
}

/* The bilateral regularizer: per control point j over the points i within
 * 3 sigma_s, B = exp(-r^2 / 2 sigma_s^2) exp(-(I_j - I_i)^2 / 2 sigma_i^2)
 * with I the fixed image at the control points (nearest voxel, 0 outside),
 * value_j = w_j sum B |c_j - c_i|^2 and d/dc_j = sum 2 B (w_j + w_i)(c_j - c_i);
 * weight x the derivative is added to deriv_soa, the sum of the values is
 * returned in *value.  Returns 0 on success. */
int skmf_bilateral(const skmf_problem *pr, double weight, double sigma_s, double sigma_i, double *value,
                   float *deriv_soa)
{
    const long long ncp = (long long)pr->cp_size[0] * pr->cp_size[1] * pr->cp_size[2];
    float *inten = malloc(sizeof(float) * ncp);
    if (!inten) return -1;
#pragma omp parallel for schedule(static)
    for (long long j = 0; j < ncp; ++j) {
        int jc[3] = {(int)(j % pr->cp_size[0]), (int)((j / pr->cp_size[0]) % pr->cp_size[1]),
                     (int)(j / ((long long)pr->cp_size[0] * pr->cp_size[1]))};
        int v[3], inside = 1;
        for (int d = 0; d < 3; ++d) {
            v[d] = (int)(pr->cp_origin[d] + pr->cp_spacing[d] * jc[d] + 0.5f);
            if (v[d] < 0 || v[d] >= pr->size[d]) inside = 0;
        }
        inten[j] = inside ? pr->fixed[v[0] + (long long)pr->size[0] * (v[1] + (long long)pr->size[1] * v[2])] : 0.0f;
    }
    const float cut = 3.0f * (float)sigma_s, inv2ss = 1.0f / (2.0f * (float)(sigma_s * sigma_s)),
                inv2si = 1.0f / (2.0f * (float)(sigma_i * sigma_i));
    double total = 0.0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total)
    for (long long j = 0; j < ncp; ++j) {
        int jc[3] = {(int)(j % pr->cp_size[0]), (int)((j / pr->cp_size[0]) % pr->cp_size[1]),
                     (int)(j / ((long long)pr->cp_size[0] * pr->cp_size[1]))};
        int lo[3], hi[3];
        for (int d = 0; d < 3; ++d) {
            lo[d] = clampi((int)(jc[d] - cut / pr->cp_spacing[d] + 1.0f), 0, pr->cp_size[d] - 1);
            hi[d] = clampi((int)(jc[d] + cut / pr->cp_spacing[d]), 0, pr->cp_size[d] - 1);
        }
        const float *cj = pr->cp_aos + 3 * j;
        const float wj = pr->cpw[j];
        float val = 0.0f, der[3] = {0, 0, 0};
        for (int k = lo[2]; k <= hi[2]; ++k)
            for (int iy = lo[1]; iy <= hi[1]; ++iy)
                for (int ix = lo[0]; ix <= hi[0]; ++ix) {
                    float v0 = pr->cp_spacing[0] * (jc[0] - ix), v1 = pr->cp_spacing[1] * (jc[1] - iy),
                          v2 = pr->cp_spacing[2] * (jc[2] - k);
                    float r2 = v0 * v0 + v1 * v1 + v2 * v2;
                    if (!(r2 < cut * cut)) continue;
                    long long i = ix + (long long)pr->cp_size[0] * (iy + (long long)pr->cp_size[1] * k);
                    float di = inten[j] - inten[i];
                    float b = expf(-r2 * inv2ss) * expf(-di * di * inv2si);
                    const float *ci = pr->cp_aos + 3 * i;
                    float e[3] = {cj[0] - ci[0], cj[1] - ci[1], cj[2] - ci[2]};
                    val += wj * b * (e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
                    float f2 = 2.0f * b * (wj + pr->cpw[i]);
                    for (int d = 0; d < 3; ++d) der[d] += f2 * e[d];
                }
        total += val;
        for (int d = 0; d < 3; ++d) deriv_soa[j + d * ncp] += (float)weight * der[d];
    }
    *value = total;
    free(inten);
    return 0;
}
