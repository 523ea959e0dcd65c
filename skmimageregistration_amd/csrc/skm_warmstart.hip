// skm_warmstart.hip -- learned warm start of the control-point coefficients:
// a small 3D convolutional encoder-decoder trained and run on gfx950 by the
// kernels below (no reference counterpart: the reference starts every
// registration from zero coefficients, src/SKMReg.cxx:375-430).
//
// Network (channels last, [batch][z][y][x][channel], FP32 activations):
//   input 2 ch (moving, fixed downsampled by 2)
//   e1 = relu(conv 2->16), relu(conv 16->16)                 full resolution
//   e2 = relu(conv/2 16->32), relu(conv 32->32)              1/2
//   e3 = relu(conv/2 32->48), relu(conv 48->48)              1/4
//   e4 = relu(conv/2 48->64), relu(conv 64->64)              1/8
//   d3 = relu(conv [up2(e4), e3] 112->48)                    1/4
//   d2 = relu(conv [up2(d3), e2] 80->32)                     1/2
//   out = head(trilinear sample of d2 at the control points) 32->3 (linear)
// Every conv is 3x3x3 with zero padding 1 (stride 1 or 2), an implicit GEMM
// on the MFMA units: BF16 weights and BF16 operands, FP32 accumulation
// (v_mfma_f32_16x16x32_bf16).  Backward: the data gradient is the same GEMM
// with the gathered output gradient and the transposed weights; the weight
// gradient is a GEMM over voxels split into chunks whose FP32 partials are
// summed in chunk order, so every gradient is deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "skm.h"
#include "skm_internal.hpp"

namespace skm {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kWsLayers = 10;      // 3x3x3 convolutions
constexpr int kHeadIn = 32, kHeadOut = 3;
constexpr int kWdwChunk = 2048;    // voxels per workgroup of the weight-gradient GEMM
constexpr int kBiasChunk = 256;    // voxels per workgroup of the bias gradient's first stage
constexpr int kMaxColTiles = 8;    // GEMM columns per wave: 8 x 16 = 128 channels (the widest is 112)

// (input channels, output channels, stride, level of the output) per conv
constexpr int kCin[kWsLayers] = {2, 16, 16, 32, 32, 48, 48, 64, 112, 80};
constexpr int kCout[kWsLayers] = {16, 16, 32, 32, 48, 48, 64, 64, 48, 32};
constexpr int kStride[kWsLayers] = {1, 1, 2, 1, 2, 1, 2, 1, 1, 1};
constexpr int kLevel[kWsLayers] = {0, 0, 1, 1, 2, 2, 3, 3, 2, 1};

__host__ __device__ inline int round_up(int v, int m) { return (v + m - 1) / m * m; }

// every GEMM's columns (cout forward, cin backward) fit one wave's tiles
constexpr bool columns_fit()
{
    for (int l = 0; l < kWsLayers; ++l)
        if (kCout[l] % 16 || kCout[l] > 16 * kMaxColTiles || (l > 0 && (kCin[l] % 16 || kCin[l] > 16 * kMaxColTiles)))
            return false;
    return true;
}
static_assert(columns_fit(), "conv channels must be multiples of 16 and at most 16 * kMaxColTiles");

// Geometry of one convolution as a GEMM over rows (voxels) and columns
// (channels): rows are the output voxels (forward) or the input voxels
// (data gradient); `src` is the tensor the A operand is gathered from.
struct ConvGemm {
    int batch;
    int rd, rh, rw;          // row voxel grid (per sample)
    int sd, sh, sw, sc;      // source grid and channels
    int ncols;               // output channels of the GEMM (multiple of 16)
    int kdim, kpad;          // 27 * sc and its multiple of 32
    int stride;
    int transposed;          // 0: forward gather i = o s + d - 1; 1: data gradient o = (i + 1 - d) / s
};

__device__ __forceinline__ __bf16 to_bf16(float v) { return (__bf16)v; }

// A-operand element: the source value for row voxel (b, z, y, x) and
// GEMM index k = tap * sc + c; 0 outside the source or beyond kdim.
__device__ __forceinline__ const float *conv_src_row(const ConvGemm &g, const float *src, int b, int z, int y, int x,
                                                     int tap)
{
    const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
    int sz, sy, sx;
    if (!g.transposed) {
        sz = z * g.stride + dz - 1;
        sy = y * g.stride + dy - 1;
        sx = x * g.stride + dx - 1;
    } else {
        const int tz = z + 1 - dz, ty = y + 1 - dy, tx = x + 1 - dx;
        if (tz < 0 || ty < 0 || tx < 0 || tz % g.stride || ty % g.stride || tx % g.stride) return nullptr;
        sz = tz / g.stride;
        sy = ty / g.stride;
        sx = tx / g.stride;
    }
    if (sz < 0 || sy < 0 || sx < 0 || sz >= g.sd || sy >= g.sh || sx >= g.sw) return nullptr;
    return src + ((((long long)b * g.sd + sz) * g.sh + sy) * g.sw + sx) * g.sc;
}

// dst[row][n] = act(sum_k A[row][k] W[n][k] + bias[n]); W is [ncols][kpad] BF16.
// One wave per 16 rows and all columns (ncols / 16 <= 8 MFMA tiles of 16x16x32);
// four waves per workgroup.  Lane l supplies A[row l%16][k 8(l/16)..+7] and
// W[column l%16][k 8(l/16)..+7]; the result of lane l is rows 4(l/16)..+3 of
// column l%16.
__global__ __launch_bounds__(256) void k_conv_gemm(ConvGemm g, const float *__restrict__ src,
                                                   const __bf16 *__restrict__ w, const float *__restrict__ bias,
                                                   float *__restrict__ dst, int relu)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long long rows = (long long)g.batch * g.rd * g.rh * g.rw;
    const long long row0 = ((long long)blockIdx.x * 4 + wave) * 16;
    if (row0 >= rows) return;
    const int r = lane & 15, kq = lane >> 4;
    const long long row = row0 + r;
    const bool rvalid = row < rows;
    int x = 0, y = 0, z = 0, b = 0;
    if (rvalid) {
        long long t = row;
        x = (int)(t % g.rw); t /= g.rw;
        y = (int)(t % g.rh); t /= g.rh;
        z = (int)(t % g.rd); t /= g.rd;
        b = (int)t;
    }
    const int ntiles = g.ncols / 16;
    f32x4 acc[kMaxColTiles];
#pragma unroll
    for (int i = 0; i < kMaxColTiles; ++i) acc[i] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const bool vec = (g.sc & 7) == 0;
    for (int k0 = 0; k0 < g.kpad; k0 += 32) {
        const int kb = k0 + kq * 8;
        bf16x8 a;
        if (vec) {
            // eight consecutive channels of one tap
            const float *p = (rvalid && kb < g.kdim) ? conv_src_row(g, src, b, z, y, x, kb / g.sc) : nullptr;
            if (p) {
                const float4 v0 = *reinterpret_cast<const float4 *>(p + kb % g.sc);
                const float4 v1 = *reinterpret_cast<const float4 *>(p + kb % g.sc + 4);
                a = bf16x8{to_bf16(v0.x), to_bf16(v0.y), to_bf16(v0.z), to_bf16(v0.w),
                           to_bf16(v1.x), to_bf16(v1.y), to_bf16(v1.z), to_bf16(v1.w)};
            } else {
                a = bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
            }
        } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int k = kb + e;
                float v = 0.0f;
                if (rvalid && k < g.kdim) {
                    const float *p = conv_src_row(g, src, b, z, y, x, k / g.sc);
                    if (p) v = p[k % g.sc];
                }
                a[e] = to_bf16(v);
            }
        }
#pragma unroll
        for (int nt = 0; nt < kMaxColTiles; ++nt) {
            if (nt >= ntiles) break;
            const bf16x8 bw = *reinterpret_cast<const bf16x8 *>(w + (long long)(nt * 16 + r) * g.kpad + kb);
            acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw, acc[nt], 0, 0, 0);
        }
    }
#pragma unroll
    for (int nt = 0; nt < kMaxColTiles; ++nt) {
        if (nt >= ntiles) break;
        const int n = nt * 16 + r;
        const float bn = bias ? bias[n] : 0.0f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long long orow = row0 + kq * 4 + i;
            if (orow < rows) {
                float v = acc[nt][i] + bn;
                if (relu) v = fmaxf(v, 0.0f);
                dst[orow * g.ncols + n] = v;
            }
        }
    }
}

// Weight gradient partials: part[chunk][co][kk] = sum over the chunk's output
// voxels o of dz[o][co] * x[i(o, tap)][ci], kk = tap * cin + ci (0 beyond
// 27 cin).  One wave per 16 x 16 tile of (co, kk), BF16 operands, FP32
// accumulation; workgroup = 4 tiles of one voxel chunk.
struct WgradGeom {
    int batch;
    int od, oh, ow, cout;    // output grid (the rows of dz)
    int id, ih, iw, cin;     // input grid
    int stride;
    int kdim, kpad;
    int chunks;
};

__global__ __launch_bounds__(256) void k_conv_wgrad(WgradGeom g, const float *__restrict__ x,
                                                    const float *__restrict__ dz, float *__restrict__ part)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int co_tiles = g.cout / 16, kk_tiles = g.kpad / 16;
    const int tile = blockIdx.x * 4 + wave;
    if (tile >= co_tiles * kk_tiles) return;
    const int chunk = blockIdx.y;
    const int co0 = (tile % co_tiles) * 16, kk0 = (tile / co_tiles) * 16;
    const int r = lane & 15, kq = lane >> 4;
    const long long nvox = (long long)g.batch * g.od * g.oh * g.ow;
    const long long v0 = (long long)chunk * kWdwChunk, v1 = std::min<long long>(nvox, v0 + kWdwChunk);
    const int kk = kk0 + r;  // this lane's B column
    const int tap = kk / g.cin, ci = kk % g.cin;
    const bool kvalid = kk < g.kdim;
    const int dzt = tap / 9, dyt = (tap / 3) % 3, dxt = tap % 3;
    f32x4 acc = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    // coordinates of this lane's first voxel of a step, decoded once and then
    // advanced: by one voxel per element, by 32 per step
    int ox, oy, oz, ob;
    {
        long long t = v0 + kq * 8;
        ox = (int)(t % g.ow); t /= g.ow;
        oy = (int)(t % g.oh); t /= g.oh;
        oz = (int)(t % g.od); t /= g.od;
        ob = (int)t;
    }
    for (long long vb = v0; vb < v1; vb += 32) {
        bf16x8 a, bw;
        int cx = ox, cy = oy, cz = oz, cb = ob;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const long long o = vb + kq * 8 + e;
            float av = 0.0f, bv = 0.0f;
            if (o < v1) {
                av = dz[o * g.cout + co0 + r];
                if (kvalid) {
                    const int sz = cz * g.stride + dzt - 1, sy = cy * g.stride + dyt - 1, sx = cx * g.stride + dxt - 1;
                    if (sz >= 0 && sy >= 0 && sx >= 0 && sz < g.id && sy < g.ih && sx < g.iw)
                        bv = x[((((long long)cb * g.id + sz) * g.ih + sy) * g.iw + sx) * g.cin + ci];
                }
            }
            a[e] = to_bf16(av);
            bw[e] = to_bf16(bv);
            if (++cx == g.ow) {
                cx = 0;
                if (++cy == g.oh) {
                    cy = 0;
                    if (++cz == g.od) {
                        cz = 0;
                        ++cb;
                    }
                }
            }
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw, acc, 0, 0, 0);
        ox += 32;
        while (ox >= g.ow) {
            ox -= g.ow;
            if (++oy == g.oh) {
                oy = 0;
                if (++oz == g.od) {
                    oz = 0;
                    ++ob;
                }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
        part[((long long)chunk * g.cout + co0 + kq * 4 + i) * g.kpad + kk0 + r] = acc[i];
}

// dw[co][k] (k < kdim, unpadded layout [cout][27 cin]) = sum of the chunk
// partials in chunk order.
__global__ __launch_bounds__(256) void k_wgrad_reduce(const float *__restrict__ part, int chunks, int cout, int kdim,
                                                      int kpad, float *__restrict__ dw)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= cout * kdim) return;
    const int co = i / kdim, k = i % kdim;
    float s = 0.0f;
    for (int c = 0; c < chunks; ++c) s += part[((long long)c * cout + co) * kpad + k];
    dw[i] = s;
}

// Bias gradient, first stage: per chunk of kBiasChunk voxels the FP32 sum of
// dz[v][c] of every channel (lanes over 64 channels at a time, reading a
// voxel's channels together; the four waves take every fourth voxel and are
// added in wave order).
__global__ __launch_bounds__(256) void k_bias_part(const float *__restrict__ dz, long long nvox, int c_count,
                                                   float *__restrict__ part)
{
    __shared__ float lds[4][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c = (int)blockIdx.y * 64 + lane;
    const long long v0 = (long long)blockIdx.x * kBiasChunk, v1 = std::min<long long>(nvox, v0 + kBiasChunk);
    float s = 0.0f;
    if (c < c_count)
        for (long long v = v0 + wave; v < v1; v += 4) s += dz[v * c_count + c];
    lds[wave][lane] = s;
    __syncthreads();
    if (wave == 0 && c < c_count)
        part[(long long)blockIdx.x * c_count + c] = ((lds[0][lane] + lds[1][lane]) + lds[2][lane]) + lds[3][lane];
}

// Bias gradient, second stage (and any per-channel column sum): db[c] =
// sum_v dz[v][c], one workgroup per channel, each thread every 256th row in
// order, FP64 wave sums in a fixed butterfly, the four waves in order.
__global__ __launch_bounds__(256) void k_bias_grad(const float *__restrict__ dz, long long nvox, int c_count,
                                                   float *__restrict__ db)
{
    __shared__ double lds[4];
    const int c = blockIdx.x;
    double s = 0.0;
    for (long long v = threadIdx.x; v < nvox; v += 256) s += (double)dz[v * c_count + c];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) db[c] = (float)(lds[0] + lds[1] + lds[2] + lds[3]);
}

// dz = dy where y > 0 (ReLU backward, in place on dy)
__global__ __launch_bounds__(256) void k_relu_grad(const float *__restrict__ y, float *__restrict__ dy, long long n)
{
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i < n && !(y[i] > 0.0f)) dy[i] = 0.0f;
}

// out[b][z][y][x] = [coarse[b][z/2][y/2][x/2] (ca), skip[b][z][y][x] (cb)]
__global__ __launch_bounds__(256) void k_upcat(const float *__restrict__ coarse, const float *__restrict__ skip,
                                               int batch, int d, int h, int w, int ca, int cb, float *__restrict__ out)
{
    const int ct = ca + cb;
    const long long n = (long long)batch * d * h * w * ct;
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i >= n) return;
    const int c = (int)(i % ct);
    long long v = i / ct;
    if (c >= ca) {
        out[i] = skip[v * cb + (c - ca)];
        return;
    }
    const int x = (int)(v % w); v /= w;
    const int y = (int)(v % h); v /= h;
    const int z = (int)(v % d);
    const int b = (int)(v / d);
    const long long cv = (((long long)b * (d / 2) + z / 2) * (h / 2) + y / 2) * (w / 2) + x / 2;
    out[i] = coarse[cv * ca + c];
}

// Backward of k_upcat: dcoarse = sum of the 8 children (overwritten); dskip += skip part.
__global__ __launch_bounds__(256) void k_upcat_grad(const float *__restrict__ dout, int batch, int d, int h, int w,
                                                    int ca, int cb, float *__restrict__ dcoarse,
                                                    float *__restrict__ dskip)
{
    const int ct = ca + cb;
    const int cd = d / 2, ch = h / 2, cw = w / 2;
    const long long ncoarse = (long long)batch * cd * ch * cw * ca;
    const long long nskip = (long long)batch * d * h * w * cb;
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i < ncoarse) {
        const int c = (int)(i % ca);
        long long v = i / ca;
        const int x = (int)(v % cw); v /= cw;
        const int y = (int)(v % ch); v /= ch;
        const int z = (int)(v % cd);
        const int b = (int)(v / cd);
        float s = 0.0f;
        for (int k = 0; k < 8; ++k) {
            const int fz = 2 * z + (k >> 2), fy = 2 * y + ((k >> 1) & 1), fx = 2 * x + (k & 1);
            s += dout[((((long long)b * d + fz) * h + fy) * w + fx) * ct + c];
        }
        dcoarse[i] = s;
    }
    if (i < nskip) {
        const int c = (int)(i % cb);
        const long long v = i / cb;
        dskip[i] += dout[v * ct + ca + c];
    }
}

// Trilinear sampling of feat ([b][fd][fh][fw][C]) at the control points:
// coordinate per axis q_j = clamp(origin + spacing j, 0, n - 1); the weight of
// feature voxel f is max(0, 1 - |q_j - f|) per axis.
struct SampleGeom {
    int batch, fd, fh, fw, c;
    int nx, ny, nz;               // control points
    float ox, oy, oz, hx, hy, hz; // control-point coordinates in feature voxels
};

__device__ __forceinline__ float sample_coord(float o, float h, int j, int n)
{
    return fminf(fmaxf(o + h * (float)j, 0.0f), (float)(n - 1));
}

__global__ __launch_bounds__(256) void k_sample(SampleGeom g, const float *__restrict__ feat, float *__restrict__ out)
{
    const long long n = (long long)g.batch * g.nx * g.ny * g.nz * g.c;
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i >= n) return;
    const int c = (int)(i % g.c);
    long long j = i / g.c;
    const int jx = (int)(j % g.nx); j /= g.nx;
    const int jy = (int)(j % g.ny); j /= g.ny;
    const int jz = (int)(j % g.nz);
    const int b = (int)(j / g.nz);
    const float qx = sample_coord(g.ox, g.hx, jx, g.fw), qy = sample_coord(g.oy, g.hy, jy, g.fh),
                qz = sample_coord(g.oz, g.hz, jz, g.fd);
    const int x0 = (int)floorf(qx), y0 = (int)floorf(qy), z0 = (int)floorf(qz);
    float s = 0.0f;
    for (int k = 0; k < 8; ++k) {
        const int fz = z0 + (k >> 2), fy = y0 + ((k >> 1) & 1), fx = x0 + (k & 1);
        if (fz >= g.fd || fy >= g.fh || fx >= g.fw) continue;
        const float wgt = fmaxf(0.0f, 1.0f - fabsf(qx - fx)) * fmaxf(0.0f, 1.0f - fabsf(qy - fy)) *
                          fmaxf(0.0f, 1.0f - fabsf(qz - fz));
        s += wgt * feat[((((long long)b * g.fd + fz) * g.fh + fy) * g.fw + fx) * g.c + c];
    }
    out[i] = s;
}

// Backward of k_sample as a gather per feature voxel: the control points
// whose coordinate lies within 1 of it along every axis (at most two per
// axis, the coordinates being at least one voxel apart), in a fixed order.
__device__ __forceinline__ void sample_range(float o, float h, int n_cp, int n_feat, int f, int &lo, int &hi)
{
    // q_j within (f - 1, f + 1); q is non-decreasing in j
    lo = n_cp;
    hi = -1;
    for (int j = 0; j < n_cp; ++j) {
        const float q = sample_coord(o, h, j, n_feat);
        if (q > (float)f - 1.0f && q < (float)f + 1.0f) {
            lo = min(lo, j);
            hi = max(hi, j);
        }
    }
}

__global__ __launch_bounds__(256) void k_sample_grad(SampleGeom g, const float *__restrict__ dout,
                                                     float *__restrict__ dfeat)
{
    const long long n = (long long)g.batch * g.fd * g.fh * g.fw * g.c;
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i >= n) return;
    const int c = (int)(i % g.c);
    long long v = i / g.c;
    const int fx = (int)(v % g.fw); v /= g.fw;
    const int fy = (int)(v % g.fh); v /= g.fh;
    const int fz = (int)(v % g.fd);
    const int b = (int)(v / g.fd);
    int xl, xh, yl, yh, zl, zh;
    sample_range(g.ox, g.hx, g.nx, g.fw, fx, xl, xh);
    sample_range(g.oy, g.hy, g.ny, g.fh, fy, yl, yh);
    sample_range(g.oz, g.hz, g.nz, g.fd, fz, zl, zh);
    float s = 0.0f;
    for (int jz = zl; jz <= zh; ++jz) {
        const float wz = fmaxf(0.0f, 1.0f - fabsf(sample_coord(g.oz, g.hz, jz, g.fd) - fz));
        for (int jy = yl; jy <= yh; ++jy) {
            const float wy = fmaxf(0.0f, 1.0f - fabsf(sample_coord(g.oy, g.hy, jy, g.fh) - fy));
            for (int jx = xl; jx <= xh; ++jx) {
                const float wx = fmaxf(0.0f, 1.0f - fabsf(sample_coord(g.ox, g.hx, jx, g.fw) - fx));
                const long long j = (((long long)b * g.nz + jz) * g.ny + jy) * g.nx + jx;
                s += wx * wy * wz * dout[j * g.c + c];
            }
        }
    }
    dfeat[i] = s;
}

// Head: out[j][k] = sum_c s[j][c] wh[k][c] + bh[k] (BF16 weights, FP32 sums).
__global__ __launch_bounds__(256) void k_head(const float *__restrict__ s, const __bf16 *__restrict__ wh,
                                              const float *__restrict__ bh, long long npts, float *__restrict__ out)
{
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i >= npts * kHeadOut) return;
    const long long j = i / kHeadOut;
    const int k = (int)(i % kHeadOut);
    float acc = bh[k];
    for (int c = 0; c < kHeadIn; ++c) acc = fmaf(s[j * kHeadIn + c], (float)wh[k * kHeadIn + c], acc);
    out[i] = acc;
}

// ds[j][c] = sum_k dout[j][k] wh[k][c]
__global__ __launch_bounds__(256) void k_head_grad_in(const float *__restrict__ dout, const __bf16 *__restrict__ wh,
                                                      long long npts, float *__restrict__ ds)
{
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i >= npts * kHeadIn) return;
    const long long j = i / kHeadIn;
    const int c = (int)(i % kHeadIn);
    float acc = 0.0f;
    for (int k = 0; k < kHeadOut; ++k) acc = fmaf(dout[j * kHeadOut + k], (float)wh[k * kHeadIn + c], acc);
    ds[i] = acc;
}

// dwh[k][c] = sum_j dout[j][k] s[j][c], dbh[k] = sum_j dout[j][k]: one
// workgroup per weight (and per bias), FP64 sums in a fixed order.
__global__ __launch_bounds__(256) void k_head_grad_w(const float *__restrict__ dout, const float *__restrict__ s,
                                                     long long npts, float *__restrict__ dwh, float *__restrict__ dbh)
{
    __shared__ double lds[4];
    const int id = blockIdx.x;  // < kHeadOut * (kHeadIn + 1)
    const int k = id / (kHeadIn + 1), c = id % (kHeadIn + 1);
    double acc = 0.0;
    for (long long j = threadIdx.x; j < npts; j += 256)
        acc += (double)dout[j * kHeadOut + k] * (c < kHeadIn ? (double)s[j * kHeadIn + c] : 1.0);
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double t = lds[0] + lds[1] + lds[2] + lds[3];
        if (c < kHeadIn) dwh[k * kHeadIn + c] = (float)t;
        else dbh[k] = (float)t;
    }
}

// Mean squared error and its gradient: dout = 2 (out - target) / n; block
// partials of the squared error (FP64).
__global__ __launch_bounds__(256) void k_mse_loss(const float *__restrict__ out, const float *__restrict__ target,
                                                  long long n, float *__restrict__ dout, double *__restrict__ part)
{
    __shared__ double lds[4];
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    double e = 0.0;
    if (i < n) {
        const float d = out[i] - target[i];
        dout[i] = 2.0f * d / (float)n;
        e = (double)d * d;
    }
    for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = lds[0] + lds[1] + lds[2] + lds[3];
}

// Adam (FP32 master weights), bias-corrected.
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                                              float *__restrict__ v, long long n, float lr, float b1, float b2,
                                              float eps, float c1, float c2)
{
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// BF16 copies of one conv's weights: w [cout][kpad] and the transpose wt
// [cin][kpad_t] with wt[ci][tap cout + co] = w[co][tap cin + ci].
__global__ __launch_bounds__(256) void k_cast_weights(const float *__restrict__ master, int cin, int cout, int kpad,
                                                      int kpad_t, __bf16 *__restrict__ w, __bf16 *__restrict__ wt)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int kdim = 27 * cin, kdim_t = 27 * cout;
    if (i < cout * kpad) {
        const int co = i / kpad, k = i % kpad;
        w[i] = to_bf16(k < kdim ? master[co * kdim + k] : 0.0f);
    }
    if (wt && i < cin * kpad_t) {
        const int ci = i / kpad_t, k = i % kpad_t;
        float v = 0.0f;
        if (k < kdim_t) {
            const int tap = k / cout, co = k % cout;
            v = master[co * kdim + tap * cin + ci];
        }
        wt[i] = to_bf16(v);
    }
}

__global__ __launch_bounds__(256) void k_cast(const float *__restrict__ src, int n, __bf16 *__restrict__ dst)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = to_bf16(src[i]);
}

__global__ __launch_bounds__(256) void k_add(float *__restrict__ dst, const float *__restrict__ src, long long n)
{
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i < n) dst[i] += src[i];
}

unsigned blocks_of(long long n) { return (unsigned)((n + 255) / 256); }

// One sample of the network input: the 2 x 2 x 2 block means of the moving
// and the fixed image (each 2 in_size) as channels 0 and 1.
__global__ __launch_bounds__(256) void k_prepare_input(const float *__restrict__ moving, const float *__restrict__ fixed,
                                                       int d, int h, int w, float *__restrict__ out)
{
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i >= (long long)d * h * w) return;
    const int x = (int)(i % w), y = (int)((i / w) % h), z = (int)(i / ((long long)w * h));
    const int W = 2 * w, H = 2 * h;
    float sm = 0.0f, sf = 0.0f;
    for (int k = 0; k < 8; ++k) {
        const long long v = ((long long)(2 * z + (k >> 2)) * H + (2 * y + ((k >> 1) & 1))) * W + 2 * x + (k & 1);
        sm += moving[v];
        sf += fixed[v];
    }
    out[2 * i] = 0.125f * sm;
    out[2 * i + 1] = 0.125f * sf;
}

}  // namespace
}  // namespace skm

using namespace skm;

struct skm_ws_net {
    skm_ws_desc desc{};
    int maxb = 1;
    int dims[4][3] = {};           // per level: z, y, x
    long long n_params = 0;
    long long w_off[kWsLayers] = {}, b_off[kWsLayers] = {}, wh_off = 0, bh_off = 0;
    int kpad[kWsLayers] = {}, kpad_t[kWsLayers] = {};
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // master, gradient, Adam moments
    __bf16 *w[kWsLayers] = {}, *wt[kWsLayers] = {}, *wh = nullptr;
    float *act[kWsLayers] = {};    // outputs of the convs (after ReLU)
    float *cat3 = nullptr, *cat2 = nullptr, *samp = nullptr, *out = nullptr;
    const float *input = nullptr;  // the last forward's input (borrowed)
    int last_batch = 0;
    float *grad_a = nullptr, *grad_b = nullptr, *grad_skip3 = nullptr, *grad_skip2 = nullptr, *dsamp = nullptr,
          *dout = nullptr;
    float *wpart = nullptr, *bpart = nullptr;
    long long wpart_n = 0, bpart_n = 0;
    double *loss_part = nullptr;
    long long adam_t = 0;
    hipStream_t stream = nullptr;

    ~skm_ws_net()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        void *ptrs[] = {p, g, m, v, wh, cat3, cat2, samp, out, grad_a, grad_b, grad_skip3, grad_skip2, dsamp, dout,
                        wpart, bpart, loss_part};
        for (void *q : ptrs)
            if (q) device_release(q);
        for (int l = 0; l < kWsLayers; ++l) {
            if (w[l]) device_release(w[l]);
            if (wt[l]) device_release(wt[l]);
            if (act[l]) device_release(act[l]);
        }
    }

    long long level_vox(int lev) const { return (long long)dims[lev][0] * dims[lev][1] * dims[lev][2]; }
    long long n_cp() const { return (long long)desc.cp_size[0] * desc.cp_size[1] * desc.cp_size[2]; }
};

namespace {

skm_status ws_fail(skm_status st, const std::string &msg) { return set_last_error(st, msg); }

#define WS_HIP(expr)                                                                                              \
    do {                                                                                                           \
        hipError_t e_ = (expr);                                                                                    \
        if (e_ != hipSuccess)                                                                                      \
            return ws_fail(e_ == hipErrorOutOfMemory ? SKM_ERROR_OUT_OF_MEMORY : SKM_ERROR_HIP,                    \
                           std::string(#expr) + ": " + hipGetErrorString(e_));                                     \
    } while (0)

template <typename T>
hipError_t ws_alloc(T **p, long long count)
{
    return device_alloc((void **)p, (size_t)std::max<long long>(count, 1) * sizeof(T));
}

// conv layer l: source tensor, its level and channels
int src_level(int l) { return l == 0 ? 0 : kLevel[l - 1]; }

void conv_forward(skm_ws_net *n, int l, const float *src, int batch)
{
    const int sl = (l == 8) ? 2 : (l == 9) ? 1 : src_level(l);
    ConvGemm g{};
    g.batch = batch;
    const int *od = n->dims[kLevel[l]], *sd = n->dims[sl];
    g.rd = od[0]; g.rh = od[1]; g.rw = od[2];
    g.sd = sd[0]; g.sh = sd[1]; g.sw = sd[2]; g.sc = kCin[l];
    g.ncols = kCout[l];
    g.kdim = 27 * kCin[l];
    g.kpad = n->kpad[l];
    g.stride = kStride[l];
    g.transposed = 0;
    const long long rows = (long long)batch * od[0] * od[1] * od[2];
    const unsigned grid = (unsigned)((rows + 63) / 64);
    k_conv_gemm<<<grid, 256, 0, n->stream>>>(g, src, n->w[l], n->p + n->b_off[l], n->act[l], 1);
}

// data gradient of conv l: dsrc (source grid, cin channels) from dz (output grid)
void conv_backward_data(skm_ws_net *n, int l, const float *dz, float *dsrc, int batch)
{
    const int sl = (l == 8) ? 2 : (l == 9) ? 1 : src_level(l);
    ConvGemm g{};
    g.batch = batch;
    const int *od = n->dims[kLevel[l]], *sd = n->dims[sl];
    g.rd = sd[0]; g.rh = sd[1]; g.rw = sd[2];      // rows: source voxels
    g.sd = od[0]; g.sh = od[1]; g.sw = od[2]; g.sc = kCout[l];
    g.ncols = kCin[l];
    g.kdim = 27 * kCout[l];
    g.kpad = n->kpad_t[l];
    g.stride = kStride[l];
    g.transposed = 1;
    const long long rows = (long long)batch * sd[0] * sd[1] * sd[2];
    k_conv_gemm<<<(unsigned)((rows + 63) / 64), 256, 0, n->stream>>>(g, dz, n->wt[l], nullptr, dsrc, 0);
}

void conv_backward_weights(skm_ws_net *n, int l, const float *src, const float *dz, int batch)
{
    const int sl = (l == 8) ? 2 : (l == 9) ? 1 : src_level(l);
    WgradGeom g{};
    g.batch = batch;
    const int *od = n->dims[kLevel[l]], *sd = n->dims[sl];
    g.od = od[0]; g.oh = od[1]; g.ow = od[2]; g.cout = kCout[l];
    g.id = sd[0]; g.ih = sd[1]; g.iw = sd[2]; g.cin = kCin[l];
    g.stride = kStride[l];
    g.kdim = 27 * kCin[l];
    g.kpad = n->kpad[l];
    const long long nvox = (long long)batch * od[0] * od[1] * od[2];
    g.chunks = (int)((nvox + kWdwChunk - 1) / kWdwChunk);
    const int tiles = (g.cout / 16) * (g.kpad / 16);
    k_conv_wgrad<<<dim3((unsigned)((tiles + 3) / 4), (unsigned)g.chunks), 256, 0, n->stream>>>(g, src, dz, n->wpart);
    k_wgrad_reduce<<<blocks_of((long long)g.cout * g.kdim), 256, 0, n->stream>>>(n->wpart, g.chunks, g.cout, g.kdim,
                                                                                g.kpad, n->g + n->w_off[l]);
    const long long bchunks = (nvox + kBiasChunk - 1) / kBiasChunk;
    k_bias_part<<<dim3((unsigned)bchunks, (unsigned)((g.cout + 63) / 64)), 256, 0, n->stream>>>(dz, nvox, g.cout,
                                                                                          n->bpart);
    k_bias_grad<<<(unsigned)g.cout, 256, 0, n->stream>>>(n->bpart, bchunks, g.cout, n->g + n->b_off[l]);
}

SampleGeom sample_geom(const skm_ws_net *n, int batch)
{
    SampleGeom s{};
    s.batch = batch;
    s.fd = n->dims[1][0]; s.fh = n->dims[1][1]; s.fw = n->dims[1][2];
    s.c = kHeadIn;
    s.nx = n->desc.cp_size[0]; s.ny = n->desc.cp_size[1]; s.nz = n->desc.cp_size[2];
    // control points in level-1 feature voxels: input voxel i -> feature i / 2
    s.ox = (float)(n->desc.cp_origin[0] * 0.5); s.hx = (float)(n->desc.cp_spacing[0] * 0.5);
    s.oy = (float)(n->desc.cp_origin[1] * 0.5); s.hy = (float)(n->desc.cp_spacing[1] * 0.5);
    s.oz = (float)(n->desc.cp_origin[2] * 0.5); s.hz = (float)(n->desc.cp_spacing[2] * 0.5);
    return s;
}

void cast_all(skm_ws_net *n)
{
    for (int l = 0; l < kWsLayers; ++l) {
        const long long cnt = std::max<long long>((long long)kCout[l] * n->kpad[l], (long long)kCin[l] * n->kpad_t[l]);
        k_cast_weights<<<blocks_of(cnt), 256, 0, n->stream>>>(n->p + n->w_off[l], kCin[l], kCout[l], n->kpad[l],
                                                              n->kpad_t[l], n->w[l], l == 0 ? nullptr : n->wt[l]);
    }
    k_cast<<<1, 256, 0, n->stream>>>(n->p + n->wh_off, kHeadIn * kHeadOut, n->wh);
}

skm_status check_batch(const skm_ws_net *n, int batch)
{
    if (batch < 1 || batch > n->maxb) return ws_fail(SKM_ERROR_INVALID_ARGUMENT, "batch outside 1..max_batch");
    return SKM_SUCCESS;
}

skm_status forward_impl(skm_ws_net *n, const float *d_input, int batch, float *d_out)
{
    skm_status st = check_batch(n, batch);
    if (st != SKM_SUCCESS) return st;
    n->input = d_input;
    n->last_batch = batch;
    const float *src = d_input;
    for (int l = 0; l < 8; ++l) {
        conv_forward(n, l, src, batch);
        src = n->act[l];
    }
    // decoder: up2(e4) ++ e3 -> conv 8; up2(d3) ++ e2 -> conv 9
    {
        const int *d2 = n->dims[2];
        const long long cnt = (long long)batch * d2[0] * d2[1] * d2[2] * kCin[8];
        k_upcat<<<blocks_of(cnt), 256, 0, n->stream>>>(n->act[7], n->act[5], batch, d2[0], d2[1], d2[2], 64, 48,
                                                       n->cat3);
        conv_forward(n, 8, n->cat3, batch);
        const int *d1 = n->dims[1];
        const long long cnt2 = (long long)batch * d1[0] * d1[1] * d1[2] * kCin[9];
        k_upcat<<<blocks_of(cnt2), 256, 0, n->stream>>>(n->act[8], n->act[3], batch, d1[0], d1[1], d1[2], 48, 32,
                                                        n->cat2);
        conv_forward(n, 9, n->cat2, batch);
    }
    const SampleGeom sg = sample_geom(n, batch);
    const long long npts = (long long)batch * n->n_cp();
    k_sample<<<blocks_of(npts * kHeadIn), 256, 0, n->stream>>>(sg, n->act[9], n->samp);
    k_head<<<blocks_of(npts * kHeadOut), 256, 0, n->stream>>>(n->samp, n->wh, n->p + n->bh_off, npts,
                                                              d_out ? d_out : n->out);
    WS_HIP(hipGetLastError());
    return SKM_SUCCESS;
}

// Gradients of all parameters into n->g from dout (device, [batch][cp][3]) of the last forward.
skm_status backward_impl(skm_ws_net *n, const float *d_dout)
{
    const int batch = n->last_batch;
    if (batch < 1 || !n->input) return ws_fail(SKM_ERROR_INVALID_ARGUMENT, "backward before forward");
    const long long npts = (long long)batch * n->n_cp();
    const SampleGeom sg = sample_geom(n, batch);
    // head
    k_head_grad_w<<<kHeadOut * (kHeadIn + 1), 256, 0, n->stream>>>(d_dout, n->samp, npts, n->g + n->wh_off,
                                                                   n->g + n->bh_off);
    k_head_grad_in<<<blocks_of(npts * kHeadIn), 256, 0, n->stream>>>(d_dout, n->wh, npts, n->dsamp);
    // sampling -> d(act9)
    float *da = n->grad_a, *db = n->grad_b;
    const long long n9 = (long long)batch * n->level_vox(1) * kCout[9];
    k_sample_grad<<<blocks_of(n9), 256, 0, n->stream>>>(sg, n->dsamp, da);
    // conv 9 (input cat2 = [up2(act8), act3])
    k_relu_grad<<<blocks_of(n9), 256, 0, n->stream>>>(n->act[9], da, n9);
    conv_backward_weights(n, 9, n->cat2, da, batch);
    conv_backward_data(n, 9, da, db, batch);  // d(cat2)
    const long long nskip2 = (long long)batch * n->level_vox(1) * 32;
    WS_HIP(hipMemsetAsync(n->grad_skip2, 0, sizeof(float) * nskip2, n->stream));
    {
        const int *d1 = n->dims[1];
        const long long cnt = std::max<long long>((long long)batch * n->level_vox(2) * 48, nskip2);
        k_upcat_grad<<<blocks_of(cnt), 256, 0, n->stream>>>(db, batch, d1[0], d1[1], d1[2], 48, 32, da,
                                                            n->grad_skip2);
    }
    // conv 8 (input cat3 = [up2(act7), act5]); da holds d(act8)
    const long long n8 = (long long)batch * n->level_vox(2) * 48;
    k_relu_grad<<<blocks_of(n8), 256, 0, n->stream>>>(n->act[8], da, n8);
    conv_backward_weights(n, 8, n->cat3, da, batch);
    conv_backward_data(n, 8, da, db, batch);  // d(cat3)
    const long long nskip3 = (long long)batch * n->level_vox(2) * 48;
    WS_HIP(hipMemsetAsync(n->grad_skip3, 0, sizeof(float) * nskip3, n->stream));
    {
        const int *d2 = n->dims[2];
        const long long cnt = std::max<long long>((long long)batch * n->level_vox(3) * 64, nskip3);
        k_upcat_grad<<<blocks_of(cnt), 256, 0, n->stream>>>(db, batch, d2[0], d2[1], d2[2], 64, 48, da,
                                                            n->grad_skip3);
    }
    // encoder, from conv 7 (da holds d(act7)) down to conv 0
    for (int l = 7; l >= 0; --l) {
        const long long nl = (long long)batch * n->level_vox(kLevel[l]) * kCout[l];
        k_relu_grad<<<blocks_of(nl), 256, 0, n->stream>>>(n->act[l], da, nl);
        const float *src = l == 0 ? n->input : n->act[l - 1];
        conv_backward_weights(n, l, src, da, batch);
        if (l == 0) break;
        conv_backward_data(n, l, da, db, batch);
        std::swap(da, db);  // da: d(act[l-1])
        // the skip connections add their decoder gradients (e3 = act5, e2 = act3)
        const long long nprev = (long long)batch * n->level_vox(kLevel[l - 1]) * kCout[l - 1];
        if (l - 1 == 5) k_add<<<blocks_of(nprev), 256, 0, n->stream>>>(da, n->grad_skip3, nprev);
        if (l - 1 == 3) k_add<<<blocks_of(nprev), 256, 0, n->stream>>>(da, n->grad_skip2, nprev);
    }
    WS_HIP(hipGetLastError());
    return SKM_SUCCESS;
}

}  // namespace

extern "C" {

SKM_API skm_status skm_ws_create(const skm_ws_desc *desc, skm_ws_net **out)
{
    if (!desc || !out) return ws_fail(SKM_ERROR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    for (int a = 0; a < 3; ++a) {
        if (desc->in_size[a] < 8 || desc->in_size[a] % 8)
            return ws_fail(SKM_ERROR_INVALID_ARGUMENT, "network input sizes must be positive multiples of 8");
        if (desc->cp_size[a] < 1) return ws_fail(SKM_ERROR_INVALID_ARGUMENT, "control-point grid must not be empty");
        if (!(desc->cp_spacing[a] >= 2.0))
            return ws_fail(SKM_ERROR_INVALID_ARGUMENT,
                           "control-point spacing must be at least 2 input voxels (one feature voxel)");
    }
    if (desc->max_batch < 1) return ws_fail(SKM_ERROR_INVALID_ARGUMENT, "max_batch must be positive");
    auto *n = new skm_ws_net();
    n->desc = *desc;
    n->maxb = desc->max_batch;
    for (int lev = 0; lev < 4; ++lev) {
        n->dims[lev][0] = desc->in_size[2] >> lev;
        n->dims[lev][1] = desc->in_size[1] >> lev;
        n->dims[lev][2] = desc->in_size[0] >> lev;
    }
    long long off = 0;
    for (int l = 0; l < kWsLayers; ++l) {
        n->w_off[l] = off;
        off += 27LL * kCin[l] * kCout[l];
        n->b_off[l] = off;
        off += kCout[l];
        n->kpad[l] = round_up(27 * kCin[l], 32);
        n->kpad_t[l] = round_up(27 * kCout[l], 32);
    }
    n->wh_off = off;
    off += kHeadIn * kHeadOut;
    n->bh_off = off;
    off += kHeadOut;
    n->n_params = off;
    auto bail = [&](hipError_t e) {
        delete n;
        return ws_fail(e == hipErrorOutOfMemory ? SKM_ERROR_OUT_OF_MEMORY : SKM_ERROR_HIP,
                       std::string("warm-start buffers: ") + hipGetErrorString(e));
    };
    hipError_t e;
    if ((e = ws_alloc(&n->p, off)) || (e = ws_alloc(&n->g, off)) || (e = ws_alloc(&n->m, off)) ||
        (e = ws_alloc(&n->v, off)) || (e = ws_alloc(&n->wh, kHeadIn * kHeadOut)))
        return bail(e);
    const long long B = n->maxb, ncp = n->n_cp();
    long long gmax = 0;
    for (int l = 0; l < kWsLayers; ++l) {
        if ((e = ws_alloc(&n->w[l], (long long)kCout[l] * n->kpad[l]))) return bail(e);
        if (l > 0 && (e = ws_alloc(&n->wt[l], (long long)kCin[l] * n->kpad_t[l]))) return bail(e);
        const long long na = B * n->level_vox(kLevel[l]) * kCout[l];
        if ((e = ws_alloc(&n->act[l], na))) return bail(e);
        const int sl = (l == 8) ? 2 : (l == 9) ? 1 : src_level(l);
        gmax = std::max(gmax, std::max(na, B * n->level_vox(sl) * kCin[l]));
        const long long chunks = (B * n->level_vox(kLevel[l]) + kWdwChunk - 1) / kWdwChunk;
        n->wpart_n = std::max(n->wpart_n, chunks * kCout[l] * n->kpad[l]);
        n->bpart_n = std::max(n->bpart_n, (B * n->level_vox(kLevel[l]) + kBiasChunk - 1) / kBiasChunk * kCout[l]);
    }
    if ((e = ws_alloc(&n->cat3, B * n->level_vox(2) * kCin[8])) || (e = ws_alloc(&n->cat2, B * n->level_vox(1) * kCin[9])) ||
        (e = ws_alloc(&n->samp, B * ncp * kHeadIn)) || (e = ws_alloc(&n->out, B * ncp * kHeadOut)) ||
        (e = ws_alloc(&n->dout, B * ncp * kHeadOut)) || (e = ws_alloc(&n->dsamp, B * ncp * kHeadIn)) ||
        (e = ws_alloc(&n->grad_a, gmax)) || (e = ws_alloc(&n->grad_b, gmax)) ||
        (e = ws_alloc(&n->grad_skip3, B * n->level_vox(2) * 48)) ||
        (e = ws_alloc(&n->grad_skip2, B * n->level_vox(1) * 32)) || (e = ws_alloc(&n->wpart, n->wpart_n)) ||
        (e = ws_alloc(&n->bpart, n->bpart_n)) ||
        (e = ws_alloc(&n->loss_part, (long long)blocks_of(B * ncp * kHeadOut))))
        return bail(e);
    if ((e = hipMemset(n->p, 0, sizeof(float) * off)) || (e = hipMemset(n->m, 0, sizeof(float) * off)) ||
        (e = hipMemset(n->v, 0, sizeof(float) * off)) || (e = hipMemset(n->g, 0, sizeof(float) * off)))
        return bail(e);
    *out = n;
    return SKM_SUCCESS;
}

SKM_API void skm_ws_destroy(skm_ws_net *net) { delete net; }

SKM_API int64_t skm_ws_num_params(const skm_ws_net *net) { return net ? net->n_params : 0; }

SKM_API skm_status skm_ws_set_stream(skm_ws_net *net, void *hip_stream)
{
    if (!net) return ws_fail(SKM_ERROR_INVALID_ARGUMENT, "null network");
    net->stream = static_cast<hipStream_t>(hip_stream);
    return SKM_SUCCESS;
}

SKM_API skm_status skm_ws_set_params(skm_ws_net *net, const float *params)
{
    if (!net || !params) return ws_fail(SKM_ERROR_INVALID_ARGUMENT, "null argument");
    WS_HIP(hipMemcpyAsync(net->p, params, sizeof(float) * net->n_params, hipMemcpyHostToDevice, net->stream));
    WS_HIP(hipMemsetAsync(net->m, 0, sizeof(float) * net->n_params, net->stream));
    WS_HIP(hipMemsetAsync(net->v, 0, sizeof(float) * net->n_params, net->stream));
    net->adam_t = 0;
    cast_all(net);
    WS_HIP(hipGetLastError());
    WS_HIP(hipStreamSynchronize(net->stream));
    return SKM_SUCCESS;
}

SKM_API skm_status skm_ws_get_params(skm_ws_net *net, float *params)
{
    if (!net || !params) return ws_fail(SKM_ERROR_INVALID_ARGUMENT, "null argument");
    WS_HIP(hipMemcpyAsync(params, net->p, sizeof(float) * net->n_params, hipMemcpyDeviceToHost, net->stream));
    WS_HIP(hipStreamSynchronize(net->stream));
    return SKM_SUCCESS;
}

SKM_API skm_status skm_ws_forward(skm_ws_net *net, const float *d_input, int32_t batch, float *d_out)
{
    if (!net || !d_input || !d_out) return ws_fail(SKM_ERROR_INVALID_ARGUMENT, "null argument");
    return forward_impl(net, d_input, batch, d_out);
}

SKM_API skm_status skm_ws_backward(skm_ws_net *net, const float *d_dout, float *d_grad)
{
    if (!net || !d_dout) return ws_fail(SKM_ERROR_INVALID_ARGUMENT, "null argument");
    skm_status st = backward_impl(net, d_dout);
    if (st != SKM_SUCCESS) return st;
    if (d_grad)
        WS_HIP(hipMemcpyAsync(d_grad, net->g, sizeof(float) * net->n_params, hipMemcpyDeviceToDevice, net->stream));
    return SKM_SUCCESS;
}

SKM_API skm_status skm_ws_prepare_input(skm_ws_net *net, const float *d_moving, const float *d_fixed, int32_t index,
                                        float *d_input)
{
    if (!net || !d_moving || !d_fixed || !d_input) return ws_fail(SKM_ERROR_INVALID_ARGUMENT, "null argument");
    if (index < 0 || index >= net->maxb) return ws_fail(SKM_ERROR_INVALID_ARGUMENT, "index outside 0..max_batch-1");
    const int d = net->dims[0][0], h = net->dims[0][1], w = net->dims[0][2];
    const long long nv = (long long)d * h * w;
    k_prepare_input<<<blocks_of(nv), 256, 0, net->stream>>>(d_moving, d_fixed, d, h, w, d_input + 2 * nv * index);
    WS_HIP(hipGetLastError());
    return SKM_SUCCESS;
}

SKM_API skm_status skm_ws_get_activation(skm_ws_net *net, int32_t layer, float *d_dst)
{
    if (!net || !d_dst) return ws_fail(SKM_ERROR_INVALID_ARGUMENT, "null argument");
    if (layer < 0 || layer >= kWsLayers) return ws_fail(SKM_ERROR_INVALID_ARGUMENT, "layer outside 0..9");
    if (net->last_batch < 1) return ws_fail(SKM_ERROR_INVALID_ARGUMENT, "no forward pass yet");
    const long long n = (long long)net->last_batch * net->level_vox(kLevel[layer]) * kCout[layer];
    WS_HIP(hipMemcpyAsync(d_dst, net->act[layer], sizeof(float) * n, hipMemcpyDeviceToDevice, net->stream));
    return SKM_SUCCESS;
}

SKM_API skm_status skm_ws_train_step(skm_ws_net *net, const float *d_input, const float *d_target, int32_t batch,
                                     double learning_rate, double *loss)
{
    if (!net || !d_input || !d_target) return ws_fail(SKM_ERROR_INVALID_ARGUMENT, "null argument");
    skm_status st = forward_impl(net, d_input, batch, net->out);
    if (st != SKM_SUCCESS) return st;
    const long long n = (long long)batch * net->n_cp() * kHeadOut;
    const unsigned nb = blocks_of(n);
    k_mse_loss<<<nb, 256, 0, net->stream>>>(net->out, d_target, n, net->dout, net->loss_part);
    if ((st = backward_impl(net, net->dout)) != SKM_SUCCESS) return st;
    net->adam_t += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - (float)std::pow((double)b1, (double)net->adam_t);
    const float c2 = 1.0f - (float)std::pow((double)b2, (double)net->adam_t);
    k_adam<<<blocks_of(net->n_params), 256, 0, net->stream>>>(net->p, net->g, net->m, net->v, net->n_params,
                                                              (float)learning_rate, b1, b2, 1e-8f, c1, c2);
    cast_all(net);
    WS_HIP(hipGetLastError());
    if (loss) {
        std::vector<double> part(nb);
        WS_HIP(hipMemcpyAsync(part.data(), net->loss_part, sizeof(double) * nb, hipMemcpyDeviceToHost, net->stream));
        WS_HIP(hipStreamSynchronize(net->stream));
        double s = 0.0;
        for (double v : part) s += v;
        *loss = s / (double)n;
    }
    return SKM_SUCCESS;
}

}  // extern "C"
