// skm_kmatrix.hip -- the kernel matrix k(x_m, p_j) assembled in device memory.
//
// The reference evaluates every kernel pair twice per GetValueAndDerivative,
// once in the value pass (GpuEvaluatorMSE.cu:61-80, GpuEvaluatorNCC.cu:61-82)
// and once in the derivative pass (GpuEvaluatorMSE.cu:134-162,
// GpuEvaluatorNCC.cu:145-176), and again on every call of the optimizer's
// loop.  With subsample 1 the samples are the fixed voxels, so k(x_m, p_j)
// depends on nothing that changes between calls: the positions, the kernel and
// the weight image / tensor field are fixed when the evaluator is created; only
// the coefficients c_j move.  Here the non-zero pairs are evaluated once per
// evaluator (two walks per layout: a count, then the values) and stored in the
// order each pass reads them; an evaluation then streams them from HBM:
//   value:  u(x) = df(x) + sum_j (w_j k) c_j over the stored pairs of x's brick;
//   gather: d_j = w_j sum_m k v_m over the stored pairs of j's tile.
// Both keep the windows of the on-the-fly passes (the value pass's per-sample
// window, the derivative's per-control-point window), so a pair that one pass
// truncates is absent from that pass's layout.  Sums stay in a fixed order:
// the results are the same bits on every call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cstdlib>
#include <vector>

#include "skm_internal.hpp"
#include "skm_walk.hpp"

namespace skm {

namespace {

// Records per batch of the two evaluations: every unit's record list is padded
// with empty records (mask 0) to a whole number of batches, so that a batch's
// loads are unconditional.  Values are allocated with kEntryPad to spare (lanes
// without a pair read their unit's first value).
constexpr int kValueBatch = 16, kGatherBatch = 2;
constexpr int kEntryPad = 16;

__host__ __device__ constexpr int round_up(int n, int b) { return (n + b - 1) / b * b; }

__device__ __forceinline__ unsigned lane_rank(unsigned long long mask)
{
    return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

__device__ __forceinline__ bool lane_bit(unsigned long long mask, int lane) { return (mask >> lane) & 1ull; }

// Loads of the layouts, which every evaluation reads once: non-temporal, so
// that they do not push the samples' and control points' data, which are read
// again by other units, out of the L2.
typedef float kmF4 __attribute__((ext_vector_type(4)));
typedef unsigned kmU4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 stream_load(const float4 *p)
{
    const kmF4 v = __builtin_nontemporal_load(reinterpret_cast<const kmF4 *>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint4 stream_load(const uint4 *p)
{
    const kmU4 v = __builtin_nontemporal_load(reinterpret_cast<const kmU4 *>(p));
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ unsigned stream_load(const unsigned *p) { return __builtin_nontemporal_load(p); }

// Where one phase of a build writes: COUNT phases store capacities per unit
// (entries, values: every lane whose pair passes the window and reach tests),
// FILL phases the entries and values at the scanned offsets and the entry
// count per unit.  A FILL unit that would pass its capacity stops and raises
// *overflow (the build then gives up; the two walks use one predicate, so it
// does not happen).
struct BuildArgs {
    int *cap_ent;
    int *cap_val;
    const long long *ent_off;
    const long long *val_off;
    int *cnt;
    int *nnz;         // FILL: non-zero pairs stored per unit
    uint4 *ent;       // value layout: two per record; gather layout: one
    float4 *val;      // one float4 of k per lane of a record's mask
    int *overflow;
};

// ------------------------------------------------------------ value layout

// One wave per brick of the value pass (Tile / tile_coords: four 4x4x4 or
// 16x4 bricks per workgroup), walking the union of its lanes' reference
// windows row by row with wave-uniform control points, as kernel_sum_wave does
// (GpuEvaluatorMSE.cu:61-80): a control point joins a lane when it lies in the
// lane's window and within the pair's reach (weighted modes: both points'
// lam; stationary modes: the support).
template <int DIM, int MODE, bool FILL>
__global__ __launch_bounds__(kThreads) void k_km_value_build(Problem pr, BuildArgs ba)
{
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    int l[3];
    tile_coords<DIM>(threadIdx.x, l);
    int m[3] = {(int)blockIdx.x * Tile<DIM>::X + l[0], (int)blockIdx.y * Tile<DIM>::Y + l[1],
                DIM == 3 ? (int)blockIdx.z * Tile<DIM>::Z + l[2] : 0};
    const bool active = m[0] < pr.nsub[0] && m[1] < pr.nsub[1] && (DIM == 2 || m[2] < pr.nsub[2]);
    if (!active) m[0] = m[1] = m[2] = 0;
    const float fp[3] = {(float)m[0], (float)m[1], (float)m[2]};
    float x[3];
    to_global<DIM>(pr.fixed, fp, x);
    PState st;
    point_state<DIM, MODE>(pr.kc, pr.weight, pr.weight_data, x, st);
    const long long unit = block_linear() * (kThreads / 64) + wave;

    float ci[3];
    to_local<DIM>(pr.cp, x, ci);
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const float r = st.support * pr.cp.ih[d];
        lo[d] = active ? clampi((int)(ci[d] - r + 1.0f), 0, pr.cp.n[d] - 1) : INT_MAX;
        hi[d] = active ? clampi((int)(ci[d] + r), 0, pr.cp.n[d] - 1) : INT_MIN;
    }
    int ulo[3] = {0, 0, 0}, uhi[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        ulo[d] = __builtin_amdgcn_readfirstlane(wave_min_i(lo[d]));
        uhi[d] = __builtin_amdgcn_readfirstlane(wave_max_i(hi[d]));
    }
    // reach: the pair's |v|^2 bound is the sample's share + the control
    // point's (weighted modes) or the support squared (stationary modes)
    float qx, qx_max;
    if constexpr (is_stationary(MODE)) {
        qx = pr.kc.smin * pr.kc.smin * 1.0001f;
        qx_max = qx;
    } else {
        qx = pair_reach2(pr.kc.smin, pstate_lam<DIM, MODE>(st), 0.0f);
        qx_max = wave_max_f(active ? qx : 0.0f);
    }
    float bmin[3] = {0.0f, 0.0f, 0.0f}, bmax[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        bmin[d] = -wave_max_f(active ? -x[d] : -FLT_MAX);
        bmax[d] = wave_max_f(active ? x[d] : -FLT_MAX);
    }
    long long ebase = 0, vbase = 0, cap_e = 0, cap_v = 0;
    if (FILL) {
        ebase = ba.ent_off[unit];
        vbase = ba.val_off[unit];
        cap_e = ba.cap_ent[unit];
        cap_v = ba.cap_val[unit];
    }
    // pending group of up to four control points (walk order): their indices,
    // each lane's k, the lanes with a pair; COUNT groups the points with a
    // pair passing the window and reach tests, FILL the same points (so that it
    // forms the same groups) and keeps the non-zero k
    int ne = 0, gn = 0;
    long long nv = 0, npairs = 0;
    bool full = false;
    unsigned gz[4] = {0u, 0u, 0u, 0u};
    float gk[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    unsigned long long gany = 0;
    auto flush = [&]() -> bool {
        const int slots = __popcll(gany);
        if (FILL && gany) {
            if (ne >= cap_e || nv + slots > cap_v) return false;
            if (lane == 0) {
                ba.ent[2 * (ebase + ne)] = make_uint4((unsigned)gany, (unsigned)(gany >> 32), gz[0], gz[1]);
                ba.ent[2 * (ebase + ne) + 1] = make_uint4(gz[2], gz[3], 0u, 0u);
            }
            if (lane_bit(gany, lane)) ba.val[vbase + nv + lane_rank(gany)] = make_float4(gk[0], gk[1], gk[2], gk[3]);
        }
        if (!FILL || gany) {
            ++ne;
            nv += slots;
        }
        gn = 0;
        gany = 0;
        gz[0] = gz[1] = gz[2] = gz[3] = 0u;
        gk[0] = gk[1] = gk[2] = gk[3] = 0.0f;
        return true;
    };
    for (int cz = ulo[2]; cz <= uhi[2] && !full; ++cz) {
        const float pz = DIM == 3 ? pr.cp.o[2] + pr.cp.h[2] * cz : 0.0f;
        const float dz = DIM == 3 ? x[2] - pz : 0.0f;
        const float dzb = DIM == 3 ? fmaxf(0.0f, fmaxf(bmin[2] - pz, pz - bmax[2])) : 0.0f;
        const bool zin = DIM == 2 || (cz >= lo[2] && cz <= hi[2]);
        for (int cy = ulo[1]; cy <= uhi[1] && !full; ++cy) {
            const float py = pr.cp.o[1] + pr.cp.h[1] * cy;
            const float dyb = fmaxf(0.0f, fmaxf(bmin[1] - py, py - bmax[1]));
            float rem = qx_max - dyb * dyb - dzb * dzb;
            if (!is_stationary(MODE))
                rem += pair_reach2(pr.kc.smin, 0.0f, pr.cp_row_lam[cy + (long long)pr.cp.n[1] * cz]);
            if (!(rem > 0.0f)) continue;
            const float rx = fast_sqrt(rem);
            const int uxa = __builtin_amdgcn_readfirstlane(
                max(ulo[0], (int)floorf((bmin[0] - rx - pr.cp.o[0]) * pr.cp.ih[0])));
            const int uxb = __builtin_amdgcn_readfirstlane(
                min(uhi[0], (int)ceilf((bmax[0] + rx - pr.cp.o[0]) * pr.cp.ih[0])));
            const bool rin = active && zin && cy >= lo[1] && cy <= hi[1];
            const int xa = rin ? lo[0] : INT_MAX, xb = rin ? hi[0] : INT_MIN;
            const float dy = x[1] - py;
            const long long row = (long long)pr.cp.n[0] * (cy + (long long)pr.cp.n[1] * cz);
            const RowTerms rt = row_terms<DIM>(dy, dz);
            float fcx = (float)uxa;
            for (int cx = uxa; cx <= uxb; ++cx, fcx += 1.0f) {
                const float qc = is_stationary(MODE) ? 0.0f : pr.cp_reach[row + cx];
                const float dx = x[0] - fmaf(pr.cp.h[0], fcx, pr.cp.o[0]);
                const bool pass = (cx >= xa) & (cx <= xb) & (fmaf(dx, dx, rt.yz2) < qx + qc);
                const unsigned long long pm = __ballot(pass);
                if (!pm) continue;
                // the control point takes slot gn of the pending group
                if (!FILL) {
                    gany |= pm;
                } else {
                    float k = 0.0f;
                    if (pass) {
                        if constexpr (is_stationary(MODE)) {
                            k = stationary_k<MODE>(fmaf(dx, dx, rt.yz2), pr.kc.inv_s4, pr.kc.inv_s0);
                        } else {
                            PState cst;
                            cp_state_load<DIM, MODE>(pr, row + cx, cst);
                            k = weighted_k_row<DIM, MODE>(pr.kc, dx, rt, st, cst);
                        }
                    }
                    const unsigned long long km = __ballot(pass && k != 0.0f);
                    if (!lane_bit(km, lane)) k = 0.0f;
                    gk[0] = gn == 0 ? k : gk[0];
                    gk[1] = gn == 1 ? k : gk[1];
                    gk[2] = gn == 2 ? k : gk[2];
                    gk[3] = gn == 3 ? k : gk[3];
                    gz[gn] = (unsigned)(row + cx);
                    gany |= km;
                    npairs += __popcll(km);
                }
                if (++gn == 4 && !flush()) {
                    full = true;
                    break;
                }
            }
        }
    }
    if (gn > 0 && !flush()) full = true;
    if (lane != 0) return;
    if (!FILL) {
        ba.cap_ent[unit] = round_up(ne, kValueBatch);
        ba.cap_val[unit] = (int)nv;
    } else {
        for (int e = ne; e < round_up(ne, kValueBatch); ++e) {
            ba.ent[2 * (ebase + e)] = make_uint4(0u, 0u, 0u, 0u);
            ba.ent[2 * (ebase + e) + 1] = make_uint4(0u, 0u, 0u, 0u);
        }
        ba.cnt[unit] = round_up(ne, kValueBatch);
        ba.nnz[unit] = (int)npairs;
        ba.cap_val[unit] = (int)nv;  // the slots stored (the capacity was read above)
        if (full) atomicOr(ba.overflow, 1);
    }
}

// Value pass from the value layout: the same per-sample prologue and tail as
// k_value (evaluateValueMSE / evaluateValueNCC), u accumulated from the brick's
// stored groups in the walk's order.  Sixteen records (64 control points) per
// batch: each lane loads one of the batch's control-point indices and gathers
// its (c, w) with one vector load, lanes 0-15 the records' masks; the values of
// all sixteen records are loaded together, and the next batch's indices and masks
// are loaded before this batch's sums.  The points' (c, w) reach every lane
// through LDS, one broadcast read per point (a v_readlane per component would
// be four VALU instructions per point, each with a wait before its SGPR is
// read).
template <int DIM, int MEASURE>
__global__ __launch_bounds__(kThreads) void k_km_value(Problem pr, const float4 *__restrict__ cpc,
                                                       const uint4 *__restrict__ ent,
                                                       const float4 *__restrict__ val,
                                                       const long long *__restrict__ ent_off,
                                                       const long long *__restrict__ val_off,
                                                       const int *__restrict__ cnt, float4 *__restrict__ sg,
                                                       float *__restrict__ sb, double *__restrict__ partials)
{
    constexpr int R = kValueBatch;
    static_assert(4 * R <= 64, "one lane per control point of a batch");
    __shared__ double lds[kPartialsK * (kThreads / 64)];
    __shared__ float4 lcw[kThreads / 64][4 * R];  // per wave: the batch's points' (c, w)
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    int l[3];
    tile_coords<DIM>(threadIdx.x, l);
    int m[3] = {(int)blockIdx.x * Tile<DIM>::X + l[0], (int)blockIdx.y * Tile<DIM>::Y + l[1],
                DIM == 3 ? (int)blockIdx.z * Tile<DIM>::Z + l[2] : 0};
    const bool valid = m[0] < pr.nsub[0] && m[1] < pr.nsub[1] && (DIM == 2 || m[2] < pr.nsub[2]);
    double acc[kPartialsK];
#pragma unroll
    for (int k = 0; k < kPartialsK; ++k) acc[k] = 0.0;
    if (!valid) m[0] = m[1] = m[2] = 0;
    const long long s = m[0] + (long long)pr.nsub[0] * (m[1] + (long long)pr.nsub[1] * m[2]);
    const int pos[3] = {m[0], m[1], m[2]};
    const float fp[3] = {(float)pos[0], (float)pos[1], (float)pos[2]};
    float x[3];
    to_global<DIM>(pr.fixed, fp, x);
    float u[3];
    df_lookup<DIM>(pr, x, u);
    const long long unit = block_linear() * (kThreads / 64) + wave;
    // record r: words 8r .. 8r+7 = mask lo, mask hi, points 0..3, 0, 0
    const unsigned *W = reinterpret_cast<const unsigned *>(ent + 2 * ent_off[unit]);
    const float4 *V = val + val_off[unit];
    const int ne = cnt[unit];  // a whole number of batches
    const int rz = lane < 4 * R ? lane >> 2 : 0, rm = lane < R ? lane : 0;
    unsigned zq = 0, mlo = 0, mhi = 0;
    if (ne > 0) {
        zq = stream_load(W + 8 * rz + 2 + (lane & 3));
        mlo = stream_load(W + 8 * rm);
        mhi = stream_load(W + 8 * rm + 1);
    }
    for (int e0 = 0; e0 < ne; e0 += R) {
        const float4 cq = cpc[zq];
        unsigned long long any[R];
        int off[R];
        int o = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            any[r] = (unsigned)__builtin_amdgcn_readlane((int)mlo, r) |
                     ((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)mhi, r) << 32);
            off[r] = o;
            o += __popcll(any[r]);
        }
        float4 kk[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const bool on = lane_bit(any[r], lane);
            const float4 kv = stream_load(V + (on ? off[r] + (int)lane_rank(any[r]) : 0));
            kk[r] = make_float4(on ? kv.x : 0.0f, on ? kv.y : 0.0f, on ? kv.z : 0.0f, on ? kv.w : 0.0f);
        }
        if (e0 + R < ne) {
            const unsigned *Wn = W + 8 * (e0 + R);
            zq = stream_load(Wn + 8 * rz + 2 + (lane & 3));
            mlo = stream_load(Wn + 8 * rm);
            mhi = stream_load(Wn + 8 * rm + 1);
        }
        if (lane < 4 * R) lcw[wave][lane] = cq;
        // a lane's k is 0 where it has no pair with the point: u + 0 c = u
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const float k4[4] = {kk[r].x, kk[r].y, kk[r].z, kk[r].w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float4 c4 = lcw[wave][4 * r + q];
                const float wk = c4.w * k4[q];
                u[0] = fmaf(wk, c4.x, u[0]);
                u[1] = fmaf(wk, c4.y, u[1]);
                if (DIM == 3) u[2] = fmaf(wk, c4.z, u[2]);
            }
        }
        V += o;
    }
    if (valid) value_sample<DIM, MEASURE>(pr, s, pos, x, u, sg, sb, acc);
    block_sum<kPartialsK>(acc, lds);
    if (threadIdx.x == 0) {
        const long long b = block_linear();
#pragma unroll
        for (int k = 0; k < kPartialsK; ++k) partials[b * kPartialsK + k] = acc[k];
    }
}

// ------------------------------------------------------------ gather layout

// Sample block of one gather step: 4x4x4 samples (3D) or 8x8 (2D), lane
// (tx, ty, tz) = (lane % BX, lane / BX % BY, lane / (BX BY)).
template <int DIM> struct GBlock;
template <> struct GBlock<2> { static constexpr int X = 8, Y = 8, Z = 1; };
template <> struct GBlock<3> { static constexpr int X = 4, Y = 4, Z = 4; };

// A gather workgroup of the kernel-matrix form takes a 2 x 2 group of the
// derivative gather's control-point tiles (CpTile: 4 points each), 16 points:
// each stored sample block then serves all 16 with one load of its samples' v.
constexpr int kGroupTiles = 4;

struct GroupMap {
    int ntx, nty, ngx, ngy, n_groups, n_tiles;
};

template <int DIM>
__host__ __device__ __forceinline__ GroupMap group_map(const Problem &pr)
{
    GroupMap g;
    g.ntx = (pr.cp.n[0] + CpTile<DIM>::X - 1) / CpTile<DIM>::X;
    g.nty = (pr.cp.n[1] + CpTile<DIM>::Y - 1) / CpTile<DIM>::Y;
    const int ntz = DIM == 3 ? (pr.cp.n[2] + CpTile<DIM>::Z - 1) / CpTile<DIM>::Z : 1;
    g.ngx = (g.ntx + 1) / 2;
    g.ngy = (g.nty + 1) / 2;
    g.n_groups = g.ngx * g.ngy * ntz;
    g.n_tiles = g.ntx * g.nty * ntz;
    return g;
}

// Workgroup -> (slab, group of tiles), slab-major.  tile[s] is the s-th tile
// of the group (x, y offsets s & 1, s >> 1) or -1 past the grid; false when
// the slab lies outside the windows of every tile of the group (k_gather_tile
// writes no partials for such a slab of a tile either).
template <int DIM>
__device__ __forceinline__ bool group_unit(const Problem &pr, const GroupMap &g, int &slab_id, int *tile, int *tc)
{
    const int grp = (int)(blockIdx.x % (unsigned)g.n_groups);
    slab_id = (int)(blockIdx.x / (unsigned)g.n_groups);
    const int gx = grp % g.ngx, gy = (grp / g.ngx) % g.ngy, gz = grp / (g.ngx * g.ngy);
    const int slab = gather_slab<DIM>(pr);
    bool live = false;
#pragma unroll
    for (int s = 0; s < kGroupTiles; ++s) {
        const int t0 = 2 * gx + (s & 1), t1 = 2 * gy + (s >> 1);
        tile[s] = -1;
        tc[3 * s] = t0;
        tc[3 * s + 1] = t1;
        tc[3 * s + 2] = gz;
        if (t0 >= g.ntx || t1 >= g.nty) continue;
        tile[s] = t0 + g.ntx * (t1 + g.nty * gz);
        int plo, phi;
        tile_plane_range<DIM>(pr, t0, t1, gz, plo, phi);
        live = live || !(plo > phi || slab_id < plo / slab || slab_id > phi / slab);
    }
    return live;
}

// Control point c (0-3) of tile (t0, t1, t2); false past the grid.
template <int DIM>
__device__ __forceinline__ bool tile_point(const Problem &pr, int t0, int t1, int t2, int c, long long &j, int *jc)
{
    constexpr int TX = CpTile<DIM>::X, TY = CpTile<DIM>::Y, TZ = CpTile<DIM>::Z;
    jc[0] = t0 * TX + c % TX;
    jc[1] = t1 * TY + (c / TX) % TY;
    jc[2] = DIM == 3 ? t2 * TZ + c / (TX * TY) : 0;
    const bool ok = jc[0] < pr.cp.n[0] && jc[1] < pr.cp.n[1] && (DIM == 2 || jc[2] < pr.cp.n[2]);
    j = ok ? jc[0] + (long long)pr.cp.n[0] * (jc[1] + (long long)pr.cp.n[1] * jc[2]) : 0;
    return ok;
}

// Walk of the derivative gather over aligned sample blocks of the union of
// the group's windows inside the slab (evaluateDerivativeMSE / NCC,
// GpuEvaluatorMSE.cu:134-162, GpuEvaluatorNCC.cu:145-176: a pair counts when
// the sample lies in the control point's window and, here, within the pair's
// reach).  Per block: one lane mask per tile (its four points' non-zero pairs)
// and per lane of a tile's mask the float4 of its four points' k.  The
// group's 16 points (position, reach share, window, state) are staged in LDS
// once; a point whose window misses a block's box, or whose reach with the
// largest lam any sample has cannot meet the box, is skipped for that block.
// sstate: packed states of the samples (weighted modes).
template <int DIM, int MODE, bool FILL>
__global__ __launch_bounds__(kThreads) void k_km_gather_build(Problem pr, GroupMap gm,
                                                              const float4 *__restrict__ sstate, BuildArgs ba)
{
    constexpr int T = CpTile<DIM>::X * CpTile<DIM>::Y * CpTile<DIM>::Z, G = kGroupTiles, NP = G * T;
    constexpr int BX = GBlock<DIM>::X, BY = GBlock<DIM>::Y, BZ = GBlock<DIM>::Z;
    static_assert(T == 4, "one float4 of k per lane, tile and block");
    __shared__ float4 lp[NP];      // position, lam (the point's reach share as in the on-the-fly gather)
    __shared__ int4 lwa[NP], lwb[NP];  // window lo (x, y, z, ok), hi (x, y, z, 0)
    __shared__ float4 lsa[NP], lsb[NP];  // packed state (weighted modes)
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const long long unit = (long long)blockIdx.x * (kThreads / 64) + wave;
    int slab_id, tile[G], tc[3 * G];
    const bool live = group_unit<DIM>(pr, gm, slab_id, tile, tc);
    if (live && threadIdx.x < NP) {
        const int sidx = threadIdx.x / T, c = threadIdx.x % T;
        long long j = 0;
        int jc[3] = {0, 0, 0};
        const bool ok = tile[sidx] >= 0 && tile_point<DIM>(pr, tc[3 * sidx], tc[3 * sidx + 1], tc[3 * sidx + 2], c, j, jc);
        const float fj[3] = {(float)jc[0], (float)jc[1], (float)jc[2]};
        float p[3] = {0.0f, 0.0f, 0.0f};
        to_global<DIM>(pr.cp, fj, p);
        float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f), b = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
        if (!is_stationary(MODE)) {
            a = pr.cp_pstate[2 * j];
            b = pr.cp_pstate[2 * j + 1];
        }
        lp[threadIdx.x] = make_float4(p[0], p[1], p[2], is_stationary(MODE) ? 1.0f : b.w);
        const int *w = pr.cp_window + j * 6;
        lwa[threadIdx.x] = make_int4(w[0], DIM > 1 ? w[1] : 0, DIM == 3 ? w[2] : 0, ok ? 1 : 0);
        lwb[threadIdx.x] = make_int4(w[3], DIM > 1 ? w[4] : 0, DIM == 3 ? w[5] : 0, 0);
        lsa[threadIdx.x] = a;
        lsb[threadIdx.x] = b;
    }
    __syncthreads();
    int ne = 0;
    long long nv = 0, npairs = 0;
    bool full = false;
    long long ebase = 0, vbase = 0, cap_e = 0, cap_v = 0;
    if (FILL) {
        ebase = ba.ent_off[unit];
        vbase = ba.val_off[unit];
        cap_e = ba.cap_ent[unit];
        cap_v = ba.cap_val[unit];
    }
    if (live) {
        const int slab = gather_slab<DIM>(pr);
        int ulo[3] = {INT_MAX, INT_MAX, DIM == 3 ? INT_MAX : 0}, uhi[3] = {INT_MIN, INT_MIN, DIM == 3 ? INT_MIN : 0};
        for (int q = 0; q < NP; ++q) {
            const int4 wa = lwa[q], wb = lwb[q];
            if (!wa.w) continue;
            const int lo3[3] = {wa.x, wa.y, wa.z}, hi3[3] = {wb.x, wb.y, wb.z};
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                ulo[d] = min(ulo[d], lo3[d]);
                uhi[d] = max(uhi[d], hi3[d]);
            }
        }
        ulo[DIM - 1] = max(ulo[DIM - 1], slab_id * slab);
        uhi[DIM - 1] = min(uhi[DIM - 1], slab_id * slab + slab - 1);
        const bool any = uhi[0] >= ulo[0] && uhi[1] >= ulo[1] && (DIM == 2 || uhi[2] >= ulo[2]);
        const int nb0 = any ? (uhi[0] - ulo[0] + BX) / BX : 0, nb1 = any ? (uhi[1] - ulo[1] + BY) / BY : 0;
        const int nb2 = any ? (DIM == 3 ? (uhi[2] - ulo[2] + BZ) / BZ : 1) : 0;
        const int tx = lane % BX, ty = (lane / BX) % BY, tz = lane / (BX * BY);
        const float smin2h = 0.5f * pr.kc.smin * pr.kc.smin * 1.0001f;
        // the largest lam a sample can have (stationary modes: 1)
        const float lam_smax = is_stationary(MODE) ? 1.0f : pr.lam_w_max;
        const int nblk = nb0 * nb1 * nb2;
        for (int blk = wave; blk < nblk && !full; blk += kThreads / 64) {
            const int bo[3] = {ulo[0] + (blk % nb0) * BX, ulo[1] + ((blk / nb0) % nb1) * BY,
                               DIM == 3 ? ulo[2] + (blk / (nb0 * nb1)) * BZ : 0};
            const int be[3] = {min(bo[0] + BX - 1, uhi[0]), min(bo[1] + BY - 1, uhi[1]),
                               DIM == 3 ? min(bo[2] + BZ - 1, uhi[2]) : 0};
            // the block's box in physical coordinates (uniform)
            float bl[3] = {0.0f, 0.0f, 0.0f}, bh[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                const float e0 = pr.fixed.o[d] + pr.fixed.h[d] * (float)bo[d];
                const float e1 = pr.fixed.o[d] + pr.fixed.h[d] * (float)be[d];
                bl[d] = fminf(e0, e1);
                bh[d] = fmaxf(e0, e1);
            }
            const int mm[3] = {bo[0] + tx, bo[1] + ty, bo[2] + tz};
            const bool inside = mm[0] <= uhi[0] && mm[1] <= uhi[1] && (DIM == 2 || mm[2] <= uhi[2]);
            const long long sidx = inside ? mm[0] + (long long)pr.nsub[0] * (mm[1] + (long long)pr.nsub[1] * mm[2]) : 0;
            PState sst;
            float lam_s = 1.0f;
            if constexpr (!is_stationary(MODE)) {
                const float4 ca = sstate[2 * sidx], cb = sstate[2 * sidx + 1];
                pstate_unpack<DIM, MODE>(ca, cb, sst);
                lam_s = cb.w;
            }
            const float fpos[3] = {(float)mm[0], (float)mm[1], (float)mm[2]};
            float xs[3];
            to_global<DIM>(pr.fixed, fpos, xs);
            unsigned long long sm[G];
            float k[G][T];
            int pairs = 0;
            // an index the compiler cannot see through: the points' LDS reads
            // stay inside the block loop instead of 16 points' data held in VGPRs
            int q0 = 0;
            asm volatile("" : "+v"(q0));
#pragma unroll
            for (int sg = 0; sg < G; ++sg) {
                sm[sg] = 0;
#pragma unroll
                for (int c = 0; c < T; ++c) {
                    k[sg][c] = 0.0f;
                    const int q = q0 + sg * T + c;
                    const int4 wa = lwa[q], wb = lwb[q];
                    const float4 pp = lp[q];
                    // the block's box against the point's window and largest reach (uniform)
                    const int wlo[3] = {wa.x, wa.y, wa.z}, whi[3] = {wb.x, wb.y, wb.z};
                    const float pc[3] = {pp.x, pp.y, pp.z};
                    bool meet = wa.w != 0;
                    float dbox = 0.0f;
#pragma unroll
                    for (int d = 0; d < DIM; ++d) {
                        meet = meet && bo[d] <= whi[d] && be[d] >= wlo[d];
                        const float g = fmaxf(0.0f, fmaxf(bl[d] - pc[d], pc[d] - bh[d]));
                        dbox = fmaf(g, g, dbox);
                    }
                    if (!meet || !(dbox < smin2h * (pp.w + lam_smax))) continue;
                    const float v[3] = {pc[0] - xs[0], pc[1] - xs[1], DIM == 3 ? pc[2] - xs[2] : 0.0f};
                    const float d2 = v[0] * v[0] + v[1] * v[1] + (DIM == 3 ? v[2] * v[2] : 0.0f);
                    bool pass = inside;
#pragma unroll
                    for (int d = 0; d < DIM; ++d) pass = pass && mm[d] >= wlo[d] && mm[d] <= whi[d];
                    pass = pass && d2 < smin2h * (pp.w + lam_s);
                    if (!FILL) {
                        sm[sg] |= __ballot(pass);
                        continue;
                    }
                    if (!__ballot(pass)) continue;
                    float kv = 0.0f;
                    if (pass) {
                        PState cst;
                        if (!is_stationary(MODE)) pstate_unpack<DIM, MODE>(lsa[q], lsb[q], cst);
                        kv = kernel_pair<DIM, MODE>(pr.kc, v, cst, sst);
                    }
                    const unsigned long long km = __ballot(pass && kv != 0.0f);
                    k[sg][c] = lane_bit(km, lane) ? kv : 0.0f;
                    sm[sg] |= km;
                    pairs += __popcll(km);
                }
            }
            int slots = 0;
#pragma unroll
            for (int sg = 0; sg < G; ++sg) slots += __popcll(sm[sg]);
            if (!slots) continue;
            if (FILL) {
                if (ne >= cap_e || nv + slots > cap_v) {
                    full = true;
                    break;
                }
                if (lane == 0) {
                    const int org = (int)(bo[0] + (long long)pr.nsub[0] * (bo[1] + (long long)pr.nsub[1] * bo[2]));
                    ba.ent[3 * (ebase + ne)] = make_uint4((unsigned)sm[0], (unsigned)(sm[0] >> 32), (unsigned)sm[1],
                                                          (unsigned)(sm[1] >> 32));
                    ba.ent[3 * (ebase + ne) + 1] = make_uint4((unsigned)sm[2], (unsigned)(sm[2] >> 32),
                                                              (unsigned)sm[3], (unsigned)(sm[3] >> 32));
                    ba.ent[3 * (ebase + ne) + 2] = make_uint4((unsigned)org, 0u, 0u, 0u);
                }
                long long vb = vbase + nv;
#pragma unroll
                for (int sg = 0; sg < G; ++sg) {
                    if (lane_bit(sm[sg], lane))
                        ba.val[vb + lane_rank(sm[sg])] = make_float4(k[sg][0], k[sg][1], k[sg][2], k[sg][3]);
                    vb += __popcll(sm[sg]);
                }
                npairs += pairs;
            }
            ++ne;
            nv += slots;
        }
    }
    if (lane != 0) return;
    if (!FILL) {
        ba.cap_ent[unit] = round_up(ne, kGatherBatch);
        ba.cap_val[unit] = (int)nv;
    } else {
        for (int e = ne; e < round_up(ne, kGatherBatch); ++e) {
            ba.ent[3 * (ebase + e)] = make_uint4(0u, 0u, 0u, 0u);
            ba.ent[3 * (ebase + e) + 1] = make_uint4(0u, 0u, 0u, 0u);
            ba.ent[3 * (ebase + e) + 2] = make_uint4(0u, 0u, 0u, 0u);
        }
        ba.cnt[unit] = round_up(ne, kGatherBatch);
        ba.nnz[unit] = (int)npairs;
        ba.cap_val[unit] = (int)nv;  // the slots stored (the capacity was read above)
        if (full) atomicOr(ba.overflow, 1);
    }
}

// Derivative gather from the gather layout: per stored block every lane with
// a pair loads its sample's v once and, per tile whose mask holds it, the
// float4 of that tile's four points' k, and adds k v to their sums; records
// are wave-uniform (scalar loads), two per batch, the next batch's loaded
// before this batch's sums.  The sums then go to the FP64 (slab, tile)
// partials that k_gather_finish adds in slab order (fixed butterfly per wave,
// the four waves in FP64), for every tile of the group.
template <int DIM>
__global__ __launch_bounds__(kThreads) void k_km_gather(Problem pr, const float4 *__restrict__ sv, GroupMap gm,
                                                        const uint4 *__restrict__ ent,
                                                        const float4 *__restrict__ val,
                                                        const long long *__restrict__ ent_off,
                                                        const long long *__restrict__ val_off,
                                                        const int *__restrict__ cnt)
{
    constexpr int T = CpTile<DIM>::X * CpTile<DIM>::Y * CpTile<DIM>::Z, G = kGroupTiles;
    constexpr int BX = GBlock<DIM>::X, BY = GBlock<DIM>::Y;
    constexpr int B = kGatherBatch;
    static_assert(T == 4, "one float4 of k per lane, tile and block");
    __shared__ double lds[G * T * 3 * 4];
    int slab_id, tile[G], tc[3 * G];
    if (!group_unit<DIM>(pr, gm, slab_id, tile, tc)) return;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const long long unit = (long long)blockIdx.x * (kThreads / 64) + wave;
    const int tx = lane % BX, ty = (lane / BX) % BY, tz = lane / (BX * BY);
    const long long loff = tx + (long long)pr.nsub[0] * (ty + (long long)pr.nsub[1] * tz);
    float acc[G][T][3];
#pragma unroll
    for (int s = 0; s < G; ++s)
#pragma unroll
        for (int c = 0; c < T; ++c) acc[s][c][0] = acc[s][c][1] = acc[s][c][2] = 0.0f;
    const uint4 *E = ent + 3 * ent_off[unit];
    const float4 *V = val + val_off[unit];
    const int ne = cnt[unit];  // a whole number of batches
    uint4 en[B][3];
#pragma unroll
    for (int i = 0; i < B; ++i)
#pragma unroll
        for (int w = 0; w < 3; ++w) en[i][w] = ne > 0 ? stream_load(E + 3 * i + w) : make_uint4(0u, 0u, 0u, 0u);
    for (int e0 = 0; e0 < ne; e0 += B) {
        // lanes without a pair read element 0 and use zeros
        float4 cv[B], kk[B][G];
        int off = 0;
#pragma unroll
        for (int i = 0; i < B; ++i) {
            const unsigned long long sm[G] = {en[i][0].x | ((unsigned long long)en[i][0].y << 32),
                                              en[i][0].z | ((unsigned long long)en[i][0].w << 32),
                                              en[i][1].x | ((unsigned long long)en[i][1].y << 32),
                                              en[i][1].z | ((unsigned long long)en[i][1].w << 32)};
            const bool on = lane_bit(sm[0] | sm[1] | sm[2] | sm[3], lane);
            const float4 a = sv[on ? (int)en[i][2].x + loff : 0];
            cv[i] = make_float4(on ? a.x : 0.0f, on ? a.y : 0.0f, on ? a.z : 0.0f, 0.0f);
#pragma unroll
            for (int s = 0; s < G; ++s) {
                const bool ons = lane_bit(sm[s], lane);
                const float4 b = stream_load(V + (ons ? off + (int)lane_rank(sm[s]) : 0));
                kk[i][s] = make_float4(ons ? b.x : 0.0f, ons ? b.y : 0.0f, ons ? b.z : 0.0f, ons ? b.w : 0.0f);
                off += __popcll(sm[s]);
            }
        }
        if (e0 + B < ne) {
#pragma unroll
            for (int i = 0; i < B; ++i)
#pragma unroll
                for (int w = 0; w < 3; ++w) en[i][w] = stream_load(E + 3 * (e0 + B + i) + w);
        }
#pragma unroll
        for (int i = 0; i < B; ++i)
#pragma unroll
            for (int s = 0; s < G; ++s) {
                const float k4[T] = {kk[i][s].x, kk[i][s].y, kk[i][s].z, kk[i][s].w};
#pragma unroll
                for (int c = 0; c < T; ++c) {
                    acc[s][c][0] = fmaf(k4[c], cv[i].x, acc[s][c][0]);
                    acc[s][c][1] = fmaf(k4[c], cv[i].y, acc[s][c][1]);
                    if (DIM == 3) acc[s][c][2] = fmaf(k4[c], cv[i].z, acc[s][c][2]);
                }
            }
        V += off;
    }
    // wave sums (fixed butterfly, FP32), then the four waves in FP64
#pragma unroll
    for (int s = 0; s < G; ++s)
#pragma unroll
        for (int c = 0; c < T; ++c)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                float v = acc[s][c][d];
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
                if (lane == 0) lds[wave * G * T * 3 + (s * T + c) * 3 + d] = v;
            }
    __syncthreads();
    if (threadIdx.x < G * T * 3) {
        const int s = threadIdx.x / (T * 3), c = (threadIdx.x / 3) % T, d = threadIdx.x % 3;
        double sum = 0.0;
        for (int w = 0; w < kThreads / 64; ++w) sum += lds[w * G * T * 3 + threadIdx.x];
        long long j;
        int jc[3];
        if (d < DIM && tile[s] >= 0 && tile_point<DIM>(pr, tc[3 * s], tc[3 * s + 1], tc[3 * s + 2], c, j, jc))
            pr.gather_partials[((long long)slab_id * gm.n_tiles + tile[s]) * (T * 3) + c * 3 + d] = sum;
    }
}

// Exclusive prefix sums of n per-unit capacities (one workgroup, in a fixed
// order), and their total.
__global__ __launch_bounds__(1024) void k_km_scan(const int *__restrict__ cap, long long n,
                                                  long long *__restrict__ off, long long *__restrict__ total)
{
    __shared__ long long part[1024];
    const int t = threadIdx.x;
    const long long chunk = (n + 1023) / 1024;
    const long long a = min(n, t * chunk), b = min(n, a + chunk);
    long long sum = 0;
    for (long long i = a; i < b; ++i) sum += cap[i];
    part[t] = sum;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const long long v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    long long run = part[t] - sum;
    for (long long i = a; i < b; ++i) {
        off[i] = run;
        run += cap[i];
    }
    if (t == 1023) *total = part[1023];
}

// Block-parallel form of the same exclusive scan: each workgroup scans 1024
// capacities (four per thread) and writes its total; the totals are scanned
// by k_km_scan; k_km_scan_add adds each workgroup's offset.
constexpr int kScanBlock = 1024;

__global__ __launch_bounds__(256) void k_km_scan_local(const int *__restrict__ cap, long long n,
                                                       long long *__restrict__ off, int *__restrict__ bsum)
{
    __shared__ long long part[256];
    const long long base = (long long)blockIdx.x * kScanBlock + 4 * threadIdx.x;
    int v[4];
    long long sum = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        v[i] = base + i < n ? cap[base + i] : 0;
        sum += v[i];
    }
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        const long long add = (int)threadIdx.x >= o ? part[threadIdx.x - o] : 0;
        __syncthreads();
        part[threadIdx.x] += add;
        __syncthreads();
    }
    long long run = part[threadIdx.x] - sum;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (base + i < n) off[base + i] = run;
        run += v[i];
    }
    if (threadIdx.x == 255) bsum[blockIdx.x] = (int)part[255];
}

__global__ __launch_bounds__(256) void k_km_scan_add(long long *__restrict__ off, long long n,
                                                     const long long *__restrict__ boff)
{
    const long long i = (long long)blockIdx.x * kScanBlock + threadIdx.x;
    const long long add = boff[blockIdx.x];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (i + 256 * k < n) off[i + 256 * k] += add;
}

// Sum of n ints (one workgroup, fixed order).
__global__ __launch_bounds__(1024) void k_km_total(const int *__restrict__ v, long long n, long long *__restrict__ total)
{
    __shared__ long long part[1024];
    long long sum = 0;
    for (long long i = threadIdx.x; i < n; i += 1024) sum += v[i];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = part[0];
}

template <int DIM>
long long gather_tiles(const Problem &pr)
{
    return (long long)((pr.cp.n[0] + CpTile<DIM>::X - 1) / CpTile<DIM>::X) *
           ((pr.cp.n[1] + CpTile<DIM>::Y - 1) / CpTile<DIM>::Y) *
           (DIM == 3 ? (pr.cp.n[2] + CpTile<DIM>::Z - 1) / CpTile<DIM>::Z : 1);
}

template <int DIM>
dim3 value_grid(const Problem &pr)
{
    return dim3((pr.nsub[0] + Tile<DIM>::X - 1) / Tile<DIM>::X, (pr.nsub[1] + Tile<DIM>::Y - 1) / Tile<DIM>::Y,
                DIM == 3 ? (pr.nsub[2] + Tile<DIM>::Z - 1) / Tile<DIM>::Z : 1);
}

template <int DIM>
long long gather_workgroups(const Problem &pr)
{
    const int slab = gather_slab<DIM>(pr);
    return (long long)group_map<DIM>(pr).n_groups * ((pr.nsub[DIM - 1] + slab - 1) / slab);
}

template <typename T>
skm_status km_alloc(T **p, long long count)
{
    if (count < 1) count = 1;
    if (device_alloc((void **)p, sizeof(T) * (size_t)count) != hipSuccess)
        return set_last_error(SKM_ERROR_OUT_OF_MEMORY, "kernel matrix: device allocation failed");
    return SKM_SUCCESS;
}

template <int DIM, int MODE>
void build_count(const Problem &pr, const float4 *sstate, const BuildArgs &bv, const BuildArgs &bg, hipStream_t s)
{
    k_km_value_build<DIM, MODE, false><<<value_grid<DIM>(pr), kThreads, 0, s>>>(pr, bv);
    k_km_gather_build<DIM, MODE, false>
        <<<(unsigned)gather_workgroups<DIM>(pr), kThreads, 0, s>>>(pr, group_map<DIM>(pr), sstate, bg);
}

template <int DIM, int MODE>
void build_fill(const Problem &pr, const float4 *sstate, const BuildArgs &bv, const BuildArgs &bg, hipStream_t s)
{
    k_km_value_build<DIM, MODE, true><<<value_grid<DIM>(pr), kThreads, 0, s>>>(pr, bv);
    k_km_gather_build<DIM, MODE, true>
        <<<(unsigned)gather_workgroups<DIM>(pr), kThreads, 0, s>>>(pr, group_map<DIM>(pr), sstate, bg);
}

#define SKM_KM_DISPATCH(pr, FN, ...)                                                                              \
    do {                                                                                                           \
        if ((pr).dim == 2) {                                                                                       \
            if ((pr).kc.mode == 0) FN<2, 0>(__VA_ARGS__);                                                          \
            else if ((pr).kc.mode == 1) FN<2, 1>(__VA_ARGS__);                                                     \
            else if ((pr).kc.mode == 2) FN<2, 2>(__VA_ARGS__);                                                     \
            else FN<2, kModeStationaryEqual>(__VA_ARGS__);                                                         \
        } else {                                                                                                   \
            if ((pr).kc.mode == 0) FN<3, 0>(__VA_ARGS__);                                                          \
            else if ((pr).kc.mode == 1) FN<3, 1>(__VA_ARGS__);                                                     \
            else if ((pr).kc.mode == 2) FN<3, 2>(__VA_ARGS__);                                                     \
            else FN<3, kModeStationaryEqual>(__VA_ARGS__);                                                         \
        }                                                                                                          \
    } while (0)

}  // namespace

// The kernel-matrix gather's problem: slabs of 4 sample planes unless the
// tuning asks for 8 (4 measured faster: 0.49 against 0.53 ms at 128x128x96,
// each slab's samples then share an XCD's L2 better).
static Problem km_gather_problem(const Problem &pr)
{
    Problem p = pr;
    if (p.dim == 3 && p.gather_split != 8) p.gather_split = 4;
    return p;
}

int kernel_matrix_slab(const Problem &pr)
{
    const Problem p = km_gather_problem(pr);
    return p.dim == 2 ? gather_slab<2>(p) : gather_slab<3>(p);
}

bool kernel_matrix_applies(const Problem &pr)
{
    return pr.sub == 1 && pr.n_sub > 0 && pr.n_sub < (1LL << 31) && pr.n_cp < (1LL << 31);
}

void kernel_matrix_release(KernelMatrix &km)
{
    void *ptrs[] = {km.v_ent, km.v_val, km.v_ent_off, km.v_val_off, km.v_cnt,
                    km.g_ent, km.g_val, km.g_ent_off, km.g_val_off, km.g_cnt};
    for (void *p : ptrs)
        if (p) device_release(p);
    km = KernelMatrix{};
}

skm_status kernel_matrix_build(const Problem &pr_in, KernelMatrix &km, hipStream_t s, long long max_bytes)
{
    const Problem pr = km_gather_problem(pr_in);
    const auto t_start = std::chrono::steady_clock::now();
    kernel_matrix_release(km);
    if (!kernel_matrix_applies(pr)) return set_last_error(SKM_ERROR_INVALID_ARGUMENT, "no kernel matrix: subsample > 1");
    const int dim = pr.dim;
    const dim3 vg = dim == 2 ? value_grid<2>(pr) : value_grid<3>(pr);
    km.units_v = (long long)vg.x * vg.y * vg.z * (kThreads / 64);
    const long long gw = dim == 2 ? gather_workgroups<2>(pr) : gather_workgroups<3>(pr);
    km.units_g = gw * (kThreads / 64);
    km.slab = dim == 2 ? gather_slab<2>(pr) : gather_slab<3>(pr);
    // per-unit capacities and counts, offsets, totals, overflow flag (released at the end)
    int *cap = nullptr, *flag = nullptr;
    long long *totals = nullptr;
    float4 *sstate = nullptr;
    auto done = [&](skm_status st) {
        (void)hipStreamSynchronize(s);
        if (cap) device_release(cap);
        if (flag) device_release(flag);
        if (totals) device_release(totals);
        if (sstate) device_release(sstate);
        if (st != SKM_SUCCESS) kernel_matrix_release(km);
        return st;
    };
    skm_status st;
    if ((st = km_alloc(&cap, 3 * (km.units_v + km.units_g))) != SKM_SUCCESS) return done(st);
    if ((st = km_alloc(&flag, 1)) != SKM_SUCCESS) return done(st);
    if ((st = km_alloc(&totals, 6)) != SKM_SUCCESS) return done(st);
    if ((st = km_alloc(&km.v_ent_off, km.units_v)) != SKM_SUCCESS) return done(st);
    if ((st = km_alloc(&km.v_val_off, km.units_v)) != SKM_SUCCESS) return done(st);
    if ((st = km_alloc(&km.v_cnt, km.units_v)) != SKM_SUCCESS) return done(st);
    if ((st = km_alloc(&km.g_ent_off, km.units_g)) != SKM_SUCCESS) return done(st);
    if ((st = km_alloc(&km.g_val_off, km.units_g)) != SKM_SUCCESS) return done(st);
    if ((st = km_alloc(&km.g_cnt, km.units_g)) != SKM_SUCCESS) return done(st);
    if (!is_stationary(pr.kc.mode)) {
        // sample states at subsample 1 = the states at every fixed voxel
        if ((st = km_alloc(&sstate, 2 * pr.n_vox)) != SKM_SUCCESS) return done(st);
        launch_voxel_states(pr, sstate, s);
    }
    int *capg = cap + 2 * km.units_v, *nnz = cap + 2 * (km.units_v + km.units_g);
    BuildArgs bv{cap, cap + km.units_v, km.v_ent_off, km.v_val_off, km.v_cnt, nnz, nullptr, nullptr, flag};
    BuildArgs bg{capg, capg + km.units_g, km.g_ent_off, km.g_val_off, km.g_cnt, nnz + km.units_v, nullptr, nullptr,
                 flag};
    SKM_KM_DISPATCH(pr, build_count, pr, sstate, bv, bg, s);
    {
        // exclusive offsets of the four capacity arrays
        const long long nbv = (km.units_v + kScanBlock - 1) / kScanBlock, nbg = (km.units_g + kScanBlock - 1) / kScanBlock;
        int *bsum = nullptr;
        long long *boff = nullptr;
        if ((st = km_alloc(&bsum, 4 * std::max(nbv, nbg))) != SKM_SUCCESS) return done(st);
        if ((st = km_alloc(&boff, 4 * std::max(nbv, nbg))) != SKM_SUCCESS) {
            device_release(bsum);
            return done(st);
        }
        const int *caps[4] = {bv.cap_ent, bv.cap_val, bg.cap_ent, bg.cap_val};
        long long *offs[4] = {km.v_ent_off, km.v_val_off, km.g_ent_off, km.g_val_off};
        const long long ns[4] = {km.units_v, km.units_v, km.units_g, km.units_g};
        const long long mb = std::max(nbv, nbg);
        for (int i = 0; i < 4; ++i) {
            const long long nb = (ns[i] + kScanBlock - 1) / kScanBlock;
            k_km_scan_local<<<(unsigned)nb, 256, 0, s>>>(caps[i], ns[i], offs[i], bsum + i * mb);
            k_km_scan<<<1, 1024, 0, s>>>(bsum + i * mb, nb, boff + i * mb, totals + i);
            k_km_scan_add<<<(unsigned)nb, 256, 0, s>>>(offs[i], ns[i], boff + i * mb);
        }
        // released with the stream's work (the counts are copied back below, after a sync)
        (void)hipStreamSynchronize(s);
        device_release(bsum);
        device_release(boff);
    }
    launch_zero_words(flag, 1, s);
    long long tot[4] = {0, 0, 0, 0};
    if (hipMemcpyAsync(tot, totals, sizeof(tot), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return done(set_last_error(SKM_ERROR_HIP, "kernel matrix: counting pass failed"));
    km.cap_ent_v = tot[0];
    km.cap_val_v = tot[1];
    km.cap_ent_g = tot[2];
    km.cap_val_g = tot[3];
    const long long bytes = km.cap_ent_v * 32 + km.cap_val_v * 16 + km.cap_ent_g * 48 + km.cap_val_g * 16 +
                            (km.units_v + km.units_g) * 20;
    km.bytes = bytes;
    if (bytes > max_bytes) {
        done(SKM_SUCCESS);
        kernel_matrix_release(km);
        km.state = 2;
        km.bytes = bytes;
        return SKM_SUCCESS;
    }
    // values with room to spare: lanes without a pair read their unit's first
    // value, which for a unit without values is one past the last
    if ((st = km_alloc(&km.v_ent, 2 * km.cap_ent_v)) != SKM_SUCCESS) return done(st);
    if ((st = km_alloc(&km.v_val, km.cap_val_v + kEntryPad)) != SKM_SUCCESS) return done(st);
    if ((st = km_alloc(&km.g_ent, 3 * km.cap_ent_g)) != SKM_SUCCESS) return done(st);
    if ((st = km_alloc(&km.g_val, km.cap_val_g + kEntryPad)) != SKM_SUCCESS) return done(st);
    bv.ent = km.v_ent;
    bv.val = km.v_val;
    bg.ent = km.g_ent;
    bg.val = km.g_val;
    SKM_KM_DISPATCH(pr, build_fill, pr, sstate, bv, bg, s);
    // stored pairs, float4 slots and records of each layout
    k_km_total<<<1, 1024, 0, s>>>(bv.nnz, km.units_v, totals + 0);
    k_km_total<<<1, 1024, 0, s>>>(bg.nnz, km.units_g, totals + 1);
    k_km_total<<<1, 1024, 0, s>>>(bv.cap_val, km.units_v, totals + 2);
    k_km_total<<<1, 1024, 0, s>>>(bg.cap_val, km.units_g, totals + 3);
    k_km_total<<<1, 1024, 0, s>>>(bv.cnt, km.units_v, totals + 4);
    k_km_total<<<1, 1024, 0, s>>>(bg.cnt, km.units_g, totals + 5);
    int overflow = 0;
    long long used[6] = {0, 0, 0, 0, 0, 0};
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&overflow, flag, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(used, totals, sizeof(used), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return done(set_last_error(SKM_ERROR_HIP, "kernel matrix: assembly failed"));
    if (overflow) return done(set_last_error(SKM_ERROR_HIP, "kernel matrix: a unit passed its capacity"));
    km.nnz_v = used[0];
    km.nnz_g = used[1];
    // bytes one evaluation reads from each layout: records and float4 slots
    km.read_v = used[4] * 32 + used[2] * 16;
    km.read_g = used[5] * 48 + used[3] * 16;
    km.state = 1;
    km.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    return done(SKM_SUCCESS);
}

void launch_value_km(const Problem &pr, const KernelMatrix &km, int measure, float4 *sg, float *sb,
                     double *partials, int *blocks_out, hipStream_t s)
{
    const dim3 g = pr.dim == 2 ? value_grid<2>(pr) : value_grid<3>(pr);
    *blocks_out = (int)(g.x * g.y * g.z);
#define SKM_KM_VALUE(D, M)                                                                                         \
    k_km_value<D, M><<<g, kThreads, 0, s>>>(pr, pr.cpc, km.v_ent, km.v_val, km.v_ent_off, km.v_val_off, km.v_cnt, \
                                            sg, sb, partials)
    if (pr.dim == 2) {
        if (measure == SKM_MEASURE_MSE) SKM_KM_VALUE(2, SKM_MEASURE_MSE);
        else SKM_KM_VALUE(2, SKM_MEASURE_NCC);
    } else {
        if (measure == SKM_MEASURE_MSE) SKM_KM_VALUE(3, SKM_MEASURE_MSE);
        else SKM_KM_VALUE(3, SKM_MEASURE_NCC);
    }
#undef SKM_KM_VALUE
}

void launch_gather_km(const Problem &pr_in, const KernelMatrix &km, const float4 *sv, float *deriv_soa,
                      hipStream_t s)
{
    const Problem pr = km_gather_problem(pr_in);
    if (pr.dim == 2) {
        k_km_gather<2><<<(unsigned)gather_workgroups<2>(pr), kThreads, 0, s>>>(
            pr, sv, group_map<2>(pr), km.g_ent, km.g_val, km.g_ent_off, km.g_val_off, km.g_cnt);
    } else {
        k_km_gather<3><<<(unsigned)gather_workgroups<3>(pr), kThreads, 0, s>>>(
            pr, sv, group_map<3>(pr), km.g_ent, km.g_val, km.g_ent_off, km.g_val_off, km.g_cnt);
    }
    launch_gather_finish(pr, deriv_soa, s);
}

}  // namespace skm
