// fz_tile.hpp -- per-workgroup device routines shared by the single-field kernels
// (fz_kernels.hip) and the batched kernels (fz_batch.hip).  Both kernel families call the same
// bodies, so a field compressed alone and the same field compressed inside a batch produce
// the same bytes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fz_capi.h"
#include "fz_device.hpp"

namespace fz {

// 16-byte global accesses at 4-byte alignment (the encoded blocks follow 4 W bytes of flags)
typedef uint32_t u32x4a4 __attribute__((ext_vector_type(4), aligned(4)));

__device__ __forceinline__ void st16(uint8_t* p, uint4 v)
{
    *reinterpret_cast<u32x4a4*>(p) = u32x4a4{v.x, v.y, v.z, v.w};
}
__device__ __forceinline__ uint4 ld16(const uint8_t* p)
{
    const u32x4a4 t = *reinterpret_cast<const u32x4a4*>(p);
    return make_uint4(t.x, t.y, t.z, t.w);
}

__device__ __forceinline__ uint32_t nonfinite(float v) { return (__float_as_uint(v) & 0x7f800000u) == 0x7f800000u; }

// ---------------------------------------------------------------------------- XCD-aware order

// Workgroups are placed round-robin over the 8 XCDs (each with its own L2).  xcd_tile maps
// workgroup b of n to a tile such that XCD b % 8 walks the (b % 8)-th contiguous eighth of
// the tile range: the neighbour rows a Lorenzo tile re-reads were fetched into the same L2
// moments before.  It is a bijection on [0, n); any placement gives the same results.
__device__ __forceinline__ uint32_t region_start(uint32_t x, uint32_t n)
{
    return x * (n >> 3) + min(x, n & 7u);
}

__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t n)
{
    return region_start(b & 7u, n) + (b >> 3);
}

// Largest i in [lo, hi) with a[i] <= v, for a nondecreasing a with a[lo] <= v.  Called by
// all 64 lanes of a wave with uniform arguments; returns a uniform value.  Each round narrows
// the range 64-fold with one load per lane (two rounds cover 4096 entries).
__device__ __forceinline__ uint32_t wave_upper(const uint32_t* __restrict__ a, uint32_t lo, uint32_t hi, uint32_t v)
{
    const uint32_t lane = lane_id();
    while (hi - lo > 64u) {
        const uint32_t step = (hi - lo + 63u) >> 6;
        const uint32_t i = lo + lane * step;
        const uint32_t c = (uint32_t)__popcll(ballot64(i < hi && a[i] <= v));
        lo += (c - 1u) * step;
        hi = min(hi, lo + step);
    }
    const uint32_t i = lo + lane;
    return __builtin_amdgcn_readfirstlane(lo + (uint32_t)__popcll(ballot64(i < hi && a[i] <= v)) - 1u);
}

// ---------------------------------------------------------------------------- value range

// min / max / non-finite flag of in[f_lo, f_hi), reduced over the NT threads of the
// workgroup; thread 0 stores (lo, hi, bad-as-bits, 0) to *out.  float4 loads over the
// 16-byte-aligned middle, scalar loads at the edges.
// FIN (the outlier mode): the min / max of the finite values only (a non-finite value enters
// as NaN, which fminf / fmaxf skip).
template <int NT, uint32_t U = 4, bool FIN = false>  // U float4 loads in flight per thread
__device__ __forceinline__ void range_span(const float* __restrict__ in, uint32_t f_lo, uint32_t f_hi,
                                           float4* __restrict__ out)
{
    auto fin = [](float v) { return (FIN && nonfinite(v)) ? __builtin_nanf("") : v; };
    __shared__ float    s_lo[NT / 64], s_hi[NT / 64];
    __shared__ uint32_t s_bad[NT / 64];
    const uint32_t tid = threadIdx.x;
    float    lo = __builtin_huge_valf(), hi = -__builtin_huge_valf();
    uint32_t bad = 0;
    const uint32_t v_lo = (f_lo + 3) / 4, v_hi = f_hi / 4;  // whole float4s inside [f_lo, f_hi)
    const float4*  in4 = reinterpret_cast<const float4*>(in);
    if (v_lo < v_hi) {
        for (uint32_t base = v_lo + tid; base < v_hi; base += NT * U) {
            float4 v[U];
#pragma unroll
            for (uint32_t j = 0; j < U; j++) v[j] = in4[min(base + j * NT, v_hi - 1)];  // re-reads are harmless
#pragma unroll
            for (uint32_t j = 0; j < U; j++) {
                bad |= nonfinite(v[j].x) | nonfinite(v[j].y) | nonfinite(v[j].z) | nonfinite(v[j].w);
                const float a = fin(v[j].x), b = fin(v[j].y), c = fin(v[j].z), d = fin(v[j].w);
                lo = fminf(lo, fminf(fminf(a, b), fminf(c, d)));
                hi = fmaxf(hi, fmaxf(fmaxf(a, b), fmaxf(c, d)));
            }
        }
    }
    const uint32_t e0 = min(4 * v_lo, f_hi), e1 = max(4 * v_hi, e0);
    for (uint32_t k = f_lo + tid; k < e0; k += NT) {
        lo = fminf(lo, fin(in[k])); hi = fmaxf(hi, fin(in[k])); bad |= nonfinite(in[k]);
    }
    for (uint32_t k = e1 + tid; k < f_hi; k += NT) {
        lo = fminf(lo, fin(in[k])); hi = fmaxf(hi, fin(in[k])); bad |= nonfinite(in[k]);
    }
    wave_minmax(lo, hi);
    bad = ballot64(bad != 0) != 0;
    const uint32_t wv = tid >> 6;
    if ((tid & 63) == 0) {
        s_lo[wv] = lo;
        s_hi[wv] = hi;
        s_bad[wv] = bad;
    }
    __syncthreads();
    if (tid < 64) {
        const bool live = tid < NT / 64;
        lo = live ? s_lo[tid] : __builtin_huge_valf();
        hi = live ? s_hi[tid] : -__builtin_huge_valf();
        wave_minmax(lo, hi);
        bad = ballot64(live && s_bad[tid] != 0) != 0;
        if (tid == 0) *out = make_float4(lo, hi, __uint_as_float(bad), 0.0f);
    }
}

// eb parameters of a field from its value range (fz.cu:137, 149; lorenzo_var.cuh:562-563):
//   eb_abs = eb_rel * (double)(max - min);  ebx2 = eb_abs * 2;  ebx2_r = 1 / ebx2  (double),
//   both narrowed to float.  `level` bounds every |s| = |round(x * ebx2_r)| (|fl(x r)| <=
//   fl(max|x| r) by monotone rounding) and picks the dual-quant paths a tile may take:
//     4  as 3, and min >= 0: packed path with the two-instruction rounding (pq_pos)
//     3  |s| <= 16383 / 8191 / 4095 (1D / 2D / 3D): packed 16-bit path
//     2  |s| < 2^21: 3D column form in float
//     1  |s| < 2^23: fast float paths in the reference's operation order
//     0  otherwise: guarded generic path
struct EbParams {
    float    lo, hi, r, ebx2;
    double   eb;
    uint32_t status;
    int      level;
};

__device__ __forceinline__ EbParams eb_params(float lo, float hi, uint32_t bad, double eb_rel, int ndim)
{
    EbParams e;
    e.lo = lo;
    e.hi = hi;
    const float range = hi - lo;
    e.eb = eb_rel * (double)range;
    const double e2 = e.eb * 2.0;
    const double e2r = 1.0 / e2;
    e.r = (float)e2r;
    e.ebx2 = (float)e2;
    e.status = bad ? ST_NONFINITE : (range == 0.0f ? ST_CONSTANT : ST_SUCCESS);
    const float m = fmaxf(__builtin_fabsf(lo), __builtin_fabsf(hi)) * e.r;
    e.level = m < (ndim == 1 ? 16383.0f : (ndim == 2 ? 8191.0f : 4095.0f)) ? (lo >= 0.0f ? 4 : 3)
              : m < 2097151.0f                                            ? 2
              : m < 8388608.0f                                            ? 1
                                                                          : 0;
    return e;
}

// The outlier mode's parameters: the range of the finite values (0, 0 when there are none);
// a zero range gives r = ebx2 = 0 (every quantised integer is 0 and the inverse fills vmin).
// Only a range that itself overflows is an error.  With no non-finite value and every
// |s| < 2^23 (level >= 1) nothing is kept verbatim and the fixed mode's float paths give the
// exact integer deltas wherever |d| < 32768 (1D and 2D: every partial sum is below 2^24 in
// magnitude; 3D needs level >= 2 for that), so tiles take them and redo only a tile with an
// overflowing delta on the integer path; otherwise (kLevelOutlier) every tile takes it.
constexpr int kLevelOutlier = -1;

__device__ __forceinline__ EbParams eb_params_o(float lo, float hi, uint32_t bad, double eb_rel, int ndim)
{
    if (!(lo <= hi)) lo = hi = 0.0f;
    const float range = hi - lo;
    EbParams    e = eb_params(lo, hi, 0u, eb_rel, ndim);
    if (range == 0.0f) {
        e.eb = 0.0;
        e.r = 0.0f;
        e.ebx2 = 0.0f;
        e.level = lo >= 0.0f ? 4 : 3;
    }
    e.status = (range - range == 0.0f) ? ST_SUCCESS : ST_NONFINITE;
    if (bad || e.level < (ndim == 3 ? 2 : 1)) e.level = kLevelOutlier;
    return e;
}

// Reduces the partials part[0, nparts) inside a 256-thread workgroup and returns the eb
// parameters to every thread.  Any nparts (a loop of 256 per round).
__device__ __forceinline__ EbParams reduce_partials(const float4* __restrict__ part, uint32_t nparts, double eb_rel,
                                                    int ndim)
{
    __shared__ float    s_lo[4], s_hi[4];
    __shared__ uint32_t s_bad[4];
    __shared__ EbParams s_eb;
    const uint32_t tid = threadIdx.x, wv = tid >> 6;
    float    lo = __builtin_huge_valf(), hi = -__builtin_huge_valf();
    uint32_t bad = 0;
    for (uint32_t i = tid; i < nparts; i += kWG) {
        const float4 p = part[i];
        lo = fminf(lo, p.x);
        hi = fmaxf(hi, p.y);
        bad |= __float_as_uint(p.z);
    }
    wave_minmax(lo, hi);
    bad = ballot64(bad != 0) != 0;
    if ((tid & 63) == 0) {
        s_lo[wv] = lo;
        s_hi[wv] = hi;
        s_bad[wv] = bad;
    }
    __syncthreads();
    if (tid == 0)
        s_eb = eb_params(fminf(fminf(s_lo[0], s_lo[1]), fminf(s_lo[2], s_lo[3])),
                         fmaxf(fmaxf(s_hi[0], s_hi[1]), fmaxf(s_hi[2], s_hi[3])),
                         s_bad[0] | s_bad[1] | s_bad[2] | s_bad[3], eb_rel, ndim);
    __syncthreads();
    return s_eb;
}

// Same result for nparts <= 256, reduced by wave 0 alone while the other waves go straight to
// the barrier: the form used by every tile workgroup.  Wave 0 loads the partials first
// (partials_load, 4 loads per lane), before the tile's raw loads, so that they return first
// (a wave's loads return in order) and the eb parameters are ready while the tile is still
// in flight; partials_reduce then consumes them.
struct Partials {
    float4 p[4];
};

__device__ __forceinline__ void partials_load(const float4* __restrict__ part, uint32_t nparts, Partials& P)
{
    if (threadIdx.x < 64) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t i = threadIdx.x + 64 * j;
            P.p[j] = i < nparts ? part[i] : make_float4(__builtin_huge_valf(), -__builtin_huge_valf(), 0.0f, 0.0f);
        }
    }
}

template <bool OUTL = false>
__device__ __forceinline__ EbParams partials_reduce(const Partials& P, double eb_rel, int ndim)
{
    __shared__ EbParams s_eb;
    const uint32_t tid = threadIdx.x;
    if (tid < 64) {
        const float4* p = P.p;
        float lo = fminf(fminf(p[0].x, p[1].x), fminf(p[2].x, p[3].x));
        float hi = fmaxf(fmaxf(p[0].y, p[1].y), fmaxf(p[2].y, p[3].y));
        const uint32_t bad =
            ballot64((__float_as_uint(p[0].z) | __float_as_uint(p[1].z) | __float_as_uint(p[2].z) |
                      __float_as_uint(p[3].z)) != 0) != 0;
        wave_minmax(lo, hi);
        if (tid == 0) s_eb = OUTL ? eb_params_o(lo, hi, bad, eb_rel, ndim) : eb_params(lo, hi, bad, eb_rel, ndim);
    }
    __syncthreads();
    return s_eb;
}

// The result record as the range fixes it; nblocks / payload_bytes are filled in by the scan.
__device__ __forceinline__ void write_header(fz_result* res, const EbParams& e, const Geom& g)
{
    res->status = e.status;
    res->ndim = (uint32_t)g.ndim;
    res->vmin = e.lo;
    res->vmax = e.hi;
    res->ebx2_r = e.r;
    res->ebx2 = e.ebx2;
    res->eb_abs = e.eb;
    res->nblocks = 0;
    res->payload_bytes = 0;
    res->tail_len = g.mode == 0 ? (uint64_t)(g.len - (uint32_t)kTileCodes * g.T) : 0;
}

// ---------------------------------------------------------------------------- compress tiles

// Which dual-quant path a tile takes (uniform over the workgroup): 2 = fast with aligned
// float4 neighbour windows (1D, or X % 4 == 0), 3 = fast with 4-byte-aligned neighbour
// windows and per-position block edges (X >= 8), 1 = unguarded generic, 0 = guarded generic.
// geom_path ignores the value range; paths 1-3 also need every |s| < 2^23 (`safe`).
template <int NDIM>
__device__ __forceinline__ int geom_path(const Geom& g, uint32_t t0)
{
    if (t0 + (uint32_t)kTileCodes > g.xyz) return 0;
    if constexpr (NDIM == 1) return 2;
    const uint32_t halo = (NDIM == 2) ? g.X + 1 : g.XY + g.X + 1;
    if (t0 < halo) return 0;
    return (g.X & 3u) == 0 ? 2 : (g.X >= 8 ? 3 : 1);
}

// issue the raw loads of a fast-path tile (gp 2 or 3): the thread at offset lo of the tile
// starting at t0 (uniform across the wave).  One load sequence for both layouts: the
// neighbour rows use 4-byte-aligned dwordx4 loads, the same instruction an aligned row gets,
// so gp 2 and gp 3 need no separate branches (whose merge would wait on the loads at once).
template <int NDIM>
__device__ __forceinline__ void raw_prefetch(const float* __restrict__ in, uint32_t t0, uint32_t lo, const Geom& g,
                                             int gp, RawTile<NDIM>& R)
{
    if (gp < 2) return;
    if constexpr (NDIM == 1) raw_load_t<1>(in, t0, lo, g, R);
    else raw_load_ut<NDIM>(in, t0, lo, g, R);
}

template <int NDIM>
__device__ __forceinline__ int tile_path(const Geom& g, uint32_t t0, bool safe)
{
    return safe ? geom_path<NDIM>(g, t0) : 0;
}

template <int NDIM>
__device__ __forceinline__ void dq8_tile(const float* __restrict__ in, uint32_t p0, const Geom& g, float r, int path,
                                         uint32_t w[4], uint32_t& ovf)
{
    if (path == 2) {
        if constexpr (NDIM == 1) dq8_fast1(in, p0, g, r, w, ovf);
        else dq8_a4<NDIM>(in, p0, g, r, w, ovf);
    } else if (path == 3) {
        if constexpr (NDIM == 2) dq8_u(in, p0, g, r, w, ovf);
        else if constexpr (NDIM == 3) dq8<3, false>(in, p0, g, r, w, ovf);
        else dq8_fast1(in, p0, g, r, w, ovf);  // not reached: 1D never takes path 3
    } else if (path == 1) {
        dq8<NDIM, false>(in, p0, g, r, w, ovf);
    } else {
        dq8<NDIM, true>(in, p0, g, r, w, ovf);
    }
}

// The unit after the last full tile: codes of the len - 2048 T positions no tile covers,
// returned in `tail` in FIXED mode (the reference never encodes them, fz.cu:173), and the
// bitflag words of the partial tile cleared (fz.cu:167 memsets them).
template <int NDIM>
__device__ __forceinline__ void tail_unit(const float* __restrict__ in, const Geom& g, float r,
                                          uint32_t* __restrict__ bwords, uint16_t* __restrict__ tail,
                                          unsigned long long* ovf_out)
{
    const uint32_t tid = threadIdx.x;
    const uint32_t t0 = (uint32_t)kTileCodes * g.T;
    const uint32_t tail_len = g.len - t0;
    uint32_t       ovf = 0;
    if (8 * tid < tail_len) {
        uint32_t w[4];
        dq8<NDIM>(in, t0 + 8 * tid, g, r, w, ovf);
        if (g.mode == 0 && tail) {
#pragma unroll
            for (int i = 0; i < 8; i++)
                if (8 * tid + i < tail_len) tail[8 * tid + i] = (uint16_t)(w[i >> 1] >> (16 * (i & 1)));
        }
    }
    for (uint32_t i = 8 * g.T + tid; i < g.W; i += kWG) bwords[i] = 0;
    ovf = wave_sum(ovf);
    if ((tid & 63) == 0 && ovf) atomicAdd(ovf_out, (unsigned long long)ovf);
}

// The 3D generic path of a fast-geometry tile (gp 2 or 3, whose raw loads R holds): level 0,
// and level 1 on the any-X layout.
__device__ __forceinline__ void generic3_raw(uint32_t p0, const Geom& g, float r, const RawTile<3>& R, uint32_t w[4],
                                             uint32_t& ovf)
{
    w[0] = w[1] = w[2] = w[3] = 0;
    if (p0 >= g.xyz) return;
    G3 q;
    g3_from_raw(R, r, q);
    g3_codes(q, p0, g, w, ovf);
}

template <int NDIM, bool AL>
__device__ __forceinline__ void pk_dispatch(const RawTile<NDIM>& R, uint32_t p0, const Geom& g, float r, int level,
                                            uint32_t w[4])
{
    if (level == 4) raw_compute_pk<NDIM, AL, true>(R, p0, g, r, w);
    else raw_compute_pk<NDIM, AL, false>(R, p0, g, r, w);
}

// Steps 2-4 of a tile (bit transpose, flags, in-tile compaction) on the thread's 4 code
// words w, whose overflow count is ovf.  With ocnt, thread 0 also stores the tile's
// (overflows, verbatim values) totals there and adds them to its run's pair at orun: the
// per-unit counts of the outlier mode.
__device__ __forceinline__ void tile_finish(uint32_t w[4], uint32_t ovf, uint32_t tile,
                                            uint32_t* __restrict__ bwords, uint8_t* __restrict__ slot,
                                            uint32_t* __restrict__ count_out, unsigned long long* ovf_out,
                                            uint32_t* run_sum, uint32_t nraw = 0, uint32_t* ocnt = nullptr,
                                            uint32_t* orun = nullptr)
{
    __shared__ uint32_t sh[32 * kLdsPitch];
    __shared__ uint32_t sh_wcnt[4], sh_wovf[4], sh_wraw[4];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;

    // 2. bit transpose of group k = t/8 (skipped for an all-zero wave); register q then holds
    //    o[32(4L+q) + k]
    const uint32_t L = tid & 7, k = tid >> 3;
    if (ballot64((w[0] | w[1] | w[2] | w[3]) != 0) != 0) transpose4(w, L);
#pragma unroll
    for (int q = 0; q < 4; q++) sh[kLdsPitch * (4 * L + q) + k] = w[q];
    if (ballot64(ovf != 0) != 0) ovf = wave_sum(ovf);  // overflows are rare: skip the sum otherwise
    if (lane == 0) sh_wovf[wv] = ovf;
    if (ocnt) {
        if (ballot64(nraw != 0) != 0) nraw = wave_sum(nraw);
        if (lane == 0) sh_wraw[wv] = nraw;
    }
    __syncthreads();

    // 3. thread i owns 16-byte block i = o[4i .. 4i+3]; flag = block != 0 (fz.cu:84-91)
    const uint32_t c = tid >> 3, k0 = 4 * (tid & 7);
    const uint4    blk = make_uint4(sh[kLdsPitch * c + k0], sh[kLdsPitch * c + k0 + 1], sh[kLdsPitch * c + k0 + 2],
                                    sh[kLdsPitch * c + k0 + 3]);
    const bool     f = (blk.x | blk.y | blk.z | blk.w) != 0;
    const uint64_t mask = ballot64(f);
    // bitflag word 8*tile + y holds flags 32y..32y+31: wave wv covers words 2wv (lanes 0-31)
    // and 2wv+1 (lanes 32-63)
    if (lane < 2) bwords[8 * tile + 2 * wv + lane] = lane ? (uint32_t)(mask >> 32) : (uint32_t)mask;
    const uint32_t rank = rank_below(mask);
    if (lane == 0) sh_wcnt[wv] = (uint32_t)__popcll(mask);
    __syncthreads();
    const uint32_t c0 = sh_wcnt[0], c1 = sh_wcnt[1], c2 = sh_wcnt[2], c3 = sh_wcnt[3];
    const uint32_t woff = (wv > 0 ? c0 : 0u) + (wv > 1 ? c1 : 0u) + (wv > 2 ? c2 : 0u);
    if (tid == 0) {
        const uint32_t total = c0 + c1 + c2 + c3;
        *count_out = total;
        if (run_sum && total) atomicAdd(run_sum, total);
        const uint32_t o = sh_wovf[0] + sh_wovf[1] + sh_wovf[2] + sh_wovf[3];
        if (o) atomicAdd(ovf_out, (unsigned long long)o);
        if (ocnt) {
            const uint32_t rw = sh_wraw[0] + sh_wraw[1] + sh_wraw[2] + sh_wraw[3];
            *reinterpret_cast<uint2*>(ocnt) = make_uint2(o, rw);
            if (o) atomicAdd(orun, o);  // the run's (overflow, verbatim) totals
            if (rw) atomicAdd(orun + 1, rw);
        }
    }
    // 4. in-tile compaction into the tile's staging slot
    if (f) *reinterpret_cast<uint4*>(slot + 16 * (woff + rank)) = blk;
}

// Step 1 of a tile: dual-quant of the thread's 8 codes into its 4 words w (overflows added
// to ovf), by the path the tile's geometry (gp) and the field's level allow.
template <int NDIM, int PATHS = 0>
__device__ __forceinline__ void tile_dualquant(const float* __restrict__ in, const Geom& g, uint32_t tile, int gp,
                                               const RawTile<NDIM>& R, float r, int level, uint32_t w[4],
                                               uint32_t& ovf)
{
    const uint32_t p0 = tile * (uint32_t)kTileCodes + 8 * threadIdx.x;
    if constexpr (NDIM == 1) {
        if (gp == 2 && level >= 3) pk_dispatch<1, true>(R, p0, g, r, level, w);
        else if (gp == 2 && level >= 1) raw_compute<1>(R, p0, g, r, w, ovf);
        else dq8_tile<1>(in, p0, g, r, level >= 1 ? gp : 0, w, ovf);
    } else if constexpr (NDIM == 2) {
        if (gp == 2 && level >= 3) pk_dispatch<2, true>(R, p0, g, r, level, w);
        else if (gp == 3 && level >= 3) pk_dispatch<2, false>(R, p0, g, r, level, w);
        else if (gp == 2 && level >= 1) raw_compute<2>(R, p0, g, r, w, ovf);
        else if (gp == 3 && level >= 1) raw_compute_u(R, p0, g, r, w, ovf);
        else dq8_tile<2>(in, p0, g, r, level >= 1 ? gp : 0, w, ovf);
    } else if constexpr (PATHS == 2) {
        dq8_tile<3>(in, p0, g, r, level >= 1 ? gp : 0, w, ovf);
    } else {
        if (gp == 2 && level >= 3) pk_dispatch<3, true>(R, p0, g, r, level, w);
        else if (gp == 3 && level >= 3) pk_dispatch<3, false>(R, p0, g, r, level, w);
        else if (gp == 2 && level >= 2) raw_compute_q3<true>(R, p0, g, r, w, ovf);
        else if (gp == 3 && level >= 2) raw_compute_q3<false>(R, p0, g, r, w, ovf);
        else if (gp == 2 && level >= 1) raw_compute<3>(R, p0, g, r, w, ovf);
        else if (gp >= 2) generic3_raw(p0, g, r, R, w, ovf);
        else dq8_tile<3>(in, p0, g, r, level >= 1 ? gp : 0, w, ovf);  // PATHS == 0 only
    }
}

// One full tile: dual-quant -> 32x32 bit transpose -> block flags -> in-tile compaction.
//   bwords[8 tile .. 8 tile + 8)  the tile's bitflag words (fz.cu:95-99, LSB first)
//   slot[0, 16 count)              its nonzero 16-byte blocks in order (fz.cu:56-57 order)
//   *count_out                     count
//   *run_sum                       += count (when run_sum is not null; relaxed atomic, read
//                                  only after the kernel ends)
// `R` holds the tile's raw loads when gp is 2 or 3 (raw_prefetch, issued by the caller before
// its other latency-bound work).  PATHS (3D): 1 compiles only the paths of gp 2 / 3 tiles, 2
// only the generic path with its own loads (gp 0 / 1); a kernel that branches between the two
// at its top keeps the register count of the larger one instead of their union (69 against
// 87 VGPRs: 7 waves per SIMD instead of 5).
template <int NDIM, int PATHS = 0>
__device__ __forceinline__ void tile_encode(const float* __restrict__ in, const Geom& g, uint32_t tile, int gp,
                                            const RawTile<NDIM>& R, float r, int level,
                                            uint32_t* __restrict__ bwords, uint8_t* __restrict__ slot,
                                            uint32_t* __restrict__ count_out, unsigned long long* ovf_out,
                                            uint32_t* run_sum = nullptr)
{
    // 1. dual-quant: words 4t..4t+3 of the tile
    uint32_t w[4];
    uint32_t ovf = 0;
    tile_dualquant<NDIM, PATHS>(in, g, tile, gp, R, r, level, w, ovf);
    tile_finish(w, ovf, tile, bwords, slot, count_out, ovf_out, run_sum);
}

// ---------------------------------------------------------------------------- outlier mode
//
// Integer form of the generic dual-quant (the side-channel format of the outlier mode): a
// value that is finite with |fl(x r)| < 2^24 quantises to s = round(x r) (exact in binary32,
// so the inverse's float of s is exact); any other value is kept verbatim and quantises to
// 0.  Deltas are the block-local Lorenzo deltas of these integers in exact int32 arithmetic
// (|d| < 2^27); |d| >= 32768 takes code 0 and a side-channel entry.  Positions >= x*y*z get
// code 0.

__device__ __forceinline__ int32_t quant_i(float v, float r, bool& raw)
{
    const float t = __fmul_rn(v, r);
    raw = !(__builtin_isfinite(v) && __builtin_fabsf(t) < 16777216.0f);
    return raw ? 0 : (int32_t)__builtin_truncf(__fadd_rn(t, __builtin_copysignf(0.49999997f, t)));
}

// quantised neighbour value; 0 outside [0, xyz) and for values kept verbatim
__device__ __forceinline__ int32_t ldq_i(const float* __restrict__ in, int64_t idx, uint32_t xyz, float r)
{
    bool raw;
    return (idx >= 0 && idx < (int64_t)xyz) ? quant_i(in[idx], r, raw) : 0;
}

struct Int8 {
    int32_t  d[8];  // deltas of p0 .. p0 + 7 (0 at positions >= x*y*z)
    uint32_t raw;   // bit i: position p0 + i is kept verbatim
    uint32_t big;   // bit i: |d[i]| >= 32768
};

template <int NDIM>
__device__ __forceinline__ void dq8_int(const float* __restrict__ in, uint32_t p0, const Geom& g, float r, Int8& o)
{
    o.raw = o.big = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) o.d[i] = 0;
    if (p0 >= g.xyz) return;
    int32_t s[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        bool raw = false;
        s[i] = p0 + i < g.xyz ? quant_i(in[p0 + i], r, raw) : 0;
        o.raw |= raw ? 1u << i : 0u;
    }
    const int32_t sw = ldq_i(in, (int64_t)p0 - 1, g.xyz, r);
    if constexpr (NDIM == 1) {
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int32_t left = (i == 0) ? (((p0 & 255u) != 0) ? sw : 0) : s[i - 1];
            o.d[i] = s[i] - left;
        }
    } else {
        const uint32_t q = fdiv(p0, g.divX);
        uint32_t x = p0 - q * g.X;
        uint32_t y, zz = 0;
        if constexpr (NDIM == 2) {
            y = q;
        } else {
            zz = fdiv(q, g.divY);
            y = q - zz * g.Y;
        }
        const int64_t pN = (int64_t)p0 - g.X;
        int32_t n[9];  // s[p - X] for p = p0-1 .. p0+7
#pragma unroll
        for (int j = 0; j < 9; j++) n[j] = ldq_i(in, pN - 1 + j, g.xyz, r);
        if constexpr (NDIM == 2) {
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const bool    bx = (x & 15u) != 0, by = (y & 15u) != 0;
                const int32_t west = (i == 0) ? sw : s[i - 1];
                o.d[i] = s[i] - (bx ? west : 0) - (by ? n[i + 1] : 0) + ((bx && by) ? n[i] : 0);
                if (++x == g.X) { x = 0; ++y; }
            }
        } else {
            const int64_t pU = (int64_t)p0 - g.XY, pUN = pU - g.X;
            int32_t u[9], un[9];  // s[p - XY], s[p - XY - X] for p = p0-1 .. p0+7
#pragma unroll
            for (int j = 0; j < 9; j++) {
                u[j] = ldq_i(in, pU - 1 + j, g.xyz, r);
                un[j] = ldq_i(in, pUN - 1 + j, g.xyz, r);
            }
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const bool    bx = (x & 7u) != 0, by = (y & 7u) != 0, bz = (zz & 7u) != 0;
                const int32_t west = (i == 0) ? sw : s[i - 1];
                o.d[i] = s[i] - (bx ? west : 0) - (by ? n[i + 1] : 0) - (bz ? u[i + 1] : 0) + ((bx && by) ? n[i] : 0) +
                         ((bx && bz) ? u[i] : 0) + ((by && bz) ? un[i + 1] : 0) - ((bx && by && bz) ? un[i] : 0);
                if (++x == g.X) {
                    x = 0;
                    if (++y == g.Y) { y = 0; ++zz; }
                }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 8; i++) {
        if (p0 + i >= g.xyz) o.d[i] = 0;
        o.big |= (o.d[i] >= 32768 || o.d[i] <= -32768) ? 1u << i : 0u;
    }
}

// sign-magnitude code of an exact integer delta; 0 for a side-channel delta
__device__ __forceinline__ uint32_t code_i(int32_t d)
{
    const uint32_t a = (uint32_t)(d < 0 ? -d : d);
    return a >= 32768u ? 0u : (a | (d < 0 ? 0x8000u : 0u));
}

// A full tile of the outlier mode: as tile_encode, plus the tile's (overflows, verbatim
// values) counts at ocnt[0..1].  At level >= 0 the fixed mode's paths run first and the
// integer path only redoes a tile in which some delta overflowed (uniform decision).
template <int NDIM, int PATHS = 0>
__device__ __forceinline__ void tile_encode_o(const float* __restrict__ in, const Geom& g, uint32_t tile, int gp,
                                              const RawTile<NDIM>& R, float r, int level,
                                              uint32_t* __restrict__ bwords, uint8_t* __restrict__ slot,
                                              uint32_t* __restrict__ count_out, unsigned long long* ovf_out,
                                              uint32_t* run_sum, uint32_t* __restrict__ ocnt, uint32_t* orun)
{
    uint32_t w[4];
    uint32_t ovf = 0, nraw = 0;
    bool     exact = level == kLevelOutlier;
    if (!exact) {
        tile_dualquant<NDIM, PATHS>(in, g, tile, gp, R, r, level, w, ovf);
        // levels 3-4 (packed paths) bound |d| by 2^(ndim-1) * max|s| < 32768: no delta can
        // overflow, so only the float paths need the workgroup's vote (uniform branch)
        if (level < 3) exact = __syncthreads_or(ovf != 0) != 0;
    }
    if (exact) {
        Int8 o;
        dq8_int<NDIM>(in, tile * (uint32_t)kTileCodes + 8 * threadIdx.x, g, r, o);
#pragma unroll
        for (int k = 0; k < 4; k++) w[k] = code_i(o.d[2 * k]) | (code_i(o.d[2 * k + 1]) << 16);
        ovf = (uint32_t)__popc(o.big);
        nraw = (uint32_t)__popc(o.raw);
    }
    tile_finish(w, ovf, tile, bwords, slot, count_out, ovf_out, run_sum, nraw, ocnt, orun);
}

// The tail unit of the outlier mode: the tail codes, the partial tile's bitflag words
// cleared, and the unit's counts at ocnt[0..1], added to its run's pair at orun.
template <int NDIM>
__device__ __forceinline__ void tail_unit_o(const float* __restrict__ in, const Geom& g, float r,
                                            uint32_t* __restrict__ bwords, uint16_t* __restrict__ tail,
                                            unsigned long long* ovf_out, uint32_t* __restrict__ ocnt, uint32_t* orun)
{
    __shared__ uint32_t sh_b[4], sh_r[4];
    const uint32_t tid = threadIdx.x, wv = tid >> 6;
    const uint32_t t0 = (uint32_t)kTileCodes * g.T;
    const uint32_t tail_len = g.len - t0;
    uint32_t       nb = 0, nr = 0;
    if (8 * tid < tail_len) {
        Int8 o;
        dq8_int<NDIM>(in, t0 + 8 * tid, g, r, o);
        nb = (uint32_t)__popc(o.big);
        nr = (uint32_t)__popc(o.raw);
        if (tail) {
#pragma unroll
            for (int i = 0; i < 8; i++)
                if (8 * tid + i < tail_len) tail[8 * tid + i] = (uint16_t)code_i(o.d[i]);
        }
    }
    for (uint32_t i = 8 * g.T + tid; i < g.W; i += kWG) bwords[i] = 0;
    nb = wave_sum(nb);
    nr = wave_sum(nr);
    if ((tid & 63) == 0) {
        sh_b[wv] = nb;
        sh_r[wv] = nr;
    }
    __syncthreads();
    if (tid == 0) {
        const uint32_t b = sh_b[0] + sh_b[1] + sh_b[2] + sh_b[3], rr = sh_r[0] + sh_r[1] + sh_r[2] + sh_r[3];
        *reinterpret_cast<uint2*>(ocnt) = make_uint2(b, rr);
        if (b) atomicAdd(ovf_out, (unsigned long long)b);
        if (b) atomicAdd(orun, b);
        if (rr) atomicAdd(orun + 1, rr);
    }
}

// ---------------------------------------------------------------------------- one wave per tile
//
// The same tile encoded by one wave instead of four, with no workgroup barrier.  Lane L
// computes the 8 codes at tile_start + 512 m + 8 L for the quarters m = 0..3 -- exactly the
// positions thread 64 m + L computes in tile_encode, through the same per-thread paths -- and
// issues the raw loads of quarter m + 1 before computing quarter m.  Each quarter is transposed
// in registers and written to the wave's own LDS region; lane L then owns blocks L + 64 j
// (j = 0..3): ballot j gives bitflag words 2j and 2j+1, and the ranks of the nonzero blocks come
// from the popcounts of ballots < j plus mbcnt.  `sh` is the wave's 32 x kLdsPitch words.

// dual-quant of one 8-code group with the raw loads in R (fast paths) or loading itself
template <int NDIM>
__device__ __forceinline__ void dq8_level(const float* __restrict__ in, uint32_t p0, const Geom& g, int gp,
                                          const RawTile<NDIM>& R, float r, int level, uint32_t w[4], uint32_t& ovf)
{
    if constexpr (NDIM == 1) {
        if (gp == 2 && level >= 3) pk_dispatch<1, true>(R, p0, g, r, level, w);
        else if (gp == 2 && level >= 1) raw_compute<1>(R, p0, g, r, w, ovf);
        else dq8_tile<1>(in, p0, g, r, level >= 1 ? gp : 0, w, ovf);
    } else if constexpr (NDIM == 2) {
        if (gp == 2 && level >= 3) pk_dispatch<2, true>(R, p0, g, r, level, w);
        else if (gp == 3 && level >= 3) pk_dispatch<2, false>(R, p0, g, r, level, w);
        else if (gp == 2 && level >= 1) raw_compute<2>(R, p0, g, r, w, ovf);
        else if (gp == 3 && level >= 1) raw_compute_u(R, p0, g, r, w, ovf);
        else dq8_tile<2>(in, p0, g, r, level >= 1 ? gp : 0, w, ovf);
    } else {
        if (gp == 2 && level >= 3) pk_dispatch<3, true>(R, p0, g, r, level, w);
        else if (gp == 3 && level >= 3) pk_dispatch<3, false>(R, p0, g, r, level, w);
        else if (gp == 2 && level >= 2) raw_compute_q3<true>(R, p0, g, r, w, ovf);
        else if (gp == 3 && level >= 2) raw_compute_q3<false>(R, p0, g, r, w, ovf);
        else if (gp == 2 && level >= 1) raw_compute<3>(R, p0, g, r, w, ovf);
        else if (gp >= 2) generic3_raw(p0, g, r, R, w, ovf);
        else dq8_tile<3>(in, p0, g, r, level >= 1 ? gp : 0, w, ovf);
    }
}

// LDS writes of the lanes of this wave visible to its other lanes
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// gp = geom_path of the tile; R[0] holds quarter 0's raw loads when gp >= 2 (issued by the
// caller with raw_prefetch at tile_start + 8 lane, before its other latency-bound work).
template <int NDIM>
__device__ __forceinline__ void tile_encode_w(const float* __restrict__ in, const Geom& g, uint32_t tile, int gp,
                                              RawTile<NDIM> (&R)[2], float r, int level,
                                              uint32_t* __restrict__ bwords, uint8_t* __restrict__ slot,
                                              uint32_t* __restrict__ count_out, unsigned long long* ovf_out,
                                              uint32_t* __restrict__ sh, uint32_t* run_sum = nullptr)
{
    const uint32_t lane = lane_id(), L8 = lane & 7, k8 = lane >> 3;
    const uint32_t t0 = tile * (uint32_t)kTileCodes;
    uint32_t       ovf = 0;
    // quarter m computes from `cur` while the loads of quarter m + 1 go to `nxt` (explicit
    // buffers: the indices stay constant even where the compiler does not unroll)
    auto quarter = [&](int m, const RawTile<NDIM>& cur, RawTile<NDIM>& nxt) {
        const uint32_t p0 = t0 + 512 * m + 8 * lane;
        // every gp >= 2 path reads R (the 3D generic one too), whatever the level
        if (gp >= 2 && m < 3) raw_prefetch<NDIM>(in, t0, 512 * (m + 1) + 8 * lane, g, gp, nxt);
        uint32_t w[4];
        dq8_level<NDIM>(in, p0, g, gp, cur, r, level, w, ovf);
        if (ballot64((w[0] | w[1] | w[2] | w[3]) != 0) != 0) transpose4(w, L8);
#pragma unroll
        for (int q = 0; q < 4; q++) sh[kLdsPitch * (4 * L8 + q) + 8 * m + k8] = w[q];
    };
    quarter(0, R[0], R[1]);
    quarter(1, R[1], R[0]);
    quarter(2, R[0], R[1]);
    quarter(3, R[1], R[0]);
    wave_lds_sync();
    uint32_t base = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint32_t i = lane + 64 * j, c = i >> 3, k0 = 4 * (i & 7);
        const uint4    blk = make_uint4(sh[kLdsPitch * c + k0], sh[kLdsPitch * c + k0 + 1], sh[kLdsPitch * c + k0 + 2],
                                        sh[kLdsPitch * c + k0 + 3]);
        const bool     f = (blk.x | blk.y | blk.z | blk.w) != 0;
        const uint64_t mask = ballot64(f);
        if (lane < 2) bwords[8 * tile + 2 * j + lane] = lane ? (uint32_t)(mask >> 32) : (uint32_t)mask;
        if (f) *reinterpret_cast<uint4*>(slot + 16 * (base + rank_below(mask))) = blk;
        base += (uint32_t)__popcll(mask);
    }
    if (lane == 0) {
        *count_out = base;
        if (run_sum && base) atomicAdd(run_sum, base);
    }
    if (ballot64(ovf != 0) != 0) {
        ovf = wave_sum(ovf);
        if (lane == 0) atomicAdd(ovf_out, (unsigned long long)ovf);
    }
    wave_lds_sync();  // the region is rewritten by this wave's next tile
}

// The unit after the last full tile, by one wave (see tail_unit).
template <int NDIM>
__device__ __forceinline__ void tail_unit_w(const float* __restrict__ in, const Geom& g, float r,
                                            uint32_t* __restrict__ bwords, uint16_t* __restrict__ tail,
                                            unsigned long long* ovf_out)
{
    const uint32_t lane = lane_id();
    const uint32_t t0 = (uint32_t)kTileCodes * g.T;
    const uint32_t tail_len = g.len - t0;
    uint32_t       ovf = 0;
    for (uint32_t j = lane; 8 * j < tail_len; j += 64) {
        uint32_t w[4];
        dq8<NDIM>(in, t0 + 8 * j, g, r, w, ovf);
        if (g.mode == 0 && tail) {
#pragma unroll
            for (int i = 0; i < 8; i++)
                if (8 * j + i < tail_len) tail[8 * j + i] = (uint16_t)(w[i >> 1] >> (16 * (i & 1)));
        }
    }
    for (uint32_t i = 8 * g.T + lane; i < g.W; i += 64) bwords[i] = 0;
    if (ballot64(ovf != 0) != 0) {
        ovf = wave_sum(ovf);
        if (lane == 0) atomicAdd(ovf_out, (unsigned long long)ovf);
    }
}

// ---------------------------------------------------------------------------- decompress tiles

// sign-magnitude code -> delta
__device__ __forceinline__ int32_t decode(uint32_t c)
{
    const int32_t m = (int32_t)(c & 0x7fffu);
    return (c & 0x8000u) ? -m : m;
}

// both codes of a word (low half first) -> deltas, on int16 pairs: m = -1 where the sign bit
// is set, d = (magnitude ^ m) - m per half (no half overflows: magnitudes are <= 32767); five
// or six instructions for the pair instead of about ten
__device__ __forceinline__ void decode2(uint32_t w, int32_t& lo, int32_t& hi)
{
    typedef short v2s __attribute__((ext_vector_type(2)));
    const v2s      m = __builtin_bit_cast(v2s, w) >> (v2s){15, 15};
    const v2s      x = __builtin_bit_cast(v2s, (w & 0x7FFF7FFFu) ^ __builtin_bit_cast(uint32_t, m));
    const uint32_t d = __builtin_bit_cast(uint32_t, (v2s)(x - m));
    lo = (int32_t)(int16_t)(d & 0xFFFFu);
    hi = (int32_t)d >> 16;
}

// Scatter a tile's nonzero blocks back (block i at payload + 4W + 16 (excl + rank)) and
// unshuffle, in one wave (64 threads): lane l ends with words 4 (l + 64 v) .. + 3 of the tile in
// words[v], v < 4.  Lane l owns blocks 4l .. 4l + 3
// (flag bits 4(l % 8) .. + 3 of word l / 8), puts them in LDS as rows of bit planes, then
// unshuffles the words of threads l + 64v, v < 4, of a 256-thread layout (8 lanes per 32×32 bit
// transpose).  A tile with no nonzero block gives zeros without touching LDS (wave-uniform).
__device__ __forceinline__ void decode_tile_w_regs(const uint8_t* __restrict__ payload, uint32_t W, uint32_t tile,
                                                   uint32_t excl, uint32_t words[4][4])
{
    __shared__ uint32_t sh[32 * kLdsPitch];
    const uint32_t l = threadIdx.x;
    const uint32_t word = reinterpret_cast<const uint32_t*>(payload)[8 * tile + (l >> 3)];
    const uint32_t nib = (word >> (4 * (l & 7))) & 0xFu;
    if (ballot64(nib != 0) == 0) {
#pragma unroll
        for (int v = 0; v < 4; v++)
#pragma unroll
            for (int q = 0; q < 4; q++) words[v][q] = 0;
        return;
    }
    const uint32_t cnt = __popc(nib);
    uint32_t       idx = excl + wave_incl_scan(cnt) - cnt;  // rank of block 4l among the tile's nonzero blocks
    const uint8_t* blocks = payload + 4ull * W;
    uint4          blk[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const bool f = (nib >> i) & 1u;
        blk[i] = f ? ld16(blocks + 16ull * idx) : make_uint4(0, 0, 0, 0);
        idx += f ? 1u : 0u;
    }
    // block b = 4l + i holds words 4 (b % 8) .. + 3 of bit plane b / 8 = l / 2
    uint32_t* row = sh + kLdsPitch * (l >> 1) + 16 * (l & 1);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        row[4 * i] = blk[i].x;
        row[4 * i + 1] = blk[i].y;
        row[4 * i + 2] = blk[i].z;
        row[4 * i + 3] = blk[i].w;
    }
    __syncthreads();
    const uint32_t L = l & 7;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const uint32_t k = (l >> 3) + 8 * v;
#pragma unroll
        for (int q = 0; q < 4; q++) words[v][q] = sh[kLdsPitch * (4 * L + q) + k];
        transpose4(words[v], L);
    }
}

// The tile's 1024 code words to out_words: words[v] of lane l are words 4 (l + 64 v) .. + 3.
__device__ __forceinline__ void decode_tile_w(const uint8_t* __restrict__ payload, uint32_t W, uint32_t tile,
                                              uint32_t excl, uint32_t* __restrict__ out_words)
{
    uint32_t words[4][4];
    decode_tile_w_regs(payload, W, tile, excl, words);
    uint4* const out4 = reinterpret_cast<uint4*>(out_words);
#pragma unroll
    for (int v = 0; v < 4; v++) out4[threadIdx.x + 64 * v] = make_uint4(words[v][0], words[v][1], words[v][2], words[v][3]);
}

// 1D fields: decode and reconstruct a unit in one wave, straight to floats (the 256-code
// Lorenzo blocks of 1D lie inside one 2048-code tile, lorenzo_var.cuh:47-108 inverse).  Lane l
// holds codes 8 (l + 64 v) .. + 7 of the unit for v < 4, so lanes 0-31 and 32-63 each cover
// one block: prefix in the thread, exclusive scan over the 32 lanes, x' = s * ebx2 for the
// positions below x*y*z.  A full tile decodes from the payload; the unit after the last full
// tile takes the trailer's codes (0 without one).
__device__ __forceinline__ void decode_recon_1d_w(const uint8_t* __restrict__ payload, const uint16_t* __restrict__ tail,
                                                  const Geom& g, uint32_t tile, uint32_t excl, float ebx2,
                                                  float* __restrict__ out)
{
    const uint32_t l = threadIdx.x, t0 = (uint32_t)kTileCodes * tile;
    uint32_t       words[4][4];
    if (tile < g.T) {
        decode_tile_w_regs(payload, g.W, tile, excl, words);
    } else {
#pragma unroll
        for (int v = 0; v < 4; v++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                uint32_t wv = 0;
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const uint32_t p = t0 + 8 * (l + 64 * v) + 2 * q + h;  // < 2^32: len <= 2^32 - 4096
                    const uint32_t c = (tail && p < g.len) ? tail[p - t0] : 0u;
                    wv |= c << (16 * h);
                }
                words[v][q] = wv;
            }
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
        int32_t s[8];
#pragma unroll
        for (int i = 0; i < 8; i += 2) decode2(words[v][i >> 1], s[i], s[i + 1]);
#pragma unroll
        for (int i = 1; i < 8; i++) s[i] += s[i - 1];
        int32_t t = s[7];
#pragma unroll
        for (int dd = 1; dd < 32; dd <<= 1) {
            const int32_t o = __shfl_up(t, dd, 32);
            if ((l & 31u) >= (uint32_t)dd) t += o;
        }
        const int32_t  ex = t - s[7];
        const uint32_t p0 = t0 + 8 * (l + 64 * v);
        if (p0 + 8 <= g.xyz) {
            float4* o4 = reinterpret_cast<float4*>(out + p0);
            o4[0] = make_float4((float)(ex + s[0]) * ebx2, (float)(ex + s[1]) * ebx2, (float)(ex + s[2]) * ebx2,
                                (float)(ex + s[3]) * ebx2);
            o4[1] = make_float4((float)(ex + s[4]) * ebx2, (float)(ex + s[5]) * ebx2, (float)(ex + s[6]) * ebx2,
                                (float)(ex + s[7]) * ebx2);
        } else {
#pragma unroll
            for (int i = 0; i < 8; i++)
                if (p0 + i < g.xyz) out[p0 + i] = (float)(ex + s[i]) * ebx2;
        }
    }
}

// Codes of the positions after the last full tile: the trailer, or 0 without one.
__device__ __forceinline__ void decode_tail(const uint16_t* __restrict__ tail, const Geom& g,
                                            uint16_t* __restrict__ codes)
{
    const uint32_t t0 = kTileCodes * g.T;
    for (uint32_t i = threadIdx.x; t0 + i < g.len; i += blockDim.x) codes[t0 + i] = tail ? tail[i] : (uint16_t)0;
}


// Deltas of the reconstruct kernels (the source SRC of recon_*): u16 sign-magnitude codes, or
// int32 deltas, or the outlier mode's codes with side-channel overrides.  *4 reads positions
// [p, p + 4) with one 8- or 16-byte load (p % 4 == 0), *16 reads [p, p + 16).
__device__ __forceinline__ int32_t delta1(const uint16_t* __restrict__ c, size_t p) { return decode(c[p]); }
__device__ __forceinline__ int32_t delta1(const int32_t* __restrict__ c, size_t p) { return c[p]; }
__device__ __forceinline__ void delta4(const uint16_t* __restrict__ c, size_t p, int32_t d[4])
{
    const uint2 v = *reinterpret_cast<const uint2*>(c + p);
    decode2(v.x, d[0], d[1]);
    decode2(v.y, d[2], d[3]);
}
__device__ __forceinline__ void delta4(const int32_t* __restrict__ c, size_t p, int32_t d[4])
{
    const int4 v = *reinterpret_cast<const int4*>(c + p);
    d[0] = v.x;
    d[1] = v.y;
    d[2] = v.z;
    d[3] = v.w;
}
__device__ __forceinline__ void delta16(const uint16_t* __restrict__ c, size_t p, int32_t d[16])
{
    const uint4 a = *reinterpret_cast<const uint4*>(c + p);
    const uint4 b = *reinterpret_cast<const uint4*>(c + p + 8);
    const uint32_t wd[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int j = 0; j < 8; j++) decode2(wd[j], d[2 * j], d[2 * j + 1]);
}
__device__ __forceinline__ void delta16(const int32_t* __restrict__ c, size_t p, int32_t d[16])
{
#pragma unroll
    for (int j = 0; j < 4; j++) delta4(c, p + 4 * j, d + 4 * j);
}

// The outlier mode's source: u16 codes, a bitmap of the positions whose delta comes from the
// side channel instead (their code is 0), and those deltas at their positions.
struct OutlierSrc {
    const uint16_t* __restrict__ codes;
    const uint32_t* __restrict__ bits;
    const int32_t* __restrict__ over;
};
__device__ __forceinline__ int32_t delta1(const OutlierSrc& s, size_t p)
{
    return ((s.bits[p >> 5] >> (p & 31)) & 1u) ? s.over[p] : decode(s.codes[p]);
}
__device__ __forceinline__ void delta4(const OutlierSrc& s, size_t p, int32_t d[4])
{
    delta4(s.codes, p, d);
    const uint32_t m = (s.bits[p >> 5] >> (p & 31)) & 0xFu;  // p % 4 == 0: one word
    if (m) {
#pragma unroll
        for (int j = 0; j < 4; j++)
            if ((m >> j) & 1u) d[j] = s.over[p + j];
    }
}
__device__ __forceinline__ void delta16(const OutlierSrc& s, size_t p, int32_t d[16])
{
    delta16(s.codes, p, d);
    const uint32_t m = (s.bits[p >> 5] >> (p & 31)) & 0xFFFFu;  // p % 16 == 0: one word
    if (m) {
#pragma unroll
        for (int j = 0; j < 16; j++)
            if ((m >> j) & 1u) d[j] = s.over[p + j];
    }
}

// 1D: workgroup wg covers codes [4096 wg, 4096 wg + 4096): 16 consecutive codes per thread,
// 16 threads per 256-block (block-local inclusive prefix sum), x = (float)s * ebx2.
template <typename SRC>
__device__ __forceinline__ void recon_1d(const SRC codes, float* __restrict__ out, const Geom& g, float ebx2,
                                         uint32_t wg)
{
    const uint32_t p0 = (wg * kWG + threadIdx.x) * 16;
    int32_t        s[16];
    if (p0 + 16 <= g.xyz) {
        delta16(codes, p0, s);
    } else {
#pragma unroll
        for (int j = 0; j < 16; j++) s[j] = (p0 + j < g.xyz) ? delta1(codes, p0 + j) : 0;
    }
#pragma unroll
    for (int j = 1; j < 16; j++) s[j] += s[j - 1];
    // exclusive prefix of the 16 thread totals of this 256-block (lanes 16-aligned)
    int32_t t = s[15];
#pragma unroll
    for (int d = 1; d < 16; d <<= 1) {
        const int32_t o = __shfl_up(t, d, 16);
        if ((threadIdx.x & 15) >= (uint32_t)d) t += o;
    }
    const int32_t ex = t - s[15];
    if (p0 + 16 <= g.xyz) {
#pragma unroll
        for (int j = 0; j < 4; j++)
            *reinterpret_cast<float4*>(out + p0 + 4 * j) =
                make_float4((float)(ex + s[4 * j]) * ebx2, (float)(ex + s[4 * j + 1]) * ebx2,
                            (float)(ex + s[4 * j + 2]) * ebx2, (float)(ex + s[4 * j + 3]) * ebx2);
    } else {
#pragma unroll
        for (int j = 0; j < 16; j++)
            if (p0 + j < g.xyz) out[p0 + j] = (float)(ex + s[j]) * ebx2;
    }
}

// 2D: workgroup (bx, by) = 16 rows x 1024 columns (64 Lorenzo blocks); thread = 4 adjacent
// columns, so 4 lanes cover a block's 16.  The 16 rows stream through: each row adds its
// decoded values to the running column sums (prefix over y), then a prefix over x in the
// thread and across its 4 lanes (two DPP steps) gives the row's values.  Loads are 8 bytes
// and stores 16 bytes per lane when X % 4 == 0 (stores at 4-byte alignment otherwise).
__device__ __forceinline__ int32_t quad_incl_scan(int32_t v)
{
    const uint32_t q = lane_id() & 3u;
    const int32_t  a = __builtin_amdgcn_mov_dpp(v, 0x90, 0xF, 0xF, false);  // quad_perm [0,0,1,2]
    v += q >= 1 ? a : 0;
    const int32_t b = __builtin_amdgcn_mov_dpp(v, 0x40, 0xF, 0xF, false);  // quad_perm [0,0,0,1]
    return v + (q >= 2 ? b : 0);
}

template <typename SRC>
__device__ __forceinline__ void recon_2d(const SRC codes, float* __restrict__ out, const Geom& g, float ebx2,
                                         uint32_t bx, uint32_t by)
{
    const uint32_t x0 = bx * kRecon2DCols + 4 * threadIdx.x, y0 = by * 16;
    const bool     full = x0 + 4 <= g.X;
    const bool     vec = full && (g.X & 3u) == 0;
    int32_t        col[4] = {0, 0, 0, 0};
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const uint32_t y = y0 + r;
        if (y >= g.Y) break;  // uniform
        const size_t row = (size_t)y * g.X;
        int32_t      dd[4];
        if (vec) {
            delta4(codes, row + x0, dd);
        } else {
#pragma unroll
            for (int j = 0; j < 4; j++) dd[j] = x0 + j < g.X ? delta1(codes, row + x0 + j) : 0;
        }
        int32_t p[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            col[j] += dd[j];
            p[j] = col[j] + (j ? p[j - 1] : 0);
        }
        const int32_t excl = quad_incl_scan(p[3]) - p[3];
        float o[4];
#pragma unroll
        for (int j = 0; j < 4; j++) o[j] = (float)(p[j] + excl) * ebx2;
        if (full) {
            typedef float f32x4a4_t __attribute__((ext_vector_type(4), aligned(4)));
            *reinterpret_cast<f32x4a4_t*>(out + row + x0) = f32x4a4_t{o[0], o[1], o[2], o[3]};
        } else {
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (x0 + j < g.X) out[row + x0 + j] = o[j];
        }
    }
}

// 3D: workgroup (bx, by, bz) = 4Q columns (x) x 256/Q rows (y) x 8 planes (z), Q in {32, 16,
// 8, 4} quads per row (narrow fields take a small Q so that fewer lanes sit past the row's
// end; the y extent is then several 8-row Lorenzo blocks).  Phase 1: thread (xq, yy) = 4
// adjacent columns of row y0 + yy over the 8 z of the strip; prefix over z in registers, over
// x in the thread and with the other half of its 8-wide Lorenzo block (the neighbour lane,
// one DPP step).  Phase 2, after an LDS transpose (16-byte accesses): thread (xq, tz, yb) runs
// the prefix over the 8 rows of y block yb in plane tz.  Loads are 8 bytes and stores 16
// bytes per lane when X % 4 == 0.
template <int Q, typename SRC>
__device__ __forceinline__ void recon_3d(const SRC codes, float* __restrict__ out, const Geom& g, float ebx2,
                                         uint32_t bx, uint32_t by, uint32_t bz, int4* sh)  // sh: 8 * kWG int4
{
    static_assert(Q == 32 || Q == 16 || Q == 8 || Q == 4, "quads per row");
    constexpr uint32_t R = kWG / Q;  // rows per workgroup (a multiple of 8)
    auto at = [&](uint32_t z, uint32_t y, uint32_t x) -> int4& { return sh[(z * R + y) * Q + x]; };  // [z][y][x quad]
    const uint32_t xq = threadIdx.x % Q, yy = threadIdx.x / Q;
    const uint32_t x0 = bx * (4 * Q) + 4 * xq, y0 = by * R, z0 = bz * 8;
    const bool     full = x0 + 4 <= g.X;
    const bool     vec = full && (g.X & 3u) == 0;
    {
        const uint32_t y = y0 + yy;
        int32_t        col[4] = {0, 0, 0, 0};
#pragma unroll
        for (int z = 0; z < 8; z++) {
            int32_t dd[4] = {0, 0, 0, 0};
            if (y < g.Y && z0 + z < g.Z) {
                const size_t row = (size_t)(z0 + z) * g.XY + (size_t)y * g.X;
                if (vec) {
                    delta4(codes, row + x0, dd);
                } else {
#pragma unroll
                    for (int j = 0; j < 4; j++) dd[j] = x0 + j < g.X ? delta1(codes, row + x0 + j) : 0;
                }
            }
            int32_t p[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                col[j] += dd[j];                     // prefix over z
                p[j] = col[j] + (j ? p[j - 1] : 0);  // prefix over the thread's 4 columns
            }
            // the odd lane of a pair holds the block's upper 4 columns: add the even lane's total
            // (xq and the lane index have the same parity, Q being even)
            const int32_t lower = __builtin_amdgcn_mov_dpp(p[3], 0xA0, 0xF, 0xF, false);  // quad_perm [0,0,2,2]
            const int32_t add = (xq & 1u) ? lower : 0;
            at(z, yy, xq) = make_int4(p[0] + add, p[1] + add, p[2] + add, p[3] + add);
        }
    }
    __syncthreads();
    const uint32_t rest = threadIdx.x / Q, tz = rest & 7u, yb = rest >> 3;
    if (z0 + tz >= g.Z) return;
    int4 acc = make_int4(0, 0, 0, 0);
#pragma unroll
    for (int y = 0; y < 8; y++) {
        const uint32_t yr = yb * 8 + y;
        if (y0 + yr >= g.Y) break;  // uniform over the lanes of one (tz, yb)
        const int4 v = at(tz, yr, xq);
        acc.x += v.x;
        acc.y += v.y;
        acc.z += v.z;
        acc.w += v.w;
        const size_t row = (size_t)(z0 + tz) * g.XY + (size_t)(y0 + yr) * g.X;
        const float  o[4] = {(float)acc.x * ebx2, (float)acc.y * ebx2, (float)acc.z * ebx2, (float)acc.w * ebx2};
        if (full) {
            typedef float f32x4a4_t __attribute__((ext_vector_type(4), aligned(4)));
            *reinterpret_cast<f32x4a4_t*>(out + row + x0) = f32x4a4_t{o[0], o[1], o[2], o[3]};
        } else {
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (x0 + j < g.X) out[row + x0 + j] = o[j];
        }
    }
}

// recon_3d of workgroup `wg` of a field whose grid is rx x ry x (Z / 8), for the field's Q
template <typename SRC>
__device__ __forceinline__ void recon_3d_any(int q, const SRC codes, float* __restrict__ out, const Geom& g,
                                             float ebx2, uint32_t wg, uint32_t rx, uint32_t ry)
{
    __shared__ int4 sh[8 * kWG];  // one 32 KB array for every Q (a static array per variant would add up)
    const uint32_t  bx = wg % rx, t = wg / rx, by = t % ry, bz = t / ry;
    switch (q) {
    case 4: recon_3d<4>(codes, out, g, ebx2, bx, by, bz, sh); break;
    case 8: recon_3d<8>(codes, out, g, ebx2, bx, by, bz, sh); break;
    case 16: recon_3d<16>(codes, out, g, ebx2, bx, by, bz, sh); break;
    default: recon_3d<32>(codes, out, g, ebx2, bx, by, bz, sh); break;
    }
}

}  // namespace fz
