// fz_device.hpp -- device building blocks of the gfx950 FZ kernels.
//
// Work decomposition used throughout: a "tile" is 2048 consecutive codes (1024 u32 words,
// the unit of the reference's bitshuffle, fz.cu:73).  A 256-thread workgroup (4 wave64s)
// owns one tile; thread t computes the 8 consecutive codes [8t, 8t+8) of it, i.e. the
// words 4t..4t+3, so the 32 words of transpose group k = t/8 sit in the 8 lanes
// 8k..8k+7 with 4 rows per lane.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fz {

constexpr int kWG = 256;           // threads per workgroup
constexpr uint32_t kRecon2DCols = 4 * kWG;  // columns per 2D reconstruct workgroup (4 per thread)
constexpr int kTileCodes = 2048;   // codes per bitshuffle tile
constexpr int kTileWords = 1024;   // u32 words per tile
constexpr int kTileBlocks = 256;   // 16-byte blocks per tile
constexpr int kLdsPitch = 33;      // padded row pitch of the transposed tile in LDS

enum : uint32_t { ST_SUCCESS = 0, ST_CONSTANT = 4, ST_NONFINITE = 5, ST_CORRUPT = 6 };

// floor(n / d) for n, d < 2^32 via m = ceil(2^64 / d): exact because n * (m - 2^64/d) < 2^64 / d.
struct FastDiv {
    uint32_t d;
    uint32_t pad;
    uint64_t m;
};

inline FastDiv make_fastdiv(uint32_t d)
{
    FastDiv f;
    f.d = d;
    f.pad = 0;
    f.m = (d <= 1) ? 0 : (~0ull) / d + 1;
    return f;
}

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv& f)
{
    if (f.d == 1) return n;
    const uint64_t lo = (uint64_t)n * (uint32_t)f.m;
    const uint64_t hi = (uint64_t)n * (uint32_t)(f.m >> 32);
    return (uint32_t)((hi + (lo >> 32)) >> 32);
}

// Field geometry + mode, passed by value to every kernel.
struct Geom {
    uint32_t len;      // floats in the input (>= xyz)
    uint32_t X, Y, Z;  // X fastest
    uint32_t xyz;      // X*Y*Z
    uint32_t XY;       // X*Y
    uint32_t T;        // full tiles
    uint32_t W;        // bitflag words
    FastDiv  divX, divY;
    int32_t  ndim;
    int32_t  mode;     // 0 fixed, 1 literal
};

// ---------------------------------------------------------------- wave helpers

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t ballot64(bool p) { return __ballot(p); }

__device__ __forceinline__ uint32_t rank_below(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

__device__ __forceinline__ float wave_min(float v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fminf(v, __shfl_xor(v, d, 64));
    return v;
}

__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fmaxf(v, __shfl_xor(v, d, 64));
    return v;
}

// min and max over the 64 lanes, both returned to every lane, through DPP (VALU-only: row
// shifts within 16-lane rows, then the row broadcasts of lanes 15 and 31, then lane 63).
// A lane whose DPP source does not exist takes the identity (+inf for min, -inf for max).
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_or(float identity, float v)
{
    return __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(identity), __float_as_int(v), CTRL, ROWS, 0xF, false));
}

__device__ __forceinline__ void wave_minmax(float& lo, float& hi)
{
    constexpr float PI = __builtin_huge_valf(), NI = -__builtin_huge_valf();
    lo = fminf(lo, dpp_or<0x111, 0xF>(PI, lo));  // row_shr:1
    hi = fmaxf(hi, dpp_or<0x111, 0xF>(NI, hi));
    lo = fminf(lo, dpp_or<0x112, 0xF>(PI, lo));  // row_shr:2
    hi = fmaxf(hi, dpp_or<0x112, 0xF>(NI, hi));
    lo = fminf(lo, dpp_or<0x114, 0xF>(PI, lo));  // row_shr:4
    hi = fmaxf(hi, dpp_or<0x114, 0xF>(NI, hi));
    lo = fminf(lo, dpp_or<0x118, 0xF>(PI, lo));  // row_shr:8: lane 15 of a row holds the row's result
    hi = fmaxf(hi, dpp_or<0x118, 0xF>(NI, hi));
    lo = fminf(lo, dpp_or<0x142, 0xA>(PI, lo));  // row_bcast:15 into rows 1 and 3
    hi = fmaxf(hi, dpp_or<0x142, 0xA>(NI, hi));
    lo = fminf(lo, dpp_or<0x143, 0xC>(PI, lo));  // row_bcast:31 into rows 2 and 3: lane 63 holds all
    hi = fmaxf(hi, dpp_or<0x143, 0xC>(NI, hi));
    lo = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(lo), 63));
    hi = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(hi), 63));
}

// inclusive scan over the 64 lanes
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    const uint32_t l = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(v, d, 64);
        if (l >= (uint32_t)d) v += o;
    }
    return v;
}

// s = roundf(x * r): round-half-away-from-zero of one binary32 product.  For every non-NaN
// float t, truncf(t + copysignf(0.5f - 2^-25, t)) == roundf(t) (checked exhaustively over
// all 2^32 bit patterns); __fmul_rn keeps the product from being contracted into an FMA.
__device__ __forceinline__ float prequant(float x, float r)
{
    const float t = __fmul_rn(x, r);
    return __builtin_truncf(__fadd_rn(t, __builtin_copysignf(0.49999997f, t)));
}

// Sign-magnitude u16 code of an integer-valued float delta (lorenzo_var.cuh:169-171):
// u16(|d|) | (d < 0) << 15.  |d| >= 65535 saturates, NaN gives 0.
__device__ __forceinline__ uint32_t code_of(float d, bool with_sign)
{
    const float a = __builtin_fabsf(d);
    const float c = (a >= 65535.0f) ? 65535.0f : a;
    const uint32_t m = (c >= 0.0f) ? (uint32_t)c : 0u;
    return m | ((with_sign && d < 0.0f) ? 0x8000u : 0u);
}

__device__ __forceinline__ uint32_t is_ovf(float d) { return !(__builtin_fabsf(d) < 32768.0f); }

// ---------------------------------------------------------------- dual-quant, 8 codes
//
// Codes of the 8 positions p0..p0+7 (p0 % 8 == 0) with the float operations of the
// reference kernels in their order:
//   1D (lorenzo_var.cuh:47-66):  d = s - (p % 256 ? s[p-1] : 0); sign only at p % 4 == 0
//                                 in LITERAL mode.
//   2D (:131-145): c1 = s - N, w1 = W - NW, d = x%16 ? c1 - w1 : c1, neighbours outside the
//                  16x16 block are 0.
//   3D (:271-283): d = s - ((((((NW3 - A) - B) - C) + D) + E) + F) in 8x8x8 blocks.
// Positions >= xyz get code 0 (the reference zero-fills quant, fz.cu:145).
// Returns the 4 packed words (code[2i] | code[2i+1] << 16) and adds overflows to *ovf.

// prequantised neighbour value; with GUARD, 0 outside [0, len)
template <bool GUARD>
__device__ __forceinline__ float ldq_t(const float* __restrict__ in, int64_t idx, uint32_t len, float r)
{
    if constexpr (GUARD) return (idx >= 0 && idx < (int64_t)len) ? prequant(in[idx], r) : 0.0f;
    else return prequant(in[idx], r);
}

// Generic path: any geometry.  GUARD = false is only used on tiles whose every neighbour
// index lies inside the field (checked per workgroup by the caller).
template <int NDIM, bool GUARD = true>
__device__ __forceinline__ void dq8(const float* __restrict__ in, uint32_t p0, const Geom& g, float r,
                                    uint32_t w[4], uint32_t& ovf)
{
    w[0] = w[1] = w[2] = w[3] = 0;
    if (p0 >= g.xyz) return;
    float s[8];
    const bool full = !GUARD || (p0 + 8 <= g.len);
    if (full) {
        const float4 a = *reinterpret_cast<const float4*>(in + p0);
        const float4 b = *reinterpret_cast<const float4*>(in + p0 + 4);
        s[0] = prequant(a.x, r); s[1] = prequant(a.y, r); s[2] = prequant(a.z, r); s[3] = prequant(a.w, r);
        s[4] = prequant(b.x, r); s[5] = prequant(b.y, r); s[6] = prequant(b.z, r); s[7] = prequant(b.w, r);
    } else {
#pragma unroll
        for (int i = 0; i < 8; i++) s[i] = (p0 + i < g.len) ? prequant(in[p0 + i], r) : 0.0f;
    }
    const float sw = ldq_t<GUARD>(in, (int64_t)p0 - 1, g.len, r);
    uint32_t c[8];

    if constexpr (NDIM == 1) {
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const float left = (i == 0) ? (((p0 & 255u) != 0) ? sw : 0.0f) : s[i - 1];
            const float d = s[i] - left;
            const bool sign = (g.mode == 0) || ((i & 3) == 0);
            const bool live = p0 + i < g.xyz;
            c[i] = live ? code_of(d, sign) : 0u;
            ovf += live ? is_ovf(d) : 0u;
        }
    } else {
        // coordinates of p0, then stepped one position at a time
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
        float n[9];  // s[p - X] for p = p0-1 .. p0+7
#pragma unroll
        for (int j = 0; j < 9; j++) n[j] = ldq_t<GUARD>(in, pN - 1 + j, g.len, r);
        if constexpr (NDIM == 2) {
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const bool bx = (x & 15u) != 0, by = (y & 15u) != 0;
                const float N = by ? n[i + 1] : 0.0f;
                const float NW = (by && bx) ? n[i] : 0.0f;
                const float Wv = bx ? ((i == 0) ? sw : s[i - 1]) : 0.0f;
                const float c1 = s[i] - N;
                const float w1 = Wv - NW;
                const float d = bx ? c1 - w1 : c1;
                const bool live = p0 + i < g.xyz;
                c[i] = live ? code_of(d, true) : 0u;
                ovf += live ? is_ovf(d) : 0u;
                if (++x == g.X) { x = 0; ++y; }
            }
        } else {
            const int64_t pU = (int64_t)p0 - g.XY, pUN = pU - g.X;
            float u[9], un[9];  // s[p - XY], s[p - XY - X] for p = p0-1 .. p0+7
#pragma unroll
            for (int j = 0; j < 9; j++) {
                u[j] = ldq_t<GUARD>(in, pU - 1 + j, g.len, r);
                un[j] = ldq_t<GUARD>(in, pUN - 1 + j, g.len, r);
            }
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const bool bx = (x & 7u) != 0, by = (y & 7u) != 0, bz = (zz & 7u) != 0;
                const float west = (i == 0) ? sw : s[i - 1];
                float p = (bz && by && bx) ? un[i] : 0.0f;  // NW3 = s[z-1][y-1][x-1]
                p = p - ((by && bx) ? n[i] : 0.0f);           // A   = s[z][y-1][x-1]
                p = p - ((bz && bx) ? u[i] : 0.0f);           // B   = s[z-1][y][x-1]
                p = p - ((bz && by) ? un[i + 1] : 0.0f);      // C   = s[z-1][y-1][x]
                p = p + (bx ? west : 0.0f);                   // D   = s[z][y][x-1]
                p = p + (by ? n[i + 1] : 0.0f);               // E   = s[z][y-1][x]
                p = p + (bz ? u[i + 1] : 0.0f);               // F   = s[z-1][y][x]
                const float d = s[i] - p;
                const bool live = p0 + i < g.xyz;
                c[i] = live ? code_of(d, true) : 0u;
                ovf += live ? is_ovf(d) : 0u;
                if (++x == g.X) {
                    x = 0;
                    if (++y == g.Y) { y = 0; ++zz; }
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) w[k] = c[2 * k] | (c[2 * k + 1] << 16);
}

// ---------------------------------------------------------------- dual-quant fast paths
//
// Used on full tiles away from the field's start when every |s| < 2^23 (checked once from the
// value range), so deltas are finite integers < 2^27 and the conversions below are exact.
// Same float operations as dq8; neighbour values outside the Lorenzo block are zero *values*,
// which gives bit-identical deltas (x - (+0) == x for every non-NaN x).

__device__ __forceinline__ uint32_t code_fast(float d, bool with_sign, uint32_t& ovf)
{
    uint32_t m = (uint32_t)__builtin_fabsf(d);
    m = m < 65535u ? m : 65535u;
    ovf += m >> 15;
    return m | ((with_sign && d < 0.0f) ? 0x8000u : 0u);
}

__device__ __forceinline__ void load8(const float* __restrict__ p, float v[8])
{
    const float4 a = *reinterpret_cast<const float4*>(p);
    const float4 b = *reinterpret_cast<const float4*>(p + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}

__device__ __forceinline__ uint32_t bits_or(const float v[8])
{
    uint32_t o = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) o |= __float_as_uint(v[i]);
    return o & 0x7fffffffu;
}

__device__ __forceinline__ void pack8(const uint32_t c[8], uint32_t w[4])
{
#pragma unroll
    for (int k = 0; k < 4; k++) w[k] = c[2 * k] | (c[2 * k + 1] << 16);
}

// Raw input of one thread's 8 positions and the neighbour windows the fast paths need,
// loaded separately from the arithmetic so a workgroup can prefetch its next tile.
template <int NDIM>
struct RawTile;
template <>
struct RawTile<1> {
    float v[8];
    float vw;
};
template <>
struct RawTile<2> {
    float v[8], vn[8];
    float vw, vnw;
};
template <>
struct RawTile<3> {
    float v[8], vn[8], vu[8], vun[8];
    float vw, vnw, vuw, vunw;
};

// v of the previous lane of the wave (DPP wave_shr:1); lane 0 gets `first`
__device__ __forceinline__ float from_prev_lane(float v, float first)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(first), __float_as_int(v), 0x138, 0xF, 0xF, false));
}

// Raw loads of the thread at offset lo (a multiple of 8, < 2048) of the tile that starts at
// t0 (uniform across the caller's wave).  The row bases are uniform, so each load is a scalar
// base, one 32-bit lane offset and a constant.  Requires the fast-path geometry (tile_path == 2
// ignoring the value range).
template <int NDIM>
__device__ __forceinline__ void raw_load_t(const float* __restrict__ in, uint32_t t0, uint32_t lo, const Geom& g,
                                           RawTile<NDIM>& R)
{
    const float* own = in + t0;
    load8(own + lo, R.v);
    if constexpr (NDIM == 1) {
        // the west value is the previous lane's last one; lane 0 of a wave starts a 256-block
        // (its offset is a multiple of 512), so its west value is never used
        R.vw = from_prev_lane(R.v[7], 0.0f);
    } else {
        const float* n = own - g.X;
        load8(n + lo, R.vn);
        R.vw = (own + lo)[-1];
        R.vnw = (n + lo)[-1];
        if constexpr (NDIM == 3) {
            const float* u = own - g.XY;
            const float* un = u - g.X;
            load8(u + lo, R.vu);
            load8(un + lo, R.vun);
            R.vuw = (u + lo)[-1];
            R.vunw = (un + lo)[-1];
        }
    }
}

// As raw_load_t at any position p0 (a multiple of 8).
template <int NDIM>
__device__ __forceinline__ void raw_load(const float* __restrict__ in, uint32_t p0, const Geom& g, RawTile<NDIM>& R)
{
    load8(in + p0, R.v);
    if constexpr (NDIM == 1) {
        R.vw = in[(p0 > 0 ? p0 : 1u) - 1];
    } else {
        load8(in + (p0 - g.X), R.vn);
        R.vw = in[p0 - 1];
        R.vnw = in[p0 - g.X - 1];
        if constexpr (NDIM == 3) {
            load8(in + (p0 - g.XY), R.vu);
            load8(in + (p0 - g.XY - g.X), R.vun);
            R.vuw = in[p0 - g.XY - 1];
            R.vunw = in[p0 - g.XY - g.X - 1];
        }
    }
}

// 1D: the 256-element blocks start at multiples of 8, so only position 0 of a thread can
// start a block.  2D / 3D when X % 4 == 0: each half (4 positions) of the thread's 8 lies in
// one row, its neighbour rows are 16-byte aligned, and only a half's first position can sit
// on an x-edge of the Lorenzo block (16 for 2D, 8 for 3D).
template <int NDIM>
__device__ __forceinline__ void raw_compute(const RawTile<NDIM>& R, uint32_t p0, const Geom& g, float r, uint32_t w[4],
                                            uint32_t& ovf)
{
    uint32_t c[8];
    if constexpr (NDIM == 1) {
        if (ballot64(((bits_or(R.v) | __float_as_uint(R.vw)) & 0x7fffffffu) != 0) == 0) {
            w[0] = w[1] = w[2] = w[3] = 0;
            return;
        }
        float s[8];
#pragma unroll
        for (int i = 0; i < 8; i++) s[i] = prequant(R.v[i], r);
        const float sw = (p0 & 255u) == 0 ? 0.0f : prequant(R.vw, r);
        c[0] = code_fast(s[0] - sw, true, ovf);
#pragma unroll
        for (int i = 1; i < 8; i++) c[i] = code_fast(s[i] - s[i - 1], g.mode == 0 || (i & 3) == 0, ovf);
    } else {
        constexpr uint32_t BM = NDIM == 2 ? 15u : 7u;
        bool bx[2], by[2], bz[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const uint32_t p = p0 + 4 * h;
            const uint32_t q = fdiv(p, g.divX);
            const uint32_t x = p - q * g.X;
            uint32_t y = q, z = 0;
            if constexpr (NDIM == 3) {
                z = fdiv(q, g.divY);
                y = q - z * g.Y;
            }
            bx[h] = (x & BM) != 0;
            by[h] = (y & BM) != 0;
            bz[h] = (z & BM) != 0;
        }
        uint32_t nz = bits_or(R.v) | bits_or(R.vn) | ((__float_as_uint(R.vw) | __float_as_uint(R.vnw)) & 0x7fffffffu);
        if constexpr (NDIM == 3)
            nz |= bits_or(R.vu) | bits_or(R.vun) | ((__float_as_uint(R.vuw) | __float_as_uint(R.vunw)) & 0x7fffffffu);
        if (ballot64(nz != 0) == 0) {
            w[0] = w[1] = w[2] = w[3] = 0;
            return;
        }
        // S[j], N[j], U[j], UN[j] for positions p0-1+j (j = 0 is the west column); rows that lie
        // on a y / z block edge contribute zero
        float S[9], N[9];
        S[0] = bx[0] ? prequant(R.vw, r) : 0.0f;
        N[0] = (bx[0] && by[0]) ? prequant(R.vnw, r) : 0.0f;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            S[i + 1] = prequant(R.v[i], r);
            N[i + 1] = by[i >> 2] ? prequant(R.vn[i], r) : 0.0f;
        }
        if constexpr (NDIM == 2) {
            float c1[8];
#pragma unroll
            for (int i = 0; i < 8; i++) c1[i] = S[i + 1] - N[i + 1];
            const float w10 = S[0] - N[0];  // west lane's c1 (0 - 0 on an x-edge)
            c[0] = code_fast(c1[0] - w10, true, ovf);
#pragma unroll
            for (int i = 1; i < 8; i++) {
                const float w1 = (i == 4) ? (bx[1] ? c1[3] : 0.0f) : c1[i - 1];
                c[i] = code_fast(c1[i] - w1, true, ovf);
            }
        } else {
            float U[9], UN[9];
            U[0] = (bx[0] && bz[0]) ? prequant(R.vuw, r) : 0.0f;
            UN[0] = (bx[0] && by[0] && bz[0]) ? prequant(R.vunw, r) : 0.0f;
#pragma unroll
            for (int i = 0; i < 8; i++) {
                U[i + 1] = bz[i >> 2] ? prequant(R.vu[i], r) : 0.0f;
                UN[i + 1] = (by[i >> 2] && bz[i >> 2]) ? prequant(R.vun[i], r) : 0.0f;
            }
#pragma unroll
            for (int i = 0; i < 8; i++) {
                // west-side terms exist unless this is a half's first position on an x-edge
                const bool wx = (i & 3) != 0 || bx[i >> 2];
                float      p = wx ? UN[i] : 0.0f;  // NW3
                p = p - (wx ? N[i] : 0.0f);        // A
                p = p - (wx ? U[i] : 0.0f);        // B
                p = p - UN[i + 1];                 // C
                p = p + (wx ? S[i] : 0.0f);        // D
                p = p + N[i + 1];                  // E
                p = p + U[i + 1];                  // F
                c[i] = code_fast(S[i + 1] - p, true, ovf);
            }
        }
    }
    pack8(c, w);
}

// ---- any X >= 8 (2D / 3D): neighbour rows at 4-byte alignment, per-position block edges
//
// With X >= 8 the 8 positions p0..p0+7 change row at most once (at position k = X - x0).
// Bit i of the masks below describes position p0+i: kx = has an in-block west neighbour
// (x % B != 0), ky = has an in-block row above (y % B != 0), kz = has an in-block plane
// below (z % B != 0); B = 16 (2D) or 8 (3D).

typedef float f32x4a4 __attribute__((ext_vector_type(4), aligned(4)));

__device__ __forceinline__ void load8u(const float* __restrict__ p, float v[8])
{
    const f32x4a4 a = *reinterpret_cast<const f32x4a4*>(p);
    const f32x4a4 b = *reinterpret_cast<const f32x4a4*>(p + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}

// Requires X >= 8 and p0 >= X + 1 (2D) / XY + X + 1 (3D).
template <int NDIM>
__device__ __forceinline__ void raw_load_ut(const float* __restrict__ in, uint32_t t0, uint32_t lo, const Geom& g,
                                            RawTile<NDIM>& R)
{
    const float* own = in + t0;
    const float* n = own - g.X;
    load8(own + lo, R.v);
    load8u(n + lo, R.vn);
    // west values: the previous lane's last value of the same row; lane 0 of the wave reads
    // the one before its row segment (a uniform address, inside the field since gp >= 2)
    // (pointer steps, not lo0 - 1: lo0 is 0 for the tile's first wave)
    const uint32_t lo0 = __builtin_amdgcn_readfirstlane(lo);
    R.vw = from_prev_lane(R.v[7], (own + lo0)[-1]);
    R.vnw = from_prev_lane(R.vn[7], (n + lo0)[-1]);
    if constexpr (NDIM == 3) {
        const float* u = own - g.XY;
        const float* un = u - g.X;
        load8u(u + lo, R.vu);
        load8u(un + lo, R.vun);
        R.vuw = from_prev_lane(R.vu[7], (u + lo0)[-1]);
        R.vunw = from_prev_lane(R.vun[7], (un + lo0)[-1]);
    }
}

// ---- generic 3D path on a RawTile
//
// The generic 3D loop of dq8<3> (lorenzo_var.cuh:271-283 order, any value range) with its
// 36 raw values taken from a RawTile<3> that the fast paths' loads filled, so that a kernel
// holding the tile's raw loads spends no extra registers on it.

// prequantized windows: s[8] own, n / u / un[9] at p - X, p - XY, p - XY - X for
// p = p0-1 .. p0+7, sw the west of p0
struct G3 {
    float s[8], n[9], u[9], un[9], sw;
};

__device__ __forceinline__ void g3_from_raw(const RawTile<3>& R, float r, G3& q)
{
#pragma unroll
    for (int i = 0; i < 8; i++) {
        q.s[i] = prequant(R.v[i], r);
        q.n[i + 1] = prequant(R.vn[i], r);
        q.u[i + 1] = prequant(R.vu[i], r);
        q.un[i + 1] = prequant(R.vun[i], r);
    }
    q.sw = prequant(R.vw, r);
    q.n[0] = prequant(R.vnw, r);
    q.u[0] = prequant(R.vuw, r);
    q.un[0] = prequant(R.vunw, r);
}

// codes of p0 .. p0+7 (p0 < xyz) from the windows, exactly as dq8<3>
__device__ __forceinline__ void g3_codes(const G3& Q, uint32_t p0, const Geom& g, uint32_t w[4], uint32_t& ovf)
{
    const uint32_t q = fdiv(p0, g.divX);
    uint32_t       x = p0 - q * g.X;
    uint32_t       zz = fdiv(q, g.divY);
    uint32_t       y = q - zz * g.Y;
    uint32_t       c[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const bool  bx = (x & 7u) != 0, by = (y & 7u) != 0, bz = (zz & 7u) != 0;
        const float west = (i == 0) ? Q.sw : Q.s[i - 1];
        float       p = (bz && by && bx) ? Q.un[i] : 0.0f;  // NW3 = s[z-1][y-1][x-1]
        p = p - ((by && bx) ? Q.n[i] : 0.0f);                // A   = s[z][y-1][x-1]
        p = p - ((bz && bx) ? Q.u[i] : 0.0f);                // B   = s[z-1][y][x-1]
        p = p - ((bz && by) ? Q.un[i + 1] : 0.0f);           // C   = s[z-1][y-1][x]
        p = p + (bx ? west : 0.0f);                          // D   = s[z][y][x-1]
        p = p + (by ? Q.n[i + 1] : 0.0f);                    // E   = s[z][y-1][x]
        p = p + (bz ? Q.u[i + 1] : 0.0f);                    // F   = s[z-1][y][x]
        const float d = Q.s[i] - p;
        const bool  live = p0 + i < g.xyz;
        c[i] = live ? code_of(d, true) : 0u;
        ovf += live ? is_ovf(d) : 0u;
        if (++x == g.X) {
            x = 0;
            if (++y == g.Y) {
                y = 0;
                ++zz;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) w[k] = c[2 * k] | (c[2 * k + 1] << 16);
}

template <int NDIM>
__device__ __forceinline__ void raw_load_u(const float* __restrict__ in, uint32_t p0, const Geom& g, RawTile<NDIM>& R)
{
    load8(in + p0, R.v);
    load8u(in + (p0 - g.X), R.vn);
    R.vw = in[p0 - 1];
    R.vnw = in[p0 - g.X - 1];
    if constexpr (NDIM == 3) {
        load8u(in + (p0 - g.XY), R.vu);
        load8u(in + (p0 - g.XY - g.X), R.vun);
        R.vuw = in[p0 - g.XY - 1];
        R.vunw = in[p0 - g.XY - g.X - 1];
    }
}

template <int NDIM>
__device__ __forceinline__ void edge_masks(uint32_t p0, const Geom& g, uint32_t& kx, uint32_t& ky, uint32_t& kz)
{
    constexpr uint32_t BM = NDIM == 2 ? 15u : 7u;
    const uint32_t q = fdiv(p0, g.divX);
    const uint32_t x0 = p0 - q * g.X;
    uint32_t       y0 = q, z0 = 0;
    if constexpr (NDIM == 3) {
        z0 = fdiv(q, g.divY);
        y0 = q - z0 * g.Y;
    }
    const uint32_t k = g.X - x0;                          // first position on the next row (>= 1)
    const uint32_t lo = k >= 8 ? 0xFFu : (1u << k) - 1u;  // positions on the row of p0
    const uint32_t hi = 0xFFu & ~lo;
    const uint32_t i0 = (BM + 1 - (x0 & BM)) & BM;        // first x-edge on the row of p0
    const uint32_t ex = (i0 < 8 && i0 < k ? 1u << i0 : 0u) | (k < 8 ? 1u << k : 0u);
    uint32_t       y1 = y0 + 1, z1 = z0;
    if constexpr (NDIM == 3) {
        if (y1 == g.Y) {
            y1 = 0;
            z1 = z0 + 1;
        }
    }
    const uint32_t ey = ((y0 & BM) == 0 ? lo : 0u) | ((y1 & BM) == 0 ? hi : 0u);
    const uint32_t ez = ((z0 & BM) == 0 ? lo : 0u) | ((z1 & BM) == 0 ? hi : 0u);
    kx = ~ex;
    ky = ~ey;
    kz = ~ez;
}

// all-ones when bit i of m is set, else 0
__device__ __forceinline__ uint32_t bitmask(uint32_t m, int i) { return (uint32_t)(-(int32_t)((m >> i) & 1u)); }

// v if the mask is all-ones, +0.0f if it is 0
__device__ __forceinline__ float keep(float v, uint32_t m) { return __uint_as_float(__float_as_uint(v) & m); }

// 2D, any X >= 8: same float operations as raw_compute (and dq8); a neighbour outside the
// block enters as the value +0.0f (masked before the prequant: prequant(+0) = +0).
__device__ __forceinline__ void raw_compute_u(const RawTile<2>& R, uint32_t p0, const Geom& g, float r, uint32_t w[4],
                                              uint32_t& ovf)
{
    const uint32_t nz = bits_or(R.v) | bits_or(R.vn) | ((__float_as_uint(R.vw) | __float_as_uint(R.vnw)) & 0x7fffffffu);
    if (ballot64(nz != 0) == 0) {
        w[0] = w[1] = w[2] = w[3] = 0;
        return;
    }
    uint32_t kx, ky, kz;
    edge_masks<2>(p0, g, kx, ky, kz);
    uint32_t       c[8];
    float          S[9], N[9];  // positions p0-1+j; the west column has the masks of position 0
    const uint32_t mx0 = bitmask(kx, 0), my0 = bitmask(ky, 0);
    S[0] = prequant(keep(R.vw, mx0), r);
    N[0] = prequant(keep(R.vnw, mx0 & my0), r);
#pragma unroll
    for (int i = 0; i < 8; i++) {
        S[i + 1] = prequant(R.v[i], r);
        N[i + 1] = prequant(keep(R.vn[i], bitmask(ky, i)), r);
    }
    float c1[8];
#pragma unroll
    for (int i = 0; i < 8; i++) c1[i] = S[i + 1] - N[i + 1];
    c[0] = code_fast(c1[0] - (S[0] - N[0]), true, ovf);
#pragma unroll
    for (int i = 1; i < 8; i++) c[i] = code_fast(c1[i] - keep(c1[i - 1], bitmask(kx, i)), true, ovf);
    pack8(c, w);
}

// ---- packed 16-bit path for small prequantised values
//
// When every |s| <= 16383 (1D), 8191 (2D) or 4095 (3D), every delta satisfies |d| <= 32764,
// the reference's float arithmetic on these integers is exact, and the exact d equals the
// result of the same sums taken modulo 2^16.  The Lorenzo deltas are then computed two
// positions per instruction on int16 pairs, in any order of the terms:
//   Q = S - N (2D),  Q = S - N - U + UN (3D),  Q = S (1D)   per column, neighbours outside the
//   block being 0;  d = Q - (has west ? Q of the west column : 0).
// (For 3D this is the reference's d = s - (NW3 - A - B - C + D + E + F) regrouped by column.)
// The code pair is |d| | sign(d) << 15 per half, i.e. the packed code word itself; no
// overflow is possible.

typedef unsigned short v2u16 __attribute__((ext_vector_type(2)));
typedef short          v2i16 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ int32_t pq_i(float x, float r)
{
    const float t = __fmul_rn(x, r);
    return (int32_t)__fadd_rn(t, __builtin_copysignf(0.49999997f, t));  // truncating conversion
}

// pq_i for x >= 0 (or -0): v_cvt_rpi_i32_f32 gives floor(t + 0.5) with the sum taken exactly,
// which is round-half-away for every t in [0, 2^23) (an exhaustive check over all such floats
// on gfx950 finds no difference from pq_i); two instructions instead of four.
__device__ __forceinline__ int32_t pq_pos(float x, float r)
{
    const float t = __fmul_rn(x, r);
    int32_t     s;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(s) : "v"(t));
    return s;
}

// (lo, hi) low 16 bits each -> one word
__device__ __forceinline__ uint32_t pk2(int32_t lo, int32_t hi)
{
    return __builtin_amdgcn_perm((uint32_t)hi, (uint32_t)lo, 0x05040100u);
}

__device__ __forceinline__ uint32_t pk_sub(uint32_t a, uint32_t b)
{
    v2u16 x, y;
    __builtin_memcpy(&x, &a, 4);
    __builtin_memcpy(&y, &b, 4);
    const v2u16 z = x - y;
    uint32_t    o;
    __builtin_memcpy(&o, &z, 4);
    return o;
}

// sign-magnitude codes of two int16 deltas (|d| < 32768); sign kept where `sign` has bit 15 / 31
__device__ __forceinline__ uint32_t pk_code(uint32_t d, uint32_t sign)
{
    v2i16 x;
    __builtin_memcpy(&x, &d, 4);
    const v2i16 zero = {0, 0};
    const v2i16 neg = __builtin_bit_cast(v2i16, __builtin_bit_cast(v2u16, zero) - __builtin_bit_cast(v2u16, x));
    const v2i16 a = __builtin_elementwise_max(x, neg);
    uint32_t    m;
    __builtin_memcpy(&m, &a, 4);
    return m | (d & sign);
}

// w[k] = codes of positions 2k, 2k+1 from the column values Q[j] of positions p0-1+j and the
// "has west" flags of positions 2k, 2k+1 (as 16-bit half masks)
__device__ __forceinline__ void pk_deltas(const int32_t Q[9], const uint32_t wmask[4], const uint32_t sign[4],
                                          uint32_t w[4])
{
    uint32_t q2[4];
#pragma unroll
    for (int k = 0; k < 4; k++) q2[k] = pk2(Q[2 * k + 1], Q[2 * k + 2]);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t west = k == 0 ? __builtin_amdgcn_perm(q2[0], (uint32_t)Q[0], 0x05040100u)
                                     : __builtin_amdgcn_alignbyte(q2[k], q2[k - 1], 2);
        w[k] = pk_code(pk_sub(q2[k], west & wmask[k]), sign[k]);
    }
}

// half masks of 8-bit position flags: pair k = positions 2k (low) and 2k+1 (high)
__device__ __forceinline__ uint32_t pair_mask(uint32_t m, int k)
{
    return __builtin_amdgcn_perm(bitmask(m, 2 * k + 1), bitmask(m, 2 * k), 0x05040100u);
}

// AL: X % 4 == 0 layout (raw_load); otherwise the any-X layout (raw_load_u, X >= 8).
// POS: every value of the field is >= 0 (pq_pos for pq_i).
template <int NDIM, bool AL, bool POS>
__device__ __forceinline__ void raw_compute_pk(const RawTile<NDIM>& R, uint32_t p0, const Geom& g, float r,
                                               uint32_t w[4])
{
    uint32_t nz = bits_or(R.v) | (__float_as_uint(R.vw) & 0x7fffffffu);
    if constexpr (NDIM >= 2) nz |= bits_or(R.vn) | (__float_as_uint(R.vnw) & 0x7fffffffu);
    if constexpr (NDIM == 3)
        nz |= bits_or(R.vu) | bits_or(R.vun) | ((__float_as_uint(R.vuw) | __float_as_uint(R.vunw)) & 0x7fffffffu);
    if (ballot64(nz != 0) == 0) {
        w[0] = w[1] = w[2] = w[3] = 0;
        return;
    }
    auto pq = [r](float x) { return POS ? pq_pos(x, r) : pq_i(x, r); };
    int32_t  Q[9];
    uint32_t wmask[4] = {~0u, ~0u, ~0u, ~0u};
    uint32_t sign[4] = {0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
    if constexpr (NDIM == 1) {
        Q[0] = (p0 & 255u) != 0 ? pq(R.vw) : 0;
#pragma unroll
        for (int i = 0; i < 8; i++) Q[i + 1] = pq(R.v[i]);
        if (g.mode != 0) {  // LITERAL: sign only at positions % 4 == 0 (lorenzo_var.cuh:51-63)
            sign[0] = sign[2] = 0x00008000u;
            sign[1] = sign[3] = 0u;
        }
    } else {
        // keep masks (all ones = keep, 0 = outside the Lorenzo block): my[i] / mz[i] for the row
        // above / plane below position p0+i, mx0 for the west of position 0
        uint32_t my[8], mz[8], mx0;
        if constexpr (AL) {
            // each half (p0, p0+4) lies in one row; x0 % 4 == 0, so only a half's first position
            // can sit on an x-edge, and the second half is the next 4 positions of the row or
            // the start of the next row
            constexpr uint32_t BM = NDIM == 2 ? 15u : 7u;
            const uint32_t q0 = fdiv(p0, g.divX);
            const uint32_t x0 = p0 - q0 * g.X;
            const bool     wrap = x0 + 4u == g.X;
            const uint32_t x1 = wrap ? 0u : x0 + 4u;
            uint32_t       y0 = q0, y1 = q0 + (wrap ? 1u : 0u), z0 = 0, z1 = 0;
            if constexpr (NDIM == 3) {
                z0 = fdiv(q0, g.divY);
                y0 = q0 - z0 * g.Y;
                const bool wrap2 = wrap && y0 + 1u == g.Y;
                y1 = wrap ? (wrap2 ? 0u : y0 + 1u) : y0;
                z1 = z0 + (wrap2 ? 1u : 0u);
            }
            const uint32_t my0 = (y0 & BM) ? ~0u : 0u, my1 = (y1 & BM) ? ~0u : 0u;
            const uint32_t mz0 = (z0 & BM) ? ~0u : 0u, mz1 = (z1 & BM) ? ~0u : 0u;
#pragma unroll
            for (int i = 0; i < 8; i++) {
                my[i] = i < 4 ? my0 : my1;
                mz[i] = NDIM == 3 ? (i < 4 ? mz0 : mz1) : ~0u;
            }
            mx0 = (x0 & BM) ? ~0u : 0u;
            wmask[0] = mx0 | 0xFFFF0000u;
            wmask[2] = ((x1 & BM) ? ~0u : 0u) | 0xFFFF0000u;
        } else {
            uint32_t kx, ky, kz = 0xFFu;
            edge_masks<NDIM>(p0, g, kx, ky, kz);
#pragma unroll
            for (int i = 0; i < 8; i++) {
                my[i] = bitmask(ky, i);
                mz[i] = bitmask(kz, i);
            }
            mx0 = bitmask(kx, 0);
#pragma unroll
            for (int k = 0; k < 4; k++) wmask[k] = pair_mask(kx, k);
        }
        Q[0] = pq(keep(R.vw, mx0)) - pq(keep(R.vnw, mx0 & my[0]));
        if constexpr (NDIM == 3) Q[0] = Q[0] - pq(keep(R.vuw, mx0 & mz[0])) + pq(keep(R.vunw, mx0 & my[0] & mz[0]));
#pragma unroll
        for (int i = 0; i < 8; i++) {
            Q[i + 1] = pq(R.v[i]) - pq(keep(R.vn[i], my[i]));
            if constexpr (NDIM == 3) Q[i + 1] = Q[i + 1] - pq(keep(R.vu[i], mz[i])) + pq(keep(R.vun[i], my[i] & mz[i]));
        }
    }
    pk_deltas(Q, wmask, sign, w);
}

// ---- 3D column form in float for |s| < 2^21
//
// Every partial sum of the reference's predictor (7 terms of |s| < 2^21) is an integer below
// 2^24, so the float arithmetic is exact in any order and equals
//   Q = ((S - N) - U) + UN per column,  d = Q - (has west ? Q of the west column : 0),
// each step exact as well (|Q| < 2^23, |d| < 2^24).  Same masks as the packed path.
template <bool AL>
__device__ __forceinline__ void raw_compute_q3(const RawTile<3>& R, uint32_t p0, const Geom& g, float r, uint32_t w[4],
                                               uint32_t& ovf)
{
    uint32_t nz = bits_or(R.v) | bits_or(R.vn) | bits_or(R.vu) | bits_or(R.vun) |
                  ((__float_as_uint(R.vw) | __float_as_uint(R.vnw) | __float_as_uint(R.vuw) | __float_as_uint(R.vunw)) &
                   0x7fffffffu);
    if (ballot64(nz != 0) == 0) {
        w[0] = w[1] = w[2] = w[3] = 0;
        return;
    }
    uint32_t kx, ky, kz;
    if constexpr (AL) {
        kx = ky = kz = 0xFFu;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const uint32_t p = p0 + 4 * h;
            const uint32_t q = fdiv(p, g.divX);
            const uint32_t x = p - q * g.X;
            const uint32_t z = fdiv(q, g.divY);
            const uint32_t y = q - z * g.Y;
            if ((x & 7u) == 0) kx &= ~(1u << (4 * h));
            if ((y & 7u) == 0) ky &= ~(0xFu << (4 * h));
            if ((z & 7u) == 0) kz &= ~(0xFu << (4 * h));
        }
    } else {
        edge_masks<3>(p0, g, kx, ky, kz);
    }
    const uint32_t mx0 = bitmask(kx, 0), my0 = bitmask(ky, 0), mz0 = bitmask(kz, 0);
    float          prev = ((prequant(keep(R.vw, mx0), r) - prequant(keep(R.vnw, mx0 & my0), r)) -
                  prequant(keep(R.vuw, mx0 & mz0), r)) +
                 prequant(keep(R.vunw, mx0 & my0 & mz0), r);
    uint32_t c[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const uint32_t my = bitmask(ky, i), mz = bitmask(kz, i);
        const float    q = ((prequant(R.v[i], r) - prequant(keep(R.vn[i], my), r)) - prequant(keep(R.vu[i], mz), r)) +
                        prequant(keep(R.vun[i], my & mz), r);
        c[i] = code_fast(q - keep(prev, bitmask(kx, i)), true, ovf);
        prev = q;
    }
    pack8(c, w);
}

__device__ __forceinline__ void dq8_u(const float* __restrict__ in, uint32_t p0, const Geom& g, float r, uint32_t w[4],
                                      uint32_t& ovf)
{
    RawTile<2> R;
    raw_load_u<2>(in, p0, g, R);
    raw_compute_u(R, p0, g, r, w, ovf);
}

__device__ __forceinline__ void dq8_fast1(const float* __restrict__ in, uint32_t p0, const Geom& g, float r,
                                          uint32_t w[4], uint32_t& ovf)
{
    RawTile<1> R;
    raw_load<1>(in, p0, g, R);
    raw_compute<1>(R, p0, g, r, w, ovf);
}

// Requires p0 >= X + 1 (2D) / XY + X + 1 (3D) and X % 4 == 0.
template <int NDIM>
__device__ __forceinline__ void dq8_a4(const float* __restrict__ in, uint32_t p0, const Geom& g, float r,
                                       uint32_t w[4], uint32_t& ovf)
{
    RawTile<NDIM> R;
    raw_load<NDIM>(in, p0, g, R);
    raw_compute<NDIM>(R, p0, g, r, w, ovf);
}

// ---------------------------------------------------------------- 32x32 bit transpose
//
// Group k of a tile is the 32 words w[32k .. 32k+31]; as a bit matrix M[j][b] = bit b of
// w[32k+j].  bitshuffleAndBitflag (fz.cu:75-82) stores column b as word o[32b + k], bit j.
// The transpose below is the recursive block swap (distance 16, 8, 4, 2, 1): rows 4L..4L+3
// live in lane L of the 8-lane group, so distances 16/8/4 are lane exchanges (xor 4/2/1)
// and 2/1 are register swaps.  Afterwards register q of lane L holds column 4L+q.  The
// same routine inverts itself.

__device__ __forceinline__ uint32_t lane_xor1(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
}
__device__ __forceinline__ uint32_t lane_xor2(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
}
__device__ __forceinline__ uint32_t lane_xor4(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x101F);  // bitmask mode, xor 4
}

template <int J>
__device__ __forceinline__ constexpr uint32_t tmask()
{
    return J == 16 ? 0x0000FFFFu : J == 8 ? 0x00FF00FFu : J == 4 ? 0x0F0F0F0Fu : J == 2 ? 0x33333333u : 0x55555555u;
}

// lane-exchange stage: `upper` lanes hold row k+J, the others row k.  J = 16 and 8 move whole
// bytes: one byte permute of (mine, other) (upper: [o2 o3 m2 m3] / [o1 m1 o3 m3], lower:
// [m0 m1 o0 o1] / [m0 o0 m2 o2], byte 0 first); J <= 4 rotates the other row and merges.
template <int J>
__device__ __forceinline__ uint32_t xstage(uint32_t mine, uint32_t other, bool upper)
{
    if constexpr (J == 16) return __builtin_amdgcn_perm(mine, other, upper ? 0x07060302u : 0x01000504u);
    if constexpr (J == 8) return __builtin_amdgcn_perm(mine, other, upper ? 0x07030501u : 0x02060004u);
    constexpr uint32_t M = tmask<J>();
    const uint32_t rot = __builtin_amdgcn_alignbit(other, other, upper ? J : 32 - J);
    const uint32_t sel = upper ? ~M : M;
    // mine where sel is set, else rot: one three-input bit operation (truth table S0 ? S1 : S2,
    // operand bytes 0xF0 / 0xCC / 0xAA)
    return __builtin_amdgcn_bitop3_b32(sel, mine, rot, 0xCA);
}

template <int J>
__device__ __forceinline__ void rstage(uint32_t& a, uint32_t& b)
{
    constexpr uint32_t M = tmask<J>();
    const uint32_t na = (a & M) | ((b << J) & ~M);
    const uint32_t nb = (b & ~M) | ((a >> J) & M);
    a = na;
    b = nb;
}

// L = lane index within the 8-lane group (threadIdx.x & 7).  All 64 lanes must be active.
__device__ __forceinline__ void transpose4(uint32_t w[4], uint32_t L)
{
#pragma unroll
    for (int q = 0; q < 4; q++) w[q] = xstage<16>(w[q], lane_xor4(w[q]), (L & 4u) != 0);
#pragma unroll
    for (int q = 0; q < 4; q++) w[q] = xstage<8>(w[q], lane_xor2(w[q]), (L & 2u) != 0);
#pragma unroll
    for (int q = 0; q < 4; q++) w[q] = xstage<4>(w[q], lane_xor1(w[q]), (L & 1u) != 0);
    rstage<2>(w[0], w[2]);
    rstage<2>(w[1], w[3]);
    rstage<1>(w[0], w[1]);
    rstage<1>(w[2], w[3]);
}

// ---------------------------------------------------------------- decoupled look-back
//
// One u64 per tile: high word = flag (0 not ready, 1 aggregate, 2 inclusive prefix), low
// word = value.  Flag and value travel in one 64-bit atomic, so relaxed agent-scope
// accesses suffice.  Tiles are claimed in order from an atomic ticket, so every
// predecessor belongs to a running workgroup and publishes without waiting: the loop ends.

constexpr uint64_t kFlagAgg = 1ull << 32;
constexpr uint64_t kFlagInc = 2ull << 32;

__device__ __forceinline__ void state_store(uint64_t* p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t state_load(uint64_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Called by all 64 lanes of one wave.  Returns the exclusive prefix of `tile`.
__device__ __forceinline__ uint32_t lookback(uint64_t* states, uint32_t tile, uint32_t aggregate)
{
    const uint32_t l = lane_id();
    if (tile == 0) {
        if (l == 0) state_store(&states[0], kFlagInc | aggregate);
        return 0;
    }
    if (l == 0) state_store(&states[tile], kFlagAgg | aggregate);
    uint32_t excl = 0;
    int64_t  base = (int64_t)tile - 1;
    while (true) {
        const int64_t idx = base - (int64_t)l;
        const uint64_t s = idx >= 0 ? state_load(&states[idx]) : kFlagInc;
        const uint32_t flag = (uint32_t)(s >> 32);
        const uint64_t inc = ballot64(flag == 2u);
        const uint64_t bad = ballot64(flag == 0u);
        const uint32_t lp = inc ? (uint32_t)__builtin_ctzll(inc) : 64u;
        const uint64_t need = lp >= 63u ? ~0ull : ((2ull << lp) - 1ull);
        if (bad & need) {
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        excl += wave_sum(l <= lp ? (uint32_t)s : 0u);
        if (lp < 64u) break;
        base -= 64;
    }
    if (l == 0) state_store(&states[tile], kFlagInc | (excl + aggregate));
    return excl;
}

// Look-back over K windows of 64 predecessors per round trip (all loads issued together), for
// grids whose workgroups finish at about the same time.  Called by all 64 lanes of one wave.
template <int K>
__device__ __forceinline__ uint32_t lookback_wide(uint64_t* states, uint32_t pos, uint32_t aggregate)
{
    const uint32_t l = lane_id();
    if (pos == 0) {
        if (l == 0) state_store(&states[0], kFlagInc | aggregate);
        return 0;
    }
    if (l == 0) state_store(&states[pos], kFlagAgg | aggregate);
    uint32_t excl = 0;
    int64_t  base = (int64_t)pos - 1;
    while (true) {
        uint64_t s[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int64_t idx = base - (int64_t)l - 64 * k;
            s[k] = idx >= 0 ? state_load(&states[idx]) : kFlagInc;
        }
        bool     retry = false, done = false;
        uint32_t add = 0;
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (!done && !retry) {
                const uint32_t flag = (uint32_t)(s[k] >> 32);
                const uint64_t inc = ballot64(flag == 2u);
                const uint64_t bad = ballot64(flag == 0u);
                const uint32_t lp = inc ? (uint32_t)__builtin_ctzll(inc) : 64u;
                const uint64_t need = lp >= 63u ? ~0ull : ((2ull << lp) - 1ull);
                if (bad & need) {
                    retry = true;
                } else {
                    add += wave_sum(l <= lp ? (uint32_t)s[k] : 0u);
                    if (lp < 64u) done = true;
                }
            }
        }
        if (retry) {
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        excl += add;
        if (done) break;
        base -= 64 * K;
    }
    if (l == 0) state_store(&states[pos], kFlagInc | (excl + aggregate));
    return excl;
}

}  // namespace fz
