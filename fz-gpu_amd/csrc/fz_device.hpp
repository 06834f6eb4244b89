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

// Requires the fast-path geometry (tile_path == 2 ignoring the value range).
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

// lane-exchange stage: `upper` lanes hold row k+J, the others row k.
template <int J>
__device__ __forceinline__ uint32_t xstage(uint32_t mine, uint32_t other, bool upper)
{
    constexpr uint32_t M = tmask<J>();
    const uint32_t rot = __builtin_amdgcn_alignbit(other, other, upper ? J : 32 - J);
    const uint32_t sel = upper ? ~M : M;
    return (mine & sel) | (rot & ~sel);
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
