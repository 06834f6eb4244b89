// fz_verify.hip -- quality metrics of a reconstruction (verify_data, verify.hh:25-83) as one
// HBM streaming pass plus a one-workgroup finish.
//
// The reference makes two sequential passes: the means first, then centred products around
// them.  Here every element is read once.  Sums are taken around a shift K = odata[0] (one
// of the values, so |mean - K| <= range) in double:
//   so = sum(o-K), sx = sum(x-K), qo = sum((o-K)^2), qx = sum((x-K)^2), c = sum((o-K)(x-K))
// and the centred sums follow as qo - so^2/n, qx - sx^2/n, c - so*sx/n.  Shifted sums add
// across threads without a merge formula because K is the same everywhere.
//
// The float quantities of the reference stay float: |x-o| (verify.hh:46), its square
// (verify.hh:58) and the pointwise ratio |x-o|/|o| (verify.hh:48).  The ratio is never
// divided per element: each thread keeps the largest one as a pair (a, b) of floats and
// compares a_i*b > a*b_i in double, where both products are exact (24-bit significands),
// so the comparison is exact; one division of the winning pair at the end gives the
// float quotient, and rounding is monotone, so that is the largest rounded quotient.
//
// The largest |x-o| keeps its first index (strict '<' scan, verify.hh:51-52): threads scan
// their elements in increasing index order and merges prefer the smaller index on ties.
// Non-numbers are skipped by every comparison, as in the reference, except element 0, which
// the reference uses as the initial value of the extrema and of the largest error.
#include <math.h>

#include "fz_internal.hpp"

namespace fz {
namespace {

constexpr uint32_t kVWG = 256;                 // threads of a partial-sum workgroup
constexpr uint32_t kVU = 4;                    // float4 loads per array, thread and chunk
constexpr uint32_t kVChunk = kVWG * 4 * kVU;   // elements of one chunk (4096)
constexpr uint32_t kVMaxParts = 1024;          // partial-sum workgroups at most
static_assert(kVMaxParts % kVWG == 0, "each finishing thread loads the same number of partials");
constexpr uint32_t kNoPos = 0xFFFFFFFFu;
constexpr uint32_t kWave = 64;

struct VAcc {
    double   so, sx, qo, qx, c, e;  // shifted sums (see above); e = sum of float |x-o|^2
    double   pa, pb;                // largest |x-o|/|o| as the ratio pa/pb of two floats
    float    omin, omax, xmin, amax;
    uint64_t aidx;                  // first index of amax
};

__device__ inline void vacc_init(VAcc& a)
{
    a.so = a.sx = a.qo = a.qx = a.c = a.e = 0.0;
    a.pa = 0.0;
    a.pb = 1.0;
    a.omin = INFINITY;
    a.omax = -INFINITY;
    a.xmin = INFINITY;
    a.amax = -1.0f;  // below every |x-o|
    a.aidx = ~0ull;
}

// One element (verify.hh:39-59).  Positions grow with the element index within a thread.
__device__ inline void vstep(VAcc& a, uint32_t& apos, float o, float x, double K, uint32_t pos)
{
    const double od = (double)o, xd = (double)x;
    const double dox = od - K, dxx = xd - K;
    a.so += dox;
    a.sx += dxx;
    a.qo = fma(dox, dox, a.qo);
    a.qx = fma(dxx, dxx, a.qx);
    a.c = fma(dox, dxx, a.c);
    const float ae = fabsf(x - o);
    const float ae2 = ae * ae;
    a.e += (double)ae2;
    a.omin = fminf(a.omin, o);
    a.omax = fmaxf(a.omax, o);
    a.xmin = fminf(a.xmin, x);
    const bool more = ae > a.amax;
    apos = more ? pos : apos;
    a.amax = more ? ae : a.amax;
    // o == 0 compares 0 * pb > pa * 0: never taken (selects, no branch)
    const double aed = (double)(o != 0.0f ? ae : 0.0f), ood = fabs(od);
    const bool   take = aed * a.pb > a.pa * ood;
    a.pa = take ? aed : a.pa;
    a.pb = take ? ood : a.pb;
}

__device__ inline void vmerge(VAcc& a, const VAcc& b)
{
    a.so += b.so;
    a.sx += b.sx;
    a.qo += b.qo;
    a.qx += b.qx;
    a.c += b.c;
    a.e += b.e;
    if (b.pa * a.pb > a.pa * b.pb) {
        a.pa = b.pa;
        a.pb = b.pb;
    }
    a.omin = fminf(a.omin, b.omin);
    a.omax = fmaxf(a.omax, b.omax);
    a.xmin = fminf(a.xmin, b.xmin);
    if (b.amax > a.amax || (b.amax == a.amax && b.aidx < a.aidx)) {
        a.amax = b.amax;
        a.aidx = b.aidx;
    }
}

constexpr int kVWords = sizeof(VAcc) / 4;
static_assert(sizeof(VAcc) == 88, "VAcc is moved between lanes as 22 dwords");

// a's dwords moved by one DPP pattern; lanes without a source (or outside ROWS) take id's
template <int CTRL, int ROWS>
__device__ inline VAcc vdpp(const VAcc& id, const VAcc& a)
{
    uint32_t iw[kVWords], aw[kVWords], bw[kVWords];
    __builtin_memcpy(iw, &id, sizeof(VAcc));
    __builtin_memcpy(aw, &a, sizeof(VAcc));
#pragma unroll
    for (int i = 0; i < kVWords; i++)
        bw[i] = (uint32_t)__builtin_amdgcn_update_dpp((int)iw[i], (int)aw[i], CTRL, ROWS, 0xF, false);
    VAcc b;
    __builtin_memcpy(&b, bw, sizeof(VAcc));
    return b;
}

// Merges the 64 lanes into lane 63 (VALU only): row shifts by 1, 2, 4, 8 leave each 16-lane
// row's result in its lane 15, then the broadcasts of lanes 15 and 31 reach lane 63.
__device__ inline void vreduce_wave(VAcc& a)
{
    VAcc id;
    vacc_init(id);
    vmerge(a, vdpp<0x111, 0xF>(id, a));
    vmerge(a, vdpp<0x112, 0xF>(id, a));
    vmerge(a, vdpp<0x114, 0xF>(id, a));
    vmerge(a, vdpp<0x118, 0xF>(id, a));
    vmerge(a, vdpp<0x142, 0xA>(id, a));
    vmerge(a, vdpp<0x143, 0xC>(id, a));
}

// Reduces a over the workgroup; the result is valid in thread 0 (in all of wave 0 with
// WAVE0).
template <uint32_t NT, bool WAVE0 = false>
__device__ inline void vreduce_block(VAcc& a)
{
    __shared__ VAcc s[NT / kWave];
    vreduce_wave(a);
    if ((threadIdx.x & (kWave - 1)) == kWave - 1) s[threadIdx.x / kWave] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        a = s[0];
        for (uint32_t i = 1; i < NT / kWave; i++) vmerge(a, s[i]);
        if constexpr (WAVE0) s[0] = a;
    }
    if constexpr (WAVE0) {
        __syncthreads();
        if (threadIdx.x < kWave) a = s[0];
    }
}

__device__ inline double lane_bcast(double v, int lane)
{
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, lane), hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), lane);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

struct VParts {
    double*   d;     // [8][G]: so, sx, qo, qx, c, e, pa, pb
    float*    f;     // [4][G]: omin, omax, xmin, amax
    uint64_t* idx;   // [G]
    uint32_t  G;
};

inline size_t al256v(size_t v) { return (v + 255) & ~size_t(255); }

inline uint32_t verify_grid(uint64_t n)
{
    const uint64_t chunks = (n + kVChunk - 1) / kVChunk;
    return (uint32_t)(chunks < kVMaxParts ? chunks : kVMaxParts);
}

inline size_t verify_temp_bytes(uint32_t G)
{
    return al256v(64ull * G) + al256v(16ull * G) + al256v(8ull * G);
}

inline VParts verify_parts(void* temp, uint32_t G)
{
    uint8_t* p = static_cast<uint8_t*>(temp);
    VParts   P;
    P.d = reinterpret_cast<double*>(p);
    P.f = reinterpret_cast<float*>(p + al256v(64ull * G));
    P.idx = reinterpret_cast<uint64_t*>(p + al256v(64ull * G) + al256v(16ull * G));
    P.G = G;
    return P;
}

// Partial sums of the chunks blockIdx.x, blockIdx.x + G, ...  A full chunk is read as float4
// (VEC: both arrays 16-byte aligned), thread t taking elements u*1024 + 4t + j (eight 16-byte
// loads in flight per thread; issuing the next chunk's loads early measured no faster).
// A partial chunk, or any chunk of unaligned arrays, is read one float per lane:
// u*1024 + 256j + t.  Either way a thread meets its elements in increasing index order.
template <bool VEC>
__global__ __launch_bounds__(kVWG) void k_verify_part(const float* __restrict__ x, const float* __restrict__ o,
                                                      uint64_t n, VParts P)
{
    const double   K = (double)o[0];
    const uint32_t t = threadIdx.x;
    const uint64_t nchunks = (n + kVChunk - 1) / kVChunk;
    VAcc           a;
    vacc_init(a);
    uint32_t apos = kNoPos;
    uint32_t k = 0;
    uint64_t c = blockIdx.x;
    if constexpr (VEC) {
        const uint64_t nfull = n / kVChunk;
        struct Buf {
            float4 o[kVU], x[kVU];
        };
        auto load = [&](Buf& B, uint64_t cc) {
            const uint64_t base = cc * kVChunk + 4 * t;
#pragma unroll
            for (uint32_t u = 0; u < kVU; u++) {
                B.o[u] = *reinterpret_cast<const float4*>(o + base + u * 1024);
                B.x[u] = *reinterpret_cast<const float4*>(x + base + u * 1024);
            }
        };
        auto step = [&](const Buf& B) {
#pragma unroll
            for (uint32_t u = 0; u < kVU; u++) {
                const uint32_t p = (k * kVU + u) * 4;
                vstep(a, apos, B.o[u].x, B.x[u].x, K, p + 0);
                vstep(a, apos, B.o[u].y, B.x[u].y, K, p + 1);
                vstep(a, apos, B.o[u].z, B.x[u].z, K, p + 2);
                vstep(a, apos, B.o[u].w, B.x[u].w, K, p + 3);
            }
            ++k;
            c += gridDim.x;
        };
        while (c < nfull) {  // step() advances c
            Buf A;
            load(A, c);
            step(A);
        }
    }
    for (; c < nchunks; c += gridDim.x, ++k) {
        const uint64_t base = c * kVChunk;
        float ov[kVU][4], xv[kVU][4];
#pragma unroll
        for (uint32_t u = 0; u < kVU; u++)
#pragma unroll
            for (uint32_t j = 0; j < 4; j++) {
                const uint64_t i = base + u * 1024 + j * 256 + t;
                ov[u][j] = i < n ? o[i] : 0.0f;
                xv[u][j] = i < n ? x[i] : 0.0f;
            }
#pragma unroll
        for (uint32_t u = 0; u < kVU; u++)
#pragma unroll
            for (uint32_t j = 0; j < 4; j++)
                if (base + u * 1024 + j * 256 + t < n) vstep(a, apos, ov[u][j], xv[u][j], K, (k * kVU + u) * 4 + j);
    }
    if (apos != kNoPos) {
        const uint32_t kk = apos / (4 * kVU), u = (apos / 4) % kVU, j = apos % 4;
        const uint64_t base = (blockIdx.x + (uint64_t)kk * gridDim.x) * kVChunk;
        const bool     vec_layout = VEC && base + kVChunk <= n;
        a.aidx = base + u * 1024 + (vec_layout ? 4 * t + j : j * 256 + t);
    }
    vreduce_block<kVWG>(a);
    if (t == 0) {
        const uint32_t b = blockIdx.x, G = P.G;
        P.d[0 * G + b] = a.so;
        P.d[1 * G + b] = a.sx;
        P.d[2 * G + b] = a.qo;
        P.d[3 * G + b] = a.qx;
        P.d[4 * G + b] = a.c;
        P.d[5 * G + b] = a.e;
        P.d[6 * G + b] = a.pa;
        P.d[7 * G + b] = a.pb;
        P.f[0 * G + b] = a.omin;
        P.f[1 * G + b] = a.omax;
        P.f[2 * G + b] = a.xmin;
        P.f[3 * G + b] = a.amax;
        P.idx[b] = a.aidx;
    }
}

// Merges the G partials and writes the fields verify_data writes (verify.hh:61-82).
__device__ void verify_finish(const float* __restrict__ x, const float* __restrict__ o, uint64_t n, const VParts& P,
                              fz_stats* __restrict__ s)
{
    // every partial is loaded before any is merged: one memory round trip
    constexpr uint32_t kPer = kVMaxParts / kVWG;
    const uint32_t     G = P.G;
    VAcc               p[kPer];
#pragma unroll
    for (uint32_t j = 0; j < kPer; j++) {
        const uint32_t b = threadIdx.x + j * kVWG;
        vacc_init(p[j]);
        if (b < G) {
            p[j].so = P.d[0 * G + b];
            p[j].sx = P.d[1 * G + b];
            p[j].qo = P.d[2 * G + b];
            p[j].qx = P.d[3 * G + b];
            p[j].c = P.d[4 * G + b];
            p[j].e = P.d[5 * G + b];
            p[j].pa = P.d[6 * G + b];
            p[j].pb = P.d[7 * G + b];
            p[j].omin = P.f[0 * G + b];
            p[j].omax = P.f[1 * G + b];
            p[j].xmin = P.f[2 * G + b];
            p[j].amax = P.f[3 * G + b];
            p[j].aidx = P.idx[b];
        }
    }
    VAcc a = p[0];
#pragma unroll
    for (uint32_t j = 1; j < kPer; j++) vmerge(a, p[j]);
    vreduce_block<kVWG, true>(a);
    if (threadIdx.x >= kWave) return;

    // wave 0 holds the merged sums; the divisions, square roots and logarithms run side by
    // side in lanes 0-3 instead of one after another
    const int    lane = (int)threadIdx.x;
    const float  o0 = o[0], x0 = x[0];
    const double nd = (double)n;
    // element 0 starts every scan (verify.hh:27-29); a non-number there is never replaced
    const double omin = isnan(o0) ? (double)o0 : (double)a.omin;
    const double omax = isnan(o0) ? (double)o0 : (double)a.omax;
    // verify.hh:44 compares odata against the xdata maximum that starts at xdata[0]
    const double xmax = isnan(x0) ? (double)x0 : fmax((double)x0, (double)a.omax);
    const double xmin = isnan(x0) ? (double)x0 : (double)a.xmin;
    const float  err0 = fabsf(x0 - o0);
    const double max_abs = isnan(err0) ? (double)err0 : (double)a.amax;
    const uint64_t max_idx = isnan(err0) ? 0 : a.aidx;

    // the reference divides by len everywhere; one reciprocal is within a rounding of that
    const double inv_n = 1.0 / nd;
    // rounding can leave a zero variance slightly negative; a non-number stays one
    double var_o = a.qo - a.so * a.so * inv_n, var_x = a.qx - a.sx * a.sx * inv_n;
    var_o = var_o < 0.0 ? 0.0 : var_o;
    var_x = var_x < 0.0 ? 0.0 : var_x;
    const double cor = a.c - a.so * a.sx * inv_n;
    const double rng_o = omax - omin;
    const double mse = a.e * inv_n;

    // lanes 0, 1, 2: std_o, std_x, sqrt(MSE)
    const double r = sqrt(lane == 0 ? var_o * inv_n : lane == 1 ? var_x * inv_n : mse);
    const double std_o = lane_bcast(r, 0), std_x = lane_bcast(r, 1), rmse = lane_bcast(r, 2);
    // lanes 0-3: coeff, NRMSE, max_err.rel (the reference's exact quotient), pwrrel's quotient
    const double num = lane == 0 ? cor * inv_n : lane == 1 ? rmse : lane == 2 ? max_abs : a.pa;
    const double den = lane == 0 ? std_o * std_x : lane == 1 ? rng_o : lane == 2 ? rng_o : a.pb;
    const double q = num / den;
    // lanes 0, 1: log10 of the range and of the MSE
    const double lg = log10(lane == 0 ? rng_o : mse);
    const double coeff = lane_bcast(q, 0), nrmse = lane_bcast(q, 1), rel = lane_bcast(q, 2), pw = lane_bcast(q, 3);
    const double psnr = 20.0 * lane_bcast(lg, 0) - 10.0 * lane_bcast(lg, 1);
    if (lane != 0) return;

    s->len = n;
    s->odata.max = omax;
    s->odata.min = omin;
    s->odata.rng = rng_o;
    s->odata.std = std_o;
    s->xdata.max = xmax;
    s->xdata.min = xmin;
    s->xdata.rng = xmax - xmin;
    s->xdata.std = std_x;
    s->max_err.idx = max_idx;
    s->max_err.abs = max_abs;
    s->max_err.rel = rel;
    s->max_err.pwrrel = (double)(float)pw;
    s->reduced.coeff = coeff;
    s->reduced.MSE = mse;
    s->reduced.NRMSE = nrmse;
    s->reduced.PSNR = psnr;
}

__global__ __launch_bounds__(kVWG) void k_verify_final(const float* __restrict__ x, const float* __restrict__ o,
                                                       uint64_t n, VParts P, fz_stats* __restrict__ s)
{
    verify_finish(x, o, n, P, s);
}

}  // namespace

size_t verify_workspace_bytes(uint64_t n) { return verify_temp_bytes(verify_grid(n)); }

hipError_t launch_verify(const float* x, const float* o, uint64_t n, void* temp, fz_stats* stats, hipStream_t s)
{
    const uint32_t G = verify_grid(n);
    const VParts   P = verify_parts(temp, G);
    const bool     vec = ((uintptr_t)x % 16) == 0 && ((uintptr_t)o % 16) == 0;
    (void)hipGetLastError();
    if (vec) k_verify_part<true><<<G, kVWG, 0, s>>>(x, o, n, P);
    else k_verify_part<false><<<G, kVWG, 0, s>>>(x, o, n, P);
    k_verify_final<<<1, kVWG, 0, s>>>(x, o, n, P, stats);
    return hipGetLastError();
}

}  // namespace fz
