// fz_outlier.hip -- the outlier mode: the FIXED-mode payload plus a side channel that makes
// every field compressible and every value recoverable.  It covers what the reference cannot
// encode: deltas with |d| >= 32768 (their code aliases the sign bit), non-finite values, values
// whose quantised integer is not exact in binary32 (|x * ebx2_r| >= 2^24) and constant fields
// (the reference divides by a zero range).  The side channel follows cuSZ's sparse
// (index, value) outlier vector (spcodec_vec.hh:29, launch_spv.cuh:22-58), as one byte string:
//
//   u64 n_delta, u64 n_raw,
//   u32 idx_d[n_delta], i32 val_d[n_delta],   exact deltas of the positions whose code is 0
//   u32 idx_r[n_raw],   u32 val_r[n_raw]      input bits of the values kept verbatim
//
// both lists in increasing index order (deterministic: per-unit counts, per-run totals, then
// each run writes its units' entries in order).
//
// Compress:   k_range_part_o     min / max of the finite values, non-finite flag
//             k_tile_compress_o  the fixed mode's paths when no value is kept verbatim (all
//                                finite, |s| < 2^23), redoing a tile with an overflowing delta
//                                on the integer path; the integer path otherwise; per-unit counts
//             k_scan_gather      placement, as in the fused path
//             k_outlier_place    one workgroup per placement run (the tile kernel adds each
//                                unit's counts to its run's pair): the run's offset from the
//                                runs before it, then its units with entries recompute their
//                                deltas and write them; with no entry anywhere it only writes
//                                the header
// Decompress: the plain inverse's codes, then k_put_deltas (the side channel's deltas at their
//             positions, flagged in a bitmap), the reconstruct kernels reading codes and the
//             flagged deltas, k_put_raw (the verbatim values; the bitmap cleared again).  A
//             constant field is a fill.
#include <hip/hip_runtime.h>

#include "fz_internal.hpp"
#include "fz_tile.hpp"

namespace fz {

namespace {

constexpr int kORangeWG = 1024;  // threads per partial-range workgroup (as the fused path)

__global__ __launch_bounds__(kORangeWG) void k_range_part_o(const float* __restrict__ in, Geom g,
                                                            float4* __restrict__ part, uint32_t* run_sums,
                                                            uint32_t nsums, uint32_t* orun, uint32_t norun,
                                                            fz_result* res)
{
    const uint32_t tid = threadIdx.x;
    if (blockIdx.x == 0) {  // clear the run sums, the runs' outlier pairs and the overflow count
        for (uint32_t i = tid; i < nsums; i += kORangeWG) run_sums[i] = 0;
        for (uint32_t i = tid; i < norun; i += kORangeWG) orun[i] = 0;
        if (tid == 0) res->overflow = 0;
    }
    const uint32_t per = (g.len + gridDim.x - 1) / gridDim.x;
    const uint32_t f_lo = min(blockIdx.x * per, g.len), f_hi = min(f_lo + per, g.len);
    range_span<kORangeWG, 8, true>(in, f_lo, f_hi, part + blockIdx.x);
}

template <int NDIM>
__global__ __launch_bounds__(kWG) void k_tile_compress_o(const float* __restrict__ in, Geom g, double eb_rel,
                                                         const float4* __restrict__ part, uint32_t nparts,
                                                         fz_result* res, uint8_t* __restrict__ payload,
                                                         uint8_t* __restrict__ staging, uint32_t* __restrict__ counts,
                                                         uint16_t* __restrict__ tail, uint32_t* run_sums, uint32_t tpw,
                                                         uint32_t* __restrict__ ocnt, uint32_t* orun, uint32_t nruns)
{
    const bool     is_tail = blockIdx.x >= g.T;
    const uint32_t tile = is_tail ? g.T : xcd_tile(blockIdx.x, g.T);
    const int      gp = is_tail ? 0 : geom_path<NDIM>(g, tile * (uint32_t)kTileCodes);
    Partials       P;
    RawTile<NDIM>  R;
    if constexpr (NDIM < 3) {  // the fused kernel's load order (partials first, then the tile)
        partials_load(part, nparts, P);
        __builtin_amdgcn_sched_barrier(0);
        raw_prefetch<NDIM>(in, tile * (uint32_t)kTileCodes, 8 * threadIdx.x, g, gp, R);
    } else {
        raw_prefetch<NDIM>(in, tile * (uint32_t)kTileCodes, 8 * threadIdx.x, g, gp, R);
        partials_load(part, nparts, P);
    }
    const EbParams e = partials_reduce<true>(P, eb_rel, NDIM);
    if (blockIdx.x == 0 && threadIdx.x == 0) write_header(res, e, g);
    if (e.status != ST_SUCCESS) return;  // uniform
    unsigned long long* ovf = reinterpret_cast<unsigned long long*>(&res->overflow);
    uint32_t*           bwords = reinterpret_cast<uint32_t*>(payload);
    if (is_tail) {
        tail_unit_o<NDIM>(in, g, e.r, bwords, tail, ovf, ocnt + 2 * g.T, orun + 2 * (nruns - 1));
        if (g.T == 0 && threadIdx.x == 0) {
            res->nblocks = 0;
            res->payload_bytes = 4ull * g.W;
        }
        return;
    }
    uint8_t* const  slot = staging + (size_t)tile * 4096;
    uint32_t* const sum = run_sums + tile / tpw;
    if constexpr (NDIM == 3) {
        if (gp < 2)
            tile_encode_o<3, 2>(in, g, tile, gp, R, e.r, e.level, bwords, slot, counts + tile, ovf, sum,
                                ocnt + 2 * tile, orun + 2 * (tile / tpw));
        else
            tile_encode_o<3, 1>(in, g, tile, gp, R, e.r, e.level, bwords, slot, counts + tile, ovf, sum,
                                ocnt + 2 * tile, orun + 2 * (tile / tpw));
    } else {
        tile_encode_o<NDIM>(in, g, tile, gp, R, e.r, e.level, bwords, slot, counts + tile, ovf, sum, ocnt + 2 * tile,
                            orun + 2 * (tile / tpw));
    }
}

// exclusive (overflow, verbatim) prefix of a pair over the workgroup (kWG threads), and the totals
__device__ __forceinline__ uint2 block_excl2(uint32_t a, uint32_t b, uint2& total)
{
    __shared__ uint32_t sa[kWG / 64], sb[kWG / 64];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint32_t ia = wave_incl_scan(a), ib = wave_incl_scan(b);
    if (lane == 63) {
        sa[wv] = ia;
        sb[wv] = ib;
    }
    __syncthreads();
    uint32_t pa = 0, pb = 0, ta = 0, tb = 0;
#pragma unroll
    for (int i = 0; i < kWG / 64; i++) {
        pa += (uint32_t)i < wv ? sa[i] : 0u;
        pb += (uint32_t)i < wv ? sb[i] : 0u;
        ta += sa[i];
        tb += sb[i];
    }
    total = make_uint2(ta, tb);
    __syncthreads();  // sa / sb are free for the next call
    return make_uint2(pa + ia - a, pb + ib - b);
}

__device__ __forceinline__ void put_u32(uint8_t* side, uint64_t cap, uint64_t off, uint32_t v)
{
    if (off + 4 <= cap) *reinterpret_cast<uint32_t*>(side + off) = v;
}

// Recompute unit u's deltas (all threads) and write its entries from the list positions kd
// (deltas) and kr (verbatim values).
template <int NDIM>
__device__ __forceinline__ void write_unit(const float* __restrict__ in, const Geom& g, float r, uint32_t u,
                                           uint64_t kd, uint64_t kr, uint64_t nd, uint64_t nr, uint8_t* side,
                                           uint64_t cap)
{
    const uint32_t tid = threadIdx.x;
    const uint32_t t0 = u * (uint32_t)kTileCodes;
    const uint32_t ulen = u == g.T ? g.len - t0 : (uint32_t)kTileCodes;
    Int8           o;
    o.big = o.raw = 0;
    if (8 * tid < ulen) dq8_int<NDIM>(in, t0 + 8 * tid, g, r, o);
    if (8 * tid + 8 > ulen) {  // positions past the unit (the tail unit only)
        const uint32_t keep = 8 * tid < ulen ? (1u << (ulen - 8 * tid)) - 1u : 0u;
        o.big &= keep;
        o.raw &= keep;
    }
    uint2       tot;
    const uint2 pre = block_excl2((uint32_t)__popc(o.big), (uint32_t)__popc(o.raw), tot);
    kd += pre.x;
    kr += pre.y;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const uint32_t p = t0 + 8 * tid + i;
        if ((o.big >> i) & 1u) {
            put_u32(side, cap, 16 + 4 * kd, p);
            put_u32(side, cap, 16 + 4 * nd + 4 * kd, (uint32_t)o.d[i]);
            ++kd;
        }
        if ((o.raw >> i) & 1u) {
            put_u32(side, cap, 16 + 8 * nd + 4 * kr, p);
            put_u32(side, cap, 16 + 8 * nd + 4 * nr + 4 * kr, __float_as_uint(in[p]));
            ++kr;
        }
    }
}

// One workgroup per placement run r (units [r tpw, (r+1) tpw), the tail unit in the last run).
// The runs' (overflow, verbatim) pairs give the run's first list positions and the totals;
// workgroup 0 writes the header and, when the entries exceed the capacity, FZ_ERR_CAPACITY.
// Only runs with entries go on: slices of 256 units are scanned, and each unit with entries is
// recomputed and written in order.
template <int NDIM>
__global__ __launch_bounds__(kWG) void k_outlier_place(const float* __restrict__ in, Geom g, fz_result* res,
                                                       const uint32_t* __restrict__ ocnt,
                                                       const uint32_t* __restrict__ orun, uint32_t nruns, uint32_t tpw,
                                                       uint8_t* __restrict__ side, uint64_t cap)
{
    __shared__ uint32_t sh_u[kWG];
    __shared__ uint2    sh_k[kWG];
    __shared__ uint32_t sh_n;
    const uint32_t r = blockIdx.x, tid = threadIdx.x;
    // the only failure before this kernel: a value range that overflows (no codes were made)
    if (res->status == ST_NONFINITE) return;
    uint32_t bd = 0, br = 0, td = 0, tr = 0;
    for (uint32_t i = tid; i < nruns; i += kWG) {
        const uint2 v = reinterpret_cast<const uint2*>(orun)[i];
        td += v.x;
        tr += v.y;
        bd += i < r ? v.x : 0u;
        br += i < r ? v.y : 0u;
    }
    uint2 before, all;
    (void)block_excl2(bd, br, before);
    (void)block_excl2(td, tr, all);
    const uint64_t nd = all.x, nr = all.y;
    if (r == 0 && tid == 0) {
        if (cap >= 16) {
            reinterpret_cast<uint64_t*>(side)[0] = nd;
            reinterpret_cast<uint64_t*>(side)[1] = nr;
        }
        if (16 + 8 * (nd + nr) > cap) res->status = FZ_ERR_CAPACITY;
    }
    const uint2 mine = reinterpret_cast<const uint2*>(orun)[r];
    if (mine.x == 0 && mine.y == 0) return;  // uniform: no entry in this run
    const float    rq = res->ebx2_r;
    const uint32_t u_lo = r * tpw, u_hi = r + 1 == nruns ? g.T + 1 : min((r + 1) * tpw, g.T + 1);
    uint64_t       kd = before.x, kr = before.y;
    for (uint32_t s0 = u_lo; s0 < u_hi; s0 += kWG) {
        const uint32_t u = s0 + tid;
        const uint2    c = u < u_hi ? reinterpret_cast<const uint2*>(ocnt)[u] : make_uint2(0u, 0u);
        uint2          slice;
        const uint2    ex = block_excl2(c.x, c.y, slice);
        // compact the units with entries (in order) with their first list positions
        const bool     has = (c.x | c.y) != 0;
        uint2          nunits;
        const uint2    at = block_excl2(has ? 1u : 0u, 0u, nunits);
        if (has) {
            sh_u[at.x] = u;
            sh_k[at.x] = ex;
        }
        if (tid == 0) sh_n = nunits.x;
        __syncthreads();
        for (uint32_t k = 0; k < sh_n; k++)
            write_unit<NDIM>(in, g, rq, sh_u[k], kd + sh_k[k].x, kr + sh_k[k].y, nd, nr, side, cap);
        kd += slice.x;
        kr += slice.y;
        __syncthreads();  // sh_u / sh_k / sh_n are read before the next slice overwrites them
    }
}

// ---------------------------------------------------------------------------- inverse

// entry counts of a side channel, clamped to what side_bytes holds (so a short or damaged
// channel never leads to an access outside it)
__device__ __forceinline__ void side_counts(const uint8_t* side, uint64_t side_bytes, uint64_t& nd, uint64_t& nr)
{
    nd = nr = 0;
    if (side_bytes < 16) return;
    const uint64_t avail = (side_bytes - 16) / 8;
    nd = reinterpret_cast<const uint64_t*>(side)[0];
    nr = reinterpret_cast<const uint64_t*>(side)[1];
    nd = nd < avail ? nd : avail;
    nr = nr < avail - nd ? nr : avail - nd;
}

// The side channel's deltas into `over` with their bitmap bits set, in the same launch as the
// clearing of the tile scan's look-back states and ticket (both must precede the scan and the
// reconstruct, and neither depends on the other).
__global__ __launch_bounds__(kWG) void k_put_deltas(const uint8_t* __restrict__ side, uint64_t side_bytes,
                                                    uint32_t* __restrict__ bits, int32_t* __restrict__ over,
                                                    uint32_t n, uint64_t* __restrict__ states, uint32_t nstates,
                                                    uint32_t* __restrict__ ticket)
{
    const uint32_t gtid = blockIdx.x * kWG + threadIdx.x;
    for (uint32_t i = gtid; i < nstates; i += gridDim.x * kWG) states[i] = 0;
    if (gtid == 0) *ticket = 0;
    uint64_t nd, nr;
    side_counts(side, side_bytes, nd, nr);
    const uint32_t* idx = reinterpret_cast<const uint32_t*>(side + 16);
    const uint32_t* val = idx + nd;
    for (uint64_t k = gtid; k < nd; k += (uint64_t)gridDim.x * kWG) {
        const uint32_t p = idx[k];
        if (p < n) {
            over[p] = (int32_t)val[k];
            atomicOr(&bits[p >> 5], 1u << (p & 31));
        }
    }
}

// the verbatim values, and the bitmap bits of the deltas cleared (the workspace's state
// between calls)
__global__ __launch_bounds__(kWG) void k_put_raw(const uint8_t* __restrict__ side, uint64_t side_bytes,
                                                 uint32_t* __restrict__ bits, float* __restrict__ out, uint32_t n)
{
    uint64_t nd, nr;
    side_counts(side, side_bytes, nd, nr);
    const uint32_t* idx_d = reinterpret_cast<const uint32_t*>(side + 16);
    const uint32_t* idx = idx_d + 2 * nd;
    const uint32_t* val = idx + nr;
    const uint64_t  step = (uint64_t)gridDim.x * kWG;
    for (uint64_t k = (uint64_t)blockIdx.x * kWG + threadIdx.x; k < nr; k += step)
        if (idx[k] < n) out[idx[k]] = __uint_as_float(val[k]);
    for (uint64_t k = (uint64_t)blockIdx.x * kWG + threadIdx.x; k < nd; k += step)
        if (idx_d[k] < n) atomicAnd(&bits[idx_d[k] >> 5], ~(1u << (idx_d[k] & 31)));
}

__global__ __launch_bounds__(kWG) void k_fill_f32(float* __restrict__ out, uint32_t n, float v)
{
    const uint32_t i = blockIdx.x * kWG + threadIdx.x;
    if (i < n) out[i] = v;
}

constexpr uint32_t kScatterWG = 256;  // workgroups of the side-channel scatters (grid-stride)

}  // namespace

hipError_t launch_compress_outlier(const float* in, const Geom& g, double eb_rel, float4* part, fz_result* res,
                                   uint8_t* payload, uint16_t* tail, uint8_t* side, uint64_t side_cap,
                                   uint8_t* staging, uint32_t* counts, uint32_t* run_sums, uint32_t nsums,
                                   uint32_t* ocnt, uint32_t* orun, hipStream_t s)
{
    const uint32_t nparts = range_grid(g.len);
    const uint32_t tpw = g.T ? gather_tpw(g.T) : 1u;
    const uint32_t nruns = outlier_runs(g.T);
    (void)hipGetLastError();
    k_range_part_o<<<nparts, kORangeWG, 0, s>>>(in, g, part, run_sums, nsums, orun, 2 * nruns, res);
    const dim3 grid(g.T + 1), blk(kWG);
    switch (g.ndim) {
        case 1: k_tile_compress_o<1><<<grid, blk, 0, s>>>(in, g, eb_rel, part, nparts, res, payload, staging, counts, tail, run_sums, tpw, ocnt, orun, nruns); break;
        case 2: k_tile_compress_o<2><<<grid, blk, 0, s>>>(in, g, eb_rel, part, nparts, res, payload, staging, counts, tail, run_sums, tpw, ocnt, orun, nruns); break;
        default: k_tile_compress_o<3><<<grid, blk, 0, s>>>(in, g, eb_rel, part, nparts, res, payload, staging, counts, tail, run_sums, tpw, ocnt, orun, nruns); break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = launch_scan_gather(counts, g.T, staging, payload, run_sums, res, g.W, s);
    if (e != hipSuccess) return e;
    (void)hipGetLastError();
    switch (g.ndim) {
        case 1: k_outlier_place<1><<<nruns, blk, 0, s>>>(in, g, res, ocnt, orun, nruns, tpw, side, side_cap); break;
        case 2: k_outlier_place<2><<<nruns, blk, 0, s>>>(in, g, res, ocnt, orun, nruns, tpw, side, side_cap); break;
        default: k_outlier_place<3><<<nruns, blk, 0, s>>>(in, g, res, ocnt, orun, nruns, tpw, side, side_cap); break;
    }
    return hipGetLastError();
}

hipError_t launch_put_deltas(const uint8_t* side, uint64_t side_bytes, uint32_t* obits, int32_t* over, const Geom& g,
                             uint64_t* states, uint32_t nstates, uint32_t* ticket, hipStream_t s)
{
    (void)hipGetLastError();
    k_put_deltas<<<kScatterWG, kWG, 0, s>>>(side, side_bytes, obits, over, g.xyz, states, nstates, ticket);
    return hipGetLastError();
}

hipError_t launch_reconstruct_outlier(const uint16_t* codes, uint32_t* obits, int32_t* over, const uint8_t* side,
                                      uint64_t side_bytes, float* out, const Geom& g, float ebx2, hipStream_t s)
{
    hipError_t e = launch_reconstruct_outlier_src(codes, obits, over, out, g, ebx2, s);
    if (e != hipSuccess || side_bytes < 24) return e;  // a channel of 16 header bytes has no entry
    k_put_raw<<<kScatterWG, kWG, 0, s>>>(side, side_bytes, obits, out, g.xyz);
    return hipGetLastError();
}

hipError_t launch_fill_outlier(const uint8_t* side, uint64_t side_bytes, uint32_t* obits, float* out, const Geom& g,
                               float vconst, hipStream_t s)
{
    (void)hipGetLastError();
    k_fill_f32<<<(g.xyz + kWG - 1) / kWG, kWG, 0, s>>>(out, g.xyz, vconst);
    k_put_raw<<<kScatterWG, kWG, 0, s>>>(side, side_bytes, obits, out, g.xyz);
    return hipGetLastError();
}

}  // namespace fz
