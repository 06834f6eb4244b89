// fz_batch.hip -- batched form: many fields of uneven shapes and dimensionality per launch.
//
// The work queue is the global unit space: every field owns T + 1 consecutive units (its
// full tiles, then the unit of the partial tile), and the hardware dispatcher hands units to
// workgroups.  A workgroup reads its field from the per-unit field table (one uniform scalar
// load, so its raw loads issue after a single short round trip), then
// runs the same per-tile bodies as the single-field kernels (fz_tile.hpp), so each field's
// bytes equal those of the per-field path.  Compress:
//
//   k_batch_set          the call's per-field pointers, from kernel arguments (no host copy,
//                        so the whole sequence can be captured into a hipGraph)
//   k_batch_range_part   min / max / non-finite of fixed-size chunks of every field
//   k_batch_range_final  one workgroup per field: eb parameters, result header
//   k_batch_tile<NDIM>   one launch per dimensionality over its fields' units: dual-quant,
//                        bit transpose, flags, in-tile compaction into the unit's staging slot
//   k_batch_scan         one workgroup per field: exclusive scan of its tile counts into the
//                        payload address of each tile's first block; payload size
//   k_batch_gather       one wave per unit: copy the staged blocks to their address
//
// Decompress: k_batch_dscan (per-field scan of the flag popcounts), k_batch_decode (one wave
// per unit: scatter, unshuffle, codes into the unit's slot; a 1D field's unit is also
// reconstructed there, straight to floats), k_batch_recon<2|3> (Lorenzo prefix sums over each
// 2D / 3D field's codes).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "fz_internal.hpp"
#include "fz_tile.hpp"

namespace fz {

template <typename K, typename... A>
static hipError_t launch_pair(K kernel, dim3 grid, dim3 block, hipStream_t s, hipEvent_t* ev, int slot, A... args)
{
    (void)hipGetLastError();
    if (ev && ev[2 * slot] && ev[2 * slot + 1])
        hipExtLaunchKernelGGL(kernel, grid, block, 0, s, ev[2 * slot], ev[2 * slot + 1], 0, args...);
    else kernel<<<grid, block, 0, s>>>(args...);
    return hipGetLastError();
}

// a skipped kernel still closes its event pair (time ~0), so slots keep their meaning
static hipError_t skip_pair(hipStream_t s, hipEvent_t* ev, int slot)
{
    if (!ev || !ev[2 * slot] || !ev[2 * slot + 1]) return hipSuccess;
    hipError_t e = hipEventRecord(ev[2 * slot], s);
    return e == hipSuccess ? hipEventRecord(ev[2 * slot + 1], s) : e;
}

// ---------------------------------------------------------------------------- pointers

__global__ __launch_bounds__(64) void k_batch_set(PtrChunk c, uint32_t first, uint32_t n, BPtrs* __restrict__ ptrs)
{
    const uint32_t i = threadIdx.x;
    if (i < n) ptrs[first + i] = c.p[i];
}

hipError_t launch_batch_set(const BPtrs* sorted, uint32_t nf, BPtrs* dev, hipStream_t s)
{
    for (uint32_t f0 = 0; f0 < nf; f0 += kPtrChunk) {
        PtrChunk       c{};
        const uint32_t n = nf - f0 < (uint32_t)kPtrChunk ? nf - f0 : (uint32_t)kPtrChunk;
        for (uint32_t i = 0; i < n; i++) c.p[i] = sorted[f0 + i];
        (void)hipGetLastError();
        k_batch_set<<<1, 64, 0, s>>>(c, f0, n, dev);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// ---------------------------------------------------------------------------- compress

constexpr int kBRangeWG = 1024;

// partial p_lo + blockIdx.x of the fields [f_lo, f_hi)
__global__ __launch_bounds__(kBRangeWG) void k_batch_range_part(const BField* __restrict__ fields,
                                                                const uint32_t* __restrict__ pstart, uint32_t f_lo,
                                                                uint32_t f_hi, uint32_t p_lo,
                                                                const BPtrs* __restrict__ ptrs, uint32_t chunk,
                                                                float4* __restrict__ part)
{
    const uint32_t b = p_lo + blockIdx.x;
    const uint32_t f = wave_upper(pstart, f_lo, f_hi, b);
    const uint32_t len = fields[f].g.len;
    const uint64_t lo = (uint64_t)(b - pstart[f]) * chunk;
    const uint32_t x_lo = (uint32_t)(lo < len ? lo : len);
    const uint32_t x_hi = (uint32_t)(lo + chunk < len ? lo + chunk : len);
    range_span<kBRangeWG>(static_cast<const float*>(ptrs[f].src), x_lo, x_hi, part + b);
}

__global__ __launch_bounds__(kWG) void k_batch_range_final(BatchDev d, uint32_t f_lo)
{
    const uint32_t f = f_lo + blockIdx.x;
    const uint32_t p0 = d.pstart[f];
    const EbParams e = reduce_partials(d.part + p0, d.pstart[f + 1] - p0, d.ptrs[f].param, d.fields[f].g.ndim);
    if (threadIdx.x == 0) {
        const BField& F = d.fields[f];
        d.prm[f] = BParams{e.r, e.status, (uint32_t)e.level, 0u};
        fz_result* res = d.res + F.caller;
        write_header(res, e, F.g);
        res->overflow = 0;
    }
}

template <int NDIM>
__global__ __launch_bounds__(kWG) void k_batch_tile(BatchDev d, uint32_t u_lo, uint32_t nunits, uint32_t f_lo,
                                                    uint32_t f_hi, uint32_t xcd)
{
    const uint32_t local = xcd ? xcd_tile(blockIdx.x, nunits) : blockIdx.x;
    const uint32_t gu = u_lo + local;
    const uint32_t f = d.ufield[gu];  // uniform: one scalar load
    const BPtrs    q = d.ptrs[f];     // the field's records, both loaded as soon as f is known
    const Geom     g = d.fields[f].g;
    const uint32_t tile = gu - d.fields[f].ustart;
    const bool     is_tail = tile >= g.T;
    const float*   in = static_cast<const float*>(q.src);
    const int      gp = is_tail ? 0 : geom_path<NDIM>(g, tile * (uint32_t)kTileCodes);
    RawTile<NDIM>  R;
    raw_prefetch<NDIM>(in, tile * (uint32_t)kTileCodes, 8 * threadIdx.x, g, gp, R);
    const BParams p = d.prm[f];
    if (p.status != ST_SUCCESS || is_tail) {
        if (threadIdx.x == 0) d.counts[gu] = 0;
        if (p.status != ST_SUCCESS) return;
    }
    unsigned long long* ovf = reinterpret_cast<unsigned long long*>(&d.res[d.fields[f].caller].overflow);
    uint32_t*           bwords = static_cast<uint32_t*>(q.dst);
    if (is_tail) {
        tail_unit<NDIM>(in, g, p.r, bwords, q.tail, ovf);
        return;
    }
    uint8_t* const slot = d.staging + (size_t)gu * 4096;
    if constexpr (NDIM == 3) {  // one branch per path family: see tile_encode
        if (gp < 2) tile_encode<3, 2>(in, g, tile, gp, R, p.r, (int)p.level, bwords, slot, d.counts + gu, ovf);
        else tile_encode<3, 1>(in, g, tile, gp, R, p.r, (int)p.level, bwords, slot, d.counts + gu, ovf);
    } else {
        tile_encode<NDIM>(in, g, tile, gp, R, p.r, (int)p.level, bwords, slot, d.counts + gu, ovf);
    }
}

// One wave per unit, four units per workgroup, no workgroup barrier: every wave finds its field,
// issues its first quarter's loads, then encodes its tile (tile_encode_w) or the tail unit.
template <int NDIM>
__global__ __launch_bounds__(kWG) void k_batch_tile_w(BatchDev d, uint32_t u_lo, uint32_t nunits, uint32_t f_lo,
                                                      uint32_t f_hi, uint32_t xcd)
{
    __shared__ uint32_t sh[4][32 * kLdsPitch];
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wave
    const uint32_t ngroups = (nunits + 3) / 4;
    const uint32_t local = 4 * (xcd ? xcd_tile(blockIdx.x, ngroups) : blockIdx.x) + wv;
    if (local >= nunits) return;
    const uint32_t gu = u_lo + local;
    const uint32_t f = d.ufield[gu];  // uniform: one scalar load
    const BPtrs    q = d.ptrs[f];     // the field's records, both loaded as soon as f is known
    const Geom     g = d.fields[f].g;
    const uint32_t tile = gu - d.fields[f].ustart;
    const bool     is_tail = tile >= g.T;
    const float*   in = static_cast<const float*>(q.src);
    const int      gp = is_tail ? 0 : geom_path<NDIM>(g, tile * (uint32_t)kTileCodes);
    RawTile<NDIM>  R[2];
    if (gp >= 2) raw_prefetch<NDIM>(in, tile * (uint32_t)kTileCodes, 8 * lane_id(), g, gp, R[0]);
    const BParams p = d.prm[f];
    if (p.status != ST_SUCCESS || is_tail) {
        if (lane_id() == 0) d.counts[gu] = 0;
        if (p.status != ST_SUCCESS) return;
    }
    unsigned long long* ovf = reinterpret_cast<unsigned long long*>(&d.res[d.fields[f].caller].overflow);
    uint32_t*           bwords = static_cast<uint32_t*>(q.dst);
    if (is_tail) {
        tail_unit_w<NDIM>(in, g, p.r, bwords, q.tail, ovf);
        return;
    }
    tile_encode_w<NDIM>(in, g, tile, gp, R, p.r, (int)p.level, bwords, d.staging + (size_t)gu * 4096, d.counts + gu,
                        ovf, sh[wv]);
}

// One workgroup per field: exclusive scan of its tile counts (16 per thread per round) into
// the payload address of every tile's first block: payload + 4 W + 16 * (blocks before it)
// (encodeDebug's placement, fz.cu:56-57); then the field's block count and payload size.
constexpr int kBScanWG = 1024;
constexpr int kBScanPer = 16;

__global__ __launch_bounds__(kBScanWG) void k_batch_scan(BatchDev d, uint32_t f_lo)
{
    __shared__ uint32_t sh_w[kBScanWG / 64];
    const uint32_t f = f_lo + blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (d.prm[f].status != ST_SUCCESS) return;
    const Geom     g = d.fields[f].g;
    const uint32_t T = g.T, u0 = d.ustart[f];
    uint8_t*       base = static_cast<uint8_t*>(d.ptrs[f].dst) + 4ull * g.W;
    uint32_t       running = 0;
    for (uint32_t r0 = 0; r0 < T; r0 += kBScanWG * kBScanPer) {
        const uint32_t t0 = r0 + kBScanPer * tid;
        uint32_t       c[kBScanPer];
#pragma unroll
        for (int j = 0; j < kBScanPer; j++) c[j] = t0 + j < T ? d.counts[u0 + t0 + j] : 0u;
        uint32_t tsum = 0;
#pragma unroll
        for (int j = 0; j < kBScanPer; j++) {
            const uint32_t x = c[j];
            c[j] = tsum;
            tsum += x;
        }
        const uint32_t incl = wave_incl_scan(tsum);
        if (lane == 63) sh_w[wv] = incl;
        __syncthreads();
        uint32_t pre = 0, tot = 0;
#pragma unroll
        for (int i = 0; i < kBScanWG / 64; i++) {
            const uint32_t v = sh_w[i];
            pre += (uint32_t)i < wv ? v : 0u;
            tot += v;
        }
        const uint32_t off = running + pre + incl - tsum;
#pragma unroll
        for (int j = 0; j < kBScanPer; j++)
            if (t0 + j < T) d.dsts[u0 + t0 + j] = (uint64_t)(uintptr_t)(base + 16ull * (off + c[j]));
        running += tot;
        __syncthreads();  // sh_w is rewritten by the next round
    }
    if (tid == 0) {
        fz_result* res = d.res + d.fields[f].caller;
        res->nblocks = running;
        res->payload_bytes = 4ull * g.W + 16ull * running;
    }
}

// One wave per unit of [u_lo, u_hi): its count (<= 256) of staged 16-byte blocks go to dsts[unit].
__global__ __launch_bounds__(1024) void k_batch_gather(const uint32_t* __restrict__ counts,
                                                       const uint64_t* __restrict__ dsts,
                                                       const uint8_t* __restrict__ staging, uint32_t u_lo,
                                                       uint32_t u_hi)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t unit = u_lo + blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (unit >= u_hi) return;
    const uint32_t c = counts[unit];
    if (c == 0) return;
    uint8_t*       dst = reinterpret_cast<uint8_t*>((uintptr_t)dsts[unit]);
    const uint8_t* src = staging + (size_t)unit * 4096;
    uint4          v[4];
#pragma unroll
    for (int u = 0; u < 4; u++)
        if (lane + 64 * u < c) v[u] = *reinterpret_cast<const uint4*>(src + 16 * (lane + 64 * u));
#pragma unroll
    for (int u = 0; u < 4; u++)
        if (lane + 64 * u < c) st16(dst + 16 * (lane + 64 * u), v[u]);
}

static hipError_t group_range(const BatchDev& d, const BatchShape& b, const BatchGroup& G, int c, hipEvent_t* ev,
                              hipStream_t s)
{
    hipError_t e = launch_pair(k_batch_range_part, dim3(G.p_hi[c] - G.p_lo[c]), dim3(kBRangeWG), s, ev, 0, d.fields,
                               d.pstart, G.f_lo, G.f_hi, G.p_lo[c], (const BPtrs*)d.ptrs, b.chunk, d.part);
    if (e == hipSuccess) e = launch_pair(k_batch_range_final, dim3(G.f_hi - G.f_lo), dim3(kWG), s, ev, 1, d, G.f_lo);
    return e;
}

static hipError_t group_tiles(const BatchDev& d, const BatchShape& b, const BatchGroup& G, hipEvent_t* ev,
                              hipStream_t s)
{
    hipError_t e = hipSuccess;
    for (int nd = 1; nd <= 3 && e == hipSuccess; nd++) {
        const uint32_t n = G.du_hi[nd] - G.du_lo[nd];
        if (n == 0) {
            e = skip_pair(s, ev, 1 + nd);
            continue;
        }
        const uint32_t ul = G.du_lo[nd], fl = G.df_lo[nd], fh = G.df_hi[nd], x = b.xcd[nd];
        const dim3     blk(kWG);
        if (b.wave_tile[nd]) {
            const dim3 grid((n + 3) / 4);
            if (nd == 1) e = launch_pair(k_batch_tile_w<1>, grid, blk, s, ev, 2, d, ul, n, fl, fh, x);
            else if (nd == 2) e = launch_pair(k_batch_tile_w<2>, grid, blk, s, ev, 3, d, ul, n, fl, fh, x);
            else e = launch_pair(k_batch_tile_w<3>, grid, blk, s, ev, 4, d, ul, n, fl, fh, x);
        } else {
            const dim3 grid(n);
            if (nd == 1) e = launch_pair(k_batch_tile<1>, grid, blk, s, ev, 2, d, ul, n, fl, fh, x);
            else if (nd == 2) e = launch_pair(k_batch_tile<2>, grid, blk, s, ev, 3, d, ul, n, fl, fh, x);
            else e = launch_pair(k_batch_tile<3>, grid, blk, s, ev, 4, d, ul, n, fl, fh, x);
        }
    }
    return e;
}

static hipError_t group_place(const BatchDev& d, const BatchShape& b, const BatchGroup& G, hipEvent_t* ev,
                              hipStream_t s)
{
    hipError_t e = launch_pair(k_batch_scan, dim3(G.f_hi - G.f_lo), dim3(kBScanWG), s, ev, 5, d, G.f_lo);
    if (e == hipSuccess) {
        const uint32_t gw = b.gather_waves, n = G.u_hi - G.u_lo;
        e = launch_pair(k_batch_gather, dim3((n + gw - 1) / gw), dim3(64 * gw), s, ev, 6, (const uint32_t*)d.counts,
                        (const uint64_t*)d.dsts, (const uint8_t*)d.staging, G.u_lo, G.u_hi);
    }
    return e;
}

hipError_t launch_batch_compress(const BatchDev& d, const BatchShape& b, const BatchGroup* groups, int ngroups,
                                 int chunk, const BatchStreams* st, hipEvent_t* ev, hipStream_t s)
{
    hipError_t e = hipSuccess;
    if (!st) {
        for (int g = 0; g < ngroups && e == hipSuccess; g++) {
            e = group_range(d, b, groups[g], chunk, ev, s);
            if (e == hipSuccess) e = group_tiles(d, b, groups[g], ev, s);
            if (e == hipSuccess) e = group_place(d, b, groups[g], ev, s);
        }
        return e;
    }
    // range stream: every group's value range; caller's stream: tiles of group g once its
    // range is known; gather stream: scan + placement of group g once its tiles are done
#define FZ_CK(x)                         \
    do {                                 \
        e = (x);                         \
        if (e != hipSuccess) return e;   \
    } while (0)
    FZ_CK(hipEventRecord(st->fork, s));
    FZ_CK(hipStreamWaitEvent(st->range, st->fork, 0));
    FZ_CK(hipStreamWaitEvent(st->gather, st->fork, 0));
    for (int g = 0; g < ngroups; g++) {
        FZ_CK(group_range(d, b, groups[g], chunk, nullptr, st->range));
        FZ_CK(hipEventRecord(st->r[g], st->range));
    }
    for (int g = 0; g < ngroups; g++) {
        FZ_CK(hipStreamWaitEvent(s, st->r[g], 0));
        FZ_CK(group_tiles(d, b, groups[g], nullptr, s));
        FZ_CK(hipEventRecord(st->t[g], s));
    }
    for (int g = 0; g < ngroups; g++) {
        FZ_CK(hipStreamWaitEvent(st->gather, st->t[g], 0));
        FZ_CK(group_place(d, b, groups[g], nullptr, st->gather));
    }
    FZ_CK(hipEventRecord(st->join_r, st->range));
    FZ_CK(hipEventRecord(st->join_g, st->gather));
    FZ_CK(hipStreamWaitEvent(s, st->join_r, 0));
    FZ_CK(hipStreamWaitEvent(s, st->join_g, 0));
#undef FZ_CK
    return hipSuccess;
}

// ---------------------------------------------------------------------------- decompress

// One workgroup per field: texcl[unit of tile t] = set flags in the field's tiles < t
// (popcount of bitflag words 8t .. 8t+7), 4 tiles per thread per round.
constexpr int kBDscanPer = 4;

__global__ __launch_bounds__(kBScanWG) void k_batch_dscan(BatchDev d)
{
    __shared__ uint32_t sh_w[kBScanWG / 64];
    const uint32_t f = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint32_t T = d.fields[f].g.T, u0 = d.ustart[f];
    const uint4*   fl = static_cast<const uint4*>(d.ptrs[f].src);  // 2 uint4 per tile
    uint32_t       running = 0;
    for (uint32_t r0 = 0; r0 < T; r0 += kBScanWG * kBDscanPer) {
        const uint32_t t0 = r0 + kBDscanPer * tid;
        uint32_t       c[kBDscanPer];
#pragma unroll
        for (int j = 0; j < kBDscanPer; j++) {
            c[j] = 0;
            if (t0 + j < T) {
                const uint4 a = fl[2 * (size_t)(t0 + j)], b2 = fl[2 * (size_t)(t0 + j) + 1];
                c[j] = __popc(a.x) + __popc(a.y) + __popc(a.z) + __popc(a.w) + __popc(b2.x) + __popc(b2.y) +
                       __popc(b2.z) + __popc(b2.w);
            }
        }
        uint32_t tsum = 0;
#pragma unroll
        for (int j = 0; j < kBDscanPer; j++) {
            const uint32_t x = c[j];
            c[j] = tsum;
            tsum += x;
        }
        const uint32_t incl = wave_incl_scan(tsum);
        if (lane == 63) sh_w[wv] = incl;
        __syncthreads();
        uint32_t pre = 0, tot = 0;
#pragma unroll
        for (int i = 0; i < kBScanWG / 64; i++) {
            const uint32_t v = sh_w[i];
            pre += (uint32_t)i < wv ? v : 0u;
            tot += v;
        }
        const uint32_t off = running + pre + incl - tsum;
#pragma unroll
        for (int j = 0; j < kBDscanPer; j++)
            if (t0 + j < T) d.texcl[u0 + t0 + j] = off + c[j];
        running += tot;
        __syncthreads();
    }
}

// One wave per unit: the tile's codes (or the tail codes) into the unit's 4096-byte slot;
// a field's slots are consecutive, so its codes form one array.  A 1D field's unit is
// reconstructed here as well, straight to floats (its Lorenzo blocks lie inside the unit), so
// no reconstruct kernel runs for 1D fields.
__global__ __launch_bounds__(64) void k_batch_decode(BatchDev d, uint32_t nf)
{
    const uint32_t gu = blockIdx.x;
    const uint32_t f = d.ufield[gu];
    const Geom     g = d.fields[f].g;
    const uint32_t u0 = d.ustart[f], tile = gu - u0;
    const BPtrs    q = d.ptrs[f];
    if (d.fields[f].wnt) return;  // decoded by a fused window kernel (uniform)
    if (g.ndim == 1) {  // uniform
        decode_recon_1d_w(static_cast<const uint8_t*>(q.src), g.mode == 0 ? q.tail : nullptr, g, tile,
                          tile < g.T ? d.texcl[gu] : 0u, (float)q.param, static_cast<float*>(q.dst));
        return;
    }
    if (tile >= g.T) {
        decode_tail(g.mode == 0 ? q.tail : nullptr, g, reinterpret_cast<uint16_t*>(d.staging + (size_t)u0 * 4096));
        return;
    }
    decode_tile_w(static_cast<const uint8_t*>(q.src), g.W, tile, d.texcl[gu],
                  reinterpret_cast<uint32_t*>(d.staging + (size_t)gu * 4096));
}

template <int NDIM>
__global__ __launch_bounds__(kWG) void k_batch_recon(BatchDev d, uint32_t r_lo)
{
    // 3D: XCD-sliced order, so that the workgroups of neighbouring narrow row segments (down
    // to 32 bytes) share one L2 and fetch each line once
    const uint32_t wg = r_lo + (NDIM == 3 ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x);
    const uint32_t f = d.rfield[wg];  // uniform: one scalar load (a search would chain two)
    const BField&  F = d.fields[f];
    const Geom     g = F.g;
    const uint32_t local = wg - d.rstart[f];
    const BPtrs    q = d.ptrs[f];
    const uint16_t* codes = reinterpret_cast<const uint16_t*>(d.staging + (size_t)F.ustart * 4096);
    float*          out = static_cast<float*>(q.dst);
    const float     ebx2 = (float)q.param;
    static_assert(NDIM == 2 || NDIM == 3, "1D fields are reconstructed by k_batch_decode");
    if constexpr (NDIM == 2) {
        const uint32_t rx = F.rdims[0];
        recon_2d(codes, out, g, ebx2, local % rx, local / rx);
    } else {
        recon_3d_any((int)F.rq, codes, out, g, ebx2, local, F.rdims[0], F.rdims[1]);
    }
}

hipError_t launch_batch_decompress(const BatchDev& d, const BatchShape& b, hipEvent_t* ev, hipStream_t s)
{
    hipError_t e = launch_pair(k_batch_dscan, dim3(b.nf), dim3(kBScanWG), s, ev, 0, d);
    if (e == hipSuccess) e = launch_pair(k_batch_decode, dim3(b.U), dim3(64), s, ev, 1, d, b.nf);
    for (int nd = 1; nd <= 3 && e == hipSuccess; nd++) {
        const uint32_t n = b.r_hi[nd] - b.r_lo[nd];
        if (n == 0) {
            e = skip_pair(s, ev, 1 + nd);
            continue;
        }
        if (nd == 1)
            e = skip_pair(s, ev, 2);  // reconstructed by k_batch_decode
        else if (nd == 2)
            e = launch_pair(k_batch_recon<2>, dim3(n), dim3(kWG), s, ev, 3, d, b.r_lo[2]);
        else
            e = launch_pair(k_batch_recon<3>, dim3(n), dim3(kWG), s, ev, 4, d, b.r_lo[3]);
    }
    // the 3D fields whose windows decode and reconstruct in one kernel (one launch per window
    // size), then the 2D fields whose rows are a multiple of 4 wide (one launch)
    const bool timed = ev && ev[10] && ev[11];
    if (e == hipSuccess && timed) e = hipEventRecord(ev[10], s);
    if (e == hipSuccess) e = launch_batch_fused3d(d, b, s);
    if (e == hipSuccess) e = launch_batch_fused2d(d, b, s);
    if (e == hipSuccess && timed) e = hipEventRecord(ev[11], s);
    return e;
}

}  // namespace fz
