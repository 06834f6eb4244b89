// fz_kernels.hip -- gfx950 kernels of the FZ path.
//
//   k_range               value range + eb parameters (fz.cu:137, 149; lorenzo_var.cuh:562-563)
//   k_tile_compress       dual-quant -> bitshuffle -> flags -> in-tile compaction, one tile per
//                         workgroup, no cross-workgroup waiting
//   k_scan_gather         scan of the tile counts + placement of the staged blocks
//   k_dualquant           plain Lorenzo stage (lorenzo_var.cuh:176-295)
//   k_bitshuffle          plain bitshuffle + flags stage (fz.cu:61-111)
//   k_scan_init / k_scan  exclusive u32 sum, single pass (the cub::DeviceScan call, fz.cu:188-192)
//   k_encode              compaction of flagged 16-byte blocks (fz.cu:49-59)
//   k_tile_scan, k_decode_unshuffle, k_reconstruct_*   the inverse
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "fz_internal.hpp"
#include "fz_tile.hpp"

namespace fz {

// A HIP error left in the thread's last-error slot by earlier, unrelated runtime calls must
// not be reported as a failure of our launch: clear it before launching, read it after.
#define FZ_LAUNCH(...)                     \
    do {                                   \
        (void)hipGetLastError();           \
        __VA_ARGS__;                       \
    } while (0)

// Launch with dispatch-timestamped start / stop events when both are given (per-kernel
// timing without extra packets between kernels), plain launch otherwise.
template <typename K, typename... A>
static hipError_t launch_timed(K kernel, dim3 grid, dim3 block, hipStream_t s, hipEvent_t t0, hipEvent_t t1, A... args)
{
    (void)hipGetLastError();
    if (t0 && t1) hipExtLaunchKernelGGL(kernel, grid, block, 0, s, t0, t1, 0, args...);
    else kernel<<<grid, block, 0, s>>>(args...);
    return hipGetLastError();
}

// ============================================================================ range

// Value range in two kernels: k_range_part (per-workgroup min / max / non-finite flag, plain
// stores, no fences) and k_range_final (one workgroup: reduce the partials, derive the eb
// parameters, clear the look-back state of the kernels that follow).  The kernel boundary
// orders the partial stores before the final reads; an in-kernel last-workgroup scheme would
// need a device-scope release per workgroup, which writes back L2 on this part.
//
//   eb_abs = eb_rel * (double)(max - min);  ebx2 = eb_abs * 2;  ebx2_r = 1 / ebx2  (double),
//   both narrowed to float (fz.cu:137, 149; lorenzo_var.cuh:562-563).
//
// Workgroup b covers the b/8-th of the 128 slices of region b%8, where region x is the
// x-th eighth of the field in the XCD-aware tile order of k_tile_compress: workgroups are
// placed round-robin over the 8 XCDs, so the slices an XCD reduces are the tiles it then
// compresses (a placement hint only; any placement gives the same result).

constexpr int kRangeWG = 1024;  // threads per partial-range workgroup

__global__ __launch_bounds__(kRangeWG) void k_range_part(const float* __restrict__ in, Geom g,
                                                         float4* __restrict__ part, uint64_t* zero_states,
                                                         uint32_t nzero, fz_result* zero_ovf)
{
    const uint32_t tid = threadIdx.x;
    if (blockIdx.x == 0) {  // clear the look-back state and overflow count of the kernels that follow
        for (uint32_t i = tid; i < nzero; i += kRangeWG) zero_states[i] = 0;
        if (tid == 0 && zero_ovf) zero_ovf->overflow = 0;
    }
    // this workgroup's span: slice (b >> 3) of region (b & 7), regions in tile units
    const uint32_t nunits = g.T + 1;  // the last unit is the partial tile
    const uint32_t x = blockIdx.x & 7u, i = blockIdx.x >> 3, per = gridDim.x >> 3;
    const uint32_t r_lo = region_start(x, nunits), r_hi = region_start(x + 1, nunits);
    const uint32_t u_lo = r_lo + (uint32_t)(((uint64_t)(r_hi - r_lo) * i) / per);
    const uint32_t u_hi = r_lo + (uint32_t)(((uint64_t)(r_hi - r_lo) * (i + 1)) / per);
    const uint32_t f_lo = min(u_lo * (uint32_t)kTileCodes, g.len), f_hi = min(u_hi * (uint32_t)kTileCodes, g.len);
    range_span<kRangeWG, 8>(in, f_lo, f_hi, part + blockIdx.x);
}

__global__ __launch_bounds__(kWG) void k_range_final(const float4* __restrict__ part, uint32_t nparts, Geom g,
                                                     double eb_rel, fz_result* res, uint64_t* zero_states,
                                                     uint32_t nzero, uint32_t* zero_ticket)
{
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < nzero; i += kWG) zero_states[i] = 0;
    if (tid == 0 && zero_ticket) *zero_ticket = 0;
    const EbParams e = reduce_partials(part, nparts, eb_rel, g.ndim);
    if (tid == 0) {
        write_header(res, e, g);
        res->overflow = 0;
    }
}

hipError_t launch_range(const float* in, const Geom& g, double eb_rel, float4* part, fz_result* res,
                        uint64_t* zero_states, uint32_t nzero, uint32_t* zero_ticket, hipStream_t s)
{
    const uint32_t nparts = range_grid(g.len);
    FZ_LAUNCH(k_range_part<<<nparts, kRangeWG, 0, s>>>(in, g, part, nullptr, 0, nullptr));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    FZ_LAUNCH(k_range_final<<<1, kWG, 0, s>>>(part, nparts, g, eb_rel, res, zero_states, nzero, zero_ticket));
    return hipGetLastError();
}

// ============================================================================ fused compress
//
// Two kernels after the range kernel, no cross-workgroup waiting in the heavy one:
//   k_tile_compress  one 2048-code tile per workgroup: dual-quant, 32x32 bit transpose, block
//                    flags (bitflag words go straight to the payload), the tile's nonzero
//                    blocks compacted in order into its own 4 KB staging slot, its block count;
//   k_scan_gather    exclusive scan of the tile counts (256 tiles per workgroup, decoupled
//                    look-back across workgroups) and the copy of every staged block to
//                    payload + 4 W + 16 * (global rank)  (encodeDebug's placement, fz.cu:56-57).

template <int NDIM>
__global__ __launch_bounds__(kWG) void k_tile_compress(const float* __restrict__ in, Geom g, double eb_rel,
                                                       const float4* __restrict__ part, uint32_t nparts,
                                                       fz_result* res, uint8_t* __restrict__ payload,
                                                       uint8_t* __restrict__ staging, uint32_t* __restrict__ counts,
                                                       uint16_t* __restrict__ tail, uint32_t* run_sums, uint32_t tpw)
{
    const bool is_tail = blockIdx.x >= g.T;
    // XCD-aware order: an XCD walks its eighth of the field
    const uint32_t tile = is_tail ? g.T : xcd_tile(blockIdx.x, g.T);
    const int      gp = is_tail ? 0 : geom_path<NDIM>(g, tile * (uint32_t)kTileCodes);
    // 1D / 2D: wave 0's partial loads first, then the tile's raw loads (they need no eb), so
    // the partials return first and the eb parameters are computed while the tile is in
    // flight.  3D issues the raw loads first: holding the partials across them would cost it
    // a wave per SIMD (86 against 69 VGPRs).
    Partials      P;
    RawTile<NDIM> R;
    if constexpr (NDIM < 3) {
        partials_load(part, nparts, P);
        __builtin_amdgcn_sched_barrier(0);
        raw_prefetch<NDIM>(in, tile * (uint32_t)kTileCodes, 8 * threadIdx.x, g, gp, R);
    } else {
        raw_prefetch<NDIM>(in, tile * (uint32_t)kTileCodes, 8 * threadIdx.x, g, gp, R);
        partials_load(part, nparts, P);
    }
    const EbParams e = partials_reduce(P, eb_rel, NDIM);
    if (blockIdx.x == 0 && threadIdx.x == 0) write_header(res, e, g);
    if (e.status != ST_SUCCESS) return;  // uniform
    unsigned long long* ovf = reinterpret_cast<unsigned long long*>(&res->overflow);
    uint32_t*           bwords = reinterpret_cast<uint32_t*>(payload);
    if (is_tail) {
        tail_unit<NDIM>(in, g, e.r, bwords, tail, ovf);
        if (g.T == 0 && threadIdx.x == 0) {
            res->nblocks = 0;
            res->payload_bytes = 4ull * g.W;
        }
        return;
    }
    uint8_t* const  slot = staging + (size_t)tile * 4096;
    uint32_t* const sum = run_sums + tile / tpw;
    if constexpr (NDIM == 3) {  // one branch per path family: see tile_encode
        if (gp < 2) tile_encode<3, 2>(in, g, tile, gp, R, e.r, e.level, bwords, slot, counts + tile, ovf, sum);
        else tile_encode<3, 1>(in, g, tile, gp, R, e.r, e.level, bwords, slot, counts + tile, ovf, sum);
    } else {
        tile_encode<NDIM>(in, g, tile, gp, R, e.r, e.level, bwords, slot, counts + tile, ovf, sum);
    }
}


// One wave per tile: a workgroup carries four consecutive units (the tail unit included) and
// reduces the range partials once for them; the waves never wait for each other afterwards.
template <int NDIM>
__global__ __launch_bounds__(kWG) void k_tile_compress_w(const float* __restrict__ in, Geom g, double eb_rel,
                                                         const float4* __restrict__ part, uint32_t nparts,
                                                         fz_result* res, uint8_t* __restrict__ payload,
                                                         uint8_t* __restrict__ staging, uint32_t* __restrict__ counts,
                                                         uint16_t* __restrict__ tail, uint32_t* run_sums, uint32_t tpw)
{
    __shared__ uint32_t sh[4][32 * kLdsPitch];
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wave
    const uint32_t unit = 4 * xcd_tile(blockIdx.x, (g.T + 4) / 4) + wv;  // units 0 .. T (T = the tail unit)
    const bool     is_tile = unit < g.T;
    const int      gp = is_tile ? geom_path<NDIM>(g, unit * (uint32_t)kTileCodes) : 0;
    RawTile<NDIM> R[2];
    if (gp >= 2) raw_prefetch<NDIM>(in, unit * (uint32_t)kTileCodes, 8 * lane_id(), g, gp, R[0]);
    Partials P;
    partials_load(part, nparts, P);
    const EbParams e = partials_reduce(P, eb_rel, NDIM);
    if (blockIdx.x == 0 && threadIdx.x == 0) write_header(res, e, g);
    if (e.status != ST_SUCCESS || unit > g.T) return;  // uniform per wave
    unsigned long long* ovf = reinterpret_cast<unsigned long long*>(&res->overflow);
    uint32_t*           bwords = reinterpret_cast<uint32_t*>(payload);
    if (!is_tile) {
        tail_unit_w<NDIM>(in, g, e.r, bwords, tail, ovf);
        if (g.T == 0 && lane_id() == 0) {
            res->nblocks = 0;
            res->payload_bytes = 4ull * g.W;
        }
        return;
    }
    tile_encode_w<NDIM>(in, g, unit, gp, R, e.r, e.level, bwords, staging + (size_t)unit * 4096, counts + unit, ovf,
                        sh[wv], run_sums + unit / tpw);
}

// Placement of the staged blocks.  Workgroup b owns the run of tiles [b tpw, (b+1) tpw) (at
// most kGatherWG runs); the tile kernel has added each tile's block count to run_sums[run], so
// the blocks before run b are the sum of run_sums[0, b): one load per thread, no chain between
// workgroups.  Within the run, a scan of the tile counts gives each staged block its rank.
constexpr int kGatherThreads = 1024;  // threads per placement workgroup
constexpr int kGW = kGatherThreads / 64;

__global__ __launch_bounds__(kGatherThreads) void k_scan_gather(const uint32_t* __restrict__ counts, uint32_t T,
                                                                uint32_t tpw, const uint8_t* __restrict__ staging,
                                                                uint8_t* __restrict__ payload,
                                                                const uint32_t* __restrict__ run_sums, fz_result* res,
                                                                uint32_t W)
{
    __shared__ uint32_t sh_off[kGatherThreads];
    __shared__ uint32_t sh_w[kGW], sh_b[kGW];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint32_t wg = blockIdx.x;
    const uint32_t t_lo = wg * tpw, t_hi = min(T, t_lo + tpw);
    // exclusive prefix over the waves' values in sh (in every thread), and their total
    auto wave_prefix = [&](const uint32_t* sh, uint32_t& total) {
        uint32_t pre = 0, tot = 0;
#pragma unroll
        for (int i = 0; i < kGW; i++) {
            const uint32_t v = sh[i];
            pre += (uint32_t)i < wv ? v : 0u;
            tot += v;
        }
        total = tot;
        return pre;
    };
    // first run of up to 1024 tiles and the blocks of the runs before this one, loaded before
    // the status is known (both buffers are always valid memory)
    uint32_t cnt = t_lo + tid < t_hi ? counts[t_lo + tid] : 0u;
    const uint32_t before = tid < wg ? run_sums[tid] : 0u;  // wg < kGatherWG <= kGatherThreads
    if (res->status != ST_SUCCESS) return;
    const uint32_t b = wave_sum(before);
    if (lane == 0) sh_b[wv] = b;
    uint32_t excl = 0, base = 0;
    for (uint32_t r0 = t_lo; r0 < t_hi; r0 += kGatherThreads) {
        if (r0 != t_lo) {
            __syncthreads();  // the previous run's sh_w / sh_off reads are done
            cnt = r0 + tid < t_hi ? counts[r0 + tid] : 0u;
        }
        const uint32_t incl = wave_incl_scan(cnt);
        if (lane == 63) sh_w[wv] = incl;
        __syncthreads();
        if (r0 == t_lo) {
            uint32_t e;
            (void)wave_prefix(sh_b, e);
            excl = e;
            if (wg == gridDim.x - 1 && tid == 0) {
                const uint64_t K = (uint64_t)excl + run_sums[wg];
                res->nblocks = K;
                res->payload_bytes = 4ull * W + 16ull * K;
            }
        }
        uint32_t run;
        sh_off[tid] = wave_prefix(sh_w, run) + incl - cnt;
        __syncthreads();
        uint8_t* out = payload + 4ull * W + 16ull * ((uint64_t)excl + base);
        // tiles of this pass (>= 1) and the halvings that narrow [0, ntl) to one tile
        const uint32_t ntl = min((uint32_t)kGatherThreads, t_hi - r0);
        const int      steps = 32 - __clz(ntl - 1);
        for (uint32_t j0 = 0; j0 < run; j0 += 4 * kGatherThreads) {
            // last tile of the pass whose first block is <= j, four searches in step
            uint32_t lo[4], hi[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                lo[u] = 0;
                hi[u] = ntl - 1;
            }
            for (int it = 0; it < steps; it++) {
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const uint32_t j = j0 + tid + u * kGatherThreads, mid = (lo[u] + hi[u] + 1) >> 1;
                    if (sh_off[mid] <= j) lo[u] = mid;
                    else hi[u] = mid - 1;
                }
            }
            uint64_t src[4];  // staging byte offsets pass 2^32 beyond 2^20 tiles
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const uint32_t j = j0 + tid + u * kGatherThreads;
                src[u] = (uint64_t)(r0 + lo[u]) * 4096u + 16u * (j - sh_off[lo[u]]);
            }
            uint4 v[4];
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (j0 + tid + u * kGatherThreads < run) v[u] = *reinterpret_cast<const uint4*>(staging + src[u]);
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const uint32_t j = j0 + tid + u * kGatherThreads;
                if (j < run) st16(out + 16ull * j, v[u]);
            }
        }
        base += run;
    }
}

uint32_t gather_tpw(uint32_t T) { return (T + kGatherWG - 1) / kGatherWG; }

uint32_t gather_grid(uint32_t T) { return T == 0 ? 0u : (T + gather_tpw(T) - 1) / gather_tpw(T); }

hipError_t launch_scan_gather(const uint32_t* counts, uint32_t T, const uint8_t* staging, uint8_t* payload,
                              const uint32_t* run_sums, fz_result* res, uint32_t W, hipStream_t s)
{
    if (T == 0) return hipSuccess;
    FZ_LAUNCH(k_scan_gather<<<gather_grid(T), kGatherThreads, 0, s>>>(counts, T, gather_tpw(T), staging, payload,
                                                                        run_sums, res, W));
    return hipGetLastError();
}

// Fused compress: range partials, tiles, scan + placement.  ev (may be null) holds start /
// stop events for the three kernels.
hipError_t launch_compress(const float* in, const Geom& g, double eb_rel, float4* part, fz_result* res,
                           uint8_t* payload, uint16_t* tail, uint8_t* staging, uint32_t* counts, uint64_t* states,
                           int wave_tile, hipEvent_t* ev, hipStream_t s)
{
    hipEvent_t none[6] = {};
    if (!ev) ev = none;
    const uint32_t nparts = range_grid(g.len);
    const uint32_t G = gather_grid(g.T);
    hipError_t e = launch_timed(k_range_part, dim3(nparts), dim3(kRangeWG), s, ev[0], ev[1], in, g, part, states,
                                G + 1, res);
    if (e != hipSuccess) return e;
    // per-run block sums of the placement kernel, in the look-back state words k_range_part
    // clears (G + 1 u64 >= G u32)
    uint32_t* const sums = reinterpret_cast<uint32_t*>(states);
    const uint32_t  tpw = gather_tpw(g.T);
    if (wave_tile) {
        const dim3 grid((g.T + 4) / 4), blk(kWG);
        switch (g.ndim) {
            case 1: e = launch_timed(k_tile_compress_w<1>, grid, blk, s, ev[2], ev[3], in, g, eb_rel, (const float4*)part, nparts, res, payload, staging, counts, tail, sums, tpw); break;
            case 2: e = launch_timed(k_tile_compress_w<2>, grid, blk, s, ev[2], ev[3], in, g, eb_rel, (const float4*)part, nparts, res, payload, staging, counts, tail, sums, tpw); break;
            default: e = launch_timed(k_tile_compress_w<3>, grid, blk, s, ev[2], ev[3], in, g, eb_rel, (const float4*)part, nparts, res, payload, staging, counts, tail, sums, tpw); break;
        }
        if (e != hipSuccess || g.T == 0) return e;
        return launch_timed(k_scan_gather, dim3(G), dim3(kGatherThreads), s, ev[4], ev[5], (const uint32_t*)counts,
                            g.T, tpw, (const uint8_t*)staging, payload, (const uint32_t*)sums, res, g.W);
    }
    const dim3 grid(g.T + 1), blk(kWG);
    switch (g.ndim) {
        case 1: e = launch_timed(k_tile_compress<1>, grid, blk, s, ev[2], ev[3], in, g, eb_rel, (const float4*)part, nparts, res, payload, staging, counts, tail, sums, tpw); break;
        case 2: e = launch_timed(k_tile_compress<2>, grid, blk, s, ev[2], ev[3], in, g, eb_rel, (const float4*)part, nparts, res, payload, staging, counts, tail, sums, tpw); break;
        default: e = launch_timed(k_tile_compress<3>, grid, blk, s, ev[2], ev[3], in, g, eb_rel, (const float4*)part, nparts, res, payload, staging, counts, tail, sums, tpw); break;
    }
    if (e != hipSuccess || g.T == 0) return e;
    return launch_timed(k_scan_gather, dim3(G), dim3(kGatherThreads), s, ev[4], ev[5], (const uint32_t*)counts, g.T,
                        tpw, (const uint8_t*)staging, payload, (const uint32_t*)sums, res, g.W);
}

// ============================================================================ plain stages

// Lorenzo stage: 8 codes per thread for every position < len (codes beyond x*y*z are 0).
template <int NDIM>
__global__ __launch_bounds__(kWG) void k_dualquant(const float* __restrict__ in, Geom g, float r_host, float r_amax,
                                                   const fz_result* res, uint16_t* __restrict__ codes,
                                                   unsigned long long* ovf_out)
{
    if (res && res->status != ST_SUCCESS) return;
    const float    r = res ? res->ebx2_r : r_host;
    const uint32_t p0 = 8 * (blockIdx.x * kWG + threadIdx.x);
    uint32_t       ovf = 0;
    // one workgroup = one 2048-position tile: same path choice as the fused kernel
    const float amax = res ? fmaxf(__builtin_fabsf(res->vmin), __builtin_fabsf(res->vmax)) : r_amax;
    const int   path = tile_path<NDIM>(g, blockIdx.x * (uint32_t)kTileCodes, amax * r < 8388608.0f);
    if (path != 0 || p0 < g.len) {
        uint32_t w[4];
        dq8_tile<NDIM>(in, p0, g, r, path, w, ovf);
        if (p0 + 8 <= g.len) {
            *reinterpret_cast<uint4*>(codes + p0) = make_uint4(w[0], w[1], w[2], w[3]);
        } else {
#pragma unroll
            for (int i = 0; i < 8; i++)
                if (p0 + i < g.len) codes[p0 + i] = (uint16_t)(w[i >> 1] >> (16 * (i & 1)));
        }
    }
    if (ovf_out) {
        ovf = wave_sum(ovf);
        if ((threadIdx.x & 63) == 0 && ovf) atomicAdd(ovf_out, (unsigned long long)ovf);
    }
}

hipError_t launch_dualquant(const float* in, const Geom& g, float r_host, float r_amax, const fz_result* res,
                            uint16_t* codes, unsigned long long* ovf, hipStream_t s)
{
    const uint32_t groups = (g.len + 7) / 8;
    const dim3     grid((groups + kWG - 1) / kWG);
    switch (g.ndim) {
        case 1: FZ_LAUNCH(k_dualquant<1><<<grid, kWG, 0, s>>>(in, g, r_host, r_amax, res, codes, ovf)); break;
        case 2: FZ_LAUNCH(k_dualquant<2><<<grid, kWG, 0, s>>>(in, g, r_host, r_amax, res, codes, ovf)); break;
        default: FZ_LAUNCH(k_dualquant<3><<<grid, kWG, 0, s>>>(in, g, r_host, r_amax, res, codes, ovf)); break;
    }
    return hipGetLastError();
}

// bitshuffleAndBitflag: one tile per workgroup.  With the extra workgroup (blockIdx == T)
// the words past the last full tile are copied unshuffled (fz.cu:154-155) and the flags
// past it cleared (fz.cu:167-168), so the stage output is complete without memsets.
__global__ __launch_bounds__(kWG) void k_bitshuffle(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                    Geom g, uint32_t* __restrict__ bitflags,
                                                    uint32_t* __restrict__ byteflags)
{
    __shared__ uint32_t sh[32 * kLdsPitch];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, tile = blockIdx.x;
    if (tile >= g.T) {
        const uint32_t w0 = kTileWords * g.T, nw = (g.len + 1) / 2;
        for (uint32_t i = w0 + tid; i < nw; i += kWG) out[i] = in[i];
        for (uint32_t i = 8 * g.T + tid; i < g.W; i += kWG) bitflags[i] = 0;
        for (uint32_t i = 256 * g.T + tid; i < 32 * g.W; i += kWG) byteflags[i] = 0;
        return;
    }
    const uint4 v = *reinterpret_cast<const uint4*>(in + (size_t)tile * kTileWords + 4 * tid);
    uint32_t    w[4] = {v.x, v.y, v.z, v.w};
    const uint32_t L = tid & 7, k = tid >> 3;
    transpose4(w, L);
#pragma unroll
    for (int q = 0; q < 4; q++) sh[kLdsPitch * (4 * L + q) + k] = w[q];
    __syncthreads();
    const uint32_t c = tid >> 3, k0 = 4 * (tid & 7);
    const uint4    blk = make_uint4(sh[kLdsPitch * c + k0], sh[kLdsPitch * c + k0 + 1], sh[kLdsPitch * c + k0 + 2],
                                    sh[kLdsPitch * c + k0 + 3]);
    *reinterpret_cast<uint4*>(out + (size_t)tile * kTileWords + 4 * tid) = blk;
    const bool f = (blk.x | blk.y | blk.z | blk.w) != 0;
    byteflags[(size_t)tile * kTileBlocks + tid] = f ? 1u : 0u;
    const uint64_t mask = ballot64(f);
    if (lane < 2) bitflags[8 * tile + 2 * wv + lane] = lane ? (uint32_t)(mask >> 32) : (uint32_t)mask;
}

hipError_t launch_bitshuffle(const uint32_t* in, uint32_t* out, const Geom& g, uint32_t* bitflags,
                             uint32_t* byteflags, bool with_tail_wg, hipStream_t s)
{
    const uint32_t nwg = g.T + (with_tail_wg ? 1u : 0u);
    if (nwg == 0) return hipSuccess;
    FZ_LAUNCH(k_bitshuffle<<<nwg, kWG, 0, s>>>(in, out, g, bitflags, byteflags));
    return hipGetLastError();
}

// Inverse of k_bitshuffle on full tiles (the transpose is its own inverse: gather the
// strided column words of group k, transpose, store rows 32k..32k+31).
__global__ __launch_bounds__(kWG) void k_unshuffle(const uint32_t* __restrict__ in, uint32_t* __restrict__ out)
{
    __shared__ uint32_t sh[32 * kLdsPitch];
    const uint32_t tid = threadIdx.x, tile = blockIdx.x;
    const uint4    v = *reinterpret_cast<const uint4*>(in + (size_t)tile * kTileWords + 4 * tid);
    const uint32_t c = tid >> 3, k0 = 4 * (tid & 7);
    sh[kLdsPitch * c + k0] = v.x;
    sh[kLdsPitch * c + k0 + 1] = v.y;
    sh[kLdsPitch * c + k0 + 2] = v.z;
    sh[kLdsPitch * c + k0 + 3] = v.w;
    __syncthreads();
    const uint32_t L = tid & 7, k = tid >> 3;
    uint32_t       w[4];
#pragma unroll
    for (int q = 0; q < 4; q++) w[q] = sh[kLdsPitch * (4 * L + q) + k];
    transpose4(w, L);
    *reinterpret_cast<uint4*>(out + (size_t)tile * kTileWords + 4 * tid) = make_uint4(w[0], w[1], w[2], w[3]);
}

hipError_t launch_unshuffle(const uint32_t* in, uint32_t* out, uint32_t ntiles, hipStream_t s)
{
    if (ntiles == 0) return hipSuccess;
    FZ_LAUNCH(k_unshuffle<<<ntiles, kWG, 0, s>>>(in, out));
    return hipGetLastError();
}

// ---------------------------------------------------------------------------- scan

__global__ void k_fill_u64(uint64_t* p, uint32_t n, uint32_t* ticket)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
    if (i == 0 && ticket) *ticket = 0;
}

hipError_t launch_scan_init(uint64_t* states, uint32_t nstates, uint32_t* ticket, hipStream_t s)
{
    const uint32_t n = nstates > 0 ? nstates : 1;
    FZ_LAUNCH(k_fill_u64<<<(n + kWG - 1) / kWG, kWG, 0, s>>>(states, nstates, ticket));
    return hipGetLastError();
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    FZ_LAUNCH(k_fill_u32<<<(unsigned)((n + kWG - 1) / kWG), kWG, 0, s>>>(p, n, v));
    return hipGetLastError();
}

// Exclusive sum of n u32 (wrapping), 4096 items per workgroup, 16 consecutive per thread.
// Optionally out[n] = total and the payload size in the result record.
__global__ __launch_bounds__(kWG) void k_scan(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t n,
                                              uint64_t* states, uint32_t* ticket, uint32_t write_total,
                                              fz_result* res, uint32_t W)
{
    __shared__ uint32_t sh_tile, sh_excl, sh_w[4];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (tid == 0) sh_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t chunk = sh_tile;
    const uint32_t base = chunk * kScanChunk + 16 * tid;
    uint32_t       v[16];
    if (base + 16 <= n) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint4 t = *reinterpret_cast<const uint4*>(in + base + 4 * j);
            v[4 * j] = t.x;
            v[4 * j + 1] = t.y;
            v[4 * j + 2] = t.z;
            v[4 * j + 3] = t.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < 16; j++) v[j] = (base + j < n) ? in[base + j] : 0u;
    }
    uint32_t tsum = 0;
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const uint32_t x = v[j];
        v[j] = tsum;
        tsum += x;
    }
    const uint32_t incl = wave_incl_scan(tsum);
    if (lane == 63) sh_w[wv] = incl;
    __syncthreads();
    const uint32_t w0 = sh_w[0], w1 = sh_w[1], w2 = sh_w[2], w3 = sh_w[3];
    const uint32_t agg = w0 + w1 + w2 + w3;
    const uint32_t woff = (wv > 0 ? w0 : 0u) + (wv > 1 ? w1 : 0u) + (wv > 2 ? w2 : 0u);
    if (wv == 0) {
        const uint32_t excl = lookback(states, chunk, agg);
        if (lane == 0) {
            sh_excl = excl;
            const uint32_t nchunks = (n + kScanChunk - 1) / kScanChunk;
            if (chunk == nchunks - 1) {
                if (write_total) out[n] = excl + agg;
                if (res) {
                    res->nblocks = (uint64_t)(excl + agg);
                    res->payload_bytes = 4ull * W + 16ull * (uint64_t)(excl + agg);
                }
            }
        }
    }
    __syncthreads();
    const uint32_t off = sh_excl + woff + incl - tsum;
    if (base + 16 <= n) {
#pragma unroll
        for (int j = 0; j < 4; j++)
            *reinterpret_cast<uint4*>(out + base + 4 * j) =
                make_uint4(off + v[4 * j], off + v[4 * j + 1], off + v[4 * j + 2], off + v[4 * j + 3]);
    } else {
#pragma unroll
        for (int j = 0; j < 16; j++)
            if (base + j < n) out[base + j] = off + v[j];
    }
}

hipError_t launch_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint64_t* states, uint32_t* ticket,
                       bool write_total, fz_result* res, uint32_t W, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    const uint32_t nchunks = (n + kScanChunk - 1) / kScanChunk;
    FZ_LAUNCH(k_scan<<<nchunks, kWG, 0, s>>>(in, out, n, states, ticket, write_total ? 1u : 0u, res, W));
    return hipGetLastError();
}

// ---------------------------------------------------------------------------- encode

__global__ __launch_bounds__(kWG) void k_encode(const uint8_t* __restrict__ in, uint8_t* __restrict__ out,
                                                const uint32_t* __restrict__ presum, uint32_t nblocks)
{
    const uint32_t i = blockIdx.x * kWG + threadIdx.x;
    if (i >= nblocks) return;
    const uint32_t a = presum[i], b = presum[i + 1];
    if (a != b) st16(out + 16ull * a, *reinterpret_cast<const uint4*>(in + 16ull * i));
}

hipError_t launch_encode(const uint8_t* in, uint8_t* out, const uint32_t* presum, uint32_t nblocks, hipStream_t s)
{
    if (nblocks == 0) return hipSuccess;
    FZ_LAUNCH(k_encode<<<(nblocks + kWG - 1) / kWG, kWG, 0, s>>>(in, out, presum, nblocks));
    return hipGetLastError();
}

// ============================================================================ inverse

// tile_excl[t] = number of set flags in tiles < t (popcount of bitflag words 8t..8t+7);
// one thread per tile, 256 tiles per workgroup, look-back across workgroups.
__global__ __launch_bounds__(kWG) void k_tile_scan(const uint32_t* __restrict__ bitflags, uint32_t T,
                                                   uint32_t* __restrict__ tile_excl, uint64_t* states,
                                                   uint32_t* ticket)
{
    __shared__ uint32_t sh_chunk, sh_excl, sh_w[4];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (tid == 0) sh_chunk = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t chunk = sh_chunk, t = chunk * kWG + tid;
    uint32_t       cnt = 0;
    if (t < T) {
        const uint4 a = *reinterpret_cast<const uint4*>(bitflags + 8 * (size_t)t);
        const uint4 b = *reinterpret_cast<const uint4*>(bitflags + 8 * (size_t)t + 4);
        cnt = __popc(a.x) + __popc(a.y) + __popc(a.z) + __popc(a.w) + __popc(b.x) + __popc(b.y) + __popc(b.z) +
              __popc(b.w);
    }
    const uint32_t incl = wave_incl_scan(cnt);
    if (lane == 63) sh_w[wv] = incl;
    __syncthreads();
    const uint32_t w0 = sh_w[0], w1 = sh_w[1], w2 = sh_w[2], w3 = sh_w[3];
    const uint32_t agg = w0 + w1 + w2 + w3;
    const uint32_t woff = (wv > 0 ? w0 : 0u) + (wv > 1 ? w1 : 0u) + (wv > 2 ? w2 : 0u);
    if (wv == 0) {
        const uint32_t excl = lookback(states, chunk, agg);
        if (lane == 0) sh_excl = excl;
    }
    __syncthreads();
    if (t < T) tile_excl[t] = sh_excl + woff + incl - cnt;
}

// Tile offsets without look-back for up to 256 chunks of 256 tiles, in one launch:
// k_tile_local writes each tile's offset within its chunk and each chunk's block count; the
// decode kernels add the counts of the chunks before the tile's own (tile_block_offset).
// Nothing to clear and no waiting between workgroups.
constexpr uint32_t kTileLocalMaxChunks = 256;

__device__ __forceinline__ uint32_t tile_popcount(const uint32_t* __restrict__ bitflags, uint32_t t)
{
    const uint4 a = *reinterpret_cast<const uint4*>(bitflags + 8 * (size_t)t);
    const uint4 b = *reinterpret_cast<const uint4*>(bitflags + 8 * (size_t)t + 4);
    return __popc(a.x) + __popc(a.y) + __popc(a.z) + __popc(a.w) + __popc(b.x) + __popc(b.y) + __popc(b.z) +
           __popc(b.w);
}

// exclusive prefix of v over the kWG threads of the workgroup, and the total
__device__ __forceinline__ uint32_t block_excl(uint32_t v, uint32_t& total)
{
    __shared__ uint32_t sh_w[kWG / 64];
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t incl = wave_incl_scan(v);
    if (lane == 63) sh_w[wv] = incl;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kWG / 64; i++) {
        pre += (uint32_t)i < wv ? sh_w[i] : 0u;
        tot += sh_w[i];
    }
    total = tot;
    return pre + incl - v;
}

__global__ __launch_bounds__(kWG) void k_tile_local(const uint32_t* __restrict__ bitflags, uint32_t T,
                                                    uint32_t* __restrict__ tile_excl, uint32_t* __restrict__ chunk_sums)
{
    const uint32_t t = blockIdx.x * kWG + threadIdx.x;
    uint32_t       total;
    const uint32_t excl = block_excl(t < T ? tile_popcount(bitflags, t) : 0u, total);
    if (t < T) tile_excl[t] = excl;
    if (threadIdx.x == 0) chunk_sums[blockIdx.x] = total;
}

hipError_t launch_tile_scan(const uint32_t* bitflags, uint32_t T, uint32_t* tile_excl, uint64_t* states,
                            uint32_t nstates, uint32_t* ticket, hipStream_t s, bool clear)
{
    if (T == 0) return hipSuccess;
    const hipError_t e = clear ? launch_scan_init(states, nstates, ticket, s) : hipSuccess;
    if (e != hipSuccess) return e;
    FZ_LAUNCH(k_tile_scan<<<(T + kWG - 1) / kWG, kWG, 0, s>>>(bitflags, T, tile_excl, states, ticket));
    return hipGetLastError();
}

hipError_t launch_tile_offsets(const uint32_t* bitflags, uint32_t T, uint32_t* tile_excl, uint64_t* states,
                               uint32_t nstates, uint32_t* ticket, hipStream_t s, const uint32_t** chunk_sums)
{
    *chunk_sums = nullptr;
    if (T == 0) return hipSuccess;
    const uint32_t nchunks = (T + kWG - 1) / kWG;
    if (nchunks > kTileLocalMaxChunks) return launch_tile_scan(bitflags, T, tile_excl, states, nstates, ticket, s, true);
    uint32_t* sums = reinterpret_cast<uint32_t*>(states);  // the state words hold the chunk counts
    FZ_LAUNCH(k_tile_local<<<nchunks, kWG, 0, s>>>(bitflags, T, tile_excl, sums));
    *chunk_sums = sums;
    return hipGetLastError();
}

// The payload block index of `tile`'s first nonzero block, by all 64 lanes of a wave:
// tile_excl[tile], plus (with chunk_sums) the counts of the 256-tile chunks before its own.
__device__ __forceinline__ uint32_t tile_block_offset(const uint32_t* __restrict__ tile_excl,
                                                      const uint32_t* __restrict__ chunk_sums, uint32_t tile)
{
    const uint32_t v = tile_excl[tile];
    if (!chunk_sums) return v;
    const uint32_t c = tile / kWG;  // < kTileLocalMaxChunks
    uint32_t       s = 0;
#pragma unroll
    for (uint32_t i = 0; i < kTileLocalMaxChunks / 64; i++) {
        const uint32_t k = lane_id() + 64 * i;
        s += k < c ? chunk_sums[k] : 0u;
    }
    return v + wave_sum(s);
}

// Inverse of the tile compress: the extra workgroup writes the tail codes (or zeros when
// there is no tail buffer); one wave per tile, so a CU holds four times as many tiles in
// flight as with a workgroup per tile, and all-zero tiles only store zeros.
__global__ __launch_bounds__(64) void k_decode_unshuffle(const uint8_t* __restrict__ payload,
                                                         const uint32_t* __restrict__ tile_excl,
                                                         const uint32_t* __restrict__ chunk_sums,
                                                         const uint16_t* __restrict__ tail, Geom g,
                                                         uint32_t* __restrict__ codes)
{
    const uint32_t tile = blockIdx.x;
    if (tile >= g.T) decode_tail(tail, g, reinterpret_cast<uint16_t*>(codes));
    else decode_tile_w(payload, g.W, tile, tile_block_offset(tile_excl, chunk_sums, tile), codes + (size_t)tile * kTileWords);
}

hipError_t launch_decode_unshuffle(const uint8_t* payload, const uint32_t* tile_excl, const uint32_t* chunk_sums,
                                   const uint16_t* tail, const Geom& g, uint32_t* codes, hipStream_t s)
{
    FZ_LAUNCH(k_decode_unshuffle<<<g.T + 1, 64, 0, s>>>(payload, tile_excl, chunk_sums, tail, g, codes));
    return hipGetLastError();
}

// 1D fields: decode and reconstruct in one wave per tile (the unit after the last full tile
// takes the trailer), straight to floats; no code array in between.
__global__ __launch_bounds__(64) void k_decode_recon_1d(const uint8_t* __restrict__ payload,
                                                        const uint32_t* __restrict__ tile_excl,
                                                        const uint32_t* __restrict__ chunk_sums,
                                                        const uint16_t* __restrict__ tail, Geom g, float ebx2,
                                                        float* __restrict__ out)
{
    const uint32_t tile = blockIdx.x;
    decode_recon_1d_w(payload, tail, g, tile, tile < g.T ? tile_block_offset(tile_excl, chunk_sums, tile) : 0u, ebx2,
                      out);
}

hipError_t launch_decode_recon_1d(const uint8_t* payload, const uint32_t* tile_excl, const uint32_t* chunk_sums,
                                  const uint16_t* tail, const Geom& g, float ebx2, float* out, hipStream_t s)
{
    FZ_LAUNCH(k_decode_recon_1d<<<g.T + 1, 64, 0, s>>>(payload, tile_excl, chunk_sums, tail, g, ebx2, out));
    return hipGetLastError();
}

// 2D fields: decode and reconstruct in one 1024-thread workgroup per window of 16 rows x 1024
// columns (one Lorenzo block row, 64 blocks), straight to floats; the codes never reach HBM.
//
// Decode.  The window's codes come into LDS by quad groups: QG q holds codes [256 q, 256 q +
// 256), the four 32-word transpose groups of tile q / 8, whose bit planes sit in blocks
// 8 b + q % 8 (b < 32) of that tile.  A half-wave (32 lanes) decodes one QG: lane b finds
// block 8 b + m's flag and rank from the tile's 8 bitflag words, loads it, stages it as bit
// plane row b in LDS, and after the in-register transpose lane j = 8 qq + L holds codes
// [256 q + 8 j, + 8), which it puts at their window positions.  The window needs the QGs its
// 16 row segments cover, each once (a QG shared by two rows of a narrow field is listed for
// the first): at most 80, and the 32 half-waves take up to 3 each, all loads in flight
// together.
//
// Reconstruct.  Thread (g, c) = (tid / 256, tid % 256) holds columns 4c .. 4c + 3 of rows
// 4g .. 4g + 3: the column sums of its 4 rows go through LDS, so each group starts from the
// sums of the groups above it (prefix over y), then its rows run as in recon_2d (prefix over
// x in the thread and its quad, two DPP steps).
constexpr uint32_t kD2Threads = 1024;
constexpr uint32_t kD2Cols = 1024;
constexpr uint32_t kD2MaxQ = 16 * (kD2Cols / 256 + 1);
constexpr uint32_t kD2Batch = 3;  // QGs per half-wave in flight together (32 x 3 >= kD2MaxQ)
// Fields whose windows find their tile offsets themselves (no tile-offset launch): at most
// 4096 tiles (the flag words before a tile fit 8 uint4 loads per thread) and X <= 4096 (a
// window's 16 rows touch at most (15 X + 1024) / 2048 + 2 <= 64 tiles)
constexpr uint32_t kD2DirectTiles = 64;
constexpr uint32_t kD2DirectMaxT = 4 * kD2Threads;
constexpr uint32_t kQgPitch = 5;

// Window wi of a 2D field (all kD2Threads threads of the workgroup); tile_excl / chunk_sums as
// launch_tile_offsets gives them, or (DIRECT) the window finds its tiles' offsets itself.
template <bool DIRECT>
__device__ __forceinline__ void d2_window(const uint8_t* __restrict__ payload, const uint32_t* __restrict__ tile_excl,
                                          const uint32_t* __restrict__ chunk_sums, const uint16_t* __restrict__ tail,
                                          const Geom& g, float ebx2, float* __restrict__ out, uint32_t rx, uint32_t wi)
{
    constexpr uint32_t NH = kD2Threads / 32;
    __shared__ __align__(16) uint16_t win[16 * kD2Cols];  // the window's codes, row-major
    __shared__ uint32_t stage[NH][32 * kQgPitch];      // one QG's bit-plane rows per half-wave
    __shared__ int4     csum[3][kD2Cols / 4];          // column sums of row groups 0-2
    __shared__ uint32_t cpre[kTileLocalMaxChunks];     // block counts of the chunks before each chunk
    __shared__ uint32_t wtot[kD2Threads / 64], qtab[kD2MaxQ], nq_s;  // the QGs to decode
    __shared__ uint32_t toff[kD2DirectTiles];         // direct offsets: blocks before tiles t_lo ..
    __shared__ uint4    tflg[2 * kD2DirectTiles];     // ... and their 8 bitflag words each
    __shared__ uint32_t twoff[8 * kD2DirectTiles];    // ... and the blocks before each of those words
    const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint32_t bx = wi % rx, by = wi / rx;
    const uint32_t xw0 = bx * kD2Cols, y0 = by * 16;
    const uint32_t cw = min(kD2Cols, g.X - xw0), rows = min(16u, g.Y - y0);
    // without tile offsets (tile_excl null: small fields) the window finds the block offsets of
    // its tiles t_lo .. from the bitflags: the popcount of every flag word before tile t_lo (at
    // most 8 uint4 loads per thread, issued together), then a scan over the window's own tiles
    constexpr bool direct = DIRECT;
    const uint32_t t_lo = min((y0 * g.X + xw0) >> 11, g.T);
    // set-up: the chunk prefix (waves 0-3) or the direct prefix (all waves), and the QG list (wave 0)
    uint32_t cv = 0, cincl = 0;
    if constexpr (direct) {
        const uint4*   bf4 = reinterpret_cast<const uint4*>(payload);
        const uint32_t n4 = 2 * t_lo;  // uint4s of flag words before tile t_lo (<= 8 kD2Threads)
        uint4          v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const uint32_t i = tid + (uint32_t)k * kD2Threads;
            v[k] = i < n4 ? bf4[i] : make_uint4(0, 0, 0, 0);
        }
        uint32_t c = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) c += __popc(v[k].x) + __popc(v[k].y) + __popc(v[k].z) + __popc(v[k].w);
        c = wave_sum(c);
        if (lane == 0) wtot[wv] = c;
    } else if (tid < kTileLocalMaxChunks) {
        cv = (chunk_sums && tid < (g.T + kWG - 1) / kWG) ? chunk_sums[tid] : 0u;
        cincl = wave_incl_scan(cv);
        if (lane == 63) wtot[wv] = cincl;
    }
    if (wv == 0) {
        // lane r < rows: QGs [a, b] of row segment r less those of row r - 1 (a and b grow with r)
        const bool     valid = lane < rows;
        const uint32_t p0 = (y0 + lane) * g.X + xw0;  // < x*y*z < 2^32 on the valid lanes
        const uint32_t a = p0 >> 8, b = (p0 + cw - 1) >> 8;
        const uint32_t bprev = __shfl_up(b, 1, 64);
        const uint32_t start = lane > 0 ? max(a, bprev + 1) : a;
        const uint32_t c = (valid && b + 1 > start) ? b + 1 - start : 0u;
        const uint32_t incl = wave_incl_scan(c);
        for (uint32_t j = 0; j < c; j++) qtab[incl - c + j] = start + j;
        if (lane == 63) nq_s = incl;
    }
    __syncthreads();
    if constexpr (direct) {
        if (wv == 0) {  // lane i: tile t_lo + i (the window's tiles span at most kD2DirectTiles)
            uint32_t base = 0;
#pragma unroll
            for (int k = 0; k < (int)(kD2Threads / 64); k++) base += wtot[k];
            const uint32_t t = t_lo + lane;
            uint4          a = make_uint4(0, 0, 0, 0), b = a;
            if (t < g.T) {
                a = reinterpret_cast<const uint4*>(payload)[2 * t];
                b = reinterpret_cast<const uint4*>(payload)[2 * t + 1];
            }
            tflg[2 * lane] = a;
            tflg[2 * lane + 1] = b;
            const uint32_t fw8[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
            uint32_t       pw[8], pc = 0;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                pw[j] = pc;
                pc += __popc(fw8[j]);
            }
            const uint32_t o = base + wave_incl_scan(pc) - pc;
            toff[lane] = o;
#pragma unroll
            for (int j = 0; j < 8; j++) twoff[8 * lane + j] = o + pw[j];
        }
    } else if (tid < kTileLocalMaxChunks) {
        uint32_t pre = 0;
        for (uint32_t k = 0; k < wv; k++) pre += wtot[k];
        cpre[tid] = pre + cincl - cv;
    }
    __syncthreads();
    const uint32_t  nq = nq_s;
    const uint32_t  half = lane >> 5, hl = lane & 31, H = 2 * wv + half;
    uint32_t* const st = stage[H];
    const uint32_t* bflags = reinterpret_cast<const uint32_t*>(payload);
    // both half-waves of a wave step together (the transpose needs all 64 lanes active)
    for (uint32_t i0 = 0; i0 < nq; i0 += kD2Batch * NH) {
        uint32_t q[kD2Batch], fw[kD2Batch], ex[kD2Batch];
        bool     live[kD2Batch];
#pragma unroll
        for (int b = 0; b < (int)kD2Batch; b++) {
            const uint32_t i = i0 + H + b * NH;
            live[b] = i < nq;
            q[b] = live[b] ? qtab[i] : 0u;
            const uint32_t t = q[b] >> 3;
            const bool     in = live[b] && t < g.T;
            // direct: the flag word of this lane's block (word hl / 4) and the blocks before that
            // word, from the window's tables; else lane hl % 8's word of the tile, scanned below
            fw[b] = !in ? 0u
                    : direct ? reinterpret_cast<const uint32_t*>(tflg)[8 * (t - t_lo) + (hl >> 2)]
                             : bflags[8 * t + (hl & 7u)];
            ex[b] = !in ? 0u : direct ? twoff[8 * (t - t_lo) + (hl >> 2)] : tile_excl[t] + (chunk_sums ? cpre[t / kWG] : 0u);
        }
        uint4 blk[kD2Batch];
        bool  nzq[kD2Batch];  // some block of the wave's two QGs is nonzero (uniform)
#pragma unroll
        for (int b = 0; b < (int)kD2Batch; b++) {
            // flag and rank of block 8 hl + m (flag word hl / 4, bit 8 (hl % 4) + m); without
            // the direct tables word j is held by the half's lanes with hl % 8 == j
            uint32_t wj = fw[b], ej = 0;
            if constexpr (!direct) {
                const uint32_t pc = __popc(fw[b]);
                uint32_t       incl = pc;
#pragma unroll
                for (int d = 1; d < 8; d <<= 1) {
                    const uint32_t o = __shfl_up(incl, d, 8);
                    if ((hl & 7u) >= (uint32_t)d) incl += o;
                }
                const uint32_t src = (lane & 32u) + (hl >> 2);
                wj = __shfl(fw[b], src, 64);
                ej = __shfl(incl - pc, src, 64);
            }
            const uint32_t bit = 8 * (hl & 3u) + (q[b] & 7u);
            const bool     f = (wj >> bit) & 1u;
            nzq[b] = ballot64(f) != 0;
            blk[b] = make_uint4(0, 0, 0, 0);
            if (f) blk[b] = ld16(payload + 4ull * g.W + 16ull * ((uint64_t)ex[b] + ej + __popc(wj & ((1u << bit) - 1u))));
        }
#pragma unroll
        for (int b = 0; b < (int)kD2Batch; b++) {
            const uint32_t t = q[b] >> 3;
            uint32_t       w[4] = {0, 0, 0, 0};
            if (nzq[b]) {  // all-zero blocks decode to zero codes without the transpose
                uint32_t* row = st + kQgPitch * hl;
                row[0] = blk[b].x;
                row[1] = blk[b].y;
                row[2] = blk[b].z;
                row[3] = blk[b].w;
                wave_lds_sync();
                const uint32_t qq = hl >> 3, L = hl & 7u;
#pragma unroll
                for (int c = 0; c < 4; c++) w[c] = st[kQgPitch * (4 * L + c) + qq];
                wave_lds_sync();  // the rows are rewritten by the next QG
                transpose4(w, L);
            }
            if (!live[b]) continue;
            const uint32_t p = 256 * q[b] + 8 * hl;
            if (t >= g.T) {  // after the last full tile: the trailer's codes (0 without one)
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    uint32_t v = 0;
#pragma unroll
                    for (int h = 0; h < 2; h++) {
                        const uint32_t pp = p + 2 * c + h;
                        if (tail && pp < g.len) v |= (uint32_t)tail[pp - kTileCodes * g.T] << (16 * h);
                    }
                    w[c] = v;
                }
            }
            uint32_t row = fdiv(p, g.divX), col = p - row * g.X;
            if ((g.X & 7u) == 0) {
                // p % 8 == 0 and X % 8 == 0: the 8 codes sit in one row, 16-byte aligned in the window
                if (row >= y0 && row < y0 + rows && col >= xw0 && col < xw0 + cw)
                    *reinterpret_cast<uint4*>(win + (row - y0) * kD2Cols + (col - xw0)) = make_uint4(w[0], w[1], w[2], w[3]);
                continue;
            }
            if ((g.X & 3u) == 0) {
                // X % 4 == 0: each half of 4 codes sits in one row, 8-byte aligned in the window
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    if (row >= y0 && row < y0 + rows && col >= xw0 && col < xw0 + cw)
                        *reinterpret_cast<uint2*>(win + (row - y0) * kD2Cols + (col - xw0)) = make_uint2(w[2 * h], w[2 * h + 1]);
                    col += 4;
                    if (col == g.X) {
                        col = 0;
                        row++;
                    }
                }
                continue;
            }
#pragma unroll
            for (int j = 0; j < 8; j++) {
                if (row >= y0 && row < y0 + rows && col >= xw0 && col < xw0 + cw)
                    win[(row - y0) * kD2Cols + (col - xw0)] = (uint16_t)(w[j >> 1] >> (16 * (j & 1)));
                if (++col == g.X) {
                    col = 0;
                    row++;
                }
            }
        }
    }
    __syncthreads();
    // reconstruct: rows 4 gr .. 4 gr + 3 (gr uniform per wave), columns x0 .. x0 + 3
    const uint32_t c = tid & (kD2Cols / 4 - 1), gr = tid / (kD2Cols / 4);
    const uint32_t x0 = xw0 + 4 * c;
    const bool     full = x0 + 4 <= g.X;
    int32_t        dd[4][4];
    int32_t        cs[4] = {0, 0, 0, 0};
#pragma unroll
    // rows past the field (r >= rows) and columns past it (x0 + j >= X) hold stale codes: the
    // sums carry them only down the column and right along the row, into positions that are
    // never stored
    for (int r4 = 0; r4 < 4; r4++) {
        const uint32_t r = 4 * gr + r4;
        const uint2    v = *reinterpret_cast<const uint2*>(win + r * kD2Cols + 4 * c);
        decode2(v.x, dd[r4][0], dd[r4][1]);
        decode2(v.y, dd[r4][2], dd[r4][3]);
#pragma unroll
        for (int j = 0; j < 4; j++) cs[j] += dd[r4][j];
    }
    if (gr < 3) csum[gr][c] = make_int4(cs[0], cs[1], cs[2], cs[3]);
    __syncthreads();
    int32_t col[4] = {0, 0, 0, 0};
    for (uint32_t k = 0; k < gr; k++) {
        const int4 v = csum[k][c];
        col[0] += v.x;
        col[1] += v.y;
        col[2] += v.z;
        col[3] += v.w;
    }
#pragma unroll
    for (int r4 = 0; r4 < 4; r4++) {
        const uint32_t r = 4 * gr + r4;
        int32_t        p[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            col[j] += dd[r4][j];
            p[j] = col[j] + (j ? p[j - 1] : 0);
        }
        const int32_t excl = quad_incl_scan(p[3]) - p[3];
        if (r >= rows) continue;  // uniform over the wave
        const size_t row = (size_t)(y0 + r) * g.X;
        float        o[4];
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

template <bool DIRECT>
__global__ __launch_bounds__(kD2Threads) void k_decode_recon_2d(const uint8_t* __restrict__ payload,
                                                                const uint32_t* __restrict__ tile_excl,
                                                                const uint32_t* __restrict__ chunk_sums,
                                                                const uint16_t* __restrict__ tail, Geom g, float ebx2,
                                                                float* __restrict__ out, uint32_t rx)
{
    // XCD-sliced order: the windows that share a tile's blocks (neighbours along x, and along y
    // for narrow fields) read them through one L2
    d2_window<DIRECT>(payload, tile_excl, chunk_sums, tail, g, ebx2, out, rx, xcd_tile(blockIdx.x, gridDim.x));
}

// With X % 4 == 0 a lane puts its codes into the window with one or two wide LDS stores, and
// the fused form beats decode + reconstruct at every size measured (3600 x 1800: 16.0 against
// 23.5 us; 16384 x 16384: 369 against 473 us).  Otherwise each code is a 2-byte store and the
// fused form wins only while the windows are resident at once or nearly (two 1024-thread
// workgroups per CU): up to 26000 x 388 (650 windows), not at 39200 x 388 (975).
constexpr uint64_t kD2MaxWindows = 768;

bool decode_recon_2d_direct(const Geom& g) { return g.T <= kD2DirectMaxT && g.X <= 4096; }

bool decode_recon_2d_pays(const Geom& g)
{
    return (g.X & 3u) == 0 || (uint64_t)((g.X + kD2Cols - 1) / kD2Cols) * ((g.Y + 15) / 16) <= kD2MaxWindows;
}

hipError_t launch_decode_recon_2d(const uint8_t* payload, const uint32_t* tile_excl, const uint32_t* chunk_sums,
                                  const uint16_t* tail, const Geom& g, float ebx2, float* out, hipStream_t s)
{
    const uint32_t rx = (g.X + kD2Cols - 1) / kD2Cols;
    const uint64_t n = (uint64_t)rx * ((g.Y + 15) / 16);
    if (n > 0xFFFFFFFFull) return hipErrorInvalidConfiguration;
    if (!tile_excl) FZ_LAUNCH(k_decode_recon_2d<true><<<(uint32_t)n, kD2Threads, 0, s>>>(payload, nullptr, nullptr, tail, g, ebx2, out, rx));
    else FZ_LAUNCH(k_decode_recon_2d<false><<<(uint32_t)n, kD2Threads, 0, s>>>(payload, tile_excl, chunk_sums, tail, g, ebx2, out, rx));
    return hipGetLastError();
}

// 3D fields: decode and reconstruct in one workgroup of R Q threads per window of 8 planes x R
// rows x 4Q columns (one Lorenzo block layer in z, R / 8 block rows in y), straight to floats;
// the codes never reach HBM.
//
// Shape.  X <= 128: Q is the fewest quads per row that hold a whole row (4 Q >= X, a power of
// two <= 32) and R = 256 / Q.  Wider rows: R = 8, and a row takes the fewest windows of at most
// 512 columns, Q the quads of its share rounded up to a multiple of 8 (whole waves); so a
// window holds within 32 columns of what it covers of the field (d3_shape: a single field with
// X <= 256 takes Q = 64).  The template NT (256, 512 or 1024) bounds R Q, and sizes the QG
// list and the staging buffers; the window's memory is dynamic (64 R Q bytes).
//
// Decode, as in k_decode_recon_2d.  The window's row segments list the quad groups they
// touch, each once: with whole rows one segment per plane (a plane's rows are consecutive in
// memory), else one per (plane, row); at most 8 (NT / 64 + 1) or 192 QGs.  Half-waves decode
// them into the LDS window, all their loads in flight together.
//
// Reconstruct.  Thread (xq, yy) = (tid % Q, tid / Q) holds columns 4 xq .. 4 xq + 3 of row yy
// over the 8 planes: prefix over z in registers, over x in the thread and with its pair lane
// (one DPP step).  The prefix over y goes through LDS four planes at a time, in the window's
// own memory once every code has been read: thread (xq, tz, yh, yb) = (tid % Q, rest % 4,
// rest / 4 % 2, rest / 8) with rest = tid / Q sums rows 8 yb .. 8 yb + 4 yh + 3 of plane tz
// and stores rows 8 yb + 4 yh .. + 3.
constexpr uint32_t d3_maxq(uint32_t nt) { return nt == 1024 ? 192u : 8u * (nt / 64 + 1); }

// Window wi of a 3D field (all R Q threads of the workgroup): cpre is the chunk prefix in LDS
// (chunk-local tile offsets) or nullptr (tile_excl holds global offsets); mq = d3_magic(Q).
template <int NT>
__device__ __forceinline__ void d3_window(const uint8_t* __restrict__ payload, const uint32_t* __restrict__ tile_excl,
                                          const uint32_t* cpre, const uint16_t* __restrict__ tail, const Geom& g,
                                          float ebx2, float* __restrict__ out, uint32_t Q, uint32_t R, uint32_t mq,
                                          uint32_t rx, uint32_t ry, uint32_t wi)
{
    static_assert(NT == 256 || NT == 512 || NT == 1024, "threads per window");
    constexpr uint32_t NH = NT / 32;                      // half-waves
    constexpr uint32_t MAXQ = d3_maxq(NT);
    // QGs per half-wave in flight together: half of them at NT threads (two rounds), which keeps
    // the kernel at 64 VGPRs without scratch
    constexpr uint32_t NB = (MAXQ + 2 * NH - 1) / (2 * NH);
    // 512 and 1024 threads: the two half-waves of a wave stage their QGs in turn through one
    // buffer (LDS 38.8 / 75.8 KB at NT threads, 4 / 2 workgroups per CU, instead of 43.3 / 85.8 KB)
    constexpr bool     SERIAL = NT >= 512;
    extern __shared__ __align__(16) uint16_t win[];       // codes [plane][row][column]; later int4 [4][R][Q]
    __shared__ uint32_t stage[SERIAL ? NT / 64 : NH][32 * kQgPitch];  // bit-plane rows of a QG
    __shared__ uint32_t qtab[MAXQ], nq_s;                 // the QGs to decode
    const uint32_t C = 4 * Q, nh = (R * Q) >> 5;
    // the lane indices are recomputed in every window from an opaque copy of the thread index:
    // values hoisted out of the loop would otherwise stay live across it (about 40 more VGPRs)
    uint32_t tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const uint32_t lane = tid & 63, wv = tid >> 6;
    // lane exchanges by the opaque lane index (the builtins' own lane index would stay live
    // across the window loop)
    auto bperm = [](uint32_t src, uint32_t v) -> uint32_t {
        return (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src << 2), (int)v);
    };
    const uint32_t bx = wi % rx, t1 = wi / rx, by = t1 % ry, bz = t1 / ry;
    const uint32_t xw0 = bx * C, y0 = by * R, z0 = bz * 8;
    const uint32_t cw = min(C, g.X - xw0), rows = min(R, g.Y - y0), nz = min(8u, g.Z - z0);
    const bool     whole = cw == g.X;
    if (wv == 0) {
        // lane i < nseg: the QGs [a, b] of segment i less those of segment i - 1 (a and b grow with i)
        const uint32_t nseg = whole ? nz : nz * rows;  // rows <= 8 when rows are not whole
        const bool     valid = lane < nseg;
        uint32_t       p0 = 0, len = 1;
        if (valid) {
            if (whole) {
                p0 = (z0 + lane) * g.XY + y0 * g.X;  // < x*y*z < 2^32
                len = rows * g.X;
            } else {
                const uint32_t pz = lane / rows, r = lane - pz * rows;
                p0 = (z0 + pz) * g.XY + (y0 + r) * g.X + xw0;
                len = cw;
            }
        }
        const uint32_t a = p0 >> 8, b = (p0 + len - 1) >> 8;
        const uint32_t bprev = bperm(lane - 1, b);
        const uint32_t start = lane > 0 ? max(a, bprev + 1) : a;
        const uint32_t c = (valid && b + 1 > start) ? b + 1 - start : 0u;
        uint32_t       incl = c;
#pragma unroll
        for (uint32_t d = 1; d < 64; d <<= 1) {
            const uint32_t o = bperm(lane - d, incl);
            if (lane >= d) incl += o;
        }
        for (uint32_t j = 0; j < c; j++) qtab[incl - c + j] = start + j;
        if (lane == 63) nq_s = incl;
    }
    // the QG list (and, before the first window, the caller's chunk prefix) is written; every
    // access to the window's memory by the workgroup's last window is done
    __syncthreads();
    const uint32_t  nq = nq_s;
    const uint32_t  half = lane >> 5, hl = lane & 31, H = 2 * wv + half;
    uint32_t* const st = stage[SERIAL ? wv : H];
    const uint32_t* bflags = reinterpret_cast<const uint32_t*>(payload);
    // a code at field position (x, y, z) inside the window goes to win[((z - z0) R + y - y0) C + x - xw0]
    auto slot = [&](uint32_t x, uint32_t y, uint32_t z) -> int {
        const uint32_t dz = z - z0, dy = y - y0, dx = x - xw0;  // wrap to large values below the window
        return (dz < nz && dy < rows && dx < cw) ? (int)((dz * R + dy) * C + dx) : -1;
    };
    // both half-waves of a wave step together (the transpose needs all 64 lanes active)
    for (uint32_t i0 = 0; i0 < nq; i0 += NB * nh) {
        uint32_t q[NB], fw[NB], ex[NB];
        bool     live[NB];
#pragma unroll
        for (int b = 0; b < (int)NB; b++) {
            const uint32_t i = i0 + H + b * nh;
            live[b] = i < nq;
            q[b] = live[b] ? qtab[i] : 0u;
            const uint32_t t = q[b] >> 3;
            const bool     in = live[b] && t < g.T;
            fw[b] = in ? bflags[8 * t + (hl & 7u)] : 0u;
            ex[b] = in ? tile_excl[t] + (cpre ? cpre[t / kWG] : 0u) : 0u;
        }
        uint4 blk[NB];
        bool  nzq[NB];  // some block of the wave's two QGs is nonzero (uniform)
#pragma unroll
        for (int b = 0; b < (int)NB; b++) {
            // flag and rank of block 8 hl + m (flag word hl / 4, bit 8 (hl % 4) + m); word j is
            // held by the half's lanes with hl % 8 == j
            const uint32_t pc = __popc(fw[b]);
            uint32_t       incl = pc;
#pragma unroll
            for (int d = 1; d < 8; d <<= 1) {
                const uint32_t o = bperm(lane - d, incl);
                if ((hl & 7u) >= (uint32_t)d) incl += o;
            }
            const uint32_t src = (lane & 32u) + (hl >> 2);
            const uint32_t wj = bperm(src, fw[b]), ej = bperm(src, incl - pc);
            const uint32_t bit = 8 * (hl & 3u) + (q[b] & 7u);
            const bool     f = (wj >> bit) & 1u;
            nzq[b] = ballot64(f) != 0;
            blk[b] = make_uint4(0, 0, 0, 0);
            if (f) blk[b] = ld16(payload + 4ull * g.W + 16ull * ((uint64_t)ex[b] + ej + __popc(wj & ((1u << bit) - 1u))));
        }
#pragma unroll
        for (int b = 0; b < (int)NB; b++) {
            const uint32_t t = q[b] >> 3;
            uint32_t       w[4] = {0, 0, 0, 0};
            if (nzq[b]) {  // all-zero blocks decode to zero codes without the transpose
                const uint32_t qq = hl >> 3, L = hl & 7u;
#pragma unroll
                for (uint32_t hh = 0; hh < (SERIAL ? 2u : 1u); hh++) {
                    if (!SERIAL || half == hh) {
                        uint32_t* row = st + kQgPitch * hl;
                        row[0] = blk[b].x;
                        row[1] = blk[b].y;
                        row[2] = blk[b].z;
                        row[3] = blk[b].w;
                    }
                    wave_lds_sync();
                    if (!SERIAL || half == hh) {
#pragma unroll
                        for (int c = 0; c < 4; c++) w[c] = st[kQgPitch * (4 * L + c) + qq];
                    }
                    wave_lds_sync();  // the rows are rewritten by the other half or the next QG
                }
                transpose4(w, L);
            }
            if (!live[b]) continue;
            const uint32_t p = 256 * q[b] + 8 * hl;
            if (t >= g.T) {  // after the last full tile: the trailer's codes (0 without one)
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    uint32_t v = 0;
#pragma unroll
                    for (int h = 0; h < 2; h++) {
                        const uint32_t pp = p + 2 * c + h;
                        if (tail && pp < g.len) v |= (uint32_t)tail[pp - kTileCodes * g.T] << (16 * h);
                    }
                    w[c] = v;
                }
            }
            const uint32_t q1 = fdiv(p, g.divX);
            uint32_t       x = p - q1 * g.X, z = fdiv(q1, g.divY), y = q1 - z * g.Y;
            if ((g.X & 7u) == 0) {
                // p % 8 == 0 and X % 8 == 0: the 8 codes sit in one row, 16-byte aligned in the window
                const int s = slot(x, y, z);
                if (s >= 0) *reinterpret_cast<uint4*>(win + s) = make_uint4(w[0], w[1], w[2], w[3]);
                continue;
            }
            if ((g.X & 3u) == 0) {
                // X % 4 == 0: each half of 4 codes sits in one row, 8-byte aligned in the window
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int s = slot(x, y, z);
                    if (s >= 0) *reinterpret_cast<uint2*>(win + s) = make_uint2(w[2 * h], w[2 * h + 1]);
                    x += 4;
                    if (x == g.X) {
                        x = 0;
                        if (++y == g.Y) {
                            y = 0;
                            z++;
                        }
                    }
                }
                continue;
            }
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int s = slot(x, y, z);
                if (s >= 0) win[s] = (uint16_t)(w[j >> 1] >> (16 * (j & 1)));
                if (++x == g.X) {
                    x = 0;
                    if (++y == g.Y) {
                        y = 0;
                        z++;
                    }
                }
            }
        }
    }
    __syncthreads();
    // reconstruct, phase 1: columns xc .. xc + 3 of row yy, the 8 planes (z prefix in registers)
    const uint32_t yy = __umulhi(tid, mq), xq = tid - yy * Q;
    const uint32_t xc = xw0 + 4 * xq;
    int32_t        r[8][4];
    {
        int32_t col[4] = {0, 0, 0, 0};
#pragma unroll
        for (int z = 0; z < 8; z++) {
            // planes, rows and columns past the field hold stale codes: the sums carry them
            // only forward in z, down in y and right in x, into positions that are never stored
            int32_t     dd[4];
            const uint2 v = *reinterpret_cast<const uint2*>(win + ((uint32_t)z * R + yy) * C + 4 * xq);
            decode2(v.x, dd[0], dd[1]);
            decode2(v.y, dd[2], dd[3]);
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
#pragma unroll
            for (int j = 0; j < 4; j++) r[z][j] = p[j] + add;
        }
    }
    // phase 2: the prefix over y, four planes at a time through the window's memory
    int4* const    ex4 = reinterpret_cast<int4*>(win);  // [4][R][Q]: 64 R Q bytes, the window's size
    const uint32_t tz = yy & 3u, yh = (yy >> 2) & 1u, yb = yy >> 3;
    const bool     full = xc + 4 <= g.X;
#pragma unroll
    for (int h = 0; h < 2; h++) {
        __syncthreads();  // every read of the window (h = 0) or of the previous planes (h = 1) is done
#pragma unroll
        for (int zz = 0; zz < 4; zz++)
            ex4[((uint32_t)zz * R + yy) * Q + xq] = make_int4(r[4 * h + zz][0], r[4 * h + zz][1], r[4 * h + zz][2], r[4 * h + zz][3]);
        __syncthreads();
        int4 acc = make_int4(0, 0, 0, 0);
        if (yh) {
#pragma unroll
            for (int y = 0; y < 4; y++) {
                const int4 v = ex4[(tz * R + 8 * yb + (uint32_t)y) * Q + xq];
                acc.x += v.x;
                acc.y += v.y;
                acc.z += v.z;
                acc.w += v.w;
            }
        }
        const uint32_t pz = 4 * h + tz;
#pragma unroll
        for (int y = 0; y < 4; y++) {
            const uint32_t yr = 8 * yb + 4 * yh + (uint32_t)y;
            const int4     v = ex4[(tz * R + yr) * Q + xq];
            acc.x += v.x;
            acc.y += v.y;
            acc.z += v.z;
            acc.w += v.w;
            if (pz >= nz || yr >= rows) continue;
            const size_t base = (size_t)(z0 + pz) * g.XY + (size_t)(y0 + yr) * g.X + xc;
            const float  o[4] = {(float)acc.x * ebx2, (float)acc.y * ebx2, (float)acc.z * ebx2, (float)acc.w * ebx2};
            if (full) {
                typedef float f32x4a4_t __attribute__((ext_vector_type(4), aligned(4)));
                *reinterpret_cast<f32x4a4_t*>(out + base) = f32x4a4_t{o[0], o[1], o[2], o[3]};
            } else {
#pragma unroll
                for (int j = 0; j < 4; j++)
                    if (xc + j < g.X) out[base + j] = o[j];
            }
        }
    }
}

template <int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NT >= 512 ? 8 : 7))) void k_decode_recon_3d(const uint8_t* __restrict__ payload,
                                                        const uint32_t* __restrict__ tile_excl,
                                                        const uint32_t* __restrict__ chunk_sums,
                                                        const uint16_t* __restrict__ tail, Geom g, float ebx2,
                                                        float* __restrict__ out, uint32_t Q, uint32_t R,
                                                        uint32_t mq, uint32_t rx, uint32_t ry, uint32_t nwin)
{
    __shared__ uint32_t cpre[kTileLocalMaxChunks];  // block counts of the chunks before each chunk
    __shared__ uint32_t wtot[4];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    // the chunk prefix, once per workgroup (waves 0-3; R Q >= 256)
    {
        uint32_t cv = 0, cincl = 0;
        if (tid < kTileLocalMaxChunks) {
            cv = (chunk_sums && tid < (g.T + kWG - 1) / kWG) ? chunk_sums[tid] : 0u;
            cincl = wave_incl_scan(cv);
            if (lane == 63) wtot[wv] = cincl;
        }
        __syncthreads();
        if (tid < kTileLocalMaxChunks) {
            uint32_t pre = 0;
            for (uint32_t k = 0; k < wv; k++) pre += wtot[k];
            cpre[tid] = pre + cincl - cv;
        }
    }
    // windows it, it + G, ... (G = gridDim.x, a multiple of 8 when G < nwin): a workgroup issues
    // its next window's loads while the stores of the last one drain.  XCD-sliced order: G % 8 ==
    // 0 keeps a workgroup on one XCD's windows, and windows that share a QG (neighbours along x,
    // or along y / z when rows or planes are whole) read its blocks through one L2.
    for (uint32_t it = blockIdx.x; it < nwin; it += gridDim.x)
        d3_window<NT>(payload, tile_excl, chunk_sums ? cpre : nullptr, tail, g, ebx2, out, Q, R, mq, rx, ry,
                      xcd_tile(it, nwin));
}

// compute units of the current device (256 on MI355X when the query fails)
static uint32_t device_cus()
{
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return 256;
    }
    return (uint32_t)n;
}

// tight = false (single field): 128 < X <= 256 takes 256-column windows of 512 threads, whose
// decode has more half-waves per QG list (one round where a tight window of 320-448 threads
// takes two); a lone field's windows mostly fit one round of the device, so their latency is
// its time (130^3: 19.4 against 21.0 us).  Batches take tight windows (more fields per round).
static void d3_shape(const Geom& g, uint32_t& Q, uint32_t& R, bool tight)
{
    if (g.X <= 128) {
        Q = 4;
        while (4 * Q < g.X) Q *= 2;  // 4 Q >= X, Q <= 32
        R = 256 / Q;
    } else if (g.X <= 256 && !tight) {
        Q = 64;
        R = 8;
    } else {
        const uint32_t nwx = (g.X + 511) / 512, w = (g.X + nwx - 1) / nwx;  // windows per row, columns of each
        Q = ((w + 3) / 4 + 7) & ~7u;                                         // <= 128
        R = 8;
    }
}

// the kernel template for windows of n = R Q threads
static int d3_nt(uint32_t n) { return n <= 256 ? 256 : n <= 512 ? 512 : 1024; }

// windows of the fused 3D decompress the device holds at once, a multiple of 8: per CU the
// fewer of what its LDS (160 KB) and its wave slots (4 SIMDs x 8 waves at 64 VGPRs, 7 for the
// 256-thread kernel) allow
static uint32_t d3_slots(uint32_t Q, uint32_t R)
{
    static uint32_t lds_static[3] = {0, 0, 0};  // the kernels' own LDS bytes
    const uint32_t  n = Q * R;
    const int       nt = d3_nt(n), c = nt == 256 ? 0 : nt == 512 ? 1 : 2;
    if (!lds_static[c]) {
        hipFuncAttributes a{};
        const void* f = nt == 256 ? reinterpret_cast<const void*>(&k_decode_recon_3d<256>)
                        : nt == 512 ? reinterpret_cast<const void*>(&k_decode_recon_3d<512>)
                                    : reinterpret_cast<const void*>(&k_decode_recon_3d<1024>);
        if (hipFuncGetAttributes(&a, f) == hipSuccess && a.sharedSizeBytes > 0) lds_static[c] = (uint32_t)a.sharedSizeBytes;
        else {
            (void)hipGetLastError();
            lds_static[c] = 12288;
        }
    }
    const uint32_t by_lds = 163840u / (lds_static[c] + 64 * n);
    const uint32_t by_waves = 4 * (nt == 256 ? 7u : 8u) / (n / 64);
    const uint32_t per_cu = by_lds < by_waves ? by_lds : by_waves;
    return device_cus() * (per_cu ? per_cu : 1u) & ~7u;
}

// Measured against decode + reconstruct (back-to-back calls, waves fields at eb 1e-3): the fused
// form wins on every field with X <= 256 (128 x 128 x 512: 18.3 against 28.0 us, 256^3: 26.3
// against 38.6 us, 130^3: 18.0 / 18.3).  Wider fields win when their windows take up to 1.75
// rounds of the device's window slots (270 x 300 x 100, 494 windows at 768 slots: 26.9 / 41.2
// us; 1000 x 1000 x 20, 1.46 rounds: 39.0 / 43.9; 500 x 500 x 100, 1.6 rounds: 48.6 / 50.3) or
// three rounds or more (2000 x 300 x 100: 101.0 / 121.2 us, 1024 x 1024 x 64: 94.1 / 127.7), and
// lose in between, where the last round holds only part of the device and the windows of a
// round run in step (300^3, 1.88 rounds: 55.4 / 51.6 us; 700 x 500 x 60, 1.97: 50.1 / 46.7;
// 600 x 600 x 100, 2.54: 80.7 / 58.9).
bool decode_recon_3d_pays(const Geom& g)
{
    if (g.ndim != 3) return false;
    if (g.X <= 256) return true;
    uint32_t Q, R;
    d3_shape(g, Q, R, false);
    const uint64_t C = 4 * Q, rx = (g.X + C - 1) / C;
    const uint64_t nwin = rx * ((g.Y + R - 1) / R) * ((g.Z + 7) / 8), slots = d3_slots(Q, R);
    return 4 * nwin <= 7 * slots || nwin >= 3 * slots;
}

hipError_t launch_decode_recon_3d(const uint8_t* payload, const uint32_t* tile_excl, const uint32_t* chunk_sums,
                                  const uint16_t* tail, const Geom& g, float ebx2, float* out, hipStream_t s)
{
    uint32_t Q, R;
    d3_shape(g, Q, R, false);
    const uint32_t C = 4 * Q, n_t = Q * R;
    const uint32_t rx = (g.X + C - 1) / C, ry = (g.Y + R - 1) / R;
    const uint64_t n = (uint64_t)rx * ry * ((g.Z + 7) / 8);
    if (n > 0xFFFFFFFFull) return hipErrorInvalidConfiguration;
    const uint32_t slots = d3_slots(Q, R);
    const dim3     grid(n <= slots ? (uint32_t)n : slots);
    const size_t   lds = 64ull * n_t;  // the window
    const int      nt = d3_nt(n_t);
    if (nt == 1024) FZ_LAUNCH(k_decode_recon_3d<1024><<<grid, n_t, lds, s>>>(payload, tile_excl, chunk_sums, tail, g, ebx2, out, Q, R, d3_magic(Q), rx, ry, (uint32_t)n));
    else if (nt == 512) FZ_LAUNCH(k_decode_recon_3d<512><<<grid, n_t, lds, s>>>(payload, tile_excl, chunk_sums, tail, g, ebx2, out, Q, R, d3_magic(Q), rx, ry, (uint32_t)n));
    else FZ_LAUNCH(k_decode_recon_3d<256><<<grid, n_t, lds, s>>>(payload, tile_excl, chunk_sums, tail, g, ebx2, out, Q, R, d3_magic(Q), rx, ry, (uint32_t)n));
    return hipGetLastError();
}

// Batched decompress: the fused 3D windows of every field of one window size (R Q threads), a
// workgroup per window; the field comes from the per-window table (one scalar load), its tile
// offsets are the per-field scan's (global within the field, no chunk prefix).  Windows in
// XCD-sliced order.
template <int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NT >= 512 ? 8 : 7))) void k_batch_fused3d(BatchDev d, uint32_t w_lo)
{
    const uint32_t w = w_lo + xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t f = d.wfield[w];
    const BField&  F = d.fields[f];
    const Geom     g = F.g;
    const BPtrs    q = d.ptrs[f];
    d3_window<NT>(static_cast<const uint8_t*>(q.src), d.texcl + F.ustart, nullptr, g.mode == 0 ? q.tail : nullptr, g,
                  (float)q.param, static_cast<float*>(q.dst), F.wq, F.wr, F.wm, F.wdims[0], F.wdims[1], w - F.wstart);
}

// In a batch the windows of many fields fill the device together, so every 3D field takes the
// fused form, in tight windows (within 32 columns of what they cover): the 256-field batch's
// 3D decode + reconstruct 3.25 -> 3.19 ms against fused windows for X <= 256 or 3/4-full
// power-of-two windows only.
bool batch_fused3d(const Geom& g, uint32_t& Q, uint32_t& R, uint32_t& rx, uint32_t& ry, uint64_t& nwin)
{
    if (g.ndim != 3) return false;
    d3_shape(g, Q, R, true);
    rx = (g.X + 4 * Q - 1) / (4 * Q);
    ry = (g.Y + R - 1) / R;
    nwin = (uint64_t)rx * ry * ((g.Z + 7) / 8);
    return true;
}

hipError_t launch_batch_fused3d(const BatchDev& d, const BatchShape& b, hipStream_t s)
{
    for (int c = 0; c < kWinClasses; c++) {
        const uint32_t n = b.w_hi[c] - b.w_lo[c], n_t = 64u * (c + 1);
        if (n == 0) continue;
        const size_t lds = 64ull * n_t;
        const int    nt = d3_nt(n_t);
        if (nt == 256) FZ_LAUNCH(k_batch_fused3d<256><<<n, n_t, lds, s>>>(d, b.w_lo[c]));
        else if (nt == 512) FZ_LAUNCH(k_batch_fused3d<512><<<n, n_t, lds, s>>>(d, b.w_lo[c]));
        else FZ_LAUNCH(k_batch_fused3d<1024><<<n, n_t, lds, s>>>(d, b.w_lo[c]));
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Batched decompress: the fused 2D windows of every field whose rows are a multiple of 4 wide (a
// lane's codes go into the window with wide LDS stores; then the fused form beats decode +
// reconstruct at every size measured), a workgroup per window over the per-window field
// table, with the per-field scan's tile offsets (global within the field).
__global__ __launch_bounds__(kD2Threads) void k_batch_fused2d(BatchDev d, uint32_t w_lo)
{
    const uint32_t w = w_lo + xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t f = d.wfield[w];
    const BField&  F = d.fields[f];
    const Geom     g = F.g;
    const BPtrs    q = d.ptrs[f];
    d2_window<false>(static_cast<const uint8_t*>(q.src), d.texcl + F.ustart, nullptr, g.mode == 0 ? q.tail : nullptr, g,
                     (float)q.param, static_cast<float*>(q.dst), F.wdims[0], w - F.wstart);
}

bool batch_fused2d(const Geom& g, uint32_t& rx, uint32_t& ry)
{
    if (g.ndim != 2 || (g.X & 3u) != 0) return false;
    rx = (g.X + kD2Cols - 1) / kD2Cols;
    ry = (g.Y + 15) / 16;
    return true;
}

hipError_t launch_batch_fused2d(const BatchDev& d, const BatchShape& b, hipStream_t s)
{
    const uint32_t n = b.w2_hi - b.w2_lo;
    if (n == 0) return hipSuccess;
    FZ_LAUNCH(k_batch_fused2d<<<n, kD2Threads, 0, s>>>(d, b.w2_lo));
    return hipGetLastError();
}

template <typename SRC>
__global__ __launch_bounds__(kWG) void k_reconstruct_1d(const SRC codes, float* __restrict__ out, Geom g, float ebx2)
{
    recon_1d(codes, out, g, ebx2, blockIdx.x);
}

// 2D and 3D grids are one-dimensional (workgroup -> (bx, by[, bz]) by division), so tall
// fields are not bound by the y / z grid limits
template <typename SRC>
__global__ __launch_bounds__(kWG) void k_reconstruct_2d(const SRC codes, float* __restrict__ out, Geom g, float ebx2,
                                                        uint32_t rx)
{
    recon_2d(codes, out, g, ebx2, blockIdx.x % rx, blockIdx.x / rx);
}

template <typename SRC>
__global__ __launch_bounds__(kWG) void k_reconstruct_3d(const SRC codes, float* __restrict__ out, Geom g, float ebx2,
                                                        int q, uint32_t rx, uint32_t ry)
{
    // XCD-sliced order: neighbouring narrow row segments (down to 32 bytes) share one L2
    recon_3d_any(q, codes, out, g, ebx2, xcd_tile(blockIdx.x, gridDim.x), rx, ry);
}

template <typename SRC>
static hipError_t launch_reconstruct_t(const SRC codes, float* out, const Geom& g, float ebx2, hipStream_t s)
{
    if (g.ndim == 1) {
        const uint32_t thr = (g.xyz + 15) / 16;
        FZ_LAUNCH(k_reconstruct_1d<SRC><<<(thr + kWG - 1) / kWG, kWG, 0, s>>>(codes, out, g, ebx2));
    } else if (g.ndim == 2) {
        const uint32_t rx = (g.X + kRecon2DCols - 1) / kRecon2DCols;
        const uint64_t n = (uint64_t)rx * ((g.Y + 15) / 16);
        if (n > 0xFFFFFFFFull) return hipErrorInvalidConfiguration;
        FZ_LAUNCH(k_reconstruct_2d<SRC><<<(uint32_t)n, kWG, 0, s>>>(codes, out, g, ebx2, rx));
    } else {
        const uint32_t q = recon3d_q(g.X, g.Y);
        const uint32_t rx = (g.X + 4 * q - 1) / (4 * q), ry = (g.Y + kWG / q - 1) / (kWG / q);
        const uint64_t n = (uint64_t)rx * ry * ((g.Z + 7) / 8);
        if (n > 0xFFFFFFFFull) return hipErrorInvalidConfiguration;
        FZ_LAUNCH(k_reconstruct_3d<SRC><<<(uint32_t)n, kWG, 0, s>>>(codes, out, g, ebx2, (int)q, rx, ry));
    }
    return hipGetLastError();
}

hipError_t launch_reconstruct(const uint16_t* codes, float* out, const Geom& g, float ebx2, hipStream_t s)
{
    return launch_reconstruct_t(codes, out, g, ebx2, s);
}

hipError_t launch_reconstruct_outlier_src(const uint16_t* codes, const uint32_t* bits, const int32_t* over,
                                          float* out, const Geom& g, float ebx2, hipStream_t s)
{
    return launch_reconstruct_t(OutlierSrc{codes, bits, over}, out, g, ebx2, s);
}

}  // namespace fz