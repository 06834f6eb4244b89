// fz_internal.hpp -- kernel launch interface shared by the C ABI (fz_capi.hip) and the
// kernels (fz_kernels.hip).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fz_capi.h"
#include "fz_device.hpp"

namespace fz {

constexpr uint32_t kRangeMaxWG = 256;  // partial-range workgroups (a multiple of 8, <= 256)

// Counters at the start of every workspace.
struct Ctr {
    uint32_t unused0;
    uint32_t unused1;
    uint32_t scan_ticket;   // chunk tickets of the flag scan (reset by the range kernel / scan init)
    uint32_t dscan_ticket;  // tickets of the decompress tile-count scan (reset by its init)
    uint32_t pad[60];
};

constexpr uint32_t kScanChunk = 4096;  // items per workgroup of the u32 scan
constexpr uint32_t kGatherWG = 256;  // max workgroups of the compress scan + gather

// partial-range workgroups (1024 threads each): a multiple of 8 (one slice per XCD region),
// at least ~16K floats each, at most 256 (the tile kernel reduces one partial per thread), so
// every CU takes a slice once the field has 4M floats
inline uint32_t range_grid(uint32_t len)
{
    uint64_t g = ((uint64_t)len + 16383) / 16384;
    g = (g + 7) & ~7ull;
    if (g < 8) g = 8;
    if (g > kRangeMaxWG) g = kRangeMaxWG;
    return (uint32_t)g;
}

// Quads per row Q of the 3D reconstruct workgroup (4Q columns x 256/Q rows x 8 planes) for an
// X x Y plane: the Q in {32, 16, 8, 4} whose workgroups leave the fewest lanes past the
// field's edges (ties go to the larger Q).
inline uint32_t recon3d_q(uint32_t X, uint32_t Y)
{
    uint32_t best = 32;
    uint64_t best_cells = ~0ull;  // lanes the grid spans for the plane, per 4 columns
    for (uint32_t q = 32; q >= 4; q /= 2) {
        const uint64_t c = 4ull * q, r = kWG / q;
        const uint64_t cells = ((X + c - 1) / c) * c * ((Y + r - 1) / r) * r;
        if (cells < best_cells) {
            best_cells = cells;
            best = q;
        }
    }
    return best;
}

// -------- launchers (return hipError_t of the launch)
hipError_t launch_range(const float* in, const Geom& g, double eb_rel, float4* part, fz_result* res,
                        uint64_t* zero_states, uint32_t nzero, uint32_t* zero_ticket, hipStream_t s);
// Fused compress (3 kernels).  wave_tile: 1 = one wave per tile (k_tile_compress_w), 0 = one
// workgroup per tile.  ev (may be null) holds start / stop events of each kernel.
hipError_t launch_compress(const float* in, const Geom& g, double eb_rel, float4* part, fz_result* res,
                           uint8_t* payload, uint16_t* tail, uint8_t* staging, uint32_t* counts, uint64_t* states,
                           int wave_tile, hipEvent_t* ev, hipStream_t s);
uint32_t   gather_grid(uint32_t T);
uint32_t   gather_tpw(uint32_t T);
// placement kernel of the fused compress on its own (run_sums filled by a tile kernel)
hipError_t launch_scan_gather(const uint32_t* counts, uint32_t T, const uint8_t* staging, uint8_t* payload,
                              const uint32_t* run_sums, fz_result* res, uint32_t W, hipStream_t s);
// r_host / r_amax: ebx2_r and max|x| when res is null (the per-stage entry point)
hipError_t launch_dualquant(const float* in, const Geom& g, float r_host, float r_amax, const fz_result* res,
                            uint16_t* codes, unsigned long long* ovf, hipStream_t s);
hipError_t launch_bitshuffle(const uint32_t* in, uint32_t* out, const Geom& g, uint32_t* bitflags,
                             uint32_t* byteflags, bool with_tail_wg, hipStream_t s);
hipError_t launch_unshuffle(const uint32_t* in, uint32_t* out, uint32_t ntiles, hipStream_t s);
hipError_t launch_scan_init(uint64_t* states, uint32_t nstates, uint32_t* ticket, hipStream_t s);
hipError_t launch_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint64_t* states, uint32_t* ticket,
                       bool write_total, fz_result* res, uint32_t W, hipStream_t s);
hipError_t launch_encode(const uint8_t* in, uint8_t* out, const uint32_t* presum, uint32_t nblocks, hipStream_t s);
// tile_excl[t] = nonzero blocks of tiles < t by the look-back scan; with `clear` after a
// launch that clears the nstates states and the ticket, else the caller has cleared them.
hipError_t launch_tile_scan(const uint32_t* bitflags, uint32_t T, uint32_t* tile_excl, uint64_t* states,
                            uint32_t nstates, uint32_t* ticket, hipStream_t s, bool clear = true);
// The tile offsets the decode kernels take, in one launch up to 65536 tiles: tile_excl[t] =
// nonzero blocks of the tiles before t in its 256-tile chunk, *chunk_sums = the chunk counts
// (in the states' memory); beyond that the look-back scan's global offsets and *chunk_sums =
// nullptr.
hipError_t launch_tile_offsets(const uint32_t* bitflags, uint32_t T, uint32_t* tile_excl, uint64_t* states,
                               uint32_t nstates, uint32_t* ticket, hipStream_t s, const uint32_t** chunk_sums);
// chunk_sums: as launch_tile_offsets gives it (nullptr: tile_excl holds global offsets)
hipError_t launch_decode_unshuffle(const uint8_t* payload, const uint32_t* tile_excl, const uint32_t* chunk_sums,
                                   const uint16_t* tail, const Geom& g, uint32_t* codes, hipStream_t s);
// 1D fields: decode + reconstruct to floats in one launch (replaces the two above)
hipError_t launch_decode_recon_1d(const uint8_t* payload, const uint32_t* tile_excl, const uint32_t* chunk_sums,
                                  const uint16_t* tail, const Geom& g, float ebx2, float* out, hipStream_t s);
// 2D fields: decode + reconstruct to floats in one launch (a workgroup per 16 x 1024 window),
// for fields whose windows are few enough that it beats decode + reconstruct
bool       decode_recon_2d_pays(const Geom& g);
// 2D fields small enough (T <= 4096, X <= 4096) that each window finds its tiles' offsets from
// the bitflags itself (launch_decode_recon_2d with tile_excl = chunk_sums = nullptr, no
// tile-offset launch)
bool       decode_recon_2d_direct(const Geom& g);
hipError_t launch_decode_recon_2d(const uint8_t* payload, const uint32_t* tile_excl, const uint32_t* chunk_sums,
                                  const uint16_t* tail, const Geom& g, float ebx2, float* out, hipStream_t s);
// 3D fields: decode + reconstruct to floats in one launch (a workgroup per window of 8 planes x
// R rows x 4Q columns)
bool       decode_recon_3d_pays(const Geom& g);
hipError_t launch_decode_recon_3d(const uint8_t* payload, const uint32_t* tile_excl, const uint32_t* chunk_sums,
                                  const uint16_t* tail, const Geom& g, float ebx2, float* out, hipStream_t s);
hipError_t launch_reconstruct(const uint16_t* codes, float* out, const Geom& g, float ebx2, hipStream_t s);
// Batched decompress: whether a 3D field is decoded and reconstructed by the fused window kernel
// (then its windows' quads per row, rows, windows along x and y, and window count)
// floor(t / Q) = umulhi(t, d3_magic(Q)) for t < 2^32 / Q (thread indices of a window)
inline uint32_t d3_magic(uint32_t Q) { return 0xFFFFFFFFu / Q + 1; }
bool       batch_fused3d(const Geom& g, uint32_t& Q, uint32_t& R, uint32_t& rx, uint32_t& ry, uint64_t& nwin);
struct BatchDev;
struct BatchShape;
hipError_t launch_batch_fused3d(const BatchDev& d, const BatchShape& b, hipStream_t s);
// Batched decompress: whether a 2D field is decoded and reconstructed by the fused window
// kernel (then its windows along x and y), and the launch over every such window
bool       batch_fused2d(const Geom& g, uint32_t& rx, uint32_t& ry);
hipError_t launch_batch_fused2d(const BatchDev& d, const BatchShape& b, hipStream_t s);
// reconstruct from u16 codes whose positions flagged in `bits` take their delta from `over`
hipError_t launch_reconstruct_outlier_src(const uint16_t* codes, const uint32_t* bits, const int32_t* over,
                                          float* out, const Geom& g, float ebx2, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s);

// Outlier mode (fz_outlier.hip).  Workspace pieces: ocnt holds 2 u32 per unit (T + 1 units),
// orun 2 u32 per placement run; for the inverse, obits one bit per position (zero between
// calls) and over the side channel's deltas at their positions.
inline uint32_t outlier_runs(uint32_t T) { return T ? gather_grid(T) : 1u; }
hipError_t launch_compress_outlier(const float* in, const Geom& g, double eb_rel, float4* part, fz_result* res,
                                   uint8_t* payload, uint16_t* tail, uint8_t* side, uint64_t side_cap,
                                   uint8_t* staging, uint32_t* counts, uint32_t* run_sums, uint32_t nsums,
                                   uint32_t* ocnt, uint32_t* orun, hipStream_t s);
// before the tile scan (launched with clear = false): the side channel's deltas into `over`,
// flagged in obits, and the scan's look-back states and ticket cleared
hipError_t launch_put_deltas(const uint8_t* side, uint64_t side_bytes, uint32_t* obits, int32_t* over, const Geom& g,
                             uint64_t* states, uint32_t nstates, uint32_t* ticket, hipStream_t s);
// after the codes of the plain inverse are in `codes`: prefix sums with the deltas flagged in
// obits, the verbatim values, obits cleared again
hipError_t launch_reconstruct_outlier(const uint16_t* codes, uint32_t* obits, int32_t* over, const uint8_t* side,
                                      uint64_t side_bytes, float* out, const Geom& g, float ebx2, hipStream_t s);
// constant field: out = vconst, then the verbatim values
hipError_t launch_fill_outlier(const uint8_t* side, uint64_t side_bytes, uint32_t* obits, float* out, const Geom& g,
                               float vconst, hipStream_t s);

// Quality metrics (fz_verify.hip): scratch bytes for n elements, and the two-kernel launch.
size_t     verify_workspace_bytes(uint64_t n);
hipError_t launch_verify(const float* x, const float* o, uint64_t n, void* temp, fz_stats* stats, hipStream_t s);

// ============================================================================ batched form
//
// Many fields of uneven shapes compressed by one launch sequence.  Fields are ordered by
// dimensionality (1D, 2D, 3D).  Every field owns T + 1 consecutive "units" of one global
// unit space (its T full tiles, then the unit that handles the partial tile); a unit's
// 4096-byte staging slot, block count and destination address sit at its unit index.

constexpr uint32_t kBatchMaxFields = 4096;  // two wave_upper rounds
constexpr int      kPtrChunk = 56;          // pointer records per k_batch_set launch (1792 B of arguments)
constexpr int      kNumChunks = 3;          // range-partial sizes the workspace holds tables for
constexpr int      kWinClasses = 16;        // fused 3D decompress window sizes: 64 (c + 1) threads
constexpr uint32_t kRangeChunks[kNumChunks] = {16384, 65536, 262144};

struct BField {        // per field, uploaded once per workspace (128 B)
    Geom     g;        // 72 B
    uint32_t caller;   // index of the field in the caller's arrays (its result record)
    uint32_t ustart;   // first global unit
    uint32_t rdims[3]; // reconstruct workgroups along x, y, z
    uint32_t rq;       // 3D: quads per row of its reconstruct workgroups (recon3d_q)
    uint32_t wnt;      // decoded and reconstructed in one kernel: threads per window (0: not)
    uint32_t wq;       // ... quads per window row
    uint32_t wr;       // ... rows per window
    uint32_t wdims[2]; // ... windows along x, y
    uint32_t wstart;   // ... first window (batch window space)
    uint32_t wm;       // ... d3_magic(wq)
    uint32_t pad;
};
static_assert(sizeof(BField) == 128, "BField layout");

struct BPtrs {         // per field and call (32 B)
    const void* src;   // compress: input floats; decompress: payload
    void*       dst;   // compress: payload; decompress: output floats
    uint16_t*   tail;  // FIXED-mode tail codes (null when the field has none)
    double      param; // compress: eb_rel; decompress: ebx2
};
static_assert(sizeof(BPtrs) == 32, "BPtrs layout");

struct BParams {       // per field, written by the range kernel of the call
    float    r;        // ebx2_r
    uint32_t status;
    uint32_t level;    // dual-quant paths allowed by the bound on |s| (EbParams::level)
    uint32_t pad;
};

struct PtrChunk {
    BPtrs p[kPtrChunk];
};

// device addresses of one batch workspace
struct BatchDev {
    const BField*   fields;
    const uint32_t* ustart;  // nf + 1, global unit of each field (sorted order)
    const uint32_t* ufield;  // per unit: its field (sorted order)
    const uint32_t* pstart;  // nf + 1, first range partial of each field (selected chunk)
    const uint32_t* rstart;  // nf + 1, first reconstruct workgroup of each field
    const uint32_t* rfield;  // per reconstruct workgroup: its field (sorted order)
    const uint32_t* wfield;  // per fused 3D window: its field (sorted order)
    BPtrs*          ptrs;
    BParams*        prm;
    float4*         part;
    uint32_t*       counts;  // per unit
    uint64_t*       dsts;    // per unit: payload address of its first block
    uint32_t*       texcl;   // per unit: blocks before the tile in its field (decompress)
    uint8_t*        staging; // 4096 B per unit (compress staging, decompress codes)
    fz_result*      res;     // caller's result array (compress)
};

struct BatchShape {
    uint32_t nf;
    uint32_t f_lo[4], f_hi[4];  // fields of dimensionality d (index 1..3)
    uint32_t u_lo[4], u_hi[4];  // their units
    uint32_t r_lo[4], r_hi[4];  // their reconstruct workgroups
    uint32_t U;                 // all units
    uint32_t P;                 // range partials
    uint32_t chunk;             // floats per range partial
    uint32_t xcd[4];            // tile order per dimensionality: 1 XCD-sliced, 0 linear
    uint32_t wave_tile[4];      // tile layout per dimensionality: 1 one wave per tile, 0 four
    uint32_t gather_waves;      // units per gather workgroup
    uint32_t w_lo[kWinClasses], w_hi[kWinClasses];  // fused 3D windows of 64 (c + 1) threads (decompress)
    uint32_t w2_lo, w2_hi;                          // fused 2D windows (decompress)
};

// A group of consecutive fields (sorted order) and the units / range partials they own.
constexpr int kMaxGroups = 16;
struct BatchGroup {
    uint32_t f_lo, f_hi, u_lo, u_hi;
    uint32_t p_lo[kNumChunks], p_hi[kNumChunks];
    uint32_t df_lo[4], df_hi[4];  // its fields of dimensionality d (index 1..3)
    uint32_t du_lo[4], du_hi[4];  // their units
};

// Streams and events of the grouped schedule: `range` and `gather` fork from the caller's
// stream and join back into it.
struct BatchStreams {
    hipStream_t range, gather;
    hipEvent_t  fork, join_r, join_g, r[kMaxGroups], t[kMaxGroups];
};

hipError_t launch_batch_set(const BPtrs* sorted, uint32_t nf, BPtrs* dev, hipStream_t s);
// Groups run in order; with st == null everything runs on s (and ev, when given, receives
// start / stop pairs of range_part, range_final, tile 1D/2D/3D, scan, gather of a single group).
hipError_t launch_batch_compress(const BatchDev& d, const BatchShape& b, const BatchGroup* groups, int ngroups,
                                 int chunk, const BatchStreams* st, hipEvent_t* ev, hipStream_t s);
// ev (may be null): start / stop pairs of dscan, decode, reconstruct 1D/2D/3D
hipError_t launch_batch_decompress(const BatchDev& d, const BatchShape& b, hipEvent_t* ev, hipStream_t s);

}  // namespace fz
