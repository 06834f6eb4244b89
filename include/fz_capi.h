/*
 * fz_capi.h -- C ABI of the MI355X (gfx950) FZ compression path.
 *
 * Path: float32 field + value-range-relative error bound
 *       -> Lorenzo dual-quant (sign-magnitude u16 codes)
 *       -> bitshuffle of 2048-code tiles + zero-block bit flags
 *       -> exclusive prefix sum over the flags
 *       -> sparse encode of the nonzero 16-byte blocks
 *       => payload = LE u32 bitflags B[0..W) || nonzero blocks (16 K bytes)
 * and its inverse.  The payload bytes equal the reference's d_out layout and the size
 * formula of fz.cu:208 (4 W + 16 K).
 *
 * Conventions (mirroring the reference's C launch layer, claunch_cuda.h:17-29):
 *   - extern "C", every call returns fz_status (0 = success), never exits;
 *   - device pointers only, the caller owns every buffer; stream argument last;
 *   - the compress/decompress entry points allocate nothing, synchronise nothing and
 *     launch kernels only, so they can be captured into a hipGraph;
 *   - workspace sizes are queried first (the CUB two-phase idiom of fz.cu:188-192).
 *
 * Geometry: x is the fastest axis (fz.cu:231-233, len3 = dim3(x, y, z) at fz.cu:116);
 * ndim = 1 if y == z == 1, 2 if z == 1, else 3 (lorenzo_var.cuh:537-544).  `len` is the
 * number of floats in the input (the reference takes it from the file size, fz.cu:117-118)
 * and must be >= x*y*z; positions >= x*y*z get code 0.
 */
#ifndef FZ_CAPI_H
#define FZ_CAPI_H

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FZ_CAPI_VERSION 1

/* Status codes.  The reference prints and exit()s on CUDA errors (cuda_err.cuh:169) and
 * throws for ndim > 3 (lorenzo_var.cuh:587); here every failure is a return code. */
typedef enum fz_status {
    FZ_SUCCESS = 0,
    FZ_ERR_INVALID_ARG = 1,         /* bad dims / eb / mode / alignment / len < x*y*z   */
    FZ_ERR_WORKSPACE_TOO_SMALL = 2, /* workspace smaller than fz_workspace_bytes()      */
    FZ_ERR_HIP = 3,                 /* a HIP runtime call failed                        */
    FZ_ERR_CONSTANT_FIELD = 4,      /* max == min: the reference divides by zero (D5)   */
    FZ_ERR_NONFINITE = 5,           /* NaN or Inf in the input                          */
    FZ_ERR_CORRUPT = 6,             /* payload inconsistent with its bitflags           */
    FZ_ERR_CAPACITY = 7,            /* request exceeds a fixed capacity                 */
    FZ_ERR_NOT_READY = 8            /* result record not yet written                    */
} fz_status;

/* FZ_MODE_FIXED: 1D codes keep the sign of every delta (the reference drops the sign of
 * 3 of 4 deltas, lorenzo_var.cuh:59-63) and the len mod 2048 codes that no tile covers
 * (never encoded by the reference, fz.cu:173) are returned in the tail buffer.
 * FZ_MODE_LITERAL: byte-identical to the reference, including both losses.
 * 2D and 3D payloads are identical in the two modes.
 * FZ_MODE_OUTLIER (opt-in, format-changing; the outlier entry points only): the FIXED
 * payload and tail plus a side channel, modelled on cuSZ's sparse outlier vector
 * (spcodec_vec.hh:29), so that every field compresses and every value comes back: deltas
 * with |d| >= 32768 (code 0 in the payload, exact delta in the channel), non-finite values and
 * values with |x * ebx2_r| >= 2^24 (kept verbatim; quantised as 0), and constant fields
 * (ebx2 = 0, the inverse fills vmin).  Deltas are exact integers, so the inverse is exact up to
 * the error bound.  Side channel (16-byte aligned, little-endian):
 *   u64 n_delta, u64 n_raw, u32 idx_d[n_delta], i32 val_d[n_delta],
 *   u32 idx_r[n_raw], u32 val_r[n_raw] (input bits),  both lists in increasing index order;
 * its size is 16 + 8 (n_delta + n_raw) bytes, at most fz_sizes.side_max_bytes.  Where no value
 * needs the channel and the reference's float arithmetic is exact (every |x * ebx2_r| < 2^21)
 * the payload equals the FIXED mode's. */
typedef enum fz_mode { FZ_MODE_FIXED = 0, FZ_MODE_LITERAL = 1, FZ_MODE_OUTLIER = 2 } fz_mode;

/* Device-resident record written by the compress kernels (64 bytes). */
typedef struct fz_result {
    uint32_t status;        /* fz_status of the device-side checks (range, finiteness)  */
    uint32_t ndim;
    float    vmin, vmax;    /* value range (fz.cu:137)                                   */
    float    ebx2_r, ebx2;  /* float kernel parameters (lorenzo_var.cuh:562-563)         */
    double   eb_abs;        /* eb_rel * (vmax - vmin)  (fz.cu:149)                        */
    uint64_t nblocks;       /* K: nonzero 16-byte blocks                                 */
    uint64_t payload_bytes; /* 4 W + 16 K  (fz.cu:208)                                   */
    uint64_t tail_len;      /* codes returned in the tail buffer (0 in LITERAL mode)     */
    uint64_t overflow;      /* deltas with |d| >= 32768 (their code aliases the sign;    */
                            /* OUTLIER mode: n_delta of the side channel)                */
} fz_result;

/* Sizes of one field's buffers. */
typedef struct fz_sizes {
    uint64_t len;               /* floats in the input                                   */
    uint64_t tiles;             /* T = floor(len / 2048)            (fz.cu:173)           */
    uint64_t flag_words;        /* W = ceil(len / 256)              (fz.cu:177-178)       */
    uint64_t byte_flags;        /* N = 32 W                                               */
    uint64_t payload_max_bytes; /* 4 W + 4096 T: every block nonzero                      */
    uint64_t tail_max_codes;    /* len - 2048 T (< 2048)                                  */
    uint64_t workspace_bytes;   /* device scratch for every entry point of this plan      */
    uint64_t side_max_bytes;    /* OUTLIER mode: 16 + 8 x*y*z (every value in the channel) */
} fz_sizes;

typedef struct fz_plan_s* fz_plan;

const char* fz_version(void);
const char* fz_status_string(fz_status s);
/* Text of the HIP error behind the calling thread's most recent FZ_ERR_HIP. */
const char* fz_last_error(void);

/* Plan = geometry + mode; host memory only.  Replaces the per-call set-up of runSzs
 * (fz.cu:113-169: sizes, cudaMalloc/cudaMemset).  len = 0 means x*y*z. */
fz_status fz_plan_create(fz_plan* plan, uint32_t x, uint32_t y, uint32_t z, uint64_t len, int mode);
fz_status fz_plan_destroy(fz_plan plan);
fz_status fz_plan_sizes(const fz_plan plan, fz_sizes* sizes);

/* Puts a freshly allocated workspace into its reset state (one fill kernel).  Call once
 * after allocating it; every entry point leaves it reset again when its kernels finish.
 * One workspace serves one stream at a time. */
fz_status fz_workspace_init(const fz_plan plan, void* d_ws, size_t ws_bytes, hipStream_t stream);

/* Fused single-field compress (replaces runSzs, fz.cu:113-209, without its file I/O and
 * host round trips), three kernels: device value range partials -> one kernel per
 * 2048-code tile doing dual-quant, bitshuffle, flags and in-tile compaction (each tile's
 * workgroup reduces the range partials to the eb itself) -> scan of the per-tile block counts
 * and placement of the compacted blocks.  Writes payload (capacity payload_max_bytes), the
 * tail codes (FIXED mode; capacity tail_max_codes) and *d_result.  d_in must be 16-byte
 * aligned. */
fz_status fz_compress_f32(const fz_plan plan, const float* d_in, double eb_rel, void* d_ws, size_t ws_bytes,
                          uint8_t* d_payload, uint16_t* d_tail, fz_result* d_result, hipStream_t stream);

/* Same output through the plain per-stage pipeline (one kernel per reference stage:
 * range, Lorenzo, bitshuffle+flags, scan init + scan, encode). */
fz_status fz_compress_plain_f32(const fz_plan plan, const float* d_in, double eb_rel, void* d_ws, size_t ws_bytes,
                                uint8_t* d_payload, uint16_t* d_tail, fz_result* d_result, hipStream_t stream);

/* Inverse: payload (+ tail codes in FIXED mode, NULL otherwise; NULL in FIXED mode is
 * refused when the plan has tail codes) -> x*y*z floats.  The payload is trusted: check an
 * archive read from outside with fz_payload_check first.
 * ebx2 = fz_result.ebx2 of the compress call.  The inverse of the reference's Lorenzo
 * (x_lorenzo_*, lorenzo_var.cuh:297-518) cannot decode sign-magnitude codes, so its block
 * structure is kept with exact int32 prefix sums and one float multiply by ebx2. */
fz_status fz_decompress_f32(const fz_plan plan, const uint8_t* d_payload, const uint16_t* d_tail, float ebx2,
                            void* d_ws, size_t ws_bytes, float* d_out, hipStream_t stream);

/* OUTLIER-mode compress (a plan created with FZ_MODE_OUTLIER; the other compress entry
 * points refuse such a plan): fz_compress_f32's payload, tail and record, and the side channel
 * in d_side (capacity side_capacity bytes, >= 16; the record's status is FZ_ERR_CAPACITY when
 * the channel does not fit, and the channel is then incomplete).  Read the channel's size from
 * its first 16 bytes.  No counterpart in fz-gpu; the side channel follows cuSZ's
 * SpcodecVec::encode (spcodec_vec.hh:47, launch_spv.cuh:22-49). */
fz_status fz_compress_outlier_f32(const fz_plan plan, const float* d_in, double eb_rel, void* d_ws,
                                  size_t ws_bytes, uint8_t* d_payload, uint16_t* d_tail, uint8_t* d_side,
                                  uint64_t side_capacity, fz_result* d_result, hipStream_t stream);

/* Its inverse: ebx2 and vconst are the record's ebx2 and vmin, side_bytes the channel's size
 * (16 + 8 (n_delta + n_raw)).  Follows SpcodecVec::decode's scatter (launch_spv.cuh:52-58)
 * into the deltas before the prefix sums, and of the verbatim values after them. */
fz_status fz_decompress_outlier_f32(const fz_plan plan, const uint8_t* d_payload, const uint16_t* d_tail, float ebx2,
                                    float vconst, const uint8_t* d_side, uint64_t side_bytes, void* d_ws,
                                    size_t ws_bytes, float* d_out, hipStream_t stream);

/* Host-side checks of an archive's pieces before any of them reaches the device (the
 * reference never reads an archive back, fz.cu:215, so these have no counterpart there).
 * fz_payload_check: h_payload (host memory, payload_bytes bytes) must hold at least the W
 * bitflag words, no flag may be set beyond the full tiles (the reference leaves those words
 * zero, fz.cu:167), payload_bytes must equal 4 W + 16 popcount(bitflags) (fz.cu:208) and so
 * be <= payload_max_bytes, and in FIXED and OUTLIER mode tail_len must equal
 * tail_max_codes (LITERAL mode ignores it).  fz_side_check: an OUTLIER-mode side channel of
 * side_bytes = 16 + 8 (n_delta + n_raw) bytes with both counts and every index below x*y*z.
 * Both return FZ_ERR_CORRUPT for an inconsistent input and read nothing past the bytes
 * given. */
fz_status fz_payload_check(const fz_plan plan, const uint8_t* h_payload, uint64_t payload_bytes, uint64_t tail_len);
fz_status fz_side_check(const fz_plan plan, const uint8_t* h_side, uint64_t side_bytes);

/* fz_compress_f32 without a workspace argument: no counterpart in the reference; for callers
 * that cannot allocate device memory (code under hipGraph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The workspace is a fixed device array inside the library,
 * one per GPU, of fz_noalloc_scratch_bytes() bytes (about 63 MiB), held from the library's load
 * on that GPU until the process ends.  Fields of at most FZ_NOALLOC_MAX_TILES full tiles
 * (len < 2048 * (FZ_NOALLOC_MAX_TILES + 1) floats) are accepted; larger ones return
 * FZ_ERR_CAPACITY.  The call allocates and frees nothing; the caller passes the input and the
 * outputs as for fz_compress_f32, whose bytes it writes.  Every call on one GPU uses that GPU's
 * array: calls that overlap on one GPU (other streams, other threads) race and must not be
 * issued. */
#define FZ_NOALLOC_MAX_TILES 16000
fz_status fz_compress_f32_noalloc(const fz_plan plan, const float* d_in, double eb_rel, uint8_t* d_payload,
                                  uint16_t* d_tail, fz_result* d_result, hipStream_t stream);
uint64_t  fz_noalloc_scratch_bytes(void);

/* Convenience: runs fz_compress_f32, waits for the stream and copies *d_result to *h_result. */
fz_status fz_compress_f32_sync(const fz_plan plan, const float* d_in, double eb_rel, void* d_ws, size_t ws_bytes,
                               uint8_t* d_payload, uint16_t* d_tail, fz_result* d_result, fz_result* h_result,
                               hipStream_t stream);

/* Launch knob of fz_compress_f32 (no counterpart in the reference); both settings give the
 * same bytes.  wave_tile = 0: one 256-thread workgroup per 2048-code tile; 1: one 64-lane wave
 * per tile, four tiles per workgroup, no workgroup barrier inside a tile. */
fz_status fz_plan_set_tile_layout(fz_plan plan, int wave_tile);

/* Kernel form of fz_decompress_f32 for 2D and 3D fields (no counterpart in the reference; the
 * reference has no inverse); every form gives the same floats.  FZ_INVERSE_AUTO (the default):
 * the form measured faster for the field's shape.  FZ_INVERSE_WINDOWS: windows of the field
 * decode into LDS and reconstruct straight to floats (one kernel after the tile offsets, none
 * before it for 2D fields of at most 4096 tiles and 4096 columns).  FZ_INVERSE_SPLIT: tile
 * offsets, decode to u16 codes in the workspace, reconstruct.  1D fields always decode and
 * reconstruct in one kernel. */
typedef enum fz_inverse_form { FZ_INVERSE_AUTO = 0, FZ_INVERSE_WINDOWS = 1, FZ_INVERSE_SPLIT = 2 } fz_inverse_form;
fz_status fz_plan_set_inverse_form(fz_plan plan, int form);

/* Per-kernel timing (the reference's cuda_timer_t around its Lorenzo launch, timer.hh:45-85,
 * extended to every kernel): when enabled, the compress/decompress calls on this plan record
 * an event before their first kernel and after each kernel; fz_plan_kernel_times waits for the
 * last one and returns the per-kernel milliseconds of the most recent call.  Off by default;
 * enabling it creates the events (host set-up, outside any hot path). */
fz_status fz_plan_profile(fz_plan plan, int enable);
fz_status fz_plan_kernel_times(const fz_plan plan, float* ms, int capacity, int* count);

/* ---- host-resident requests: pinned, asynchronous staging (fz.cu:143, 195-201) ---- */

/* Pinned host memory (hipHostMalloc): a buffer the stager can copy from or to asynchronously. */
fz_status fz_host_alloc(void** p, size_t bytes);
fz_status fz_host_free(void* p);

#define FZ_STAGER_MAX_DEPTH 8
typedef struct fz_stager_s* fz_stager;

/* A stager compresses (FIXED mode) and decompresses fields of one geometry held in host
 * memory, replacing the reference's blocking pageable copies around its kernels (fz.cu:143,
 * :195-201).  It keeps `depth` requests (1..8) in flight on three streams, so the copy in of
 * one request, the kernels of the one before and the copy out of the one before that
 * overlap.  It owns its device buffers, one workspace and, per slot, a pinned output buffer. */
fz_status fz_stager_create(fz_stager* stager, uint32_t x, uint32_t y, uint32_t z, uint64_t len, int depth);
fz_status fz_stager_destroy(fz_stager stager);
fz_status fz_stager_sizes(const fz_stager stager, fz_sizes* sizes);

/* Enqueue a compress of h_in (len floats; pinned memory, e.g. from fz_host_alloc, gives an
 * asynchronous copy, pageable memory a blocking one).  h_in must stay unchanged until the
 * request completes.  Returns at once unless all `depth` slots are in flight, in which case
 * it first waits for the oldest.  The results of a request stay readable until `depth` newer
 * requests have been made. */
fz_status fz_stager_compress(fz_stager stager, const float* h_in, double eb_rel, uint64_t* ticket);

/* Enqueue a decompress of a payload (and the tail codes, or NULL) into h_out (x*y*z floats;
 * pinned for an asynchronous copy). */
fz_status fz_stager_decompress(fz_stager stager, const uint8_t* h_payload, uint64_t payload_bytes,
                               const uint16_t* h_tail, float ebx2, float* h_out, uint64_t* ticket);

/* Wait for a request.  For a compress: *result is its record, *payload and *tail point into the
 * slot's pinned buffer (payload_bytes and tail_len from the record).  Returns the record's
 * status (FZ_SUCCESS for a decompress). */
fz_status fz_stager_wait(fz_stager stager, uint64_t ticket, fz_result* result, const uint8_t** payload,
                         const uint16_t** tail);

/* ---- quality metrics of a reconstruction (verify.hh:25-83, quality_viewer.hh:97-110) ---- */

/* cusz_stats (type.h:185-194), size_t as uint64_t. */
typedef struct fz_stats_range {
    double min, max, rng, std;
} fz_stats_range;
typedef struct fz_stats {
    fz_stats_range odata, xdata;
    struct {
        double PSNR, MSE, NRMSE, coeff;
    } reduced;
    struct {
        double   abs, rel, pwrrel;
        uint64_t idx;
    } max_err;
    struct {
        double lag_one, lag_two;
    } autocor;
    double   user_eb;
    uint64_t len;
} fz_stats;

/* verify_data (verify.hh:25) on the device: the metrics of reconstruction d_xdata against
 * original d_odata, written to *d_stats (device memory) with the reference's definitions:
 * population std, max_err.rel = max |x-o| / odata range, pwrrel = max over odata != 0 of the
 * float |x-o|/|o|, idx = first index of the largest |x-o|, PSNR = 20 log10(range) -
 * 10 log10(MSE), coeff = Pearson correlation.  xdata.max repeats the reference's scan of
 * odata (verify.hh:44): max(xdata[0], max odata).  autocor and user_eb are not written.
 * Call with d_temp == NULL to get *temp_bytes (at most 90 KiB), then again to run: one
 * streaming pass and one finishing workgroup.  Arrays need 4-byte alignment (16 for the
 * vector path), so the lag-1/lag-2 autocorrelation calls of quality_viewer.hh:105-107
 * (origin vs origin+1, len-1) work as they are. */
fz_status fz_verify_f32(void* d_temp, size_t* temp_bytes, fz_stats* d_stats, const float* d_xdata,
                        const float* d_odata, uint64_t len, hipStream_t stream);

/* fz_decompress_f32 followed on the same stream by fz_verify_f32 of d_out against d_odata
 * (x*y*z floats; the plan's workspace holds the metric partials). */
fz_status fz_decompress_verify_f32(const fz_plan plan, const uint8_t* d_payload, const uint16_t* d_tail,
                                   float ebx2, void* d_ws, size_t ws_bytes, float* d_out, const float* d_odata,
                                   fz_stats* d_stats, hipStream_t stream);

/* ---- per-stage entry points (the reference's kernel boundaries) ---- */

/* launch_construct_LorenzoI_var<float, uint16_t, float> (lorenzo_var.cuh:520-597).
 * eb_abs = eb_rel * range (fz.cu:149).  Writes codes for all len positions (0 beyond
 * x*y*z), so d_delta needs no pre-zeroing.  time_elapsed (ms) may be NULL; when given,
 * the call times itself with events and waits for the stream, as the reference does. */
fz_status fz_lorenzo_construct_f32(const float* d_data, uint16_t* d_delta, uint32_t x, uint32_t y, uint32_t z,
                                   uint64_t len, double eb_abs, int mode, float* time_elapsed, hipStream_t stream);

/* launch_reconstruct_LorenzoI_var (lorenzo_var.cuh:599-673) for sign-magnitude codes. */
fz_status fz_lorenzo_reconstruct_f32(const uint16_t* d_delta, float* d_xdata, uint32_t x, uint32_t y, uint32_t z,
                                     double eb_abs, float* time_elapsed, hipStream_t stream);

/* bitshuffleAndBitflag (fz.cu:61-111), grid floor(len/2048): tiles of 1024 u32 words.
 * Writes the shuffled words, byte flags [256 T] and bit flags [8 T] of the full tiles only,
 * like the reference kernel. */
fz_status fz_bitshuffle_bitflag(const uint32_t* d_in, uint32_t* d_out, uint64_t len, uint32_t* d_bitflags,
                                uint32_t* d_byteflags, hipStream_t stream);

/* Inverse of fz_bitshuffle_bitflag over the same tiles. */
fz_status fz_bitunshuffle(const uint32_t* d_in, uint32_t* d_out, uint64_t len, hipStream_t stream);

/* cub::DeviceScan::ExclusiveSum (fz.cu:188-192): call with d_temp == NULL to get
 * *temp_bytes, then again to run.  Single pass, decoupled look-back. */
fz_status fz_exclusive_sum_u32(void* d_temp, size_t* temp_bytes, const uint32_t* d_in, uint32_t* d_out,
                               uint64_t num_items, hipStream_t stream);

/* encodeDebug (fz.cu:49-59) with its intended semantics: block i (16 bytes) of d_in is
 * copied to d_out + 16*presum[i] iff presum[i+1] != presum[i]; presum has num_blocks+1
 * entries. */
fz_status fz_encode_blocks(const uint8_t* d_in, uint8_t* d_out, const uint32_t* d_presum, uint64_t num_blocks,
                           hipStream_t stream);

/* ---- batched form ----
 * No counterpart in the reference, which compresses one field per process (fz.cu:226-237).
 * It runs runSzs (fz.cu:113-209) over many fields of uneven shapes and dimensionality in
 * one launch sequence: each field's payload, tail codes and result equal those of
 * fz_compress_f32 / fz_compress_plain_f32 on that field alone, byte for byte.
 * Host arrays (d_in, eb_rel, d_payload, d_tail, ebx2, d_out) hold one entry per field, in
 * the order of the shapes given to fz_batch_plan_create; d_results is a device array of
 * nfields records in that order.  Per-field buffer sizes are those of fz_batch_field_sizes. */
typedef struct fz_batch_plan_s* fz_batch_plan;

typedef struct fz_field_shape {
    uint32_t x, y, z; /* x fastest, as fz_plan_create */
    uint32_t reserved;
    uint64_t len;     /* floats in the field, >= x*y*z; 0 means x*y*z */
} fz_field_shape;

typedef struct fz_batch_sizes {
    uint64_t workspace_bytes;
    uint64_t total_len;  /* floats over all fields */
    uint64_t units;      /* sum over fields of (full tiles + 1) */
    uint32_t nfields;
    uint32_t range_parts;
} fz_batch_sizes;

/* Launch knobs; any setting gives the same bytes. */
typedef struct fz_batch_tuning {
    uint32_t range_chunk;  /* floats per range partial: 16384, 65536 or 262144 */
    uint32_t xcd_order[3]; /* tile order of 1D, 2D, 3D fields: 1 XCD-sliced, 0 linear */
    uint32_t gather_waves; /* units per gather workgroup: 1, 2, 4, 8 or 16 */
    uint32_t groups;       /* field groups pipelined over three streams: 1, 2, 4, 8 or 16 */
    uint32_t wave_tile[3]; /* tile layout of 1D, 2D, 3D fields: 1 one wave per tile, 0 four waves per tile */
} fz_batch_tuning;

/* At most 4096 fields; every field as fz_plan_create accepts it. */
fz_status fz_batch_plan_create(fz_batch_plan* plan, const fz_field_shape* shapes, uint32_t nfields, int mode);
fz_status fz_batch_plan_destroy(fz_batch_plan plan);
fz_status fz_batch_plan_sizes(const fz_batch_plan plan, fz_batch_sizes* sizes);
fz_status fz_batch_field_sizes(const fz_batch_plan plan, uint32_t field, fz_sizes* sizes);
fz_status fz_batch_get_tuning(const fz_batch_plan plan, fz_batch_tuning* tuning);
fz_status fz_batch_set_tuning(fz_batch_plan plan, const fz_batch_tuning* tuning);

/* Writes the plan's field tables into a freshly allocated workspace (one host-to-device
 * copy; waits for the stream).  Call once per workspace, outside any graph capture. */
fz_status fz_batch_workspace_init(const fz_batch_plan plan, void* d_ws, size_t ws_bytes, hipStream_t stream);

/* Compress every field.  Launches kernels only (the per-field pointers travel as kernel
 * arguments), so the call can be captured into a hipGraph.  One workspace serves one
 * stream at a time.  With tuning.groups > 1 the fields are split into that many groups and
 * the call forks onto two streams owned by the plan (joined back into `stream` before it
 * returns): the value range of one group runs beside the tiles of the previous one and the
 * placement of a group's blocks beside the tiles of the next.  Calls on one plan are then
 * issued from one host thread at a time. */
fz_status fz_batch_compress_f32(const fz_batch_plan plan, const float* const* d_in, const double* eb_rel,
                                void* d_ws, size_t ws_bytes, uint8_t* const* d_payload, uint16_t* const* d_tail,
                                fz_result* d_results, hipStream_t stream);

/* Decompress every field (ebx2[i] = fz_result.ebx2 of field i; d_tail entries may be NULL). */
fz_status fz_batch_decompress_f32(const fz_batch_plan plan, const uint8_t* const* d_payload,
                                  const uint16_t* const* d_tail, const float* ebx2, void* d_ws, size_t ws_bytes,
                                  float* const* d_out, hipStream_t stream);

/* OUTLIER-mode batches (a batch plan created with FZ_MODE_OUTLIER; the two calls above refuse
 * such a plan): each field's payload, tail, side channel and record equal
 * fz_compress_outlier_f32 on that field alone (fz_sizes.side_max_bytes from
 * fz_batch_field_sizes bounds its channel; a smaller capacity may give that field
 * FZ_ERR_CAPACITY).  No counterpart in the reference; the side channel follows cuSZ's
 * SpcodecVec (spcodec_vec.hh:29).  The fields run one after another on `stream`, in one
 * workspace sized for the largest field, through the single-field outlier kernels, so a batch
 * of fields with large deltas, constant values or NaN / Inf compresses where the FIXED batch
 * reports them as failed.  Kernels only: the calls can be captured into a hipGraph.  For the
 * inverse, side_bytes[i] is the channel's size (16 + 8 (n_delta + n_raw)) or any capacity at
 * least that large; ebx2 and vconst come from each field's record (ebx2, vmin). */
fz_status fz_batch_compress_outlier_f32(const fz_batch_plan plan, const float* const* d_in, const double* eb_rel,
                                        void* d_ws, size_t ws_bytes, uint8_t* const* d_payload,
                                        uint16_t* const* d_tail, uint8_t* const* d_side,
                                        const uint64_t* side_capacity, fz_result* d_results, hipStream_t stream);
fz_status fz_batch_decompress_outlier_f32(const fz_batch_plan plan, const uint8_t* const* d_payload,
                                          const uint16_t* const* d_tail, const float* ebx2, const float* vconst,
                                          const uint8_t* const* d_side, const uint64_t* side_bytes, void* d_ws,
                                          size_t ws_bytes, float* const* d_out, hipStream_t stream);

/* Times the tuning candidates on these buffers (reps runs each, per-kernel events), keeps
 * the fastest in the plan and returns it in *chosen (may be NULL).  Waits for the stream. */
fz_status fz_batch_autotune(fz_batch_plan plan, const float* const* d_in, const double* eb_rel, void* d_ws,
                            size_t ws_bytes, uint8_t* const* d_payload, uint16_t* const* d_tail, fz_result* d_results,
                            int reps, fz_batch_tuning* chosen, hipStream_t stream);

/* Per-kernel timing of the most recent batched call, as fz_plan_profile / _kernel_times.
 * Compress slots: range, range final, tiles 1D, 2D, 3D, scan, gather.  Decompress slots:
 * flag scan, decode, reconstruct 1D, 2D, 3D.  A kernel with no work reports ~0.  While
 * profiling is on, compress runs as one group on the caller's stream. */
fz_status fz_batch_profile(fz_batch_plan plan, int enable);
fz_status fz_batch_kernel_times(const fz_batch_plan plan, float* ms, int capacity, int* count);

/* hipGraph of one batched call on fixed buffers, captured once; launching it replays the
 * call's kernels on those buffers. */
typedef struct fz_graph_s* fz_graph;
fz_status fz_batch_compress_graph(const fz_batch_plan plan, const float* const* d_in, const double* eb_rel,
                                  void* d_ws, size_t ws_bytes, uint8_t* const* d_payload, uint16_t* const* d_tail,
                                  fz_result* d_results, fz_graph* graph);
fz_status fz_batch_decompress_graph(const fz_batch_plan plan, const uint8_t* const* d_payload,
                                    const uint16_t* const* d_tail, const float* ebx2, void* d_ws, size_t ws_bytes,
                                    float* const* d_out, fz_graph* graph);
fz_status fz_graph_launch(fz_graph graph, hipStream_t stream);
fz_status fz_graph_destroy(fz_graph graph);

#ifdef __cplusplus
}
#endif
#endif /* FZ_CAPI_H */
