/*
 * fz_oracle.h -- CPU oracle for the FZ compression path (TEST INFRASTRUCTURE).
 *
 * This is the checker, never the product: only tests/, the smoke check and the
 * bench's CPU-baseline leg may load it.  It is a single-threaded C restatement of
 * the reference kernels, executed workgroup by workgroup and thread by thread, with
 * each __syncthreads phase as its own loop and each warp ballot/shuffle as a loop
 * over lanes.  Every function cites the reference lines it restates.
 *
 * The *_omp entry points are the same code with OpenMP work-sharing over
 * independent blocks/tiles (used as the reported CPU baseline); their output is
 * byte-identical to the single-threaded path.
 *
 * Two modes:
 *   FZO_MODE_FIXED   (0): the 1D kernel keeps the sign bit of every delta
 *                         (reference defect D1 fixed) and the n mod 2048 tail codes,
 *                         which the reference drops (D3), are returned as a trailer.
 *   FZO_MODE_LITERAL (1): reproduces the reference bytes exactly, including D1 and D3.
 * 2D and 3D payloads are identical in both modes.
 */
#ifndef FZ_ORACLE_H
#define FZ_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FZO_MODE_FIXED 0
#define FZO_MODE_LITERAL 1
#define FZO_MODE_OUTLIER 2 /* FIXED format + side channel (an extension of the format; no reference form) */

/* status codes: same values as fz_status in include/fz_capi.h */
#define FZO_SUCCESS 0
#define FZO_ERR_INVALID_ARG 1
#define FZO_ERR_CONSTANT_FIELD 4
#define FZO_ERR_NONFINITE 5
#define FZO_ERR_CORRUPT 6
#define FZO_ERR_CAPACITY 7

typedef struct {
    int32_t  status;
    int32_t  ndim;
    float    vmin, vmax, range;
    float    ebx2_r, ebx2;
    double   eb_abs;
    uint64_t len;           /* floats in the input (the reference's file length / 4) */
    uint64_t T;             /* bitshuffle tiles = floor(len / 2048) */
    uint64_t W;             /* bitflag words   = ceil(len / 256)    */
    uint64_t N;             /* byte flags      = 32 W               */
    uint64_t K;             /* nonzero 16-byte blocks               */
    uint64_t payload_bytes; /* 4 W + 16 K                           */
    uint64_t tail_len;      /* len - 2048 T                         */
    uint64_t overflow;      /* deltas with |d| >= 32768 (code aliases the sign bit) */
    uint64_t n_raw;         /* OUTLIER mode: values kept verbatim                    */
    uint64_t side_bytes;    /* OUTLIER mode: 16 + 8 (overflow + n_raw)               */
} fzo_info;

/* geometry helpers (fz.cu:157-161, 173, 177-178; lorenzo_var.cuh:537-544) */
int      fzo_ndim(uint32_t x, uint32_t y, uint32_t z);
uint64_t fzo_tiles(uint64_t len);
uint64_t fzo_flag_words(uint64_t len);
uint64_t fzo_payload_max_bytes(uint64_t len);

/* stages */
int  fzo_range(const float* data, uint64_t len, float* vmin, float* vmax);
void fzo_eb_params(float range, double eb_rel, double* eb_abs, float* ebx2_r, float* ebx2);
uint64_t fzo_dualquant(const float* data, uint16_t* delta, uint32_t x, uint32_t y, uint32_t z,
                       float ebx2_r, int mode, int nthreads);
void fzo_bitshuffle_flags(const uint32_t* in, uint32_t* out, uint64_t ntiles, uint32_t* bitflags,
                          uint32_t* byteflags, int nthreads);
void fzo_exclusive_sum(const uint32_t* in, uint32_t* out, uint64_t n, int nthreads);
void fzo_encode(const uint8_t* in, uint8_t* out, const uint32_t* presum, uint64_t nblocks, int nthreads);

/* inverse stages */
void fzo_unshuffle(const uint32_t* in, uint32_t* out, uint64_t ntiles, int nthreads);
void fzo_reconstruct(const uint16_t* delta, float* xdata, uint32_t x, uint32_t y, uint32_t z, float ebx2,
                     int nthreads);

/* OUTLIER mode side channel, one byte string:
 *   u64 n_delta, u64 n_raw,
 *   u32 idx_d[n_delta], i32 val_d[n_delta],   deltas with |d| >= 32768 (their code is 0)
 *   u32 idx_r[n_raw],   u32 val_r[n_raw]      input bits of values kept verbatim
 * both lists in increasing index order.  A value is kept verbatim when it is not finite or
 * when |fl(x * ebx2_r)| >= 2^24 (its quantised integer would not be exact in binary32); its
 * quantised integer is then 0.  Deltas are exact integers (block-local Lorenzo on the
 * quantised integers), so the inverse is exact wherever the error bound is defined. */
uint64_t fzo_outlier_max_bytes(uint64_t len);

/* whole pipeline.  payload capacity must be >= fzo_payload_max_bytes(len); tail has room
 * for 2047 codes.  The intermediate arrays of the reference (codes, shuffled words,
 * byte flags, prefix sums) are written to the optional dbg_* pointers when non-NULL:
 * dbg_codes u16[len], dbg_shuffled u32[ceil(len/2)], dbg_byteflags u32[N],
 * dbg_presum u32[N+1]. */
int fzo_compress(const float* data, uint64_t len, uint32_t x, uint32_t y, uint32_t z, double eb_rel, int mode,
                 int nthreads, uint8_t* payload, uint16_t* tail, fzo_info* info, uint16_t* dbg_codes,
                 uint32_t* dbg_shuffled, uint32_t* dbg_byteflags, uint32_t* dbg_presum);

int fzo_decompress(const uint8_t* payload, uint64_t payload_bytes, const uint16_t* tail, uint64_t tail_len,
                   uint64_t len, uint32_t x, uint32_t y, uint32_t z, float ebx2, int nthreads, float* xdata,
                   uint16_t* dbg_codes);

/* OUTLIER mode: fzo_compress with mode FZO_MODE_OUTLIER also writes the side channel
 * (capacity side_cap bytes; FZO_ERR_CAPACITY when it does not fit).  The inverse takes the
 * side channel and vconst (the compress record's vmin), used when ebx2 == 0 (constant
 * field: every finite value equals vmin). */
int fzo_compress_outlier(const float* data, uint64_t len, uint32_t x, uint32_t y, uint32_t z, double eb_rel,
                         uint8_t* payload, uint16_t* tail, uint8_t* side, uint64_t side_cap, fzo_info* info,
                         uint16_t* dbg_codes);
int fzo_decompress_outlier(const uint8_t* payload, uint64_t payload_bytes, const uint16_t* tail, uint64_t tail_len,
                           const uint8_t* side, uint64_t side_bytes, uint64_t len, uint32_t x, uint32_t y,
                           uint32_t z, float ebx2, float vconst, float* xdata);

/* Quality metrics of a reconstruction (cusz_stats, type.h:185-194; size_t -> uint64_t). */
typedef struct {
    struct { double min, max, rng, std; } odata, xdata;
    struct { double PSNR, MSE, NRMSE, coeff; } reduced;
    struct { double abs, rel, pwrrel; uint64_t idx; } max_err;
    struct { double lag_one, lag_two; } autocor;
    double   user_eb;
    uint64_t len;
} fzo_stats;

/* verify_data (verify.hh:25-83), sequential, in the reference's order and types: double
 * sums and two passes; float |x-o| and float |x-o|/|o|; max_err.idx is the first index of
 * the largest error.  xdata.max repeats the reference's scan over odata (verify.hh:44).
 * autocor and user_eb are not written. */
void fzo_verify(const float* xdata, const float* odata, uint64_t len, fzo_stats* s);

#ifdef __cplusplus
}
#endif
#endif
