/*
 * fz_oracle.c -- CPU oracle for the FZ compression path (TEST INFRASTRUCTURE ONLY).
 *
 * Restates, in plain C and single-threaded, the reference kernels
 *   lorenzo_var.cuh:47-295   (c_lorenzo_1d1l / 2d1l_16x16data_mapto16x2 / 3d1l_32x8x8data_mapto32x1x8)
 *   lorenzo_var.cuh:520-597  (launch_construct_LorenzoI_var: ndim, eb -> ebx2 / ebx2_r)
 *   fz.cu:61-111             (bitshuffleAndBitflag)
 *   fz.cu:187-201            (cub::DeviceScan::ExclusiveSum + host total)
 *   fz.cu:49-59              (encodeDebug)
 *   fz.cu:113-209            (runSzs orchestration, sizes)
 *   verify.hh:25-83          (verify_data: quality metrics of a reconstruction)
 * Each CUDA workgroup is a loop iteration, each thread a nested loop, each
 * __syncthreads a phase boundary, each __ballot_sync / __shfl_*_sync a loop over lanes.
 * The inverse (absent from the reference binary) follows the block structure of
 * lorenzo_var.cuh:297-518 adapted to sign-magnitude codes.
 *
 * Build with -ffp-contract=off and without fast-math: every float operation below is a
 * single IEEE binary32 operation in the reference's order.
 */
#include "fz_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define OMP_FOR _Pragma("omp parallel for schedule(static) num_threads(nthreads) if(nthreads > 1)")

/* ------------------------------------------------------------------ geometry */

/* lorenzo_var.cuh:537-544 */
int fzo_ndim(uint32_t x, uint32_t y, uint32_t z)
{
    (void)x;
    if (z == 1 && y == 1) return 1;
    if (z == 1 && y != 1) return 2;
    return 3;
}

/* fz.cu:173  grid(floor(len / 2048)) */
uint64_t fzo_tiles(uint64_t len) { return len / 2048; }

/* fz.cu:177-178  dataChunkSize = ceil(2 len / (16 * 32)) */
uint64_t fzo_flag_words(uint64_t len)
{
    uint64_t nb = 2 * len;
    return nb % 512 == 0 ? nb / 512 : nb / 512 + 1;
}

uint64_t fzo_payload_max_bytes(uint64_t len) { return 4 * fzo_flag_words(len) + 4096 * fzo_tiles(len); }

/* static_cast<uint16_t>(float) on a fabs() result (lorenzo_var.cuh:53, 62, 169, 289).
 * In-range values truncate; values >= 65536 saturate to 65535 and NaN gives 0, which is
 * what the GPU's float->unsigned conversion does (parity unpinned against nvcc for those). */
static inline uint16_t f2u16(float f)
{
    if (!(f >= 0.0f)) return 0;
    if (f >= 65535.0f) return 65535;
    return (uint16_t)f;
}

static inline int is_overflow(float d) { return !(fabsf(d) < 32768.0f); }

/* ------------------------------------------------------------------ range + eb */

/* fz.cu:137  range = *max_element - *min_element (first extreme, '<' comparisons). */
int fzo_range(const float* data, uint64_t len, float* vmin, float* vmax)
{
    uint64_t lo = 0, hi = 0;
    int finite = 1;
    for (uint64_t i = 0; i < len; i++) {
        if (!isfinite(data[i])) finite = 0;
        if (data[i] < data[lo]) lo = i;  /* std::min_element */
        if (data[hi] < data[i]) hi = i;  /* std::max_element */
    }
    *vmin = data[lo];
    *vmax = data[hi];
    return finite ? FZO_SUCCESS : FZO_ERR_NONFINITE;
}

/* fz.cu:149 eb * range (double);  lorenzo_var.cuh:562-563 ebx2 = eb*2, ebx2_r = 1/ebx2 in
 * double, narrowed to float at the kernel's FP parameter. */
void fzo_eb_params(float range, double eb_rel, double* eb_abs, float* ebx2_r, float* ebx2)
{
    double eb = eb_rel * (double)range;
    double e2 = eb * 2;
    double e2r = 1 / e2;
    *eb_abs = eb;
    *ebx2_r = (float)e2r;
    *ebx2 = (float)e2;
}

/* ------------------------------------------------------------------ dual-quant */

/* c_lorenzo_1d1l<float,u16,float,256,4> (lorenzo_var.cuh:182-214) with load1d :68-89,
 * pred1d :47-66, write1d :91-108.  Grid ceil(dimx/256), 64 threads x SEQ 4. */
static uint64_t dq_1d(const float* data, uint16_t* delta, uint32_t dimx, float ebx2_r, int mode, int nthreads)
{
    enum { BLOCK = 256, SEQ = 4, NTHREAD = BLOCK / SEQ };
    const int64_t nblk = ((int64_t)dimx - 1) / BLOCK + 1;
    uint64_t ovf = 0;
    OMP_FOR
    for (int64_t bix = 0; bix < nblk; bix++) {
        float    shmem_data[BLOCK];
        uint16_t shmem_delta[BLOCK];
        float    thread_scope[NTHREAD][SEQ];
        float    from_last_stripe[NTHREAD];
        uint64_t id_base = (uint64_t)bix * BLOCK;
        uint64_t local_ovf = 0;
        memset(shmem_data, 0, sizeof(shmem_data)); /* unset slots are never read for in-range ids */
        memset(shmem_delta, 0, sizeof(shmem_delta));
        /* load1d phase 1 */
        for (int tix = 0; tix < NTHREAD; tix++)
            for (int i = 0; i < SEQ; i++) {
                uint64_t id = id_base + tix + i * NTHREAD;
                if (id < dimx) shmem_data[tix + i * NTHREAD] = roundf(data[id] * ebx2_r);
            }
        /* load1d phase 2 */
        for (int tix = 0; tix < NTHREAD; tix++) {
            for (int i = 0; i < SEQ; i++) thread_scope[tix][i] = shmem_data[tix * SEQ + i];
            from_last_stripe[tix] = tix > 0 ? shmem_data[tix * SEQ - 1] : 0.0f;
        }
        /* pred1d<FIRST_POINT=true> */
        for (int tix = 0; tix < NTHREAD; tix++) {
            float    d = thread_scope[tix][0] - from_last_stripe[tix];
            uint16_t tmp = f2u16(fabsf(d));
            tmp |= (uint16_t)((d < 0) << 15);
            shmem_delta[0 + tix * SEQ] = tmp;
            if (id_base + tix * SEQ < dimx) local_ovf += is_overflow(d);
        }
        /* pred1d<FIRST_POINT=false>: the reference keeps the sign only in the discarded
         * signum[] (D1); FIXED mode ORs it into the code like the first point. */
        for (int tix = 0; tix < NTHREAD; tix++)
            for (int i = 1; i < SEQ; i++) {
                float    d = thread_scope[tix][i] - thread_scope[tix][i - 1];
                uint16_t tmp = f2u16(fabsf(d));
                if (mode == FZO_MODE_FIXED) tmp |= (uint16_t)((d < 0) << 15);
                shmem_delta[i + tix * SEQ] = tmp;
                if (id_base + tix * SEQ + i < dimx) local_ovf += is_overflow(d);
            }
        /* write1d */
        for (int tix = 0; tix < NTHREAD; tix++)
            for (int i = 0; i < SEQ; i++) {
                uint64_t id = id_base + tix + i * NTHREAD;
                if (id < dimx) delta[id] = shmem_delta[tix + i * NTHREAD];
            }
#pragma omp atomic
        ovf += local_ovf;
    }
    return ovf;
}

/* c_lorenzo_2d1l_16x16data_mapto16x2 (lorenzo_var.cuh:216-238) with load2d_prequant
 * :110-129, pred2d :131-145, postquant_write2d :147-174.  Grid (ceil(X/16), ceil(Y/16)),
 * block (16,2): lane = TIY*16 + TIX, 8 rows per lane. */
static uint64_t dq_2d(const float* data, uint16_t* delta, uint32_t dimx, uint32_t dimy, float ebx2_r,
                      int nthreads)
{
    enum { BLOCK = 16, YSEQ = 8, LANES = 32 };
    const uint32_t stridey = dimx;
    const int64_t  gx = ((int64_t)dimx - 1) / BLOCK + 1, gy = ((int64_t)dimy - 1) / BLOCK + 1;
    uint64_t       ovf = 0;
    OMP_FOR
    for (int64_t b = 0; b < gx * gy; b++) {
        const uint64_t bix = (uint64_t)(b % gx), biy = (uint64_t)(b / gx);
        float          center[LANES][YSEQ + 1];
        float          xfer[LANES];
        uint64_t       local_ovf = 0;
        /* load2d_prequant */
        for (int lane = 0; lane < LANES; lane++) {
            const uint64_t tix = lane % 16, tiy = lane / 16;
            const uint64_t gix = bix * BLOCK + tix, giy_base = biy * BLOCK + tiy * YSEQ;
            for (int i = 0; i < YSEQ + 1; i++) center[lane][i] = 0.0f;
            for (int i = 0; i < YSEQ; i++)
                if (gix < dimx && giy_base + i < dimy)
                    center[lane][i + 1] = roundf(data[(giy_base + i) * stridey + gix] * ebx2_r);
        }
        /* __shfl_up_sync(center[YSEQ], 16); lanes with TIY == 1 keep it as their north row */
        for (int lane = 0; lane < LANES; lane++) xfer[lane] = lane >= 16 ? center[lane - 16][YSEQ] : center[lane][YSEQ];
        for (int lane = 16; lane < LANES; lane++) center[lane][0] = xfer[lane];
        /* pred2d: top-down over the 8 rows; the west lane's value is taken after its own
         * north subtraction (shfl_up by 1, width 16) */
        for (int i = YSEQ; i > 0; i--) {
            for (int lane = 0; lane < LANES; lane++) center[lane][i] -= center[lane][i - 1];
            for (int lane = 0; lane < LANES; lane++) xfer[lane] = (lane % 16) >= 1 ? center[lane - 1][i] : center[lane][i];
            for (int lane = 0; lane < LANES; lane++)
                if (lane % 16 > 0) center[lane][i] -= xfer[lane];
        }
        /* postquant_write2d */
        for (int lane = 0; lane < LANES; lane++) {
            const uint64_t tix = lane % 16, tiy = lane / 16;
            const uint64_t gix = bix * BLOCK + tix, giy_base = biy * BLOCK + tiy * YSEQ;
            for (int i = 1; i < YSEQ + 1; i++) {
                if (gix < dimx && giy_base + i - 1 < dimy) {
                    const float d = center[lane][i];
                    uint16_t    tmp = f2u16(fabsf(d));
                    tmp |= (uint16_t)((d < 0) << 15);
                    delta[(giy_base + i - 1) * stridey + gix] = tmp;
                    local_ovf += is_overflow(d);
                }
            }
        }
#pragma omp atomic
        ovf += local_ovf;
    }
    return ovf;
}

/* c_lorenzo_3d1l_32x8x8data_mapto32x1x8 (lorenzo_var.cuh:240-295).  Grid
 * (ceil(X/32), ceil(Y/8), ceil(Z/8)), block (32,1,8); 8x8x8 Lorenzo blocks (x = TIX % 8). */
static uint64_t dq_3d(const float* data, uint16_t* delta, uint32_t dimx, uint32_t dimy, uint32_t dimz,
                      float ebx2_r, int nthreads)
{
    enum { BLOCK = 8 };
    const uint64_t stridey = dimx, stridez = (uint64_t)dimx * dimy;
    const int64_t  gx = ((int64_t)dimx - 1) / 32 + 1, gy = ((int64_t)dimy - 1) / BLOCK + 1,
                  gz = ((int64_t)dimz - 1) / BLOCK + 1;
    uint64_t ovf = 0;
    OMP_FOR
    for (int64_t b = 0; b < gx * gy * gz; b++) {
        const uint64_t bix = (uint64_t)(b % gx), biy = (uint64_t)((b / gx) % gy), biz = (uint64_t)(b / (gx * gy));
        float          shmem[8][8][32];
        uint64_t       local_ovf = 0;
        memset(shmem, 0, sizeof(shmem)); /* unset cells are never read for in-range outputs */
        for (int z = 0; z < 8; z++)
            for (int tix = 0; tix < 32; tix++) {
                const uint64_t gix = bix * (BLOCK * 4) + tix, giy_base = biy * BLOCK, giz = biz * BLOCK + z;
                const uint64_t base_id = gix + giy_base * stridey + giz * stridez;
                if (gix < dimx && giz < dimz)
                    for (int y = 0; y < BLOCK; y++)
                        if (giy_base + y < dimy) shmem[z][y][tix] = roundf(data[base_id + y * stridey] * ebx2_r);
            }
        /* __syncthreads */
        for (int z = 0; z < 8; z++)
            for (int tix = 0; tix < 32; tix++) {
                const uint64_t gix = bix * (BLOCK * 4) + tix, giy_base = biy * BLOCK, giz = biz * BLOCK + z;
                const uint64_t base_id = gix + giy_base * stridey + giz * stridez;
                const int      x = tix % 8;
                for (int y = 0; y < BLOCK; y++) {
                    /* ((((((NW3 - A) - B) - C) + D) + E) + F), each term 0 across the block edge */
                    float p = (z > 0 && y > 0 && x > 0) ? shmem[z - 1][y - 1][tix - 1] : 0.0f;
                    p = p - ((y > 0 && x > 0) ? shmem[z][y - 1][tix - 1] : 0.0f);
                    p = p - ((z > 0 && x > 0) ? shmem[z - 1][y][tix - 1] : 0.0f);
                    p = p - ((z > 0 && y > 0) ? shmem[z - 1][y - 1][tix] : 0.0f);
                    p = p + (x > 0 ? shmem[z][y][tix - 1] : 0.0f);
                    p = p + (y > 0 ? shmem[z][y - 1][tix] : 0.0f);
                    p = p + (z > 0 ? shmem[z - 1][y][tix] : 0.0f);
                    const float d = shmem[z][y][tix] - p;
                    if (gix < dimx && giy_base + y < dimy && giz < dimz) {
                        uint16_t tmp = f2u16(fabsf(d));
                        tmp |= (uint16_t)((d < 0) << 15);
                        delta[base_id + y * stridey] = tmp;
                        local_ovf += is_overflow(d);
                    }
                }
            }
#pragma omp atomic
        ovf += local_ovf;
    }
    return ovf;
}

/* launch_construct_LorenzoI_var dispatch (lorenzo_var.cuh:571-588).  delta must be
 * pre-zeroed (fz.cu:145).  Returns the number of deltas with |d| >= 32768. */
uint64_t fzo_dualquant(const float* data, uint16_t* delta, uint32_t x, uint32_t y, uint32_t z, float ebx2_r,
                       int mode, int nthreads)
{
    switch (fzo_ndim(x, y, z)) {
        case 1: return dq_1d(data, delta, x, ebx2_r, mode, nthreads);
        case 2: return dq_2d(data, delta, x, y, ebx2_r, nthreads);
        default: return dq_3d(data, delta, x, y, z, ebx2_r, nthreads);
    }
}

/* ------------------------------------------------------------------ bitshuffle + flags */

/* bitshuffleAndBitflag (fz.cu:61-111), grid = ntiles, block (32,32): warp k = threadIdx.y,
 * lane j = threadIdx.x.  The byte-flag OR (fz.cu:84-91) reuses smem after the transposed
 * words were read at fz.cu:82; here the reads complete first (the intended order, D6). */
void fzo_bitshuffle_flags(const uint32_t* in, uint32_t* out, uint64_t ntiles, uint32_t* bitflags,
                          uint32_t* byteflags, int nthreads)
{
    OMP_FOR
    for (int64_t t = 0; t < (int64_t)ntiles; t++) {
        uint32_t smem[32][33];
        uint32_t bitflagArr[8];
        uint32_t byteFlagArray[256];
        const uint64_t base = (uint64_t)t * 1024;
        /* 32 ballots per warp (fz.cu:73-79) */
        for (int k = 0; k < 32; k++)
            for (int i = 0; i < 32; i++) {
                uint32_t ballot = 0;
                for (int j = 0; j < 32; j++)
                    if (in[base + j + 32 * k] & (1u << i)) ballot |= 1u << j;
                smem[k][i] = ballot;
            }
        /* __syncthreads; fz.cu:82 */
        for (int k = 0; k < 32; k++)
            for (int j = 0; j < 32; j++) out[base + j + 32 * k] = smem[j][k];
        /* fz.cu:84-91 */
        for (int k = 0; k < 32; k++)
            for (int j = 0; j < 8; j++) {
                for (int i = 1; i < 4; i++) smem[j * 4][k] |= smem[j * 4 + i][k];
                byteFlagArray[k * 8 + j] = smem[j * 4][k] > 0;
            }
        /* __syncthreads; fz.cu:94-99 */
        for (int y = 0; y < 8; y++) {
            uint32_t ballot = 0;
            for (int x = 0; x < 32; x++)
                if (byteFlagArray[y * 32 + x]) ballot |= 1u << x;
            bitflagArr[y] = ballot;
        }
        /* __syncthreads; fz.cu:102-110 */
        for (int y = 0; y < 8; y++)
            for (int x = 0; x < 32; x++) byteflags[(uint64_t)t * 256 + y * 32 + x] = byteFlagArray[y * 32 + x];
        for (int x = 0; x < 8; x++) bitflags[(uint64_t)t * 8 + x] = bitflagArr[x];
    }
}

/* ------------------------------------------------------------------ scan */

/* cub::DeviceScan::ExclusiveSum (fz.cu:190-192): out[i] = sum in[0..i), i < n.
 * u32 arithmetic (wraps like the device sum). */
void fzo_exclusive_sum(const uint32_t* in, uint32_t* out, uint64_t n, int nthreads)
{
    if (nthreads <= 1 || n < 65536) {
        uint32_t acc = 0;
        for (uint64_t i = 0; i < n; i++) {
            out[i] = acc;
            acc += in[i];
        }
        return;
    }
    /* two-level form for the OpenMP baseline: chunk totals, serial scan of totals, chunk scans */
    const int64_t chunk = 16384, nchunk = (int64_t)((n + chunk - 1) / chunk);
    uint32_t*     tot = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(nchunk + 1));
    OMP_FOR
    for (int64_t c = 0; c < nchunk; c++) {
        uint64_t lo = (uint64_t)c * chunk, hi = lo + chunk < n ? lo + chunk : n;
        uint32_t s = 0;
        for (uint64_t i = lo; i < hi; i++) s += in[i];
        tot[c] = s;
    }
    uint32_t acc = 0;
    for (int64_t c = 0; c < nchunk; c++) {
        uint32_t v = tot[c];
        tot[c] = acc;
        acc += v;
    }
    OMP_FOR
    for (int64_t c = 0; c < nchunk; c++) {
        uint64_t lo = (uint64_t)c * chunk, hi = lo + chunk < n ? lo + chunk : n;
        uint32_t s = tot[c];
        for (uint64_t i = lo; i < hi; i++) {
            out[i] = s;
            s += in[i];
        }
    }
    free(tot);
}

/* ------------------------------------------------------------------ encode */

/* encodeDebug (fz.cu:49-59), grid nblocks/32, block 32, blockSize 16.  fz.cu:53 lets every
 * lane store sumArr[0] (D2); the intended value, preSum[32*blockIdx.x], is used here, so
 * block i is copied iff presum[i+1] != presum[i], to byte offset 16*presum[i]. */
void fzo_encode(const uint8_t* in, uint8_t* out, const uint32_t* presum, uint64_t nblocks, int nthreads)
{
    const int64_t nwg = (int64_t)(nblocks / 32);
    OMP_FOR
    for (int64_t b = 0; b < nwg; b++) {
        uint32_t sumArr[33];
        sumArr[0] = presum[(uint64_t)b * 32];
        for (int lane = 0; lane < 32; lane++) sumArr[lane + 1] = presum[(uint64_t)b * 32 + lane + 1];
        for (int lane = 0; lane < 32; lane++) {
            const uint64_t tid = (uint64_t)b * 32 + lane;
            if (sumArr[lane + 1] != sumArr[lane])
                for (int i = 0; i < 16; i++) out[(uint64_t)sumArr[lane] * 16 + i] = in[tid * 16 + i];
        }
    }
}

/* ------------------------------------------------------------------ compress */

/* The stages after the dual-quant (fz.cu:154-205), shared by every mode: bitshuffle + flags
 * of the full tiles, exclusive scan of the byte flags, payload = B || compacted blocks. */
static void encode_codes(const uint16_t* codes, uint64_t len, uint8_t* payload, fzo_info* info, int nthreads,
                         uint32_t* dbg_shuffled, uint32_t* dbg_byteflags, uint32_t* dbg_presum)
{
    const uint64_t T = info->T, W = info->W, N = info->N;
    /* fz.cu:154-155: shuffled buffer starts as a copy of the codes (tail stays raw) */
    uint8_t* shuffled = (uint8_t*)malloc((size_t)(2 * len + 4));
    memcpy(shuffled, codes, (size_t)(2 * len));
    uint32_t* bitflags = (uint32_t*)calloc((size_t)W, sizeof(uint32_t));
    uint32_t* byteflags = (uint32_t*)calloc((size_t)N, sizeof(uint32_t));
    fzo_bitshuffle_flags((const uint32_t*)codes, (uint32_t*)shuffled, T, bitflags, byteflags, nthreads);

    /* fz.cu:180-201 */
    uint32_t* presum = (uint32_t*)calloc((size_t)N + 1, sizeof(uint32_t));
    fzo_exclusive_sum(byteflags, presum, N, nthreads);
    presum[N] = presum[N - 1] + byteflags[N - 1];
    info->K = presum[N];
    info->payload_bytes = 4 * W + 16 * info->K;

    /* payload = B || compacted blocks */
    for (uint64_t i = 0; i < W; i++) {
        const uint32_t v = bitflags[i];
        payload[4 * i + 0] = (uint8_t)(v);
        payload[4 * i + 1] = (uint8_t)(v >> 8);
        payload[4 * i + 2] = (uint8_t)(v >> 16);
        payload[4 * i + 3] = (uint8_t)(v >> 24);
    }
    fzo_encode(shuffled, payload + 4 * W, presum, N, nthreads);

    if (dbg_shuffled) memcpy(dbg_shuffled, shuffled, (size_t)(2 * len));
    if (dbg_byteflags) memcpy(dbg_byteflags, byteflags, (size_t)(4 * N));
    if (dbg_presum) memcpy(dbg_presum, presum, (size_t)(4 * (N + 1)));
    free(shuffled);
    free(bitflags);
    free(byteflags);
    free(presum);
}

/* runSzs (fz.cu:113-209) without file I/O.  The payload is the reference's output bytes:
 * LE u32 bitflags B[0..W) followed by d_out[0..16K) (size formula fz.cu:208). */
int fzo_compress(const float* data, uint64_t len, uint32_t x, uint32_t y, uint32_t z, double eb_rel, int mode,
                 int nthreads, uint8_t* payload, uint16_t* tail, fzo_info* info, uint16_t* dbg_codes,
                 uint32_t* dbg_shuffled, uint32_t* dbg_byteflags, uint32_t* dbg_presum)
{
    memset(info, 0, sizeof(*info));
    info->len = len;
    if (len == 0 || x == 0 || y == 0 || z == 0 || (uint64_t)x * y * z > len || !(eb_rel > 0) || !isfinite(eb_rel) ||
        (mode != FZO_MODE_FIXED && mode != FZO_MODE_LITERAL))
        return info->status = FZO_ERR_INVALID_ARG;
    info->ndim = fzo_ndim(x, y, z);
    const uint64_t T = fzo_tiles(len), W = fzo_flag_words(len), N = 32 * W;
    info->T = T;
    info->W = W;
    info->N = N;
    info->tail_len = len - 2048 * T;

    float vmin, vmax;
    int   rc = fzo_range(data, len, &vmin, &vmax);
    info->vmin = vmin;
    info->vmax = vmax;
    if (rc) return info->status = rc;
    const float range = vmax - vmin;
    info->range = range;
    if (range == 0.0f) return info->status = FZO_ERR_CONSTANT_FIELD; /* D5: 1/0 in the reference */
    fzo_eb_params(range, eb_rel, &info->eb_abs, &info->ebx2_r, &info->ebx2);

    /* fz.cu:140, 145: codes zero-initialised, then the Lorenzo kernel */
    uint16_t* codes = (uint16_t*)calloc((size_t)len + 2, sizeof(uint16_t));
    info->overflow = fzo_dualquant(data, codes, x, y, z, info->ebx2_r, mode, nthreads);

    encode_codes(codes, len, payload, info, nthreads, dbg_shuffled, dbg_byteflags, dbg_presum);
    if (mode == FZO_MODE_FIXED && tail)
        for (uint64_t i = 0; i < info->tail_len; i++) tail[i] = codes[2048 * T + i];
    if (dbg_codes) memcpy(dbg_codes, codes, (size_t)(2 * len));
    free(codes);
    return info->status = FZO_SUCCESS;
}

/* ------------------------------------------------------------------ outlier mode */

uint64_t fzo_outlier_max_bytes(uint64_t len) { return 16 + 8 * len; }

/* quantised integer of one value; raw = kept verbatim (not finite, or |fl(v r)| >= 2^24) */
static inline int32_t quant_i(float v, float r, int* raw)
{
    const float t = v * r;
    if (!isfinite(v) || !(fabsf(t) < 16777216.0f)) {
        *raw = 1;
        return 0;
    }
    *raw = 0;
    return (int32_t)roundf(t);
}

static void put_u32(uint8_t* p, uint32_t v)
{
    p[0] = (uint8_t)v;
    p[1] = (uint8_t)(v >> 8);
    p[2] = (uint8_t)(v >> 16);
    p[3] = (uint8_t)(v >> 24);
}

static uint32_t get_u32(const uint8_t* p)
{
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

static void put_u64(uint8_t* p, uint64_t v)
{
    put_u32(p, (uint32_t)v);
    put_u32(p + 4, (uint32_t)(v >> 32));
}

static uint64_t get_u64(const uint8_t* p) { return (uint64_t)get_u32(p) | ((uint64_t)get_u32(p + 4) << 32); }

/* Range of the finite values (0, 0 when there are none), eb parameters, then the
 * block-local Lorenzo deltas of the quantised integers in exact integer arithmetic, in the
 * reference's block shapes (256 / 16x16 / 8x8x8, neighbours outside the block are 0). */
int fzo_compress_outlier(const float* data, uint64_t len, uint32_t x, uint32_t y, uint32_t z, double eb_rel,
                         uint8_t* payload, uint16_t* tail, uint8_t* side, uint64_t side_cap, fzo_info* info,
                         uint16_t* dbg_codes)
{
    memset(info, 0, sizeof(*info));
    info->len = len;
    if (len == 0 || x == 0 || y == 0 || z == 0 || (uint64_t)x * y * z > len || !(eb_rel > 0) || !isfinite(eb_rel))
        return info->status = FZO_ERR_INVALID_ARG;
    const int      nd = fzo_ndim(x, y, z);
    const uint64_t xyz = (uint64_t)x * y * z;
    info->ndim = nd;
    info->T = fzo_tiles(len);
    info->W = fzo_flag_words(len);
    info->N = 32 * info->W;
    info->tail_len = len - 2048 * info->T;

    float lo = INFINITY, hi = -INFINITY;
    for (uint64_t i = 0; i < len; i++)
        if (isfinite(data[i])) {
            lo = data[i] < lo ? data[i] : lo;
            hi = data[i] > hi ? data[i] : hi;
        }
    if (lo > hi) lo = hi = 0.0f;
    info->vmin = lo;
    info->vmax = hi;
    const float range = hi - lo;
    info->range = range;
    if (!isfinite(range)) return info->status = FZO_ERR_NONFINITE; /* the range itself overflows */
    if (range == 0.0f) {
        info->eb_abs = 0.0;
        info->ebx2_r = 0.0f;
        info->ebx2 = 0.0f; /* constant: every quantised integer is 0, the inverse fills vmin */
    } else {
        fzo_eb_params(range, eb_rel, &info->eb_abs, &info->ebx2_r, &info->ebx2);
    }

    int32_t*  q = (int32_t*)calloc((size_t)xyz, sizeof(int32_t));
    uint8_t*  raw = (uint8_t*)calloc((size_t)xyz, 1);
    uint16_t* codes = (uint16_t*)calloc((size_t)len + 2, sizeof(uint16_t));
    for (uint64_t i = 0; i < xyz; i++) {
        int r;
        q[i] = quant_i(data[i], info->ebx2_r, &r);
        raw[i] = (uint8_t)r;
        info->n_raw += (uint64_t)r;
    }
    uint32_t* idx_d = (uint32_t*)malloc((size_t)(4 * xyz + 4));
    int32_t*  val_d = (int32_t*)malloc((size_t)(4 * xyz + 4));
    uint64_t  nd_out = 0;
    const uint64_t sy = x, sz = (uint64_t)x * y;
    for (uint64_t p = 0; p < xyz; p++) {
        int64_t d;
        if (nd == 1) {
            d = (int64_t)q[p] - ((p % 256) ? q[p - 1] : 0);
        } else if (nd == 2) {
            const uint64_t ix = p % x, iy = p / x;
            const int bx = (ix % 16) != 0, by = (iy % 16) != 0;
            d = (int64_t)q[p] - (bx ? q[p - 1] : 0) - (by ? q[p - sy] : 0) + ((bx && by) ? q[p - sy - 1] : 0);
        } else {
            const uint64_t ix = p % x, iy = (p / x) % y, iz = p / sz;
            const int bx = (ix % 8) != 0, by = (iy % 8) != 0, bz = (iz % 8) != 0;
            d = (int64_t)q[p] - (bx ? q[p - 1] : 0) - (by ? q[p - sy] : 0) - (bz ? q[p - sz] : 0) +
                ((bx && by) ? q[p - sy - 1] : 0) + ((bx && bz) ? q[p - sz - 1] : 0) +
                ((by && bz) ? q[p - sz - sy] : 0) - ((bx && by && bz) ? q[p - sz - sy - 1] : 0);
        }
        if (d >= 32768 || d <= -32768) {
            idx_d[nd_out] = (uint32_t)p;
            val_d[nd_out] = (int32_t)d;
            nd_out++;
            codes[p] = 0;
        } else {
            codes[p] = (uint16_t)((d < 0 ? -d : d) | (d < 0 ? 0x8000 : 0));
        }
    }
    info->overflow = nd_out;
    info->side_bytes = 16 + 8 * (nd_out + info->n_raw);

    encode_codes(codes, len, payload, info, 1, NULL, NULL, NULL);
    if (tail)
        for (uint64_t i = 0; i < info->tail_len; i++) tail[i] = codes[2048 * info->T + i];
    if (dbg_codes) memcpy(dbg_codes, codes, (size_t)(2 * len));

    int rc = FZO_SUCCESS;
    if (info->side_bytes > side_cap) {
        rc = FZO_ERR_CAPACITY;
    } else if (side) {
        const uint64_t nr = info->n_raw;
        uint8_t*       o = side;
        put_u64(o, nd_out);
        put_u64(o + 8, nr);
        for (uint64_t k = 0; k < nd_out; k++) {
            put_u32(o + 16 + 4 * k, idx_d[k]);
            put_u32(o + 16 + 4 * nd_out + 4 * k, (uint32_t)val_d[k]);
        }
        uint64_t k = 0;
        for (uint64_t p = 0; p < xyz; p++)
            if (raw[p]) {
                uint32_t bits;
                memcpy(&bits, &data[p], 4);
                put_u32(o + 16 + 8 * nd_out + 4 * k, (uint32_t)p);
                put_u32(o + 16 + 8 * nd_out + 4 * nr + 4 * k, bits);
                k++;
            }
    }
    free(q);
    free(raw);
    free(codes);
    free(idx_d);
    free(val_d);
    return info->status = rc;
}

/* ------------------------------------------------------------------ inverse */

/* Inverse of fz.cu:73-82: bit b of w[1024t+32k+j] = bit j of o[1024t+32b+k]. */
void fzo_unshuffle(const uint32_t* in, uint32_t* out, uint64_t ntiles, int nthreads)
{
    OMP_FOR
    for (int64_t t = 0; t < (int64_t)ntiles; t++) {
        const uint64_t base = (uint64_t)t * 1024;
        for (int k = 0; k < 32; k++)
            for (int j = 0; j < 32; j++) {
                uint32_t w = 0;
                for (int b = 0; b < 32; b++) w |= ((in[base + 32 * b + k] >> j) & 1u) << b;
                out[base + 32 * k + j] = w;
            }
    }
}

static inline int32_t decode(uint16_t c)
{
    const int32_t m = c & 0x7fff;
    return (c & 0x8000) ? -m : m;
}

/* Block-local inclusive prefix sums that undo the Lorenzo delta, in the block shapes of
 * x_lorenzo_1d1l (256), x_lorenzo_2d1l (16x16: y then x, lorenzo_var.cuh:408-428) and
 * x_lorenzo_3d1l (8x8x8: y, x, z, :473-508), in exact int32, then one float multiply by
 * ebx2 (:357, :428, :515). */
static void reconstruct_i32(const int32_t* delta, float* xdata, uint32_t x, uint32_t y, uint32_t z, float ebx2,
                            int nthreads)
{
    const int nd = fzo_ndim(x, y, z);
    if (nd == 1) {
        const int64_t nblk = ((int64_t)x - 1) / 256 + 1;
        OMP_FOR
        for (int64_t b = 0; b < nblk; b++) {
            int32_t acc = 0;
            for (uint64_t id = (uint64_t)b * 256; id < (uint64_t)b * 256 + 256 && id < x; id++) {
                acc += delta[id];
                xdata[id] = (float)acc * ebx2;
            }
        }
    } else if (nd == 2) {
        const int64_t gx = ((int64_t)x - 1) / 16 + 1, gy = ((int64_t)y - 1) / 16 + 1;
        OMP_FOR
        for (int64_t b = 0; b < gx * gy; b++) {
            const uint64_t bx = (uint64_t)(b % gx) * 16, by = (uint64_t)(b / gx) * 16;
            int32_t        s[16][16];
            for (int iy = 0; iy < 16; iy++)
                for (int ix = 0; ix < 16; ix++) {
                    const uint64_t gxi = bx + ix, gyi = by + iy;
                    s[iy][ix] = (gxi < x && gyi < y) ? delta[gyi * x + gxi] : 0;
                }
            for (int iy = 1; iy < 16; iy++)
                for (int ix = 0; ix < 16; ix++) s[iy][ix] += s[iy - 1][ix];
            for (int iy = 0; iy < 16; iy++)
                for (int ix = 1; ix < 16; ix++) s[iy][ix] += s[iy][ix - 1];
            for (int iy = 0; iy < 16; iy++)
                for (int ix = 0; ix < 16; ix++) {
                    const uint64_t gxi = bx + ix, gyi = by + iy;
                    if (gxi < x && gyi < y) xdata[gyi * x + gxi] = (float)s[iy][ix] * ebx2;
                }
        }
    } else {
        const int64_t  gx = ((int64_t)x - 1) / 8 + 1, gy = ((int64_t)y - 1) / 8 + 1, gz = ((int64_t)z - 1) / 8 + 1;
        const uint64_t sy = x, sz = (uint64_t)x * y;
        OMP_FOR
        for (int64_t b = 0; b < gx * gy * gz; b++) {
            const uint64_t bx = (uint64_t)(b % gx) * 8, by = (uint64_t)((b / gx) % gy) * 8,
                           bz = (uint64_t)(b / (gx * gy)) * 8;
            int32_t s[8][8][8];
            for (int iz = 0; iz < 8; iz++)
                for (int iy = 0; iy < 8; iy++)
                    for (int ix = 0; ix < 8; ix++) {
                        const uint64_t gxi = bx + ix, gyi = by + iy, gzi = bz + iz;
                        s[iz][iy][ix] = (gxi < x && gyi < y && gzi < z) ? delta[gzi * sz + gyi * sy + gxi] : 0;
                    }
            for (int iz = 0; iz < 8; iz++)
                for (int iy = 1; iy < 8; iy++)
                    for (int ix = 0; ix < 8; ix++) s[iz][iy][ix] += s[iz][iy - 1][ix];
            for (int iz = 0; iz < 8; iz++)
                for (int iy = 0; iy < 8; iy++)
                    for (int ix = 1; ix < 8; ix++) s[iz][iy][ix] += s[iz][iy][ix - 1];
            for (int iz = 1; iz < 8; iz++)
                for (int iy = 0; iy < 8; iy++)
                    for (int ix = 0; ix < 8; ix++) s[iz][iy][ix] += s[iz - 1][iy][ix];
            for (int iz = 0; iz < 8; iz++)
                for (int iy = 0; iy < 8; iy++)
                    for (int ix = 0; ix < 8; ix++) {
                        const uint64_t gxi = bx + ix, gyi = by + iy, gzi = bz + iz;
                        if (gxi < x && gyi < y && gzi < z) xdata[gzi * sz + gyi * sy + gxi] = (float)s[iz][iy][ix] * ebx2;
                    }
        }
    }
}

void fzo_reconstruct(const uint16_t* delta, float* xdata, uint32_t x, uint32_t y, uint32_t z, float ebx2,
                     int nthreads)
{
    const uint64_t xyz = (uint64_t)x * y * z;
    int32_t*       d = (int32_t*)malloc((size_t)(4 * xyz));
    for (uint64_t i = 0; i < xyz; i++) d[i] = decode(delta[i]);
    reconstruct_i32(d, xdata, x, y, z, ebx2, nthreads);
    free(d);
}

/* Build-defined inverse of runSzs, up to the codes: unpack B, exclusive-scan the flags,
 * scatter the compacted blocks back into the shuffled tiles, unshuffle, append the tail
 * trailer.  codes has room for len + 2. */
static int decode_codes(const uint8_t* payload, uint64_t payload_bytes, const uint16_t* tail, uint64_t tail_len,
                        uint64_t len, int nthreads, uint16_t* codes)
{
    const uint64_t T = fzo_tiles(len), W = fzo_flag_words(len), N = 32 * W;
    if (payload_bytes < 4 * W) return FZO_ERR_CORRUPT;
    if (tail && tail_len != len - 2048 * T) return FZO_ERR_INVALID_ARG;
    uint32_t* flags = (uint32_t*)calloc((size_t)N, sizeof(uint32_t));
    uint32_t* presum = (uint32_t*)calloc((size_t)N + 1, sizeof(uint32_t));
    int       rc = FZO_SUCCESS;
    for (uint64_t i = 0; i < N; i++) {
        const uint64_t wi = i / 32;
        const uint32_t v = (uint32_t)payload[4 * wi] | ((uint32_t)payload[4 * wi + 1] << 8) |
                           ((uint32_t)payload[4 * wi + 2] << 16) | ((uint32_t)payload[4 * wi + 3] << 24);
        flags[i] = (v >> (i % 32)) & 1u;
        if (flags[i] && i >= 256 * T) rc = FZO_ERR_CORRUPT;
    }
    fzo_exclusive_sum(flags, presum, N, nthreads);
    presum[N] = presum[N - 1] + flags[N - 1];
    if (rc == FZO_SUCCESS && payload_bytes != 4 * W + 16 * (uint64_t)presum[N]) rc = FZO_ERR_CORRUPT;
    if (rc != FZO_SUCCESS) {
        free(flags);
        free(presum);
        return rc;
    }
    uint32_t* shuffled = (uint32_t*)calloc((size_t)(1024 * T + 1), sizeof(uint32_t));
    const uint8_t* blocks = payload + 4 * W;
    for (uint64_t i = 0; i < 256 * T; i++)
        if (flags[i]) memcpy((uint8_t*)shuffled + 16 * i, blocks + 16 * (uint64_t)presum[i], 16);
    memset(codes, 0, (size_t)(2 * len + 4));
    fzo_unshuffle(shuffled, (uint32_t*)codes, T, nthreads);
    if (tail)
        for (uint64_t i = 0; i < tail_len; i++) codes[2048 * T + i] = tail[i];
    free(flags);
    free(presum);
    free(shuffled);
    return FZO_SUCCESS;
}

int fzo_decompress(const uint8_t* payload, uint64_t payload_bytes, const uint16_t* tail, uint64_t tail_len,
                   uint64_t len, uint32_t x, uint32_t y, uint32_t z, float ebx2, int nthreads, float* xdata,
                   uint16_t* dbg_codes)
{
    if (len == 0 || x == 0 || y == 0 || z == 0 || (uint64_t)x * y * z > len) return FZO_ERR_INVALID_ARG;
    uint16_t* codes = (uint16_t*)calloc((size_t)len + 2, sizeof(uint16_t));
    const int rc = decode_codes(payload, payload_bytes, tail, tail_len, len, nthreads, codes);
    if (rc == FZO_SUCCESS) {
        fzo_reconstruct(codes, xdata, x, y, z, ebx2, nthreads);
        if (dbg_codes) memcpy(dbg_codes, codes, (size_t)(2 * len));
    }
    free(codes);
    return rc;
}

/* Inverse of the OUTLIER mode: codes -> int32 deltas, the side channel's deltas in place of
 * their (zero) codes, block prefix sums times ebx2 (vconst everywhere when ebx2 == 0), then
 * the verbatim values. */
int fzo_decompress_outlier(const uint8_t* payload, uint64_t payload_bytes, const uint16_t* tail, uint64_t tail_len,
                           const uint8_t* side, uint64_t side_bytes, uint64_t len, uint32_t x, uint32_t y,
                           uint32_t z, float ebx2, float vconst, float* xdata)
{
    if (len == 0 || x == 0 || y == 0 || z == 0 || (uint64_t)x * y * z > len) return FZO_ERR_INVALID_ARG;
    if (side_bytes < 16) return FZO_ERR_CORRUPT;
    const uint64_t nd = get_u64(side), nr = get_u64(side + 8), xyz = (uint64_t)x * y * z;
    if (nd > xyz || nr > xyz || side_bytes != 16 + 8 * (nd + nr)) return FZO_ERR_CORRUPT;
    for (uint64_t k = 0; k < nd; k++)
        if (get_u32(side + 16 + 4 * k) >= xyz) return FZO_ERR_CORRUPT;
    for (uint64_t k = 0; k < nr; k++)
        if (get_u32(side + 16 + 8 * nd + 4 * k) >= xyz) return FZO_ERR_CORRUPT;
    if (ebx2 == 0.0f) {
        for (uint64_t i = 0; i < xyz; i++) xdata[i] = vconst;
    } else {
        uint16_t* codes = (uint16_t*)calloc((size_t)len + 2, sizeof(uint16_t));
        const int rc = decode_codes(payload, payload_bytes, tail, tail_len, len, 1, codes);
        if (rc != FZO_SUCCESS) {
            free(codes);
            return rc;
        }
        int32_t* d = (int32_t*)malloc((size_t)(4 * xyz));
        for (uint64_t i = 0; i < xyz; i++) d[i] = decode(codes[i]);
        for (uint64_t k = 0; k < nd; k++) d[get_u32(side + 16 + 4 * k)] = (int32_t)get_u32(side + 16 + 4 * nd + 4 * k);
        reconstruct_i32(d, xdata, x, y, z, ebx2, 1);
        free(d);
        free(codes);
    }
    for (uint64_t k = 0; k < nr; k++) {
        const uint32_t bits = get_u32(side + 16 + 8 * nd + 4 * nr + 4 * k);
        memcpy(&xdata[get_u32(side + 16 + 8 * nd + 4 * k)], &bits, 4);
    }
    return FZO_SUCCESS;
}

/* verify.hh:25-83.  T = float: abserr is a float (fabs of a float difference, :46), so is
 * its quotient by fabs(odata[i]) (:48) and its square (:58); every other quantity is a
 * double.  Two passes as in the reference: the means first (:31-34), then one loop over
 * extrema, errors and centred products (:39-59). */
void fzo_verify(const float* xdata, const float* odata, uint64_t len, fzo_stats* s)
{
    double max_odata = odata[0], min_odata = odata[0];
    double max_xdata = xdata[0], min_xdata = xdata[0];
    double max_abserr = fabsf(xdata[0] - odata[0]);

    double sum_0 = 0, sum_x = 0;
    for (uint64_t i = 0; i < len; i++) {
        sum_0 += odata[i];
        sum_x += xdata[i];
    }
    const double mean_odata = sum_0 / (double)len, mean_xdata = sum_x / (double)len;
    double sum_var_odata = 0, sum_var_xdata = 0, sum_err2 = 0, sum_corr = 0;
    double max_pwrrel_abserr = 0;
    uint64_t max_abserr_index = 0;
    for (uint64_t i = 0; i < len; i++) {
        const float o = odata[i], x = xdata[i];
        max_odata = max_odata < o ? o : max_odata;
        min_odata = min_odata > o ? o : min_odata;
        /* verify.hh:44 scans odata for the xdata maximum; kept for identical output */
        max_xdata = max_xdata < o ? o : max_xdata;
        min_xdata = min_xdata > x ? x : min_xdata;

        const float abserr = fabsf(x - o);
        if (o != 0) {
            const double rel_abserr = (float)(abserr / fabsf(o));
            max_pwrrel_abserr = max_pwrrel_abserr < rel_abserr ? rel_abserr : max_pwrrel_abserr;
        }
        max_abserr_index = max_abserr < abserr ? i : max_abserr_index;
        max_abserr = max_abserr < abserr ? abserr : max_abserr;
        sum_corr += (o - mean_odata) * (x - mean_xdata);
        sum_var_odata += (o - mean_odata) * (o - mean_odata);
        sum_var_xdata += (x - mean_xdata) * (x - mean_xdata);
        sum_err2 += (float)(abserr * abserr);
    }
    const double std_odata = sqrt(sum_var_odata / (double)len);
    const double std_xdata = sqrt(sum_var_xdata / (double)len);
    const double ee = sum_corr / (double)len;

    s->len = len;
    s->odata.max = max_odata;
    s->odata.min = min_odata;
    s->odata.rng = max_odata - min_odata;
    s->odata.std = std_odata;
    s->xdata.max = max_xdata;
    s->xdata.min = min_xdata;
    s->xdata.rng = max_xdata - min_xdata;
    s->xdata.std = std_xdata;
    s->max_err.idx = max_abserr_index;
    s->max_err.abs = max_abserr;
    s->max_err.rel = max_abserr / s->odata.rng;
    s->max_err.pwrrel = max_pwrrel_abserr;
    s->reduced.coeff = ee / std_odata / std_xdata;
    s->reduced.MSE = sum_err2 / (double)len;
    s->reduced.NRMSE = sqrt(s->reduced.MSE) / s->odata.rng;
    s->reduced.PSNR = 20 * log10(s->odata.rng) - 10 * log10(s->reduced.MSE);
}
