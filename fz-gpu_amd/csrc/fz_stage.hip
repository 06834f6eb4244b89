// fz_stage.hip -- host-resident requests: pinned, asynchronous host <-> device staging around
// the fused compress, overlapped across requests.
//
// The reference copies the field in with a blocking pageable cudaMemcpy (fz.cu:143) and reads
// the scan total back with another (fz.cu:195-201), so transfers and kernels of one field and
// of the next never overlap.  A stager keeps `depth` request slots in flight on three streams
// (all at the default priority):
//   copy-in   H2D of request i from the caller's host buffer (pinned memory gives an
//             asynchronous DMA; fz_host_alloc allocates it)
//   compute   fz_compress_f32 of request i into its slot's device buffers, after its H2D
//   copy-out  k_stage_out of request i: the record, the payload_bytes of payload and the tail
//             codes written by the GPU straight into the slot's pinned host buffer (their size
//             is only known on the device, so no host round trip is needed to size a copy)
// so request i + 1 copies in while request i computes and request i - 1 copies out.  The
// compute stream runs one call at a time, so the slots share one workspace.  Decompress
// requests take the same slots: payload and tail in, fz_decompress_f32, the floats out to the
// caller's host buffer.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>

#include "fz_internal.hpp"

struct fz_stager_s {
    struct Slot {
        float*     d_in = nullptr;
        uint8_t*   d_payload = nullptr;
        uint16_t*  d_tail = nullptr;
        fz_result* d_res = nullptr;
        uint8_t*   h_out = nullptr;  // pinned: record (64 B) | payload (payload_max) | tail codes
        hipEvent_t e_in = nullptr, e_comp = nullptr, e_done = nullptr;
        uint64_t   ticket = 0;
        bool       busy = false;
        bool       compress = true;
    };
    fz_plan     plan = nullptr;
    fz_sizes    sz{};
    uint64_t    xyz = 0;
    int         depth = 0;
    uint64_t    next = 0;  // ticket of the next request
    hipStream_t s_in = nullptr, s_comp = nullptr, s_out = nullptr;
    void*       d_ws = nullptr;
    Slot        slot[FZ_STAGER_MAX_DEPTH];
};

namespace fz {
namespace {

constexpr uint32_t kOutWG = 256;
constexpr uint32_t kOutGrid = 512;

// record, payload and tail of one compress into pinned host memory (grid-stride, 16-byte
// stores where the payload allows)
__global__ __launch_bounds__(kOutWG) void k_stage_out(const fz_result* __restrict__ res,
                                                      const uint8_t* __restrict__ payload,
                                                      const uint16_t* __restrict__ tail, uint8_t* __restrict__ h_out,
                                                      uint64_t payload_cap)
{
    const uint64_t gtid = (uint64_t)blockIdx.x * kOutWG + threadIdx.x, nthr = (uint64_t)gridDim.x * kOutWG;
    if (gtid < sizeof(fz_result) / 4)
        reinterpret_cast<uint32_t*>(h_out)[gtid] = reinterpret_cast<const uint32_t*>(res)[gtid];
    if (res->status != 0) return;
    const uint64_t pb = res->payload_bytes < payload_cap ? res->payload_bytes : payload_cap;
    uint8_t*       hp = h_out + sizeof(fz_result);
    for (uint64_t i = gtid; i < pb / 16; i += nthr)
        reinterpret_cast<uint4*>(hp)[i] = reinterpret_cast<const uint4*>(payload)[i];
    for (uint64_t i = (pb / 16) * 4 + gtid; i < pb / 4; i += nthr)
        reinterpret_cast<uint32_t*>(hp)[i] = reinterpret_cast<const uint32_t*>(payload)[i];
    uint16_t* ht = reinterpret_cast<uint16_t*>(hp + payload_cap);
    for (uint64_t i = gtid; i < res->tail_len; i += nthr) ht[i] = tail[i];
}

}  // namespace
}  // namespace fz

using namespace fz;

namespace {

fz_status hipst(hipError_t e) { return e == hipSuccess ? FZ_SUCCESS : FZ_ERR_HIP; }

#define ST_TRY(expr)                                     \
    do {                                                 \
        const fz_status s_ = (expr);                     \
        if (s_ != FZ_SUCCESS) return s_;                 \
    } while (0)

void stager_free(fz_stager s)
{
    for (auto& sl : s->slot) {
        if (sl.d_in) (void)hipFree(sl.d_in);
        if (sl.d_payload) (void)hipFree(sl.d_payload);
        if (sl.d_tail) (void)hipFree(sl.d_tail);
        if (sl.d_res) (void)hipFree(sl.d_res);
        if (sl.h_out) (void)hipHostFree(sl.h_out);
        for (hipEvent_t e : {sl.e_in, sl.e_comp, sl.e_done})
            if (e) (void)hipEventDestroy(e);
    }
    if (s->d_ws) (void)hipFree(s->d_ws);
    for (hipStream_t st : {s->s_in, s->s_comp, s->s_out})
        if (st) (void)hipStreamDestroy(st);
    if (s->plan) (void)fz_plan_destroy(s->plan);
    delete s;
}

uint64_t out_bytes(const fz_sizes& sz) { return sizeof(fz_result) + sz.payload_max_bytes + 2 * sz.tail_max_codes + 16; }

}  // namespace

extern "C" {

fz_status fz_host_alloc(void** p, size_t bytes)
{
    if (!p || bytes == 0) return FZ_ERR_INVALID_ARG;
    return hipst(hipHostMalloc(p, bytes, hipHostMallocDefault));
}

fz_status fz_host_free(void* p) { return p ? hipst(hipHostFree(p)) : FZ_SUCCESS; }

fz_status fz_stager_create(fz_stager* out, uint32_t x, uint32_t y, uint32_t z, uint64_t len, int depth)
{
    if (!out || depth < 1 || depth > FZ_STAGER_MAX_DEPTH) return FZ_ERR_INVALID_ARG;
    *out = nullptr;
    fz_stager s = new (std::nothrow) fz_stager_s;
    if (!s) return FZ_ERR_CAPACITY;
    s->depth = depth;
    s->xyz = (uint64_t)x * y * z;
    fz_status st = fz_plan_create(&s->plan, x, y, z, len, FZ_MODE_FIXED);
    if (st == FZ_SUCCESS) st = fz_plan_sizes(s->plan, &s->sz);
    if (st != FZ_SUCCESS) {
        stager_free(s);
        return st;
    }
    const fz_sizes& sz = s->sz;
    auto fail = [&](hipError_t e) {
        stager_free(s);
        return hipst(e);
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&s->s_in, hipStreamNonBlocking)) != hipSuccess) return fail(e);
    if ((e = hipStreamCreateWithFlags(&s->s_comp, hipStreamNonBlocking)) != hipSuccess) return fail(e);
    if ((e = hipStreamCreateWithFlags(&s->s_out, hipStreamNonBlocking)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&s->d_ws, sz.workspace_bytes)) != hipSuccess) return fail(e);
    for (int i = 0; i < depth; i++) {
        auto& sl = s->slot[i];
        if ((e = hipMalloc(&sl.d_in, 4 * sz.len + 16)) != hipSuccess) return fail(e);
        if ((e = hipMalloc(&sl.d_payload, sz.payload_max_bytes + 16)) != hipSuccess) return fail(e);
        if ((e = hipMalloc(&sl.d_tail, 2 * sz.tail_max_codes + 16)) != hipSuccess) return fail(e);
        if ((e = hipMalloc(&sl.d_res, sizeof(fz_result))) != hipSuccess) return fail(e);
        if ((e = hipHostMalloc(reinterpret_cast<void**>(&sl.h_out), out_bytes(sz), hipHostMallocDefault)) != hipSuccess)
            return fail(e);
        for (hipEvent_t* ev : {&sl.e_in, &sl.e_comp, &sl.e_done})
            if ((e = hipEventCreateWithFlags(ev, hipEventDisableTiming)) != hipSuccess) return fail(e);
    }
    st = fz_workspace_init(s->plan, s->d_ws, sz.workspace_bytes, s->s_comp);
    if (st == FZ_SUCCESS) st = hipst(hipStreamSynchronize(s->s_comp));
    if (st != FZ_SUCCESS) {
        stager_free(s);
        return st;
    }
    *out = s;
    return FZ_SUCCESS;
}

fz_status fz_stager_destroy(fz_stager s)
{
    if (!s) return FZ_SUCCESS;
    for (hipStream_t st : {s->s_in, s->s_comp, s->s_out}) (void)hipStreamSynchronize(st);
    stager_free(s);
    return FZ_SUCCESS;
}

fz_status fz_stager_sizes(const fz_stager s, fz_sizes* sizes)
{
    if (!s || !sizes) return FZ_ERR_INVALID_ARG;
    *sizes = s->sz;
    return FZ_SUCCESS;
}

fz_status fz_stager_compress(fz_stager s, const float* h_in, double eb_rel, uint64_t* ticket)
{
    if (!s || !h_in || !ticket || !(eb_rel > 0.0) || !std::isfinite(eb_rel)) return FZ_ERR_INVALID_ARG;
    const uint64_t t = s->next;
    auto&          sl = s->slot[t % (uint64_t)s->depth];
    // the slot's previous request (depth requests ago) must have left it: its results are
    // overwritten from here on
    if (sl.busy) ST_TRY(hipst(hipEventSynchronize(sl.e_done)));
    const fz_sizes& sz = s->sz;
    ST_TRY(hipst(hipMemcpyAsync(sl.d_in, h_in, 4 * sz.len, hipMemcpyHostToDevice, s->s_in)));
    ST_TRY(hipst(hipEventRecord(sl.e_in, s->s_in)));
    ST_TRY(hipst(hipStreamWaitEvent(s->s_comp, sl.e_in, 0)));
    ST_TRY(fz_compress_f32(s->plan, sl.d_in, eb_rel, s->d_ws, sz.workspace_bytes, sl.d_payload, sl.d_tail, sl.d_res,
                           s->s_comp));
    ST_TRY(hipst(hipEventRecord(sl.e_comp, s->s_comp)));
    ST_TRY(hipst(hipStreamWaitEvent(s->s_out, sl.e_comp, 0)));
    (void)hipGetLastError();
    k_stage_out<<<kOutGrid, kOutWG, 0, s->s_out>>>(sl.d_res, sl.d_payload, sl.d_tail, sl.h_out, sz.payload_max_bytes);
    ST_TRY(hipst(hipGetLastError()));
    ST_TRY(hipst(hipEventRecord(sl.e_done, s->s_out)));
    sl.ticket = t;
    sl.busy = true;
    sl.compress = true;
    s->next = t + 1;
    *ticket = t;
    return FZ_SUCCESS;
}

fz_status fz_stager_decompress(fz_stager s, const uint8_t* h_payload, uint64_t payload_bytes, const uint16_t* h_tail,
                               float ebx2, float* h_out, uint64_t* ticket)
{
    if (!s || !h_payload || !h_out || !ticket || payload_bytes > s->sz.payload_max_bytes) return FZ_ERR_INVALID_ARG;
    const uint64_t t = s->next;
    auto&          sl = s->slot[t % (uint64_t)s->depth];
    if (sl.busy) ST_TRY(hipst(hipEventSynchronize(sl.e_done)));
    const fz_sizes& sz = s->sz;
    const uint64_t  n = s->xyz;
    ST_TRY(hipst(hipMemcpyAsync(sl.d_payload, h_payload, payload_bytes, hipMemcpyHostToDevice, s->s_in)));
    if (h_tail && sz.tail_max_codes)
        ST_TRY(hipst(hipMemcpyAsync(sl.d_tail, h_tail, 2 * sz.tail_max_codes, hipMemcpyHostToDevice, s->s_in)));
    ST_TRY(hipst(hipEventRecord(sl.e_in, s->s_in)));
    ST_TRY(hipst(hipStreamWaitEvent(s->s_comp, sl.e_in, 0)));
    ST_TRY(fz_decompress_f32(s->plan, sl.d_payload, h_tail ? sl.d_tail : nullptr, ebx2, s->d_ws, sz.workspace_bytes,
                             sl.d_in, s->s_comp));
    ST_TRY(hipst(hipEventRecord(sl.e_comp, s->s_comp)));
    ST_TRY(hipst(hipStreamWaitEvent(s->s_out, sl.e_comp, 0)));
    ST_TRY(hipst(hipMemcpyAsync(h_out, sl.d_in, 4 * n, hipMemcpyDeviceToHost, s->s_out)));
    ST_TRY(hipst(hipEventRecord(sl.e_done, s->s_out)));
    sl.ticket = t;
    sl.busy = true;
    sl.compress = false;
    s->next = t + 1;
    *ticket = t;
    return FZ_SUCCESS;
}

fz_status fz_stager_wait(fz_stager s, uint64_t ticket, fz_result* result, const uint8_t** payload,
                         const uint16_t** tail)
{
    if (!s) return FZ_ERR_INVALID_ARG;
    auto& sl = s->slot[ticket % (uint64_t)s->depth];
    if (!sl.busy || sl.ticket != ticket) return FZ_ERR_INVALID_ARG;  // never submitted, or its slot reused
    ST_TRY(hipst(hipEventSynchronize(sl.e_done)));
    if (!sl.compress) {
        if (result) std::memset(result, 0, sizeof(fz_result));
        if (payload) *payload = nullptr;
        if (tail) *tail = nullptr;
        return FZ_SUCCESS;
    }
    fz_result r;
    std::memcpy(&r, sl.h_out, sizeof(fz_result));
    if (result) *result = r;
    if (payload) *payload = sl.h_out + sizeof(fz_result);
    if (tail) *tail = reinterpret_cast<const uint16_t*>(sl.h_out + sizeof(fz_result) + s->sz.payload_max_bytes);
    return (fz_status)r.status;
}

}  // extern "C"
