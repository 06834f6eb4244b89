// fz_capi.hip -- the C ABI (include/fz_capi.h): plans, workspace layout, launch sequences.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>

#include "fz_internal.hpp"

using namespace fz;

namespace {

constexpr uint64_t kMaxLen = 0xFFFFF000ull;  // positions are 32-bit inside the kernels

inline size_t al256(size_t v) { return (v + 255) & ~(size_t)255; }

struct Layout {
    size_t ctr, part, tstate, sstate, dstate, codes, shuf, F, P, texcl, counts, stage, total;
    uint32_t n_tstate, n_sstate, n_dstate;
};

Layout make_layout(const Geom& g)
{
    Layout L{};
    const uint64_t N = 32ull * g.W;
    L.n_tstate = gather_grid(g.T) + 2;
    L.n_sstate = (uint32_t)((N + kScanChunk - 1) / kScanChunk) + 1;
    L.n_dstate = (g.T + kWG - 1) / kWG + 1;
    size_t o = 0;
    L.ctr = o;    o = al256(o + sizeof(Ctr));
    L.part = o;   o = al256(o + 16ull * kRangeMaxWG);
    L.tstate = o; o = al256(o + 8ull * L.n_tstate);
    L.sstate = o; o = al256(o + 8ull * L.n_sstate);
    L.dstate = o; o = al256(o + 8ull * L.n_dstate);
    L.codes = o;  o = al256(o + 2ull * g.len + 64);
    L.shuf = o;   o = al256(o + 2ull * g.len + 64);
    L.F = o;      o = al256(o + 4ull * N + 64);
    L.P = o;      o = al256(o + 4ull * (N + 1) + 64);
    L.texcl = o;  o = al256(o + 4ull * (g.T + 1));
    L.counts = o; o = al256(o + 4ull * (g.T + 1));
    L.stage = o;  o = al256(o + 4096ull * g.T + 64);
    L.total = o;
    return L;
}

Geom make_geom(uint32_t x, uint32_t y, uint32_t z, uint64_t len, int mode)
{
    Geom g{};
    g.len = (uint32_t)len;
    g.X = x;
    g.Y = y;
    g.Z = z;
    g.xyz = (uint32_t)((uint64_t)x * y * z);
    g.XY = (uint32_t)((uint64_t)x * y);
    g.T = (uint32_t)(len / kTileCodes);
    g.W = (uint32_t)((2 * len + 511) / 512);
    g.divX = make_fastdiv(x);
    g.divY = make_fastdiv(y);
    g.ndim = (z == 1 && y == 1) ? 1 : (z == 1 ? 2 : 3);
    g.mode = mode;
    return g;
}

bool geom_ok(uint32_t x, uint32_t y, uint32_t z, uint64_t len)
{
    if (x == 0 || y == 0 || z == 0) return false;
    const uint64_t xyz = (uint64_t)x * y * z;
    return len >= xyz && len >= 1 && len <= kMaxLen;
}

inline bool aligned(const void* p, size_t a) { return ((uintptr_t)p % a) == 0; }

thread_local const char* g_last_error = "no error";

inline fz_status hip_st(hipError_t e)
{
    if (e == hipSuccess) return FZ_SUCCESS;
    g_last_error = hipGetErrorString(e);
    return FZ_ERR_HIP;
}

#define FZ_TRY(expr)                                   \
    do {                                               \
        hipError_t e_ = (expr);                        \
        if (e_ != hipSuccess) return hip_st(e_);       \
    } while (0)

}  // namespace

constexpr int kMaxProfEvents = 8;

struct fz_plan_s {
    Geom       g;
    Layout     L;
    int        mode;
    int        prof = 0;                    // per-kernel events enabled
    int        nev = 0;                     // events recorded by the last call
    int        paired = 0;                  // last call recorded (start, stop) pairs per kernel
    hipEvent_t ev[kMaxProfEvents] = {};
};

namespace {
// records event i of the current call when profiling is on
inline hipError_t prof_mark(fz_plan p, hipStream_t s, bool first)
{
    if (!p->prof) return hipSuccess;
    if (first) {
        p->nev = 0;
        p->paired = 0;
    }
    if (p->nev >= kMaxProfEvents) return hipSuccess;
    return hipEventRecord(p->ev[p->nev++], s);
}
}  // namespace

extern "C" {

const char* fz_version(void) { return "fz-mi355x 0.1 (gfx950, C ABI v1)"; }

const char* fz_last_error(void) { return g_last_error; }

const char* fz_status_string(fz_status s)
{
    switch (s) {
        case FZ_SUCCESS: return "success";
        case FZ_ERR_INVALID_ARG: return "invalid argument";
        case FZ_ERR_WORKSPACE_TOO_SMALL: return "workspace too small";
        case FZ_ERR_HIP: return "HIP runtime error";
        case FZ_ERR_CONSTANT_FIELD: return "constant field (value range is 0)";
        case FZ_ERR_NONFINITE: return "non-finite value in the input";
        case FZ_ERR_CORRUPT: return "corrupt payload";
        case FZ_ERR_CAPACITY: return "capacity exceeded";
        case FZ_ERR_NOT_READY: return "result not ready";
    }
    return "unknown status";
}

fz_status fz_plan_create(fz_plan* plan, uint32_t x, uint32_t y, uint32_t z, uint64_t len, int mode)
{
    if (!plan) return FZ_ERR_INVALID_ARG;
    *plan = nullptr;
    if (len == 0) len = (uint64_t)x * y * z;
    if (!geom_ok(x, y, z, len) || (mode != FZ_MODE_FIXED && mode != FZ_MODE_LITERAL)) return FZ_ERR_INVALID_ARG;
    fz_plan p = new (std::nothrow) fz_plan_s;
    if (!p) return FZ_ERR_CAPACITY;
    p->g = make_geom(x, y, z, len, mode);
    p->L = make_layout(p->g);
    p->mode = mode;
    *plan = p;
    return FZ_SUCCESS;
}

fz_status fz_plan_destroy(fz_plan plan)
{
    if (!plan) return FZ_SUCCESS;
    for (int i = 0; i < kMaxProfEvents; i++)
        if (plan->ev[i]) (void)hipEventDestroy(plan->ev[i]);
    delete plan;
    return FZ_SUCCESS;
}

fz_status fz_plan_profile(fz_plan plan, int enable)
{
    if (!plan) return FZ_ERR_INVALID_ARG;
    if (enable && !plan->ev[0])
        for (int i = 0; i < kMaxProfEvents; i++) FZ_TRY(hipEventCreate(&plan->ev[i]));
    plan->prof = enable ? 1 : 0;
    plan->nev = 0;
    return FZ_SUCCESS;
}

fz_status fz_plan_kernel_times(const fz_plan plan, float* ms, int capacity, int* count)
{
    if (!plan || !count) return FZ_ERR_INVALID_ARG;
    *count = 0;
    if (!plan->prof || plan->nev < 2) return FZ_ERR_NOT_READY;
    FZ_TRY(hipEventSynchronize(plan->ev[plan->nev - 1]));
    const int n = plan->paired ? plan->nev / 2 : plan->nev - 1;
    for (int i = 0; i < n && i < capacity; i++) {
        if (plan->paired) FZ_TRY(hipEventElapsedTime(&ms[i], plan->ev[2 * i], plan->ev[2 * i + 1]));
        else FZ_TRY(hipEventElapsedTime(&ms[i], plan->ev[i], plan->ev[i + 1]));
    }
    *count = n < capacity ? n : capacity;
    return FZ_SUCCESS;
}

fz_status fz_plan_sizes(const fz_plan plan, fz_sizes* s)
{
    if (!plan || !s) return FZ_ERR_INVALID_ARG;
    const Geom& g = plan->g;
    s->len = g.len;
    s->tiles = g.T;
    s->flag_words = g.W;
    s->byte_flags = 32ull * g.W;
    s->payload_max_bytes = 4ull * g.W + 4096ull * g.T;
    s->tail_max_codes = (uint64_t)g.len - (uint64_t)kTileCodes * g.T;
    s->workspace_bytes = plan->L.total;
    return FZ_SUCCESS;
}

fz_status fz_workspace_init(const fz_plan plan, void* d_ws, size_t ws_bytes, hipStream_t stream)
{
    if (!plan || !d_ws || !aligned(d_ws, 256)) return FZ_ERR_INVALID_ARG;
    if (ws_bytes < plan->L.total) return FZ_ERR_WORKSPACE_TOO_SMALL;
    // counters, partials and all look-back states sit before the code buffers
    return hip_st(launch_fill_u32(static_cast<uint32_t*>(d_ws), plan->L.codes / 4, 0u, stream));
}

static fz_status check_compress_args(const fz_plan plan, const float* d_in, double eb_rel, void* d_ws,
                                     size_t ws_bytes, uint8_t* d_payload, uint16_t* d_tail, fz_result* d_result)
{
    if (!plan || !d_in || !d_ws || !d_payload || !d_result) return FZ_ERR_INVALID_ARG;
    if (!(eb_rel > 0.0) || !std::isfinite(eb_rel)) return FZ_ERR_INVALID_ARG;
    if (!aligned(d_in, 16) || !aligned(d_payload, 16) || !aligned(d_ws, 256) || !aligned(d_result, 8))
        return FZ_ERR_INVALID_ARG;
    const Geom& g = plan->g;
    if (plan->mode == FZ_MODE_FIXED && !d_tail && g.len != (uint64_t)kTileCodes * g.T) return FZ_ERR_INVALID_ARG;
    if (ws_bytes < plan->L.total) return FZ_ERR_WORKSPACE_TOO_SMALL;
    return FZ_SUCCESS;
}

fz_status fz_compress_f32(const fz_plan plan, const float* d_in, double eb_rel, void* d_ws, size_t ws_bytes,
                          uint8_t* d_payload, uint16_t* d_tail, fz_result* d_result, hipStream_t stream)
{
    fz_status st = check_compress_args(plan, d_in, eb_rel, d_ws, ws_bytes, d_payload, d_tail, d_result);
    if (st != FZ_SUCCESS) return st;
    const Geom&   g = plan->g;
    const Layout& L = plan->L;
    uint8_t*      ws = static_cast<uint8_t*>(d_ws);
    uint64_t*     tstate = reinterpret_cast<uint64_t*>(ws + L.tstate);
    hipEvent_t* ev = nullptr;
    if (plan->prof) {  // dispatch-timestamped start / stop of each kernel
        ev = plan->ev;
        plan->nev = g.T == 0 ? 4 : 6;
        plan->paired = 1;
    }
    FZ_TRY(launch_compress(d_in, g, eb_rel, reinterpret_cast<float4*>(ws + L.part), d_result, d_payload, d_tail,
                           ws + L.stage, reinterpret_cast<uint32_t*>(ws + L.counts), tstate, ev, stream));
    return FZ_SUCCESS;
}

fz_status fz_compress_plain_f32(const fz_plan plan, const float* d_in, double eb_rel, void* d_ws, size_t ws_bytes,
                                uint8_t* d_payload, uint16_t* d_tail, fz_result* d_result, hipStream_t stream)
{
    fz_status st = check_compress_args(plan, d_in, eb_rel, d_ws, ws_bytes, d_payload, d_tail, d_result);
    if (st != FZ_SUCCESS) return st;
    const Geom&   g = plan->g;
    const Layout& L = plan->L;
    uint8_t*      ws = static_cast<uint8_t*>(d_ws);
    Ctr*          ctr = reinterpret_cast<Ctr*>(ws + L.ctr);
    uint64_t*     sstate = reinterpret_cast<uint64_t*>(ws + L.sstate);
    uint16_t*     codes = reinterpret_cast<uint16_t*>(ws + L.codes);
    uint32_t*     shuf = reinterpret_cast<uint32_t*>(ws + L.shuf);
    uint32_t*     F = reinterpret_cast<uint32_t*>(ws + L.F);
    uint32_t*     P = reinterpret_cast<uint32_t*>(ws + L.P);
    const uint32_t N = 32u * g.W;
    FZ_TRY(prof_mark(plan, stream, true));
    FZ_TRY(launch_range(d_in, g, eb_rel, reinterpret_cast<float4*>(ws + L.part), d_result, sstate, L.n_sstate,
                        &ctr->scan_ticket, stream));
    FZ_TRY(prof_mark(plan, stream, false));
    FZ_TRY(launch_dualquant(d_in, g, 0.0f, 0.0f, d_result, codes,
                            reinterpret_cast<unsigned long long*>(&d_result->overflow), stream));
    FZ_TRY(prof_mark(plan, stream, false));
    FZ_TRY(launch_bitshuffle(reinterpret_cast<const uint32_t*>(codes), shuf, g, reinterpret_cast<uint32_t*>(d_payload),
                             F, true, stream));
    FZ_TRY(prof_mark(plan, stream, false));
    FZ_TRY(launch_scan(F, P, N, sstate, &ctr->scan_ticket, true, d_result, g.W, stream));
    FZ_TRY(prof_mark(plan, stream, false));
    FZ_TRY(launch_encode(reinterpret_cast<const uint8_t*>(shuf), d_payload + 4ull * g.W, P, N, stream));
    FZ_TRY(prof_mark(plan, stream, false));
    const uint64_t tail_len = (uint64_t)g.len - (uint64_t)kTileCodes * g.T;
    if (plan->mode == FZ_MODE_FIXED && tail_len)
        FZ_TRY(hipMemcpyAsync(d_tail, codes + (uint64_t)kTileCodes * g.T, 2 * tail_len, hipMemcpyDeviceToDevice,
                              stream));
    return FZ_SUCCESS;
}

fz_status fz_decompress_f32(const fz_plan plan, const uint8_t* d_payload, const uint16_t* d_tail, float ebx2,
                            void* d_ws, size_t ws_bytes, float* d_out, hipStream_t stream)
{
    if (!plan || !d_payload || !d_ws || !d_out) return FZ_ERR_INVALID_ARG;
    if (!aligned(d_payload, 16) || !aligned(d_ws, 256) || !aligned(d_out, 16)) return FZ_ERR_INVALID_ARG;
    if (!(ebx2 > 0.0f) || !std::isfinite(ebx2)) return FZ_ERR_INVALID_ARG;
    if (ws_bytes < plan->L.total) return FZ_ERR_WORKSPACE_TOO_SMALL;
    const Geom&   g = plan->g;
    const Layout& L = plan->L;
    uint8_t*      ws = static_cast<uint8_t*>(d_ws);
    Ctr*          ctr = reinterpret_cast<Ctr*>(ws + L.ctr);
    uint64_t*     dstate = reinterpret_cast<uint64_t*>(ws + L.dstate);
    uint32_t*     texcl = reinterpret_cast<uint32_t*>(ws + L.texcl);
    uint32_t*     codes = reinterpret_cast<uint32_t*>(ws + L.codes);
    const uint16_t* tail = plan->mode == FZ_MODE_FIXED ? d_tail : nullptr;
    FZ_TRY(prof_mark(plan, stream, true));
    FZ_TRY(launch_scan_init(dstate, L.n_dstate, &ctr->dscan_ticket, stream));
    FZ_TRY(prof_mark(plan, stream, false));
    FZ_TRY(launch_tile_scan(reinterpret_cast<const uint32_t*>(d_payload), g.T, texcl, dstate, &ctr->dscan_ticket,
                            stream));
    FZ_TRY(prof_mark(plan, stream, false));
    FZ_TRY(launch_decode_unshuffle(d_payload, texcl, tail, g, codes, stream));
    FZ_TRY(prof_mark(plan, stream, false));
    FZ_TRY(launch_reconstruct(reinterpret_cast<const uint16_t*>(codes), d_out, g, ebx2, stream));
    FZ_TRY(prof_mark(plan, stream, false));
    return FZ_SUCCESS;
}

fz_status fz_compress_f32_sync(const fz_plan plan, const float* d_in, double eb_rel, void* d_ws, size_t ws_bytes,
                               uint8_t* d_payload, uint16_t* d_tail, fz_result* d_result, fz_result* h_result,
                               hipStream_t stream)
{
    if (!h_result) return FZ_ERR_INVALID_ARG;
    fz_status st = fz_compress_f32(plan, d_in, eb_rel, d_ws, ws_bytes, d_payload, d_tail, d_result, stream);
    if (st != FZ_SUCCESS) return st;
    FZ_TRY(hipMemcpyAsync(h_result, d_result, sizeof(fz_result), hipMemcpyDeviceToHost, stream));
    FZ_TRY(hipStreamSynchronize(stream));
    return (fz_status)h_result->status;
}

// ---------------------------------------------------------------- per-stage entry points

namespace {
struct EventTimer {
    hipEvent_t a = nullptr, b = nullptr;
    float*     out;
    hipStream_t s;
    hipError_t  err = hipSuccess;
    EventTimer(float* o, hipStream_t st) : out(o), s(st)
    {
        if (out) {
            err = hipEventCreate(&a);
            if (err == hipSuccess) err = hipEventCreate(&b);
            if (err == hipSuccess) err = hipEventRecord(a, s);
        }
    }
    hipError_t finish()
    {
        if (!out) return hipSuccess;
        if (err != hipSuccess) return err;
        hipError_t e = hipEventRecord(b, s);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        if (e == hipSuccess) e = hipEventElapsedTime(out, a, b);
        return e;
    }
    ~EventTimer()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};
}  // namespace

fz_status fz_lorenzo_construct_f32(const float* d_data, uint16_t* d_delta, uint32_t x, uint32_t y, uint32_t z,
                                   uint64_t len, double eb_abs, int mode, float* time_elapsed, hipStream_t stream)
{
    if (len == 0) len = (uint64_t)x * y * z;
    if (!d_data || !d_delta || !geom_ok(x, y, z, len) || !(eb_abs > 0) || !std::isfinite(eb_abs) ||
        (mode != FZ_MODE_FIXED && mode != FZ_MODE_LITERAL) || !aligned(d_data, 16) || !aligned(d_delta, 16))
        return FZ_ERR_INVALID_ARG;
    const Geom  g = make_geom(x, y, z, len, mode);
    const double ebx2 = eb_abs * 2;        // lorenzo_var.cuh:562
    const float  ebx2_r = (float)(1 / ebx2);  // :563, narrowed at the kernel parameter
    EventTimer  tm(time_elapsed, stream);
    // the fast paths need max|x| * ebx2_r < 2^23; without the range at hand take the guarded path
    FZ_TRY(launch_dualquant(d_data, g, ebx2_r, __builtin_huge_valf(), nullptr, d_delta, nullptr, stream));
    FZ_TRY(tm.finish());
    return FZ_SUCCESS;
}

fz_status fz_lorenzo_reconstruct_f32(const uint16_t* d_delta, float* d_xdata, uint32_t x, uint32_t y, uint32_t z,
                                     double eb_abs, float* time_elapsed, hipStream_t stream)
{
    const uint64_t len = (uint64_t)x * y * z;
    if (!d_delta || !d_xdata || !geom_ok(x, y, z, len) || !(eb_abs > 0) || !std::isfinite(eb_abs) ||
        !aligned(d_delta, 16) || !aligned(d_xdata, 16))
        return FZ_ERR_INVALID_ARG;
    const Geom g = make_geom(x, y, z, len, FZ_MODE_FIXED);
    EventTimer tm(time_elapsed, stream);
    FZ_TRY(launch_reconstruct(d_delta, d_xdata, g, (float)(eb_abs * 2), stream));
    FZ_TRY(tm.finish());
    return FZ_SUCCESS;
}

fz_status fz_bitshuffle_bitflag(const uint32_t* d_in, uint32_t* d_out, uint64_t len, uint32_t* d_bitflags,
                                uint32_t* d_byteflags, hipStream_t stream)
{
    if (!d_in || !d_out || !d_bitflags || !d_byteflags || len > kMaxLen || !aligned(d_in, 16) || !aligned(d_out, 16))
        return FZ_ERR_INVALID_ARG;
    const Geom g = make_geom((uint32_t)(len ? len : 1), 1, 1, len ? len : 1, FZ_MODE_FIXED);
    if (len < (uint64_t)kTileCodes) return FZ_SUCCESS;
    return hip_st(launch_bitshuffle(d_in, d_out, g, d_bitflags, d_byteflags, false, stream));
}

fz_status fz_bitunshuffle(const uint32_t* d_in, uint32_t* d_out, uint64_t len, hipStream_t stream)
{
    if (!d_in || !d_out || len > kMaxLen || !aligned(d_in, 16) || !aligned(d_out, 16)) return FZ_ERR_INVALID_ARG;
    return hip_st(launch_unshuffle(d_in, d_out, (uint32_t)(len / kTileCodes), stream));
}

fz_status fz_exclusive_sum_u32(void* d_temp, size_t* temp_bytes, const uint32_t* d_in, uint32_t* d_out,
                               uint64_t num_items, hipStream_t stream)
{
    if (!temp_bytes || num_items > kMaxLen) return FZ_ERR_INVALID_ARG;
    const uint32_t nchunks = (uint32_t)((num_items + kScanChunk - 1) / kScanChunk);
    const size_t   need = 256 + 8ull * (nchunks + 1);
    if (!d_temp) {
        *temp_bytes = need;
        return FZ_SUCCESS;
    }
    if (*temp_bytes < need) return FZ_ERR_WORKSPACE_TOO_SMALL;
    if (!d_in || !d_out || !aligned(d_in, 16) || !aligned(d_out, 16) || !aligned(d_temp, 256))
        return FZ_ERR_INVALID_ARG;
    if (num_items == 0) return FZ_SUCCESS;
    uint32_t* ticket = static_cast<uint32_t*>(d_temp);
    uint64_t* states = reinterpret_cast<uint64_t*>(static_cast<uint8_t*>(d_temp) + 256);
    FZ_TRY(launch_scan_init(states, nchunks + 1, ticket, stream));
    FZ_TRY(launch_scan(d_in, d_out, (uint32_t)num_items, states, ticket, false, nullptr, 0, stream));
    return FZ_SUCCESS;
}

fz_status fz_encode_blocks(const uint8_t* d_in, uint8_t* d_out, const uint32_t* d_presum, uint64_t num_blocks,
                           hipStream_t stream)
{
    if (!d_in || !d_out || !d_presum || num_blocks > kMaxLen || !aligned(d_in, 16) || !aligned(d_out, 4))
        return FZ_ERR_INVALID_ARG;
    return hip_st(launch_encode(d_in, d_out, d_presum, (uint32_t)num_blocks, stream));
}

}  // extern "C"
