// fz_capi.hip -- the C ABI (include/fz_capi.h): plans, workspace layout, launch sequences.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <new>

#include "fz_internal.hpp"

using namespace fz;

namespace {

constexpr uint64_t kMaxLen = 0xFFFFF000ull;  // positions are 32-bit inside the kernels

inline size_t al256(size_t v) { return (v + 255) & ~(size_t)255; }

struct Layout {
    size_t ctr, part, tstate, sstate, dstate, obits, codes, shuf, F, P, texcl, counts, stage, vtemp, ocnt, orun, over,
        total;
    uint32_t n_tstate, n_sstate, n_dstate;
};

// Workspace regions, 256-byte aligned.  State words are shared between kernels of one call:
//   tstate  fused compress: per-placement-run block sums (u32), cleared by k_range_part;
//   sstate  u32 flag scan of the plain pipeline: look-back states, cleared by the range kernel;
//   dstate  decompress tile offsets: look-back states (u64 flag|value), or (<= 65536 tiles,
//           launch_tile_offsets) raw u32 chunk counts.  Its contents are undefined after any
//           decompress, so every look-back user clears it first (launch_tile_scan with clear,
//           or k_put_deltas in the outlier inverse).
Layout make_layout(const Geom& g, int mode)
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
    if (mode == FZ_MODE_OUTLIER) {  // inside the zero-filled prefix: no bit set between calls
        L.obits = o; o = al256(o + 4ull * ((g.xyz + 31) / 32));
    }
    L.codes = o;  o = al256(o + 2ull * g.len + 64);
    L.shuf = o;   o = al256(o + 2ull * g.len + 64);
    L.F = o;      o = al256(o + 4ull * N + 64);
    L.P = o;      o = al256(o + 4ull * (N + 1) + 64);
    L.texcl = o;  o = al256(o + 4ull * (g.T + 1));
    L.counts = o; o = al256(o + 4ull * (g.T + 1));
    L.stage = o;  o = al256(o + 4096ull * g.T + 64);
    L.vtemp = o;  o = al256(o + verify_workspace_bytes((uint64_t)g.xyz));
    if (mode == FZ_MODE_OUTLIER) {  // per-unit side-channel counts and offsets; int32 deltas
        L.ocnt = o;   o = al256(o + 8ull * (g.T + 1));
        L.orun = o;   o = al256(o + 8ull * outlier_runs(g.T));
        L.over = o;   o = al256(o + 4ull * g.xyz + 64);
    }
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
    g.mode = mode == FZ_MODE_LITERAL ? 1 : 0;  // kernels: 1 = the reference's 1D sign loss and no tail
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
    int        wave_tile = 0;               // tile-kernel layout of the fused compress
    int        inverse_form = FZ_INVERSE_AUTO;  // kernel form of the 2D / 3D decompress
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
    if (!geom_ok(x, y, z, len) || (mode != FZ_MODE_FIXED && mode != FZ_MODE_LITERAL && mode != FZ_MODE_OUTLIER))
        return FZ_ERR_INVALID_ARG;
    fz_plan p = new (std::nothrow) fz_plan_s;
    if (!p) return FZ_ERR_CAPACITY;
    p->g = make_geom(x, y, z, len, mode);
    p->L = make_layout(p->g, mode);
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

fz_status fz_plan_set_tile_layout(fz_plan plan, int wave_tile)
{
    if (!plan || (wave_tile != 0 && wave_tile != 1)) return FZ_ERR_INVALID_ARG;
    plan->wave_tile = wave_tile;
    return FZ_SUCCESS;
}

fz_status fz_plan_set_inverse_form(fz_plan plan, int form)
{
    if (!plan || form < FZ_INVERSE_AUTO || form > FZ_INVERSE_SPLIT) return FZ_ERR_INVALID_ARG;
    plan->inverse_form = form;
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
    s->side_max_bytes = plan->mode == FZ_MODE_OUTLIER ? 16ull + 8ull * g.xyz : 0;
    return FZ_SUCCESS;
}

fz_status fz_workspace_init(const fz_plan plan, void* d_ws, size_t ws_bytes, hipStream_t stream)
{
    if (!plan || !d_ws || !aligned(d_ws, 256)) return FZ_ERR_INVALID_ARG;
    if (ws_bytes < plan->L.total) return FZ_ERR_WORKSPACE_TOO_SMALL;
    // counters, partials, all look-back states and the outlier bitmap sit before the code buffers
    return hip_st(launch_fill_u32(static_cast<uint32_t*>(d_ws), plan->L.codes / 4, 0u, stream));
}

static fz_status check_compress_args(const fz_plan plan, const float* d_in, double eb_rel, void* d_ws,
                                     size_t ws_bytes, uint8_t* d_payload, uint16_t* d_tail, fz_result* d_result,
                                     bool outlier = false)
{
    if (!plan || !d_in || !d_ws || !d_payload || !d_result) return FZ_ERR_INVALID_ARG;
    if ((plan->mode == FZ_MODE_OUTLIER) != outlier) return FZ_ERR_INVALID_ARG;  // the mode picks the entry point
    if (!(eb_rel > 0.0) || !std::isfinite(eb_rel)) return FZ_ERR_INVALID_ARG;
    if (!aligned(d_in, 16) || !aligned(d_payload, 16) || !aligned(d_ws, 256) || !aligned(d_result, 8))
        return FZ_ERR_INVALID_ARG;
    const Geom& g = plan->g;
    if (plan->mode != FZ_MODE_LITERAL && !d_tail && g.len != (uint64_t)kTileCodes * g.T) return FZ_ERR_INVALID_ARG;
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
                           ws + L.stage, reinterpret_cast<uint32_t*>(ws + L.counts), tstate, plan->wave_tile, ev,
                           stream));
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
    if (plan->mode != FZ_MODE_LITERAL && tail_len)
        FZ_TRY(hipMemcpyAsync(d_tail, codes + (uint64_t)kTileCodes * g.T, 2 * tail_len, hipMemcpyDeviceToDevice,
                              stream));
    return FZ_SUCCESS;
}

fz_status fz_decompress_f32(const fz_plan plan, const uint8_t* d_payload, const uint16_t* d_tail, float ebx2,
                            void* d_ws, size_t ws_bytes, float* d_out, hipStream_t stream)
{
    if (!plan || !d_payload || !d_ws || !d_out || plan->mode == FZ_MODE_OUTLIER) return FZ_ERR_INVALID_ARG;
    if (!aligned(d_payload, 16) || !aligned(d_ws, 256) || !aligned(d_out, 16)) return FZ_ERR_INVALID_ARG;
    if (!(ebx2 > 0.0f) || !std::isfinite(ebx2)) return FZ_ERR_INVALID_ARG;
    if (ws_bytes < plan->L.total) return FZ_ERR_WORKSPACE_TOO_SMALL;
    const Geom&   g = plan->g;
    if (plan->mode != FZ_MODE_LITERAL && !d_tail && g.len != (uint64_t)kTileCodes * g.T) return FZ_ERR_INVALID_ARG;
    const Layout& L = plan->L;
    uint8_t*      ws = static_cast<uint8_t*>(d_ws);
    Ctr*          ctr = reinterpret_cast<Ctr*>(ws + L.ctr);
    uint64_t*     dstate = reinterpret_cast<uint64_t*>(ws + L.dstate);
    uint32_t*     texcl = reinterpret_cast<uint32_t*>(ws + L.texcl);
    uint32_t*     codes = reinterpret_cast<uint32_t*>(ws + L.codes);
    const uint16_t* tail = plan->mode != FZ_MODE_LITERAL ? d_tail : nullptr;
    const int  form = plan->inverse_form;
    const bool win2 = g.ndim == 2 && (form == FZ_INVERSE_WINDOWS || (form == FZ_INVERSE_AUTO && decode_recon_2d_pays(g)));
    const bool win3 = g.ndim == 3 && (form == FZ_INVERSE_WINDOWS || (form == FZ_INVERSE_AUTO && decode_recon_3d_pays(g)));
    FZ_TRY(prof_mark(plan, stream, true));
    if (win2 && decode_recon_2d_direct(g)) {  // small 2D fields: no tile-offset launch either
        FZ_TRY(launch_decode_recon_2d(d_payload, nullptr, nullptr, tail, g, ebx2, d_out, stream));
        FZ_TRY(prof_mark(plan, stream, false));
        return FZ_SUCCESS;
    }
    const uint32_t* csums = nullptr;
    FZ_TRY(launch_tile_offsets(reinterpret_cast<const uint32_t*>(d_payload), g.T, texcl, dstate, L.n_dstate,
                               &ctr->dscan_ticket, stream, &csums));
    FZ_TRY(prof_mark(plan, stream, false));
    if (g.ndim == 1) {  // 1D Lorenzo blocks lie inside a tile: one fused launch
        FZ_TRY(launch_decode_recon_1d(d_payload, texcl, csums, tail, g, ebx2, d_out, stream));
        FZ_TRY(prof_mark(plan, stream, false));
        return FZ_SUCCESS;
    }
    if (win2) {  // codes decoded into LDS per window: one launch
        FZ_TRY(launch_decode_recon_2d(d_payload, texcl, csums, tail, g, ebx2, d_out, stream));
        FZ_TRY(prof_mark(plan, stream, false));
        return FZ_SUCCESS;
    }
    if (win3) {  // codes decoded into LDS per window: one launch
        FZ_TRY(launch_decode_recon_3d(d_payload, texcl, csums, tail, g, ebx2, d_out, stream));
        FZ_TRY(prof_mark(plan, stream, false));
        return FZ_SUCCESS;
    }
    FZ_TRY(launch_decode_unshuffle(d_payload, texcl, csums, tail, g, codes, stream));
    FZ_TRY(prof_mark(plan, stream, false));
    FZ_TRY(launch_reconstruct(reinterpret_cast<const uint16_t*>(codes), d_out, g, ebx2, stream));
    FZ_TRY(prof_mark(plan, stream, false));
    return FZ_SUCCESS;
}

fz_status fz_compress_outlier_f32(const fz_plan plan, const float* d_in, double eb_rel, void* d_ws,
                                  size_t ws_bytes, uint8_t* d_payload, uint16_t* d_tail, uint8_t* d_side,
                                  uint64_t side_capacity, fz_result* d_result, hipStream_t stream)
{
    fz_status st = check_compress_args(plan, d_in, eb_rel, d_ws, ws_bytes, d_payload, d_tail, d_result, true);
    if (st != FZ_SUCCESS) return st;
    if (!d_side || !aligned(d_side, 16) || side_capacity < 16) return FZ_ERR_INVALID_ARG;
    const Geom&   g = plan->g;
    const Layout& L = plan->L;
    uint8_t*      ws = static_cast<uint8_t*>(d_ws);
    FZ_TRY(prof_mark(plan, stream, true));
    FZ_TRY(launch_compress_outlier(d_in, g, eb_rel, reinterpret_cast<float4*>(ws + L.part), d_result, d_payload,
                                   d_tail, d_side, side_capacity, ws + L.stage,
                                   reinterpret_cast<uint32_t*>(ws + L.counts), reinterpret_cast<uint32_t*>(ws + L.tstate),
                                   L.n_tstate, reinterpret_cast<uint32_t*>(ws + L.ocnt),
                                   reinterpret_cast<uint32_t*>(ws + L.orun), stream));
    FZ_TRY(prof_mark(plan, stream, false));
    return FZ_SUCCESS;
}

fz_status fz_decompress_outlier_f32(const fz_plan plan, const uint8_t* d_payload, const uint16_t* d_tail, float ebx2,
                                    float vconst, const uint8_t* d_side, uint64_t side_bytes, void* d_ws,
                                    size_t ws_bytes, float* d_out, hipStream_t stream)
{
    if (!plan || plan->mode != FZ_MODE_OUTLIER || !d_payload || !d_side || !d_ws || !d_out) return FZ_ERR_INVALID_ARG;
    if (!aligned(d_payload, 16) || !aligned(d_side, 16) || !aligned(d_ws, 256) || !aligned(d_out, 16))
        return FZ_ERR_INVALID_ARG;
    if (!(ebx2 >= 0.0f) || !std::isfinite(ebx2) || !std::isfinite(vconst) || side_bytes < 16)
        return FZ_ERR_INVALID_ARG;
    if (ws_bytes < plan->L.total) return FZ_ERR_WORKSPACE_TOO_SMALL;
    const Geom&   g = plan->g;
    if (ebx2 != 0.0f && !d_tail && g.len != (uint64_t)kTileCodes * g.T) return FZ_ERR_INVALID_ARG;
    const Layout& L = plan->L;
    uint8_t*      ws = static_cast<uint8_t*>(d_ws);
    FZ_TRY(prof_mark(plan, stream, true));
    if (ebx2 == 0.0f) {  // constant field
        FZ_TRY(launch_fill_outlier(d_side, side_bytes, reinterpret_cast<uint32_t*>(ws + L.obits), d_out, g, vconst,
                                   stream));
        FZ_TRY(prof_mark(plan, stream, false));
        return FZ_SUCCESS;
    }
    Ctr*      ctr = reinterpret_cast<Ctr*>(ws + L.ctr);
    uint64_t* dstate = reinterpret_cast<uint64_t*>(ws + L.dstate);
    uint32_t* texcl = reinterpret_cast<uint32_t*>(ws + L.texcl);
    uint32_t* codes = reinterpret_cast<uint32_t*>(ws + L.codes);
    uint32_t* obits = reinterpret_cast<uint32_t*>(ws + L.obits);
    int32_t*  over = reinterpret_cast<int32_t*>(ws + L.over);
    FZ_TRY(launch_put_deltas(d_side, side_bytes, obits, over, g, dstate, L.n_dstate, &ctr->dscan_ticket, stream));
    FZ_TRY(launch_tile_scan(reinterpret_cast<const uint32_t*>(d_payload), g.T, texcl, dstate, L.n_dstate,
                            &ctr->dscan_ticket, stream, false));
    FZ_TRY(launch_decode_unshuffle(d_payload, texcl, nullptr, d_tail, g, codes, stream));
    FZ_TRY(launch_reconstruct_outlier(reinterpret_cast<const uint16_t*>(codes), obits, over, d_side, side_bytes, d_out,
                                      g, ebx2, stream));
    FZ_TRY(prof_mark(plan, stream, false));
    return FZ_SUCCESS;
}

// Host checks of an archive's payload and side channel: the oracle's corrupt-input rules
// (fz_oracle.c decode_codes / fzo_decompress_outlier), run before anything is copied to the
// device.
fz_status fz_payload_check(const fz_plan plan, const uint8_t* h_payload, uint64_t payload_bytes, uint64_t tail_len)
{
    if (!plan || (!h_payload && payload_bytes)) return FZ_ERR_INVALID_ARG;
    const Geom&    g = plan->g;
    const uint64_t W = g.W, T = g.T;
    if (payload_bytes < 4 * W) return FZ_ERR_CORRUPT;
    if (payload_bytes > 4 * W + 4096 * T) return FZ_ERR_CORRUPT;
    if (plan->mode != FZ_MODE_LITERAL && tail_len != (uint64_t)g.len - (uint64_t)kTileCodes * T) return FZ_ERR_CORRUPT;
    uint64_t K = 0;
    for (uint64_t i = 0; i < W; i++) {
        uint32_t v;
        std::memcpy(&v, h_payload + 4 * i, 4);  // little-endian host
        if (i >= 8 * T && v) return FZ_ERR_CORRUPT;  // flags of the partial tile are never set
        K += (uint64_t)__builtin_popcount(v);
    }
    return payload_bytes == 4 * W + 16 * K ? FZ_SUCCESS : FZ_ERR_CORRUPT;
}

fz_status fz_side_check(const fz_plan plan, const uint8_t* h_side, uint64_t side_bytes)
{
    if (!plan || plan->mode != FZ_MODE_OUTLIER || (!h_side && side_bytes)) return FZ_ERR_INVALID_ARG;
    if (side_bytes < 16) return FZ_ERR_CORRUPT;
    uint64_t n[2];
    std::memcpy(n, h_side, 16);
    const uint64_t xyz = plan->g.xyz;
    if (n[0] > xyz || n[1] > xyz || side_bytes != 16 + 8 * (n[0] + n[1])) return FZ_ERR_CORRUPT;
    const uint8_t* idx_d = h_side + 16;
    const uint8_t* idx_r = h_side + 16 + 8 * n[0];
    for (uint64_t k = 0; k < n[0] + n[1]; k++) {
        uint32_t v;
        std::memcpy(&v, (k < n[0] ? idx_d + 4 * k : idx_r + 4 * (k - n[0])), 4);
        if (v >= xyz) return FZ_ERR_CORRUPT;
    }
    return FZ_SUCCESS;
}

fz_status fz_verify_f32(void* d_temp, size_t* temp_bytes, fz_stats* d_stats, const float* d_xdata,
                        const float* d_odata, uint64_t len, hipStream_t stream)
{
    if (!temp_bytes || len == 0) return FZ_ERR_INVALID_ARG;
    const size_t need = verify_workspace_bytes(len);
    if (!d_temp) {
        *temp_bytes = need;
        return FZ_SUCCESS;
    }
    if (!d_stats || !d_xdata || !d_odata || !aligned(d_temp, 256) || !aligned(d_stats, 8) ||
        !aligned(d_xdata, 4) || !aligned(d_odata, 4))
        return FZ_ERR_INVALID_ARG;
    if (*temp_bytes < need) return FZ_ERR_WORKSPACE_TOO_SMALL;
    FZ_TRY(launch_verify(d_xdata, d_odata, len, d_temp, d_stats, stream));
    return FZ_SUCCESS;
}

fz_status fz_decompress_verify_f32(const fz_plan plan, const uint8_t* d_payload, const uint16_t* d_tail,
                                   float ebx2, void* d_ws, size_t ws_bytes, float* d_out, const float* d_odata,
                                   fz_stats* d_stats, hipStream_t stream)
{
    if (!plan || !d_odata || !d_stats || !aligned(d_odata, 4) || !aligned(d_stats, 8)) return FZ_ERR_INVALID_ARG;
    const fz_status st = fz_decompress_f32(plan, d_payload, d_tail, ebx2, d_ws, ws_bytes, d_out, stream);
    if (st != FZ_SUCCESS) return st;
    uint8_t* ws = static_cast<uint8_t*>(d_ws);
    FZ_TRY(launch_verify(d_out, d_odata, plan->g.xyz, ws + plan->L.vtemp, d_stats, stream));
    FZ_TRY(prof_mark(plan, stream, false));
    return FZ_SUCCESS;
}

}  // extern "C"

// ---------------------------------------------------------------- compress without a workspace
//
// fz_compress_f32_noalloc runs fz_compress_f32's kernels on a workspace that is a file-scope
// device array instead of a caller's buffer, so the call allocates nothing.  The array is one
// per device, allocated when the library's code object is loaded on that device and held until
// the process ends; its device address is looked up once per device (hipGetSymbolAddress, with
// that device current) and kept, since host code that names the array gets a host-side stand-in,
// not a device address.  The fused compress needs nothing zeroed on entry: its range kernel
// clears the per-run block sums the tile kernel adds to, and every other region is written
// before it is read.  Calls that overlap on one device share the array and race.
namespace {

constexpr uint32_t kNoallocMaxTiles = FZ_NOALLOC_MAX_TILES;

struct ScratchLayout {
    size_t part, tstate, counts, stage, total;
};

ScratchLayout scratch_layout(uint32_t T)
{
    ScratchLayout S{};
    size_t o = 0;
    S.part = o;   o = al256(o + 16ull * kRangeMaxWG);
    S.tstate = o; o = al256(o + 8ull * (gather_grid(T) + 2));
    S.counts = o; o = al256(o + 4ull * (T + 1));
    S.stage = o;  o = al256(o + 4096ull * T + 64);
    S.total = o;
    return S;
}

// room for the largest field the entry point accepts (gather_grid(T) <= 256)
constexpr size_t kScratchBytes = 256 + 16ull * kRangeMaxWG + 8ull * 258 + 256 + 4ull * (kNoallocMaxTiles + 1) + 256 +
                                 4096ull * kNoallocMaxTiles + 64 + 256;

}  // namespace

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex    g_scratch_mu;
void*         g_scratch_dev[kMaxDevices] = {};

// device address of the current device's copy of `scratch`
hipError_t scratch_address(void** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return hipSuccess;
}

}  // namespace

extern "C" {

uint64_t fz_noalloc_scratch_bytes(void) { return kScratchBytes; }

fz_status fz_compress_f32_noalloc(const fz_plan plan, const float* d_in, double eb_rel, uint8_t* d_payload,
                                  uint16_t* d_tail, fz_result* d_result, hipStream_t stream)
{
    if (!plan || !d_in || !d_payload || !d_result || plan->mode == FZ_MODE_OUTLIER) return FZ_ERR_INVALID_ARG;
    if (!(eb_rel > 0.0) || !std::isfinite(eb_rel)) return FZ_ERR_INVALID_ARG;
    if (!aligned(d_in, 16) || !aligned(d_payload, 16) || !aligned(d_result, 8)) return FZ_ERR_INVALID_ARG;
    const Geom& g = plan->g;
    if (plan->mode != FZ_MODE_LITERAL && !d_tail && g.len != (uint64_t)kTileCodes * g.T) return FZ_ERR_INVALID_ARG;
    if (g.T > kNoallocMaxTiles) return FZ_ERR_CAPACITY;
    const ScratchLayout S = scratch_layout(g.T);
    if (S.total > kScratchBytes) return FZ_ERR_CAPACITY;
    void* base = nullptr;
    FZ_TRY(scratch_address(&base));
    uint8_t*    ws = static_cast<uint8_t*>(base);
    hipEvent_t* ev = nullptr;
    if (plan->prof) {
        ev = plan->ev;
        plan->nev = g.T == 0 ? 4 : 6;
        plan->paired = 1;
    }
    FZ_TRY(launch_compress(d_in, g, eb_rel, reinterpret_cast<float4*>(ws + S.part), d_result, d_payload, d_tail,
                           ws + S.stage, reinterpret_cast<uint32_t*>(ws + S.counts),
                           reinterpret_cast<uint64_t*>(ws + S.tstate), plan->wave_tile, ev, stream));
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

// ============================================================================ batched form

#include <algorithm>
#include <vector>

constexpr int kBatchSlots = 7;  // event pairs: compress kernels (decompress uses the first 5)

struct fz_batch_plan_s {
    uint32_t              nf = 0;
    int                   mode = 0;
    std::vector<BField>   fields;   // sorted by dimensionality (stable)
    std::vector<uint32_t> ustart, rstart, pstart[kNumChunks];
    uint32_t              P[kNumChunks] = {};
    uint64_t              total_len = 0;
    BatchShape            shape{};
    fz_batch_tuning       tune{};
    struct {
        size_t fields, ustart, ufield, pstart[kNumChunks], rstart, rfield, wfield, tables_end, ptrs, prm, part, counts, dsts, texcl,
            staging, total;
    } L{};
    std::vector<BPtrs>      host_ptrs;  // staging of the call's pointer records (sorted order)
    BatchGroup              all{};      // every field as one group
    std::vector<BatchGroup> groups;     // the tuned grouping
    BatchStreams            st{};       // created at the first grouped call
    bool                    st_ready = false;
    int                prof = 0;
    int                nslots = 0;  // event pairs recorded by the last call
    hipEvent_t         ev[2 * kBatchSlots] = {};
    // OUTLIER mode: the single-field outlier launches field after field in one shared
    // workspace, laid out per field (caller order) as a single-field plan's
    std::vector<Geom>   ogeo;
    std::vector<Layout> olay;
    size_t              o_prefix = 0;  // largest zero-filled prefix over the fields
};

struct fz_graph_s {
    hipGraph_t     graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

bool pow2_upto16(uint32_t v) { return v == 1 || v == 2 || v == 4 || v == 8 || v == 16; }

bool tuning_ok(const fz_batch_tuning& t)
{
    bool chunk = false;
    for (int c = 0; c < kNumChunks; c++) chunk |= t.range_chunk == kRangeChunks[c];
    return chunk && pow2_upto16(t.gather_waves) && pow2_upto16(t.groups) && t.xcd_order[0] <= 1 &&
           t.xcd_order[1] <= 1 && t.xcd_order[2] <= 1 && t.wave_tile[0] <= 1 && t.wave_tile[1] <= 1 &&
           t.wave_tile[2] <= 1;
}

// fields [f_lo, f_hi) of the sorted order as one group
BatchGroup make_group(const fz_batch_plan p, uint32_t f_lo, uint32_t f_hi)
{
    BatchGroup G{};
    G.f_lo = f_lo;
    G.f_hi = f_hi;
    G.u_lo = p->ustart[f_lo];
    G.u_hi = p->ustart[f_hi];
    for (int c = 0; c < kNumChunks; c++) {
        G.p_lo[c] = p->pstart[c][f_lo];
        G.p_hi[c] = p->pstart[c][f_hi];
    }
    for (int nd = 1; nd <= 3; nd++) {
        const uint32_t lo = std::max(f_lo, p->shape.f_lo[nd]), hi = std::min(f_hi, p->shape.f_hi[nd]);
        if (lo < hi) {
            G.df_lo[nd] = lo;
            G.df_hi[nd] = hi;
            G.du_lo[nd] = p->ustart[lo];
            G.du_hi[nd] = p->ustart[hi];
        }
    }
    return G;
}

// up to n groups of consecutive fields with about equal units
void build_groups(fz_batch_plan p, uint32_t n)
{
    p->groups.clear();
    const uint64_t U = p->shape.U;
    uint32_t       f = 0;
    for (uint32_t g = 0; g < n && f < p->nf; g++) {
        const uint64_t target = U * (g + 1) / n;
        uint32_t       e = f + 1;
        while (e < p->nf && (g == n - 1 || p->ustart[e] < target)) e++;
        p->groups.push_back(make_group(p, f, e));
        f = e;
    }
}

hipError_t ensure_streams(fz_batch_plan p)
{
    if (p->st_ready) return hipSuccess;
    BatchStreams& st = p->st;
    hipError_t    e = hipStreamCreateWithFlags(&st.range, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st.gather, hipStreamNonBlocking);
    hipEvent_t* evs[3] = {&st.fork, &st.join_r, &st.join_g};
    for (int i = 0; i < 3 && e == hipSuccess; i++) e = hipEventCreateWithFlags(evs[i], hipEventDisableTiming);
    for (int i = 0; i < kMaxGroups && e == hipSuccess; i++) {
        e = hipEventCreateWithFlags(&st.r[i], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&st.t[i], hipEventDisableTiming);
    }
    p->st_ready = e == hipSuccess;
    return e;
}

void destroy_streams(fz_batch_plan p)
{
    BatchStreams& st = p->st;
    if (st.range) (void)hipStreamDestroy(st.range);
    if (st.gather) (void)hipStreamDestroy(st.gather);
    hipEvent_t evs[3] = {st.fork, st.join_r, st.join_g};
    for (hipEvent_t e : evs)
        if (e) (void)hipEventDestroy(e);
    for (int i = 0; i < kMaxGroups; i++) {
        if (st.r[i]) (void)hipEventDestroy(st.r[i]);
        if (st.t[i]) (void)hipEventDestroy(st.t[i]);
    }
    st = BatchStreams{};
    p->st_ready = false;
}

void apply_tuning(fz_batch_plan p)
{
    int c = 0;
    while (kRangeChunks[c] != p->tune.range_chunk) c++;
    p->shape.chunk = kRangeChunks[c];
    p->shape.P = p->P[c];
    for (int nd = 1; nd <= 3; nd++) {
        p->shape.xcd[nd] = p->tune.xcd_order[nd - 1];
        p->shape.wave_tile[nd] = p->tune.wave_tile[nd - 1];
    }
    p->shape.gather_waves = p->tune.gather_waves;
    build_groups(p, p->tune.groups);
}

int chunk_index(const fz_batch_plan p)
{
    int c = 0;
    while (kRangeChunks[c] != p->tune.range_chunk) c++;
    return c;
}

BatchDev batch_dev(const fz_batch_plan p, void* d_ws, fz_result* res)
{
    uint8_t* ws = static_cast<uint8_t*>(d_ws);
    BatchDev d{};
    d.fields = reinterpret_cast<const BField*>(ws + p->L.fields);
    d.ustart = reinterpret_cast<const uint32_t*>(ws + p->L.ustart);
    d.ufield = reinterpret_cast<const uint32_t*>(ws + p->L.ufield);
    d.pstart = reinterpret_cast<const uint32_t*>(ws + p->L.pstart[chunk_index(p)]);
    d.rstart = reinterpret_cast<const uint32_t*>(ws + p->L.rstart);
    d.rfield = reinterpret_cast<const uint32_t*>(ws + p->L.rfield);
    d.wfield = reinterpret_cast<const uint32_t*>(ws + p->L.wfield);
    d.ptrs = reinterpret_cast<BPtrs*>(ws + p->L.ptrs);
    d.prm = reinterpret_cast<BParams*>(ws + p->L.prm);
    d.part = reinterpret_cast<float4*>(ws + p->L.part);
    d.counts = reinterpret_cast<uint32_t*>(ws + p->L.counts);
    d.dsts = reinterpret_cast<uint64_t*>(ws + p->L.dsts);
    d.texcl = reinterpret_cast<uint32_t*>(ws + p->L.texcl);
    d.staging = ws + p->L.staging;
    d.res = res;
    return d;
}

inline uint64_t tail_codes(const Geom& g) { return (uint64_t)g.len - (uint64_t)kTileCodes * g.T; }

fz_status batch_compress_impl(const fz_batch_plan plan, const float* const* d_in, const double* eb_rel, void* d_ws,
                              size_t ws_bytes, uint8_t* const* d_payload, uint16_t* const* d_tail,
                              fz_result* d_results, hipStream_t stream, bool timed)
{
    if (!plan || !d_in || !eb_rel || !d_ws || !d_payload || !d_results) return FZ_ERR_INVALID_ARG;
    if (plan->mode == FZ_MODE_OUTLIER) return FZ_ERR_INVALID_ARG;  // fz_batch_compress_outlier_f32
    if (!aligned(d_ws, 256) || !aligned(d_results, 8)) return FZ_ERR_INVALID_ARG;
    if (ws_bytes < plan->L.total) return FZ_ERR_WORKSPACE_TOO_SMALL;
    for (uint32_t k = 0; k < plan->nf; k++) {
        const BField&  F = plan->fields[k];
        const uint32_t c = F.caller;
        const bool     want_tail = plan->mode == FZ_MODE_FIXED && tail_codes(F.g) > 0;
        if (!d_in[c] || !aligned(d_in[c], 16) || !d_payload[c] || !aligned(d_payload[c], 16)) return FZ_ERR_INVALID_ARG;
        if (!(eb_rel[c] > 0.0) || !std::isfinite(eb_rel[c])) return FZ_ERR_INVALID_ARG;
        if (want_tail && (!d_tail || !d_tail[c] || !aligned(d_tail[c], 2))) return FZ_ERR_INVALID_ARG;
        plan->host_ptrs[k] = BPtrs{d_in[c], d_payload[c], want_tail ? d_tail[c] : nullptr, eb_rel[c]};
    }
    const BatchDev d = batch_dev(plan, d_ws, d_results);
    FZ_TRY(launch_batch_set(plan->host_ptrs.data(), plan->nf, d.ptrs, stream));
    const int c = chunk_index(plan);
    if (timed && plan->prof) {  // one group on the caller's stream, an event pair per kernel
        plan->nslots = kBatchSlots;
        FZ_TRY(launch_batch_compress(d, plan->shape, &plan->all, 1, c, nullptr, plan->ev, stream));
        return FZ_SUCCESS;
    }
    if (plan->groups.size() <= 1) {
        FZ_TRY(launch_batch_compress(d, plan->shape, &plan->all, 1, c, nullptr, nullptr, stream));
        return FZ_SUCCESS;
    }
    FZ_TRY(ensure_streams(plan));
    FZ_TRY(launch_batch_compress(d, plan->shape, plan->groups.data(), (int)plan->groups.size(), c, &plan->st, nullptr,
                                 stream));
    return FZ_SUCCESS;
}

fz_status batch_decompress_impl(const fz_batch_plan plan, const uint8_t* const* d_payload,
                                const uint16_t* const* d_tail, const float* ebx2, void* d_ws, size_t ws_bytes,
                                float* const* d_out, hipStream_t stream, bool timed)
{
    if (!plan || !d_payload || !ebx2 || !d_ws || !d_out) return FZ_ERR_INVALID_ARG;
    if (plan->mode == FZ_MODE_OUTLIER) return FZ_ERR_INVALID_ARG;  // fz_batch_decompress_outlier_f32
    if (!aligned(d_ws, 256)) return FZ_ERR_INVALID_ARG;
    if (ws_bytes < plan->L.total) return FZ_ERR_WORKSPACE_TOO_SMALL;
    for (uint32_t k = 0; k < plan->nf; k++) {
        const BField&  F = plan->fields[k];
        const uint32_t c = F.caller;
        if (!d_payload[c] || !aligned(d_payload[c], 16) || !d_out[c] || !aligned(d_out[c], 16))
            return FZ_ERR_INVALID_ARG;
        if (!(ebx2[c] > 0.0f) || !std::isfinite(ebx2[c])) return FZ_ERR_INVALID_ARG;
        const uint16_t* t = (plan->mode == FZ_MODE_FIXED && d_tail) ? d_tail[c] : nullptr;
        if (t && !aligned(t, 2)) return FZ_ERR_INVALID_ARG;
        plan->host_ptrs[k] = BPtrs{d_payload[c], d_out[c], const_cast<uint16_t*>(t), (double)ebx2[c]};
    }
    const BatchDev d = batch_dev(plan, d_ws, nullptr);
    FZ_TRY(launch_batch_set(plan->host_ptrs.data(), plan->nf, d.ptrs, stream));
    hipEvent_t* ev = nullptr;
    if (timed && plan->prof) {
        ev = plan->ev;
        plan->nslots = 6;
    }
    FZ_TRY(launch_batch_decompress(d, plan->shape, ev, stream));
    return FZ_SUCCESS;
}

// Captures body(stream) on a private stream into an instantiated graph.
template <typename F>
fz_status capture_graph(F body, fz_graph* out)
{
    if (!out) return FZ_ERR_INVALID_ARG;
    *out = nullptr;
    hipStream_t s = nullptr;
    FZ_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(s);
        return hip_st(e);
    }
    const fz_status st = body(s);
    hipGraph_t      g = nullptr;
    e = hipStreamEndCapture(s, &g);
    (void)hipStreamDestroy(s);
    if (st != FZ_SUCCESS || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return st != FZ_SUCCESS ? st : hip_st(e);
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return hip_st(e);
    }
    fz_graph gr = new (std::nothrow) fz_graph_s;
    if (!gr) {
        (void)hipGraphExecDestroy(x);
        (void)hipGraphDestroy(g);
        return FZ_ERR_CAPACITY;
    }
    gr->graph = g;
    gr->exec = x;
    *out = gr;
    return FZ_SUCCESS;
}

}  // namespace

extern "C" {

fz_status fz_batch_plan_create(fz_batch_plan* plan, const fz_field_shape* shapes, uint32_t nfields, int mode)
{
    if (!plan) return FZ_ERR_INVALID_ARG;
    *plan = nullptr;
    if (!shapes || nfields == 0 || nfields > kBatchMaxFields) return FZ_ERR_INVALID_ARG;
    if (mode != FZ_MODE_FIXED && mode != FZ_MODE_LITERAL && mode != FZ_MODE_OUTLIER) return FZ_ERR_INVALID_ARG;
    std::vector<Geom> geo(nfields);
    for (uint32_t i = 0; i < nfields; i++) {
        const fz_field_shape& s = shapes[i];
        const uint64_t        len = s.len ? s.len : (uint64_t)s.x * s.y * s.z;
        if (!geom_ok(s.x, s.y, s.z, len)) return FZ_ERR_INVALID_ARG;
        geo[i] = make_geom(s.x, s.y, s.z, len, mode);
    }
    fz_batch_plan p = new (std::nothrow) fz_batch_plan_s;
    if (!p) return FZ_ERR_CAPACITY;
    p->nf = nfields;
    p->mode = mode;
    std::vector<uint32_t> order(nfields);
    for (uint32_t i = 0; i < nfields; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return geo[a].ndim < geo[b].ndim; });
    p->fields.resize(nfields);
    p->ustart.resize(nfields + 1);
    p->rstart.resize(nfields + 1);
    for (int c = 0; c < kNumChunks; c++) p->pstart[c].resize(nfields + 1);
    uint64_t u = 0, r = 0, parts[kNumChunks] = {};
    std::vector<std::pair<uint32_t, uint64_t>> fused[kWinClasses];  // (field, windows) per window size
    std::vector<std::pair<uint32_t, uint64_t>> fused2d;  // (field, windows): fused 2D
    BatchShape& b = p->shape;
    b.nf = nfields;
    for (uint32_t k = 0; k < nfields; k++) {
        const Geom& g = geo[order[k]];
        BField      F{};
        F.g = g;
        F.caller = order[k];
        F.ustart = (uint32_t)u;
        if (g.ndim == 1) {
            F.rdims[0] = (g.xyz + 4095) / 4096;
            F.rdims[1] = F.rdims[2] = 1;
        } else if (g.ndim == 2) {
            F.rdims[0] = (g.X + kRecon2DCols - 1) / kRecon2DCols;
            F.rdims[1] = (g.Y + 15) / 16;
            F.rdims[2] = 1;
            uint32_t rx, ry;
            if (mode != FZ_MODE_OUTLIER && batch_fused2d(g, rx, ry)) {
                // decoded and reconstructed by the fused 2D window kernel: no reconstruct workgroups
                F.wnt = 1024;
                F.wdims[0] = rx;
                F.wdims[1] = ry;
                F.rdims[0] = F.rdims[1] = F.rdims[2] = 0;
                fused2d.push_back({k, (uint64_t)rx * ry});
            }
        } else {
            F.rq = recon3d_q(g.X, g.Y);
            F.rdims[0] = (g.X + 4 * F.rq - 1) / (4 * F.rq);
            F.rdims[1] = (g.Y + kWG / F.rq - 1) / (kWG / F.rq);
            F.rdims[2] = (g.Z + 7) / 8;
            uint32_t wq, wr, rx, ry;
            uint64_t nwin;
            if (mode != FZ_MODE_OUTLIER && batch_fused3d(g, wq, wr, rx, ry, nwin)) {
                // decoded and reconstructed by the fused window kernel: no reconstruct workgroups
                F.wnt = wq * wr;  // a multiple of 64, at most 1024
                F.wq = wq;
                F.wr = wr;
                F.wm = d3_magic(wq);
                F.wdims[0] = rx;
                F.wdims[1] = ry;
                F.rdims[0] = F.rdims[1] = F.rdims[2] = 0;
                fused[F.wnt / 64 - 1].push_back({k, nwin});
            }
        }
        const int nd = g.ndim;
        if (k == 0 || geo[order[k - 1]].ndim != nd) {
            b.f_lo[nd] = k;
            b.u_lo[nd] = (uint32_t)u;
            b.r_lo[nd] = (uint32_t)r;
        }
        p->ustart[k] = (uint32_t)u;
        p->rstart[k] = (uint32_t)r;
        for (int c = 0; c < kNumChunks; c++) {
            p->pstart[c][k] = (uint32_t)parts[c];
            parts[c] += ((uint64_t)g.len + kRangeChunks[c] - 1) / kRangeChunks[c];  // len + chunk passes 2^32
        }
        u += (uint64_t)g.T + 1;
        r += (uint64_t)F.rdims[0] * F.rdims[1] * F.rdims[2];
        p->total_len += g.len;
        b.f_hi[nd] = k + 1;
        b.u_hi[nd] = (uint32_t)u;
        b.r_hi[nd] = (uint32_t)r;
        p->fields[k] = F;
        if (u > 0xFFFFFFF0ull || r > 0xFFFFFFF0ull || parts[0] > 0x7FFFFFFFull) {
            delete p;
            return FZ_ERR_CAPACITY;
        }
    }
    p->ustart[nfields] = (uint32_t)u;
    p->rstart[nfields] = (uint32_t)r;
    uint64_t w = 0;  // the fused 3D windows, by window size, then the fused 2D ones
    for (int c = 0; c < kWinClasses; c++) {
        b.w_lo[c] = (uint32_t)w;
        for (const auto& fw : fused[c]) {
            p->fields[fw.first].wstart = (uint32_t)w;
            w += fw.second;
        }
        b.w_hi[c] = (uint32_t)w;
    }
    b.w2_lo = (uint32_t)w;  // the fused 2D windows after them
    for (const auto& fw : fused2d) {
        p->fields[fw.first].wstart = (uint32_t)w;
        w += fw.second;
    }
    b.w2_hi = (uint32_t)w;
    if (w > 0xFFFFFFF0ull) {
        delete p;
        return FZ_ERR_CAPACITY;
    }
    for (int c = 0; c < kNumChunks; c++) {
        p->pstart[c][nfields] = (uint32_t)parts[c];
        p->P[c] = (uint32_t)parts[c];
    }
    b.U = (uint32_t)u;
    // workspace: tables (uploaded by fz_batch_workspace_init), then per-call regions
    auto& L = p->L;
    size_t o = 0;
    L.fields = o; o = al256(o + sizeof(BField) * nfields);
    L.ustart = o; o = al256(o + 4ull * (nfields + 1));
    L.ufield = o; o = al256(o + 4ull * u);
    for (int c = 0; c < kNumChunks; c++) {
        L.pstart[c] = o;
        o = al256(o + 4ull * (nfields + 1));
    }
    L.rstart = o; o = al256(o + 4ull * (nfields + 1));
    L.rfield = o; o = al256(o + 4ull * r);
    L.wfield = o; o = al256(o + 4ull * w);
    L.tables_end = o;
    L.ptrs = o;    o = al256(o + sizeof(BPtrs) * nfields);
    L.prm = o;     o = al256(o + sizeof(BParams) * nfields);
    L.part = o;    o = al256(o + 16ull * p->P[0]);
    L.counts = o;  o = al256(o + 4ull * u);
    L.dsts = o;    o = al256(o + 8ull * u);
    L.texcl = o;   o = al256(o + 4ull * u);
    L.staging = o; o = al256(o + 4096ull * u);
    L.total = o;
    if (mode == FZ_MODE_OUTLIER) {  // one single-field workspace, sized for the largest field
        p->ogeo = geo;
        p->olay.resize(nfields);
        size_t big = 0;
        for (uint32_t i = 0; i < nfields; i++) {
            p->olay[i] = make_layout(geo[i], mode);
            big = std::max(big, p->olay[i].total);
            p->o_prefix = std::max(p->o_prefix, p->olay[i].codes);
        }
        L.total = big;
    }
    p->host_ptrs.resize(nfields);
    p->tune.range_chunk = 65536;
    p->tune.xcd_order[0] = p->tune.xcd_order[1] = p->tune.xcd_order[2] = 1;
    p->tune.gather_waves = 16;
    p->tune.groups = 1;
    p->all = make_group(p, 0, nfields);
    apply_tuning(p);
    *plan = p;
    return FZ_SUCCESS;
}

fz_status fz_batch_plan_destroy(fz_batch_plan plan)
{
    if (!plan) return FZ_SUCCESS;
    for (int i = 0; i < 2 * kBatchSlots; i++)
        if (plan->ev[i]) (void)hipEventDestroy(plan->ev[i]);
    destroy_streams(plan);
    delete plan;
    return FZ_SUCCESS;
}

fz_status fz_batch_plan_sizes(const fz_batch_plan plan, fz_batch_sizes* s)
{
    if (!plan || !s) return FZ_ERR_INVALID_ARG;
    s->workspace_bytes = plan->L.total;
    s->total_len = plan->total_len;
    s->units = plan->shape.U;
    s->nfields = plan->nf;
    s->range_parts = plan->shape.P;
    return FZ_SUCCESS;
}

fz_status fz_batch_field_sizes(const fz_batch_plan plan, uint32_t field, fz_sizes* s)
{
    if (!plan || !s || field >= plan->nf) return FZ_ERR_INVALID_ARG;
    for (const BField& F : plan->fields) {
        if (F.caller != field) continue;
        const Geom& g = F.g;
        s->len = g.len;
        s->tiles = g.T;
        s->flag_words = g.W;
        s->byte_flags = 32ull * g.W;
        s->payload_max_bytes = 4ull * g.W + 4096ull * g.T;
        s->tail_max_codes = tail_codes(g);
        s->workspace_bytes = plan->L.total;
        s->side_max_bytes = plan->mode == FZ_MODE_OUTLIER ? 16ull + 8ull * g.xyz : 0;
        return FZ_SUCCESS;
    }
    return FZ_ERR_INVALID_ARG;
}

fz_status fz_batch_get_tuning(const fz_batch_plan plan, fz_batch_tuning* t)
{
    if (!plan || !t) return FZ_ERR_INVALID_ARG;
    *t = plan->tune;
    return FZ_SUCCESS;
}

fz_status fz_batch_set_tuning(fz_batch_plan plan, const fz_batch_tuning* t)
{
    if (!plan || !t || !tuning_ok(*t)) return FZ_ERR_INVALID_ARG;
    plan->tune = *t;
    apply_tuning(plan);
    return FZ_SUCCESS;
}

fz_status fz_batch_workspace_init(const fz_batch_plan plan, void* d_ws, size_t ws_bytes, hipStream_t stream)
{
    if (!plan || !d_ws || !aligned(d_ws, 256)) return FZ_ERR_INVALID_ARG;
    if (ws_bytes < plan->L.total) return FZ_ERR_WORKSPACE_TOO_SMALL;
    if (plan->mode == FZ_MODE_OUTLIER)  // every field's zero-filled prefix
        return hip_st(launch_fill_u32(static_cast<uint32_t*>(d_ws), plan->o_prefix / 4, 0u, stream));
    const auto&          L = plan->L;
    std::vector<uint8_t> t(L.tables_end, 0);
    std::memcpy(t.data() + L.fields, plan->fields.data(), sizeof(BField) * plan->nf);
    std::memcpy(t.data() + L.ustart, plan->ustart.data(), 4ull * (plan->nf + 1));
    {
        uint32_t* uf = reinterpret_cast<uint32_t*>(t.data() + L.ufield);
        for (uint32_t f = 0; f < plan->nf; f++)
            for (uint32_t u = plan->ustart[f]; u < plan->ustart[f + 1]; u++) uf[u] = f;
    }
    for (int c = 0; c < kNumChunks; c++)
        std::memcpy(t.data() + L.pstart[c], plan->pstart[c].data(), 4ull * (plan->nf + 1));
    std::memcpy(t.data() + L.rstart, plan->rstart.data(), 4ull * (plan->nf + 1));
    {
        uint32_t* rf = reinterpret_cast<uint32_t*>(t.data() + L.rfield);
        for (uint32_t f = 0; f < plan->nf; f++)
            for (uint32_t w = plan->rstart[f]; w < plan->rstart[f + 1]; w++) rf[w] = f;
    }
    {
        uint32_t* wf = reinterpret_cast<uint32_t*>(t.data() + L.wfield);
        for (uint32_t f = 0; f < plan->nf; f++) {
            const BField& F = plan->fields[f];
            if (!F.wnt) continue;
            const uint64_t nw = (uint64_t)F.wdims[0] * F.wdims[1] * ((F.g.Z + 7) / 8);
            for (uint64_t w = 0; w < nw; w++) wf[F.wstart + w] = f;
        }
    }
    FZ_TRY(hipMemcpyAsync(d_ws, t.data(), t.size(), hipMemcpyHostToDevice, stream));
    FZ_TRY(hipStreamSynchronize(stream));
    return FZ_SUCCESS;
}

fz_status fz_batch_compress_f32(const fz_batch_plan plan, const float* const* d_in, const double* eb_rel,
                                void* d_ws, size_t ws_bytes, uint8_t* const* d_payload, uint16_t* const* d_tail,
                                fz_result* d_results, hipStream_t stream)
{
    return batch_compress_impl(plan, d_in, eb_rel, d_ws, ws_bytes, d_payload, d_tail, d_results, stream, true);
}

fz_status fz_batch_decompress_f32(const fz_batch_plan plan, const uint8_t* const* d_payload,
                                  const uint16_t* const* d_tail, const float* ebx2, void* d_ws, size_t ws_bytes,
                                  float* const* d_out, hipStream_t stream)
{
    return batch_decompress_impl(plan, d_payload, d_tail, ebx2, d_ws, ws_bytes, d_out, stream, true);
}

// OUTLIER-mode batches: the single-field outlier launches (fz_compress_outlier_f32 /
// fz_decompress_outlier_f32) field after field on the caller's stream, one shared workspace
// laid out for each field in turn.  Compress needs nothing zeroed on entry; the inverse needs
// its field's position bitmap clear, which a previous field's other regions may have written,
// so each field's bitmap is cleared first.
fz_status fz_batch_compress_outlier_f32(const fz_batch_plan plan, const float* const* d_in, const double* eb_rel,
                                        void* d_ws, size_t ws_bytes, uint8_t* const* d_payload,
                                        uint16_t* const* d_tail, uint8_t* const* d_side,
                                        const uint64_t* side_capacity, fz_result* d_results, hipStream_t stream)
{
    if (!plan || plan->mode != FZ_MODE_OUTLIER) return FZ_ERR_INVALID_ARG;
    if (!d_in || !eb_rel || !d_ws || !d_payload || !d_side || !side_capacity || !d_results) return FZ_ERR_INVALID_ARG;
    if (!aligned(d_ws, 256) || !aligned(d_results, 8)) return FZ_ERR_INVALID_ARG;
    if (ws_bytes < plan->L.total) return FZ_ERR_WORKSPACE_TOO_SMALL;
    for (uint32_t i = 0; i < plan->nf; i++) {
        const Geom& g = plan->ogeo[i];
        if (!d_in[i] || !aligned(d_in[i], 16) || !d_payload[i] || !aligned(d_payload[i], 16)) return FZ_ERR_INVALID_ARG;
        if (!(eb_rel[i] > 0.0) || !std::isfinite(eb_rel[i])) return FZ_ERR_INVALID_ARG;
        if (tail_codes(g) > 0 && (!d_tail || !d_tail[i])) return FZ_ERR_INVALID_ARG;
        if (!d_side[i] || !aligned(d_side[i], 16) || side_capacity[i] < 16) return FZ_ERR_INVALID_ARG;
    }
    uint8_t* ws = static_cast<uint8_t*>(d_ws);
    for (uint32_t i = 0; i < plan->nf; i++) {
        const Geom&   g = plan->ogeo[i];
        const Layout& L = plan->olay[i];
        FZ_TRY(launch_compress_outlier(d_in[i], g, eb_rel[i], reinterpret_cast<float4*>(ws + L.part), d_results + i,
                                       d_payload[i], d_tail ? d_tail[i] : nullptr, d_side[i], side_capacity[i],
                                       ws + L.stage, reinterpret_cast<uint32_t*>(ws + L.counts),
                                       reinterpret_cast<uint32_t*>(ws + L.tstate), L.n_tstate,
                                       reinterpret_cast<uint32_t*>(ws + L.ocnt), reinterpret_cast<uint32_t*>(ws + L.orun),
                                       stream));
    }
    return FZ_SUCCESS;
}

fz_status fz_batch_decompress_outlier_f32(const fz_batch_plan plan, const uint8_t* const* d_payload,
                                          const uint16_t* const* d_tail, const float* ebx2, const float* vconst,
                                          const uint8_t* const* d_side, const uint64_t* side_bytes, void* d_ws,
                                          size_t ws_bytes, float* const* d_out, hipStream_t stream)
{
    if (!plan || plan->mode != FZ_MODE_OUTLIER) return FZ_ERR_INVALID_ARG;
    if (!d_payload || !ebx2 || !vconst || !d_side || !side_bytes || !d_ws || !d_out) return FZ_ERR_INVALID_ARG;
    if (!aligned(d_ws, 256)) return FZ_ERR_INVALID_ARG;
    if (ws_bytes < plan->L.total) return FZ_ERR_WORKSPACE_TOO_SMALL;
    for (uint32_t i = 0; i < plan->nf; i++) {
        const Geom& g = plan->ogeo[i];
        if (!d_payload[i] || !aligned(d_payload[i], 16) || !d_out[i] || !aligned(d_out[i], 16)) return FZ_ERR_INVALID_ARG;
        if (!d_side[i] || !aligned(d_side[i], 16) || side_bytes[i] < 16) return FZ_ERR_INVALID_ARG;
        if (!(ebx2[i] >= 0.0f) || !std::isfinite(ebx2[i]) || !std::isfinite(vconst[i])) return FZ_ERR_INVALID_ARG;
        if (ebx2[i] != 0.0f && tail_codes(g) > 0 && (!d_tail || !d_tail[i])) return FZ_ERR_INVALID_ARG;
    }
    uint8_t* ws = static_cast<uint8_t*>(d_ws);
    for (uint32_t i = 0; i < plan->nf; i++) {
        const Geom&   g = plan->ogeo[i];
        const Layout& L = plan->olay[i];
        uint32_t*     obits = reinterpret_cast<uint32_t*>(ws + L.obits);
        FZ_TRY(launch_fill_u32(obits, (g.xyz + 31) / 32, 0u, stream));
        if (ebx2[i] == 0.0f) {  // constant field
            FZ_TRY(launch_fill_outlier(d_side[i], side_bytes[i], obits, d_out[i], g, vconst[i], stream));
            continue;
        }
        Ctr*      ctr = reinterpret_cast<Ctr*>(ws + L.ctr);
        uint64_t* dstate = reinterpret_cast<uint64_t*>(ws + L.dstate);
        uint32_t* texcl = reinterpret_cast<uint32_t*>(ws + L.texcl);
        uint32_t* codes = reinterpret_cast<uint32_t*>(ws + L.codes);
        int32_t*  over = reinterpret_cast<int32_t*>(ws + L.over);
        const uint16_t* tail = d_tail ? d_tail[i] : nullptr;
        FZ_TRY(launch_put_deltas(d_side[i], side_bytes[i], obits, over, g, dstate, L.n_dstate, &ctr->dscan_ticket,
                                 stream));
        FZ_TRY(launch_tile_scan(reinterpret_cast<const uint32_t*>(d_payload[i]), g.T, texcl, dstate, L.n_dstate,
                                &ctr->dscan_ticket, stream, false));
        FZ_TRY(launch_decode_unshuffle(d_payload[i], texcl, nullptr, tail, g, codes, stream));
        FZ_TRY(launch_reconstruct_outlier(reinterpret_cast<const uint16_t*>(codes), obits, over, d_side[i],
                                          side_bytes[i], d_out[i], g, ebx2[i], stream));
    }
    return FZ_SUCCESS;
}

fz_status fz_batch_profile(fz_batch_plan plan, int enable)
{
    if (!plan) return FZ_ERR_INVALID_ARG;
    if (enable && !plan->ev[0])
        for (int i = 0; i < 2 * kBatchSlots; i++) FZ_TRY(hipEventCreate(&plan->ev[i]));
    plan->prof = enable ? 1 : 0;
    plan->nslots = 0;
    return FZ_SUCCESS;
}

fz_status fz_batch_kernel_times(const fz_batch_plan plan, float* ms, int capacity, int* count)
{
    if (!plan || !count) return FZ_ERR_INVALID_ARG;
    *count = 0;
    if (!plan->prof || plan->nslots == 0) return FZ_ERR_NOT_READY;
    FZ_TRY(hipEventSynchronize(plan->ev[2 * plan->nslots - 1]));
    int n = 0;
    for (; n < plan->nslots && n < capacity; n++) FZ_TRY(hipEventElapsedTime(&ms[n], plan->ev[2 * n], plan->ev[2 * n + 1]));
    *count = n;
    return FZ_SUCCESS;
}

fz_status fz_batch_autotune(fz_batch_plan plan, const float* const* d_in, const double* eb_rel, void* d_ws,
                            size_t ws_bytes, uint8_t* const* d_payload, uint16_t* const* d_tail, fz_result* d_results,
                            int reps, fz_batch_tuning* chosen, hipStream_t stream)
{
    if (!plan || reps < 1 || plan->mode == FZ_MODE_OUTLIER) return FZ_ERR_INVALID_ARG;
    const int             prof0 = plan->prof;
    const fz_batch_tuning start = plan->tune;
    fz_status             st = fz_batch_profile(plan, 1);
    // mean per-slot ms over reps runs at the current tuning
    auto run = [&](float out[kBatchSlots]) -> fz_status {
        for (int i = 0; i < kBatchSlots; i++) out[i] = 0;
        for (int r = 0; r < reps; r++) {
            fz_status s = batch_compress_impl(plan, d_in, eb_rel, d_ws, ws_bytes, d_payload, d_tail, d_results,
                                              stream, true);
            if (s != FZ_SUCCESS) return s;
            float ms[kBatchSlots];
            int   n = 0;
            s = fz_batch_kernel_times(plan, ms, kBatchSlots, &n);
            if (s != FZ_SUCCESS) return s;
            for (int i = 0; i < n; i++) out[i] += ms[i] / reps;
        }
        return FZ_SUCCESS;
    };
    float t[kBatchSlots];
    if (st == FZ_SUCCESS) st = run(t);  // warm-up
    fz_batch_tuning best = start;
    // range chunk: range + range-final time
    float best_ms = 1e30f;
    for (int c = 0; c < kNumChunks && st == FZ_SUCCESS; c++) {
        plan->tune.range_chunk = kRangeChunks[c];
        apply_tuning(plan);
        st = run(t);
        if (st == FZ_SUCCESS && t[0] + t[1] < best_ms) {
            best_ms = t[0] + t[1];
            best.range_chunk = kRangeChunks[c];
        }
    }
    plan->tune.range_chunk = best.range_chunk;
    // tile order x tile layout per dimensionality: tile-kernel time of each (v = 2 layout + order)
    float xt[4][3] = {};
    for (uint32_t v = 0; v < 4 && st == FZ_SUCCESS; v++) {
        for (int nd = 0; nd < 3; nd++) {
            plan->tune.xcd_order[nd] = v & 1u;
            plan->tune.wave_tile[nd] = v >> 1;
        }
        apply_tuning(plan);
        st = run(t);
        for (int nd = 0; nd < 3; nd++) xt[v][nd] = t[2 + nd];
    }
    for (int nd = 0; nd < 3; nd++) {
        uint32_t bv = 1;  // XCD-sliced, one workgroup per tile on ties
        for (uint32_t v = 0; v < 4; v++)
            if (xt[v][nd] < xt[bv][nd]) bv = v;
        best.xcd_order[nd] = plan->tune.xcd_order[nd] = bv & 1u;
        best.wave_tile[nd] = plan->tune.wave_tile[nd] = bv >> 1;
    }
    // gather width
    best_ms = 1e30f;
    const uint32_t gws[3] = {4, 8, 16};
    for (int i = 0; i < 3 && st == FZ_SUCCESS; i++) {
        plan->tune.gather_waves = gws[i];
        apply_tuning(plan);
        st = run(t);
        if (st == FZ_SUCCESS && t[6] < best_ms) {
            best_ms = t[6];
            best.gather_waves = gws[i];
        }
    }
    // group count: whole-call time (profiling off, so the grouped schedule runs)
    plan->tune.gather_waves = best.gather_waves;
    (void)fz_batch_profile(plan, 0);
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (st == FZ_SUCCESS) st = hip_st(hipEventCreate(&t0));
    if (st == FZ_SUCCESS) st = hip_st(hipEventCreate(&t1));
    best_ms = 1e30f;
    const uint32_t gs[5] = {1, 2, 4, 8, 16};
    for (int i = 0; i < 5 && st == FZ_SUCCESS; i++) {
        plan->tune.groups = gs[i];
        apply_tuning(plan);
        for (int r = 0; r <= reps && st == FZ_SUCCESS; r++) {  // the first run warms the streams up
            if (r == 1) st = hip_st(hipEventRecord(t0, stream));
            if (st == FZ_SUCCESS)
                st = batch_compress_impl(plan, d_in, eb_rel, d_ws, ws_bytes, d_payload, d_tail, d_results, stream,
                                         false);
        }
        if (st == FZ_SUCCESS) st = hip_st(hipEventRecord(t1, stream));
        if (st == FZ_SUCCESS) st = hip_st(hipEventSynchronize(t1));
        float ms = 0;
        if (st == FZ_SUCCESS) st = hip_st(hipEventElapsedTime(&ms, t0, t1));
        if (st == FZ_SUCCESS && ms < best_ms * 0.98f) {  // a group count must win by 2 %
            best_ms = ms;
            best.groups = gs[i];
        }
    }
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    plan->tune = st == FZ_SUCCESS ? best : start;
    apply_tuning(plan);
    (void)fz_batch_profile(plan, prof0);
    if (st == FZ_SUCCESS && chosen) *chosen = plan->tune;
    return st;
}

fz_status fz_batch_compress_graph(const fz_batch_plan plan, const float* const* d_in, const double* eb_rel,
                                  void* d_ws, size_t ws_bytes, uint8_t* const* d_payload, uint16_t* const* d_tail,
                                  fz_result* d_results, fz_graph* graph)
{
    return capture_graph(
        [&](hipStream_t s) {
            return batch_compress_impl(plan, d_in, eb_rel, d_ws, ws_bytes, d_payload, d_tail, d_results, s, false);
        },
        graph);
}

fz_status fz_batch_decompress_graph(const fz_batch_plan plan, const uint8_t* const* d_payload,
                                    const uint16_t* const* d_tail, const float* ebx2, void* d_ws, size_t ws_bytes,
                                    float* const* d_out, fz_graph* graph)
{
    return capture_graph(
        [&](hipStream_t s) {
            return batch_decompress_impl(plan, d_payload, d_tail, ebx2, d_ws, ws_bytes, d_out, s, false);
        },
        graph);
}

fz_status fz_graph_launch(fz_graph graph, hipStream_t stream)
{
    if (!graph || !graph->exec) return FZ_ERR_INVALID_ARG;
    FZ_TRY(hipGraphLaunch(graph->exec, stream));
    return FZ_SUCCESS;
}

fz_status fz_graph_destroy(fz_graph graph)
{
    if (!graph) return FZ_SUCCESS;
    if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
    if (graph->graph) (void)hipGraphDestroy(graph->graph);
    delete graph;
    return FZ_SUCCESS;
}

}  // extern "C"
