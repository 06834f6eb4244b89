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

// partial-range workgroups: a multiple of 8 (one slice per XCD region), ~16K floats each
// partial-range workgroups (1024 threads each): a multiple of 8 (one slice per XCD region),
// ~64K floats each, at most 256 (the tile kernel reduces one partial per thread)
inline uint32_t range_grid(uint32_t len)
{
    uint64_t g = ((uint64_t)len + 65535) / 65536;
    g = (g + 7) & ~7ull;
    if (g < 8) g = 8;
    if (g > kRangeMaxWG) g = kRangeMaxWG;
    return (uint32_t)g;
}

// -------- launchers (return hipError_t of the launch)
hipError_t launch_range(const float* in, const Geom& g, double eb_rel, float4* part, fz_result* res,
                        uint64_t* zero_states, uint32_t nzero, uint32_t* zero_ticket, hipStream_t s);
// Fused compress (3 kernels).  ev (may be null) holds start / stop events of each kernel.
hipError_t launch_compress(const float* in, const Geom& g, double eb_rel, float4* part, fz_result* res,
                           uint8_t* payload, uint16_t* tail, uint8_t* staging, uint32_t* counts, uint64_t* states,
                           hipEvent_t* ev, hipStream_t s);
uint32_t   gather_grid(uint32_t T);
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
hipError_t launch_tile_scan(const uint32_t* bitflags, uint32_t T, uint32_t* tile_excl, uint64_t* states,
                            uint32_t* ticket, hipStream_t s);
hipError_t launch_decode_unshuffle(const uint8_t* payload, const uint32_t* tile_excl, const uint16_t* tail,
                                   const Geom& g, uint32_t* codes, hipStream_t s);
hipError_t launch_reconstruct(const uint16_t* codes, float* out, const Geom& g, float ebx2, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s);

}  // namespace fz
