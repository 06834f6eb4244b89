// fz_archive.hpp -- on-disk container around the FZ payload.
//
// The reference never writes its output (d_out is freed, fz.cu:215); this container keeps
// the payload bytes exactly as produced (bitflags || nonzero blocks) and adds what the
// inverse needs:
//   offset  size  field
//        0     4  magic "FZMI"
//        4     2  version (1)
//        6     1  mode (0 fixed, 1 literal, 2 outlier)
//        7     1  ndim
//        8    12  x, y, z (u32, x fastest)
//       20     4  reserved (0)
//       24     8  len (floats in the input)
//       32     8  eb_rel (f64)
//       40     4  vmin (f32)
//       44     4  vmax (f32)
//       48     4  ebx2 (f32, the dequantisation factor)
//       52     4  reserved (0)
//       56     8  payload_bytes
//       64     8  tail_len (u16 codes after the payload; 0 in literal mode)
//       72     8  overflow count
//       80        payload, then tail codes (LE u16), then in outlier mode the side channel
//                 (u64 n_delta, u64 n_raw, then 8 (n_delta + n_raw) bytes of lists)
// All fields little-endian.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace fzar {

struct Header {
    char     magic[4];
    uint16_t version;
    uint8_t  mode;
    uint8_t  ndim;
    uint32_t x, y, z;
    uint32_t reserved0;
    uint64_t len;
    double   eb_rel;
    float    vmin, vmax, ebx2;
    uint32_t reserved1;
    uint64_t payload_bytes;
    uint64_t tail_len;
    uint64_t overflow;
};
static_assert(sizeof(Header) == 80, "archive header is 80 bytes");

constexpr uint8_t kModeOutlier = 2;

inline bool write_archive(const std::string& path, const Header& h, const uint8_t* payload, const uint16_t* tail,
                          const std::vector<uint8_t>* side = nullptr)
{
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1;
    if (ok && h.payload_bytes) ok = std::fwrite(payload, 1, h.payload_bytes, f) == h.payload_bytes;
    if (ok && h.tail_len) ok = std::fwrite(tail, 2, h.tail_len, f) == h.tail_len;
    if (ok && h.mode == kModeOutlier) ok = side && std::fwrite(side->data(), 1, side->size(), f) == side->size();
    return std::fclose(f) == 0 && ok;
}

inline bool read_archive(const std::string& path, Header& h, std::vector<uint8_t>& payload,
                         std::vector<uint16_t>& tail, std::vector<uint8_t>* side = nullptr)
{
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    long fsize = -1;
    if (std::fseek(f, 0, SEEK_END) == 0) fsize = std::ftell(f);
    std::rewind(f);
    bool ok = fsize >= (long)sizeof(h) && std::fread(&h, sizeof(h), 1, f) == 1 &&
              std::memcmp(h.magic, "FZMI", 4) == 0 && h.version == 1 && h.mode <= kModeOutlier;
    // sizes the file cannot hold are refused before anything is allocated for them
    const uint64_t rest = ok ? (uint64_t)fsize - sizeof(h) : 0;
    ok = ok && h.payload_bytes <= rest && h.tail_len <= (rest - h.payload_bytes) / 2;
    if (ok) {
        payload.resize(h.payload_bytes);
        tail.resize(h.tail_len);
        if (h.payload_bytes) ok = std::fread(payload.data(), 1, h.payload_bytes, f) == h.payload_bytes;
        if (ok && h.tail_len) ok = std::fread(tail.data(), 2, h.tail_len, f) == h.tail_len;
    }
    if (ok && h.mode == kModeOutlier) {  // side channel: its 16-byte header gives its size
        uint64_t n[2];
        ok = side && std::fread(n, 8, 2, f) == 2 && n[0] <= h.len && n[1] <= h.len;
        const long at = ok ? std::ftell(f) : -1;
        ok = ok && at >= 0 && 8 * (n[0] + n[1]) <= (uint64_t)(fsize - at);
        if (ok) {
            side->resize(16 + 8 * (n[0] + n[1]));
            std::memcpy(side->data(), n, 16);
            const size_t lists = side->size() - 16;
            ok = lists == 0 || std::fread(side->data() + 16, 1, lists, f) == lists;
        }
    }
    std::fclose(f);
    return ok;
}

}  // namespace fzar
