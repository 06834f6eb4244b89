// fz-gpu -- command line drop-in for the reference's `fz-gpu <file> <x> <y> <z> <eb>`
// (fz.cu:226-237).  With the five positional arguments it prints exactly the three lines
// of fz.cu:207-209.  Extra behaviour sits behind trailing options:
//   -o <archive>   write the compressed archive (header + payload + tail)
//   --literal      byte-identical reference mode (1D sign loss and dropped tail included)
//   --outliers     outlier mode: FIXED payload + side channel (large deltas, NaN/Inf, values
//                  too large to quantise, constant fields); the archive carries the channel
//   --plain        run the plain per-stage pipeline instead of the fused kernel
//   --json         print one JSON line of statistics after the three lines
// and a decompress mode, with the quality metrics of the reconstruction when the original
// is given (quality_viewer.hh:97-110 layout, computed on the GPU):
//   fz-gpu -d <archive> <out.f32> [original.f32]
#include <hip/hip_runtime.h>
#include <sys/stat.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fz_archive.hpp"
#include "fz_capi.h"

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) {                                                                \
            std::fprintf(stderr, "fz-gpu: %s failed: %s\n", #x, hipGetErrorString(e_));        \
            return 2;                                                                          \
        }                                                                                      \
    } while (0)

#define FZCHK(x)                                                                               \
    do {                                                                                       \
        fz_status s_ = (x);                                                                    \
        if (s_ != FZ_SUCCESS) {                                                                \
            std::fprintf(stderr, "fz-gpu: %s: %s\n", #x, fz_status_string(s_));               \
            return 3;                                                                          \
        }                                                                                      \
    } while (0)

static long file_size(const std::string& name)  // fz.cu:19-24
{
    struct stat st;
    return stat(name.c_str(), &st) == 0 ? (long)st.st_size : -1;
}

static int usage()
{
    std::fprintf(stderr,
                 "usage: fz-gpu <file> <x> <y> <z> <eb> [-o archive] [--literal | --outliers] [--plain] [--json]\n"
                 "       fz-gpu -d <archive> <out.f32> [original.f32]\n");
    return 1;
}

// print_metrics_cross + print_metrics_auto (quality_viewer.hh:53-95) for float data
static void print_metrics(const fz_stats& s, double compressed_bytes, double lag1, double lag2)
{
    auto head = [](const char* a, const char* b, const char* c, const char* d, const char* e) {
        std::printf("  \033[1m\033[31m%-10s %16s %16s %16s %16s\033[0m\n", a, b, c, d, e);
    };
    auto row = [](const char* a, double b, double c, double d, double e) {
        std::printf("  %-10s %16.8g %16.8g %16.8g %16.8g\n", a, b, c, d, e);
    };
    std::printf("\nquality metrics (using GPU checker):\n");
    head("", "data-len", "data-byte", "fp-type?", "");
    std::printf("  %-10s %16llu %16d %16s\n", "", (unsigned long long)s.len, 4, "yes");
    head("", "min", "max", "rng", "std");
    row("origin", s.odata.min, s.odata.max, s.odata.rng, s.odata.std);
    row("eb-lossy", s.xdata.min, s.xdata.max, s.xdata.rng, s.xdata.std);
    head("", "abs-val", "abs-idx", "pw-rel", "VS-RNG");
    row("max-error", s.max_err.abs, (double)s.max_err.idx, s.max_err.pwrrel, s.max_err.rel);
    head("", "CR", "NRMSE", "cross-cor", "PSNR");
    row("metrics", s.len * 4.0 / compressed_bytes, s.reduced.NRMSE, s.reduced.coeff, s.reduced.PSNR);
    head("", "lag1-cor", "lag2-cor", "", "");
    std::printf("  %-10s %16lf %16lf\n\n", "auto", lag1, lag2);
}

static int decompress_main(const std::string& arc, const std::string& out, const std::string& original)
{
    fzar::Header          h;
    std::vector<uint8_t>  payload;
    std::vector<uint16_t> tail;
    std::vector<uint8_t>  side;
    if (!fzar::read_archive(arc, h, payload, tail, &side)) {
        std::fprintf(stderr, "fz-gpu: cannot read archive %s\n", arc.c_str());
        return 2;
    }
    fz_plan plan;
    FZCHK(fz_plan_create(&plan, h.x, h.y, h.z, h.len, h.mode));
    fz_sizes sz;
    FZCHK(fz_plan_sizes(plan, &sz));
    // the header's sizes and the payload's bitflags must agree before anything reaches the GPU
    FZCHK(fz_payload_check(plan, payload.data(), h.payload_bytes, h.tail_len));
    if (h.mode == FZ_MODE_OUTLIER) FZCHK(fz_side_check(plan, side.data(), side.size()));
    const uint64_t n = (uint64_t)h.x * h.y * h.z;
    uint8_t*       d_payload;
    uint16_t*      d_tail = nullptr;
    float*         d_out;
    void*          d_ws;
    HIPCHK(hipMalloc(&d_payload, sz.payload_max_bytes + 16));
    HIPCHK(hipMalloc(&d_out, 4 * n + 16));
    HIPCHK(hipMalloc(&d_ws, sz.workspace_bytes));
    if (h.tail_len) {
        HIPCHK(hipMalloc(&d_tail, 2 * h.tail_len));
        HIPCHK(hipMemcpy(d_tail, tail.data(), 2 * h.tail_len, hipMemcpyHostToDevice));
    }
    if (h.payload_bytes) HIPCHK(hipMemcpy(d_payload, payload.data(), h.payload_bytes, hipMemcpyHostToDevice));
    hipStream_t st;
    HIPCHK(hipStreamCreate(&st));
    FZCHK(fz_workspace_init(plan, d_ws, sz.workspace_bytes, st));
    const bool outlier = h.mode == FZ_MODE_OUTLIER;
    uint8_t*   d_side = nullptr;
    if (outlier) {
        HIPCHK(hipMalloc(&d_side, side.size()));
        HIPCHK(hipMemcpy(d_side, side.data(), side.size(), hipMemcpyHostToDevice));
        FZCHK(fz_decompress_outlier_f32(plan, d_payload, d_tail, h.ebx2, h.vmin, d_side, side.size(), d_ws,
                                        sz.workspace_bytes, d_out, st));
    }
    if (original.empty()) {
        if (!outlier) FZCHK(fz_decompress_f32(plan, d_payload, d_tail, h.ebx2, d_ws, sz.workspace_bytes, d_out, st));
    } else {
        if (file_size(original) != (long)(4 * n)) {
            std::fprintf(stderr, "fz-gpu: %s does not hold %llu floats\n", original.c_str(), (unsigned long long)n);
            return 2;
        }
        std::vector<float> orig(n);
        FILE*              fo = std::fopen(original.c_str(), "rb");
        if (!fo || std::fread(orig.data(), 4, n, fo) != n) {
            std::fprintf(stderr, "fz-gpu: cannot read %s\n", original.c_str());
            return 2;
        }
        std::fclose(fo);
        float*    d_orig;
        fz_stats* d_stats;
        HIPCHK(hipMalloc(&d_orig, 4 * n + 16));
        HIPCHK(hipMalloc(&d_stats, 3 * sizeof(fz_stats)));
        HIPCHK(hipMemcpy(d_orig, orig.data(), 4 * n, hipMemcpyHostToDevice));
        size_t tb = 0;
        FZCHK(fz_verify_f32(nullptr, &tb, nullptr, nullptr, nullptr, n, st));
        void* d_temp;
        HIPCHK(hipMalloc(&d_temp, tb));
        if (outlier)  // already decompressed: the metrics alone
            FZCHK(fz_verify_f32(d_temp, &tb, d_stats, d_out, d_orig, n, st));
        else
            FZCHK(fz_decompress_verify_f32(plan, d_payload, d_tail, h.ebx2, d_ws, sz.workspace_bytes, d_out, d_orig,
                                           d_stats, st));
        // lag-1 and lag-2 autocorrelation of the original (quality_viewer.hh:105-108)
        double lag[2] = {NAN, NAN};
        for (int l = 1; l <= 2; l++)
            if (n > (uint64_t)l) FZCHK(fz_verify_f32(d_temp, &tb, d_stats + l, d_orig, d_orig + l, n - l, st));
        fz_stats hs[3];
        HIPCHK(hipMemcpyAsync(hs, d_stats, sizeof(hs), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        for (int l = 1; l <= 2; l++)
            if (n > (uint64_t)l) lag[l - 1] = hs[l].reduced.coeff;
        const double archive_bytes = (double)file_size(arc);
        print_metrics(hs[0], archive_bytes, lag[0], lag[1]);
        (void)hipFree(d_temp);
        (void)hipFree(d_orig);
        (void)hipFree(d_stats);
    }
    HIPCHK(hipStreamSynchronize(st));
    std::vector<float> host(n);
    HIPCHK(hipMemcpy(host.data(), d_out, 4 * n, hipMemcpyDeviceToHost));
    FILE* f = std::fopen(out.c_str(), "wb");
    if (!f || std::fwrite(host.data(), 4, n, f) != n) {
        std::fprintf(stderr, "fz-gpu: cannot write %s\n", out.c_str());
        return 2;
    }
    std::fclose(f);
    (void)hipFree(d_payload);
    (void)hipFree(d_out);
    (void)hipFree(d_ws);
    if (d_tail) (void)hipFree(d_tail);
    if (d_side) (void)hipFree(d_side);
    (void)hipStreamDestroy(st);
    fz_plan_destroy(plan);
    return 0;
}

int main(int argc, char* argv[])
{
    if (argc >= 2 && std::strcmp(argv[1], "-d") == 0) {
        if (argc != 4 && argc != 5) return usage();
        return decompress_main(argv[2], argv[3], argc == 5 ? argv[4] : "");
    }
    if (argc < 6) return usage();
    const std::string fileName(argv[1]);
    const int         x = std::stoi(argv[2]), y = std::stoi(argv[3]), z = std::stoi(argv[4]);
    const double      eb = std::stod(argv[5]);
    std::string       out_path;
    int               mode = FZ_MODE_FIXED;
    bool              plain = false, json = false;
    for (int i = 6; i < argc; i++) {
        if (!std::strcmp(argv[i], "-o") && i + 1 < argc)
            out_path = argv[++i];
        else if (!std::strcmp(argv[i], "--literal"))
            mode = FZ_MODE_LITERAL;
        else if (!std::strcmp(argv[i], "--outliers"))
            mode = FZ_MODE_OUTLIER;
        else if (!std::strcmp(argv[i], "--plain"))
            plain = true;
        else if (!std::strcmp(argv[i], "--json"))
            json = true;
        else
            return usage();
    }

    // fz.cu:117-118, 136: the float count comes from the file size
    const long fileSizeL = file_size(fileName);
    if (fileSizeL < 4) {
        std::fprintf(stderr, "fail to open %s\n", fileName.c_str());
        return 1;
    }
    const int      fileSize = (int)fileSizeL;
    const uint64_t len = (uint64_t)fileSizeL / sizeof(float);
    std::vector<float> h_data(len);
    {
        FILE* f = std::fopen(fileName.c_str(), "rb");
        if (!f || std::fread(h_data.data(), 4, len, f) != len) {
            std::fprintf(stderr, "fail to open %s\n", fileName.c_str());
            return 1;
        }
        std::fclose(f);
    }

    fz_plan plan;
    FZCHK(fz_plan_create(&plan, (uint32_t)x, (uint32_t)y, (uint32_t)z, len, mode));
    fz_sizes sz;
    FZCHK(fz_plan_sizes(plan, &sz));
    float*     d_data;
    uint8_t*   d_payload;
    uint16_t*  d_tail;
    fz_result* d_res;
    void*      d_ws;
    HIPCHK(hipMalloc(&d_data, 4 * len + 16));
    HIPCHK(hipMalloc(&d_payload, sz.payload_max_bytes + 16));
    HIPCHK(hipMalloc(&d_tail, 2 * 2048));
    HIPCHK(hipMalloc(&d_res, sizeof(fz_result)));
    HIPCHK(hipMalloc(&d_ws, sz.workspace_bytes));
    HIPCHK(hipMemcpy(d_data, h_data.data(), 4 * len, hipMemcpyHostToDevice));
    hipStream_t st;
    HIPCHK(hipStreamCreate(&st));
    FZCHK(fz_workspace_init(plan, d_ws, sz.workspace_bytes, st));

    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, st));
    uint8_t* d_side = nullptr;
    if (mode == FZ_MODE_OUTLIER) {
        if (plain) return usage();
        HIPCHK(hipMalloc(&d_side, sz.side_max_bytes));
        FZCHK(fz_compress_outlier_f32(plan, d_data, eb, d_ws, sz.workspace_bytes, d_payload, d_tail, d_side,
                                      sz.side_max_bytes, d_res, st));
    } else if (plain) {
        FZCHK(fz_compress_plain_f32(plan, d_data, eb, d_ws, sz.workspace_bytes, d_payload, d_tail, d_res, st));
    } else {
        FZCHK(fz_compress_f32(plan, d_data, eb, d_ws, sz.workspace_bytes, d_payload, d_tail, d_res, st));
    }
    HIPCHK(hipEventRecord(e1, st));
    fz_result r;
    HIPCHK(hipMemcpyAsync(&r, d_res, sizeof(r), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    if (r.status != FZ_SUCCESS) {
        std::fprintf(stderr, "fz-gpu: %s\n", fz_status_string((fz_status)r.status));
        return 3;
    }

    // fz.cu:207-209 (the size is 4 W + 16 K)
    const uint64_t csize = r.payload_bytes;
    std::printf("original size: %d\n", fileSize);
    std::printf("compressed size: %d\n", (int)csize);
    std::printf("compression ratio: %f\n", float(fileSize) / float(csize));
    if (json)
        std::printf("{\"len\": %llu, \"K\": %llu, \"payload_bytes\": %llu, \"tail_len\": %llu, \"overflow\": %llu, "
                    "\"vmin\": %.9g, \"vmax\": %.9g, \"ms\": %.6f, \"input_GBps\": %.3f}\n",
                    (unsigned long long)len, (unsigned long long)r.nblocks, (unsigned long long)r.payload_bytes,
                    (unsigned long long)r.tail_len, (unsigned long long)r.overflow, r.vmin, r.vmax, ms,
                    4.0 * len / (ms * 1e-3) / 1e9);

    if (!out_path.empty()) {
        std::vector<uint8_t>  payload(r.payload_bytes);
        std::vector<uint16_t> tail(r.tail_len);
        HIPCHK(hipMemcpy(payload.data(), d_payload, r.payload_bytes, hipMemcpyDeviceToHost));
        if (r.tail_len) HIPCHK(hipMemcpy(tail.data(), d_tail, 2 * r.tail_len, hipMemcpyDeviceToHost));
        fzar::Header h{};
        std::memcpy(h.magic, "FZMI", 4);
        h.version = 1;
        h.mode = (uint8_t)mode;
        h.ndim = (uint8_t)r.ndim;
        h.x = (uint32_t)x;
        h.y = (uint32_t)y;
        h.z = (uint32_t)z;
        h.len = len;
        h.eb_rel = eb;
        h.vmin = r.vmin;
        h.vmax = r.vmax;
        h.ebx2 = r.ebx2;
        h.payload_bytes = r.payload_bytes;
        h.tail_len = r.tail_len;
        h.overflow = r.overflow;
        std::vector<uint8_t> side;
        if (d_side) {  // the channel's 16-byte header gives its size
            uint64_t cnt[2];
            HIPCHK(hipMemcpy(cnt, d_side, 16, hipMemcpyDeviceToHost));
            side.resize(16 + 8 * (cnt[0] + cnt[1]));
            HIPCHK(hipMemcpy(side.data(), d_side, side.size(), hipMemcpyDeviceToHost));
        }
        if (!fzar::write_archive(out_path, h, payload.data(), tail.data(), &side)) {
            std::fprintf(stderr, "fz-gpu: cannot write %s\n", out_path.c_str());
            return 2;
        }
    }
    (void)hipFree(d_data);
    (void)hipFree(d_payload);
    if (d_side) (void)hipFree(d_side);
    (void)hipFree(d_tail);
    (void)hipFree(d_res);
    (void)hipFree(d_ws);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    fz_plan_destroy(plan);
    return 0;
}
