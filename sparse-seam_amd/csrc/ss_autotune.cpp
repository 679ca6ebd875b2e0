// ss_autotune.cpp -- on-device autotuner for (tile, threads).
//
// Replaces the reference's offline configuration study: nine preregistered
// (T, classify-threads) builds timed in rotated rounds, each scored by
// sum(pixels) / sum(per-image median) (harness/configuration_study/
// run_config_study_v0_8.ps1:85-109, sparse_config_benchmark_v0_8.cu:17-30).
// Here every candidate is a runtime template instance, so one process times
// them all on the caller's resident images and keeps the best as the default.
// ss_autotune follows that rule (per-image launches); ss_autotune_batch
// scores each candidate on the batched launch it will be used for (one launch
// set over all the images), where per-image launch overhead does not apply.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "ss_internal.hpp"

namespace {

const ss_config kCandidates[] = {
    {16, 64}, {16, 128}, {16, 256}, {32, 64}, {32, 128}, {32, 256}, {32, 512},
    {64, 64}, {64, 128}, {64, 256}, {64, 512},
};

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

}  // namespace

extern "C" ss_status ss_autotune(const ss_image_desc* images, int n, int reps, hipStream_t stream,
                                 ss_config* best, double* throughput) {
    if (!images || n <= 0 || !best) return SS_ERR_ARG;
    if (reps <= 0) reps = 5;
    size_t ws_bytes = 0;
    for (const ss_config& c : kCandidates) {
        for (int i = 0; i < n; ++i) {
            size_t b = 0;
            const ss_status st = ss_workspace_size(images[i].width, images[i].height, &c, &b);
            if (st != SS_OK) return st;
            ws_bytes = std::max(ws_bytes, b);
        }
    }
    DevBuf ws;
    if (hipMalloc(&ws.p, ws_bytes) != hipSuccess) return SS_ERR_HIP;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return SS_ERR_HIP;
    double best_rate = -1.0;
    ss_config best_cfg = kCandidates[0];
    ss_status result = SS_OK;
    double total_px = 0.0;
    for (int i = 0; i < n; ++i) total_px += (double)images[i].width * images[i].height;
    for (const ss_config& c : kCandidates) {
        double sum_median_ms = 0.0;
        for (int i = 0; i < n && result == SS_OK; ++i) {
            const ss_image_desc& d = images[i];
            std::vector<float> t;
            for (int r = 0; r < reps + 1 && result == SS_OK; ++r) {
                (void)hipEventRecord(e0, stream);
                result = ss_extrema_u8(d.img, d.pitch, d.width, d.height, d.max, d.min, d.out_pitch, ws.p,
                                       ws_bytes, &c, stream);
                (void)hipEventRecord(e1, stream);
                if (hipEventSynchronize(e1) != hipSuccess) result = SS_ERR_HIP;
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                if (r > 0) t.push_back(ms);  // first run is a warm-up
            }
            if (result != SS_OK) break;
            std::sort(t.begin(), t.end());
            const size_t m = t.size();
            sum_median_ms += m % 2 ? t[m / 2] : 0.5 * (t[m / 2 - 1] + t[m / 2]);
        }
        if (result != SS_OK) break;
        const double rate = total_px / (sum_median_ms * 1e-3);
        if (rate > best_rate) {
            best_rate = rate;
            best_cfg = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (result != SS_OK) return result;
    *best = best_cfg;
    if (throughput) *throughput = best_rate;
    return ss_set_default_config(&best_cfg);
}

extern "C" ss_status ss_autotune_batch(const ss_image_desc* images, int n, int reps, hipStream_t stream,
                                       ss_config* best, double* throughput) {
    if (!images || n <= 0 || !best) return SS_ERR_ARG;
    if (reps <= 0) reps = 5;
    size_t ws_bytes = 0;
    for (const ss_config& c : kCandidates) {
        size_t b = 0;
        const ss_status st = ss_batch_workspace_size(images, n, &c, &b);
        if (st != SS_OK) return st;
        ws_bytes = std::max(ws_bytes, b);
    }
    DevBuf ws;
    if (hipMalloc(&ws.p, ws_bytes) != hipSuccess) return SS_ERR_HIP;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return SS_ERR_HIP;
    double total_px = 0.0;
    for (int i = 0; i < n; ++i) total_px += (double)images[i].width * images[i].height;
    double best_rate = -1.0;
    ss_config best_cfg = kCandidates[0];
    ss_status result = SS_OK;
    for (const ss_config& c : kCandidates) {
        std::vector<float> t;
        for (int r = 0; r < reps + 1 && result == SS_OK; ++r) {
            (void)hipEventRecord(e0, stream);
            result = ss_extrema_batch_u8(images, n, ws.p, ws_bytes, &c, stream);
            (void)hipEventRecord(e1, stream);
            if (hipEventSynchronize(e1) != hipSuccess) result = SS_ERR_HIP;
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (r > 0) t.push_back(ms);  // first run is a warm-up
        }
        if (result != SS_OK) break;
        std::sort(t.begin(), t.end());
        const size_t m = t.size();
        const double med = m % 2 ? t[m / 2] : 0.5 * (t[m / 2 - 1] + t[m / 2]);
        const double rate = total_px / (med * 1e-3);
        if (rate > best_rate) {
            best_rate = rate;
            best_cfg = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (result != SS_OK) return result;
    *best = best_cfg;
    if (throughput) *throughput = best_rate;
    return ss_set_default_config(&best_cfg);
}
