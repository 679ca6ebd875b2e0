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
//
// Admission (the study's oracle gate, run_config_study_v0_8.ps1:57-68): the
// warm-up run of every candidate is checked before its timings count -- its
// device invariant word must be zero and the digest of all its masks must
// equal the digest of the largest group of agreeing candidates (ss_admit.hpp).
// Only admitted candidates can win.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "ss_admit.hpp"
#include "ss_internal.hpp"

namespace {

const ss_config kCandidates[] = {
    {16, 64}, {16, 128}, {16, 256}, {32, 64}, {32, 128}, {32, 256}, {32, 512},
    {64, 64}, {64, 128}, {64, 256}, {64, 512},
};
constexpr int kNumCandidates = (int)(sizeof(kCandidates) / sizeof(kCandidates[0]));

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

struct Events {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Events() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

double median(std::vector<float> t) {
    std::sort(t.begin(), t.end());
    const size_t m = t.size();
    return m % 2 ? t[m / 2] : 0.5 * (t[m / 2 - 1] + t[m / 2]);
}

// Digest salts: image i's max mask and min mask each get their own.
uint64_t salt_of(int image, int which) { return 0x9E3779B97F4A7C15ull * (uint64_t)(2 * image + which + 1); }

// Adds the digests of the two masks of images[0 .. n) to *d_dig; images[i]
// is image first + i of the caller's list.
hipError_t digest_masks(const ss_image_desc* images, int first, int n, uint64_t* d_dig, hipStream_t s) {
    for (int i = 0; i < n; ++i) {
        const ss_image_desc& d = images[i];
        const int k = first + i;
        hipError_t e = ssx::launch_digest(d.max, (int64_t)d.out_pitch, d.width, d.height, salt_of(k, 0), d_dig, s);
        if (e == hipSuccess)
            e = ssx::launch_digest(d.min, (int64_t)d.out_pitch, d.width, d.height, salt_of(k, 1), d_dig, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Picks the fastest admitted candidate; SS_ERR_INVARIANT if none is admitted.
ss_status pick(const bool* ok, const double* rate, uint64_t* d_dig, hipStream_t s, ss_config* best,
               double* throughput) {
    uint64_t dig[kNumCandidates];
    if (hipMemcpyAsync(dig, d_dig, sizeof(dig), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return SS_ERR_HIP;
    bool adm[kNumCandidates];
    if (ssx::admit_candidates(ok, dig, kNumCandidates, adm) == 0) return SS_ERR_INVARIANT;
    int bi = -1;
    for (int i = 0; i < kNumCandidates; ++i)
        if (adm[i] && (bi < 0 || rate[i] > rate[bi])) bi = i;
    *best = kCandidates[bi];
    if (throughput) *throughput = rate[bi];
    return ss_set_default_config(best);
}

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
    DevBuf ws, dig;
    if (hipMalloc(&ws.p, ws_bytes) != hipSuccess) return SS_ERR_HIP;
    if (hipMalloc(&dig.p, kNumCandidates * sizeof(uint64_t)) != hipSuccess) return SS_ERR_HIP;
    if (hipMemsetAsync(dig.p, 0, kNumCandidates * sizeof(uint64_t), stream) != hipSuccess) return SS_ERR_HIP;
    Events ev;
    if (hipEventCreate(&ev.e0) != hipSuccess || hipEventCreate(&ev.e1) != hipSuccess) return SS_ERR_HIP;
    double total_px = 0.0;
    for (int i = 0; i < n; ++i) total_px += (double)images[i].width * images[i].height;
    bool ok[kNumCandidates];
    double rate[kNumCandidates];
    for (int ci = 0; ci < kNumCandidates; ++ci) {
        const ss_config& c = kCandidates[ci];
        uint64_t* d_dig = static_cast<uint64_t*>(dig.p) + ci;
        ok[ci] = true;
        rate[ci] = 0.0;
        double sum_median_ms = 0.0;
        for (int i = 0; i < n && ok[ci]; ++i) {
            const ss_image_desc& d = images[i];
            std::vector<float> t;
            for (int r = 0; r < reps + 1; ++r) {
                (void)hipEventRecord(ev.e0, stream);
                ss_status st = ss_extrema_u8(d.img, d.pitch, d.width, d.height, d.max, d.min, d.out_pitch, ws.p,
                                             ws_bytes, &c, stream);
                if (st != SS_OK) return st;
                (void)hipEventRecord(ev.e1, stream);
                if (hipEventSynchronize(ev.e1) != hipSuccess) return SS_ERR_HIP;
                if (r == 0) {
                    // the warm-up run is the admission run
                    ss_stats s{};
                    st = ss_read_stats(ws.p, ws_bytes, &s, stream);
                    if (st == SS_ERR_INVARIANT) {
                        ok[ci] = false;
                        break;
                    }
                    if (st != SS_OK) return st;
                    if (digest_masks(&d, i, 1, d_dig, stream) != hipSuccess) return SS_ERR_HIP;
                    continue;
                }
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, ev.e0, ev.e1);
                t.push_back(ms);
            }
            if (ok[ci]) sum_median_ms += median(t);
        }
        if (ok[ci]) rate[ci] = total_px / (sum_median_ms * 1e-3);
    }
    return pick(ok, rate, static_cast<uint64_t*>(dig.p), stream, best, throughput);
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
    DevBuf ws, dig;
    if (hipMalloc(&ws.p, ws_bytes) != hipSuccess) return SS_ERR_HIP;
    if (hipMalloc(&dig.p, kNumCandidates * sizeof(uint64_t)) != hipSuccess) return SS_ERR_HIP;
    if (hipMemsetAsync(dig.p, 0, kNumCandidates * sizeof(uint64_t), stream) != hipSuccess) return SS_ERR_HIP;
    Events ev;
    if (hipEventCreate(&ev.e0) != hipSuccess || hipEventCreate(&ev.e1) != hipSuccess) return SS_ERR_HIP;
    double total_px = 0.0;
    for (int i = 0; i < n; ++i) total_px += (double)images[i].width * images[i].height;
    bool ok[kNumCandidates];
    double rate[kNumCandidates];
    for (int ci = 0; ci < kNumCandidates; ++ci) {
        const ss_config& c = kCandidates[ci];
        ok[ci] = true;
        rate[ci] = 0.0;
        std::vector<float> t;
        for (int r = 0; r < reps + 1; ++r) {
            (void)hipEventRecord(ev.e0, stream);
            ss_status st = ss_extrema_batch_u8(images, n, ws.p, ws_bytes, &c, stream);
            if (st != SS_OK) return st;
            (void)hipEventRecord(ev.e1, stream);
            if (hipEventSynchronize(ev.e1) != hipSuccess) return SS_ERR_HIP;
            if (r == 0) {
                ss_stats s{};
                st = ss_read_stats(ws.p, ws_bytes, &s, stream);
                if (st == SS_ERR_INVARIANT) {
                    ok[ci] = false;
                    break;
                }
                if (st != SS_OK) return st;
                if (digest_masks(images, 0, n, static_cast<uint64_t*>(dig.p) + ci, stream) != hipSuccess)
                    return SS_ERR_HIP;
                continue;
            }
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, ev.e0, ev.e1);
            t.push_back(ms);
        }
        if (ok[ci]) rate[ci] = total_px / (median(t) * 1e-3);
    }
    return pick(ok, rate, static_cast<uint64_t*>(dig.p), stream, best, throughput);
}
