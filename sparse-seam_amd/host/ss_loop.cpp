// ss_loop.cpp -- the reference's per-image loop in C++ over the C ABI, for
// timing the plain path on many images without any binding overhead.
//
// The reference's harness processes one image per iteration of its main loop
// (harness/primary/gpu_common_v0_6.cu:584-745), one run_sparse_resident per
// image.  ss_loop_per_image does the same back to back on one stream: every
// pass enqueues ss_extrema_u8 for each image in order, and a hipEvent pair
// around the pass gives its device time.  One untimed pass comes first.
#include <hip/hip_runtime.h>

#include "ss_extrema.h"

extern "C" ss_status ss_loop_per_image(const ss_image_desc* images, int n, void* d_ws, size_t ws_bytes,
                                       const ss_config* cfg, hipStream_t stream, int reps, float* ms) {
    if (!images || n <= 0 || reps <= 0 || !ms) return SS_ERR_ARG;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess) return SS_ERR_HIP;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return SS_ERR_HIP;
    }
    ss_status st = SS_OK;
    for (int r = -1; r < reps && st == SS_OK; ++r) {
        if (hipEventRecord(e0, stream) != hipSuccess) st = SS_ERR_HIP;
        for (int i = 0; i < n && st == SS_OK; ++i) {
            const ss_image_desc& d = images[i];
            st = ss_extrema_u8(d.img, d.pitch, d.width, d.height, d.max, d.min, d.out_pitch, d_ws, ws_bytes, cfg,
                               stream);
        }
        if (st != SS_OK) break;
        if (hipEventRecord(e1, stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
            st = SS_ERR_HIP;
            break;
        }
        float t = 0.f;
        if (hipEventElapsedTime(&t, e0, e1) != hipSuccess) st = SS_ERR_HIP;
        if (r >= 0) ms[r] = t;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return st;
}
