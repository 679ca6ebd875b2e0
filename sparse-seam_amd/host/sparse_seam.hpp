// sparse_seam.hpp -- C++ host API over the C ABI, mirroring the reference's
// source-level operator interface (harness/primary/snapshot_sparse_seam_cuda_v1_0.cu:
// DeviceBuffers :57-68, allocate_buffers :710-737, free_buffers :739-751,
// run_sparse_seam :753-878; harness/primary/gpu_common_v0_6.cu:
// run_sparse_resident :137-206).  Same names, argument meaning and error
// behaviour: failures throw std::runtime_error, and a non-zero device
// invariant word throws "Sparse-seam device invariant failed, flag=N".
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdexcept>
#include <string>

#include "ss_extrema.h"

namespace sparse_seam {

#define SS_HIP_CHECK(call)                                                          \
    do {                                                                            \
        const hipError_t ec_ = (call);                                              \
        if (ec_ != hipSuccess)                                                      \
            throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(ec_)); \
    } while (0)

inline void check(ss_status s, const char* what) {
    if (s != SS_OK) throw std::runtime_error(std::string(what) + ": " + ss_status_str(s));
}

// Device state of one image: the image, both masks and the operator workspace.
struct DeviceBuffers {
    uint8_t* image = nullptr;
    uint8_t* maxima = nullptr;
    uint8_t* minima = nullptr;
    void* workspace = nullptr;
    size_t workspace_bytes = 0;
    int width = 0;
    int height = 0;
    ss_config config{32, 256};
};

// Allocates image, masks and workspace for a width x height image.
inline void allocate_buffers(DeviceBuffers& b, int width, int height, const ss_config* cfg = nullptr) {
    ss_config c{};
    if (cfg) c = *cfg; else check(ss_get_default_config(&c), "ss_get_default_config");
    size_t ws = 0;
    check(ss_workspace_size(width, height, &c, &ws), "ss_workspace_size");
    const size_t px = (size_t)width * (size_t)height;
    SS_HIP_CHECK(hipMalloc(&b.image, px));
    SS_HIP_CHECK(hipMalloc(&b.maxima, px));
    SS_HIP_CHECK(hipMalloc(&b.minima, px));
    SS_HIP_CHECK(hipMalloc(&b.workspace, ws));
    b.workspace_bytes = ws;
    b.width = width;
    b.height = height;
    b.config = c;
}

inline void free_buffers(DeviceBuffers& b) {
    if (b.image) SS_HIP_CHECK(hipFree(b.image));
    if (b.maxima) SS_HIP_CHECK(hipFree(b.maxima));
    if (b.minima) SS_HIP_CHECK(hipFree(b.minima));
    if (b.workspace) SS_HIP_CHECK(hipFree(b.workspace));
    b = DeviceBuffers{};
}

// Enqueues the operator on the image already resident in b.image.
inline int run_sparse_resident(DeviceBuffers& b, hipStream_t stream = nullptr) {
    check(ss_extrema_u8(b.image, (size_t)b.width, b.width, b.height, b.maxima, b.minima, (size_t)b.width,
                        b.workspace, b.workspace_bytes, &b.config, stream),
          "ss_extrema_u8");
    return 1;
}

// Co-reference route on the image resident in b.image: the dense Global-UF
// operator writes its masks into b.maxima / b.minima, using its own
// workspace `ws` of ss_global_uf_workspace_size bytes.  Mirrors run_global_uf
// (snapshot_gpu_benchmark_globaluf_ole_v0_1.cu:341-385), which returns 1.
inline int run_global_uf(DeviceBuffers& b, void* ws, size_t ws_bytes, hipStream_t stream = nullptr) {
    check(ss_global_uf_u8(b.image, (size_t)b.width, b.width, b.height, b.maxima, b.minima, (size_t)b.width, ws,
                          ws_bytes, stream),
          "ss_global_uf_u8");
    return 1;
}

// Reads the run's counters; throws on a device invariant failure.
inline ss_stats read_stats(const DeviceBuffers& b, hipStream_t stream = nullptr) {
    ss_stats st{};
    const ss_status s = ss_read_stats(b.workspace, b.workspace_bytes, &st, stream);
    if (s == SS_ERR_INVARIANT)
        throw std::runtime_error("Sparse-seam device invariant failed, flag=" + std::to_string(st.invariant));
    check(s, "ss_read_stats");
    return st;
}

// H2D image -> operator -> D2H masks.  Returns the device time of the
// operator in ms (hipEvents) and the number of seam fragments.
inline float run_sparse_seam(DeviceBuffers& b, const uint8_t* host_image, uint8_t* host_max,
                             uint8_t* host_min, uint64_t& active_fragments, hipStream_t stream = nullptr) {
    const size_t px = (size_t)b.width * (size_t)b.height;
    SS_HIP_CHECK(hipMemcpyAsync(b.image, host_image, px, hipMemcpyHostToDevice, stream));
    hipEvent_t start = nullptr, stop = nullptr;
    SS_HIP_CHECK(hipEventCreate(&start));
    SS_HIP_CHECK(hipEventCreate(&stop));
    SS_HIP_CHECK(hipEventRecord(start, stream));
    run_sparse_resident(b, stream);
    SS_HIP_CHECK(hipEventRecord(stop, stream));
    SS_HIP_CHECK(hipEventSynchronize(stop));
    float ms = 0.f;
    SS_HIP_CHECK(hipEventElapsedTime(&ms, start, stop));
    SS_HIP_CHECK(hipEventDestroy(start));
    SS_HIP_CHECK(hipEventDestroy(stop));
    const ss_stats st = read_stats(b, stream);
    active_fragments = st.fragments;
    SS_HIP_CHECK(hipMemcpyAsync(host_max, b.maxima, px, hipMemcpyDeviceToHost, stream));
    SS_HIP_CHECK(hipMemcpyAsync(host_min, b.minima, px, hipMemcpyDeviceToHost, stream));
    SS_HIP_CHECK(hipStreamSynchronize(stream));
    return ms;
}

}  // namespace sparse_seam
