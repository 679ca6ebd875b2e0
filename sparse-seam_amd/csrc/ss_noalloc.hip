// ss_noalloc.hip -- the allocation-free form of the single-image operator.
//
// For callers that cannot allocate device memory (code under hipGraph
// capture, tight serving loops, harnesses that forbid hipMalloc): the
// pipeline's workspace lives in `scratch`, a file-scope device array of
// SS_NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds until the
// process ends.  The caller passes only the image, the two masks and,
// optionally, a device word that receives the run's invariant code.
//
// In host code `scratch` names a host-side stand-in, not the device array: its
// device address is looked up once per device with hipGetSymbolAddress (with
// that device current) and only that address is handed to kernels.  Every call
// on a GPU uses that GPU's one array: calls that overlap on one GPU (other
// streams, other threads) race and must be ordered by the caller.  The array
// keeps the last call's contents; the pipeline needs nothing in it cleared
// (every region is written before it is read within one call).
#include <atomic>
#include <mutex>

#include "ss_internal.hpp"

namespace ssx {

namespace {

static __device__ __align__(256) uint8_t scratch[SS_NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
std::atomic<void*> g_scratch_ptr[MAX_DEVICES];

// Device address of the current device's scratch array (looked up on the
// first call on that device), or null on failure.
void* scratch_ptr() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return nullptr;
    void* p = g_scratch_ptr[dev].load(std::memory_order_acquire);
    if (p) return p;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    p = g_scratch_ptr[dev].load(std::memory_order_relaxed);
    if (!p) {
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        g_scratch_ptr[dev].store(p, std::memory_order_release);
    }
    return p;
}

// Copies the run's invariant word out of the workspace header.
__global__ void k_copy_invariant(const WsHeader* h, uint32_t* out) {
    if (threadIdx.x == 0) *out = h->error;
}

}  // namespace

}  // namespace ssx

using namespace ssx;

extern "C" ss_status ss_extrema_u8_noalloc(const uint8_t* d_img, size_t pitch, int width, int height,
                                           uint8_t* d_max, uint8_t* d_min, size_t out_pitch, const ss_config* cfg,
                                           uint32_t* d_invariant, hipStream_t stream) {
    void* ws = scratch_ptr();
    if (!ws) return SS_ERR_NO_DEVICE;
    size_t need = 0;
    const ss_status st = extrema_single(d_img, pitch, width, height, d_max, d_min, out_pitch, ws,
                                        SS_NOALLOC_SCRATCH_BYTES, cfg, stream, &need);
    if (st == SS_ERR_WORKSPACE && need > SS_NOALLOC_SCRATCH_BYTES) return SS_ERR_CAPACITY;
    if (st != SS_OK || !d_invariant) return st;
    k_copy_invariant<<<1, 64, 0, stream>>>(static_cast<const WsHeader*>(ws), d_invariant);
    return hipGetLastError() == hipSuccess ? SS_OK : SS_ERR_HIP;
}
