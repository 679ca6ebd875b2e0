// ss_global_uf.hip -- the dense Global-UF route on gfx950: the reference's
// co-reference route, kept as an independent cross-check of the Sparse-Seam
// path (the reference compares every timed image's masks against it,
// harness/primary/gpu_common_v0_6.cu:659-669).
//
// Route (snapshot_gpu_benchmark_globaluf_ole_v0_1.cu:105-259, sequence
// :341-385): one union-find node per pixel; every pixel is united with its
// equal left, up-left, up and up-right neighbours (min-root CAS); the forest is
// fully compressed; each pixel's "a strictly greater / smaller 8-neighbour
// exists" flags are ORed into its root; every pixel takes its root's flags.
// Here the root's flags live in the two top bits of its own parent word (so
// the route needs 4 bytes per pixel of workspace instead of the reference's
// 14), walks halve paths, and the flag pass reads the 3 x 3 neighbourhood
// straight from the image.
#include "ss_device.hpp"
#include "ss_internal.hpp"

namespace ssx {

namespace {

constexpr uint32_t GUF_MASK = 0x3FFFFFFFu;   // node ids: images below 2^30 pixels
constexpr int GUF_SHIFT = 30;                // bit 30: greater exists, bit 31: smaller exists
constexpr int GUF_BX = 64, GUF_BY = 4;

struct GufImg {
    const uint8_t* img;
    uint8_t* mx;
    uint8_t* mn;
    int64_t pitch, opitch;
    int w, h;
};

__global__ __launch_bounds__(GUF_BX* GUF_BY) void k_guf_init(uint32_t* parent, GufImg im) {
    const int x = blockIdx.x * GUF_BX + (int)(threadIdx.x % GUF_BX);
    const int y = blockIdx.y * GUF_BY + (int)(threadIdx.x / GUF_BX);
    if (x >= im.w || y >= im.h) return;
    const uint32_t i = (uint32_t)y * (uint32_t)im.w + (uint32_t)x;
    parent[i] = i;
}

__device__ __forceinline__ uint32_t guf_find(uint32_t* parent, uint32_t x) {
    for (;;) {
        const uint32_t p = parent[x];
        if (p == x) return x;
        const uint32_t gp = parent[p];
        if (gp == p) return p;
        parent[x] = gp;
        x = gp;
    }
}

__device__ __forceinline__ void guf_union(uint32_t* parent, uint32_t a, uint32_t b) {
    for (;;) {
        a = guf_find(parent, a);
        b = guf_find(parent, b);
        if (a == b) return;
        const uint32_t hi = a > b ? a : b, lo = a > b ? b : a;
        uint32_t expected = hi;
        if (__hip_atomic_compare_exchange_strong(parent + hi, &expected, lo, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return;
        a = expected;
        b = lo;
    }
}

// Unions with the equal left, up-left, up and up-right neighbours.
__global__ __launch_bounds__(GUF_BX* GUF_BY) void k_guf_union(uint32_t* parent, GufImg im) {
    const int x = blockIdx.x * GUF_BX + (int)(threadIdx.x % GUF_BX);
    const int y = blockIdx.y * GUF_BY + (int)(threadIdx.x / GUF_BX);
    if (x >= im.w || y >= im.h) return;
    const uint8_t* row = im.img + (int64_t)y * im.pitch;
    const uint8_t v = row[x];
    const uint32_t i = (uint32_t)y * (uint32_t)im.w + (uint32_t)x;
    if (x > 0 && row[x - 1] == v) guf_union(parent, i, i - 1);
    if (y > 0) {
        const uint8_t* up = row - im.pitch;
        const uint32_t u = i - (uint32_t)im.w;
        if (x > 0 && up[x - 1] == v) guf_union(parent, i, u - 1);
        if (up[x] == v) guf_union(parent, i, u);
        if (x + 1 < im.w && up[x + 1] == v) guf_union(parent, i, u + 1);
    }
}

__global__ __launch_bounds__(256) void k_guf_compress(uint32_t* parent, uint32_t count) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= count) return;
    uint32_t r = parent[i];
    while (parent[r] != r) r = parent[r];
    parent[i] = r;
}

// Flags of every pixel ORed into its root (one atomic only for a pixel whose
// flags the root word may lack).  Neighbours outside the image are absent.
__global__ __launch_bounds__(GUF_BX* GUF_BY) void k_guf_flags(uint32_t* parent, GufImg im) {
    const int x = blockIdx.x * GUF_BX + (int)(threadIdx.x % GUF_BX);
    const int y = blockIdx.y * GUF_BY + (int)(threadIdx.x / GUF_BX);
    if (x >= im.w || y >= im.h) return;
    const uint8_t* row = im.img + (int64_t)y * im.pitch;
    const uint32_t v = row[x];
    uint32_t hi = 0, lo = 255;
    for (int dy = -1; dy <= 1; ++dy) {
        const int ny = y + dy;
        if (ny < 0 || ny >= im.h) continue;
        const uint8_t* r = row + (int64_t)dy * im.pitch;
        for (int dx = -1; dx <= 1; ++dx) {
            const int nx = x + dx;
            if ((dx == 0 && dy == 0) || nx < 0 || nx >= im.w) continue;
            const uint32_t n = r[nx];
            hi = n > hi ? n : hi;
            lo = n < lo ? n : lo;
        }
    }
    const uint32_t f = (hi > v ? 1u : 0u) | (lo < v ? 2u : 0u);
    if (!f) return;
    const uint32_t i = (uint32_t)y * (uint32_t)im.w + (uint32_t)x;
    const uint32_t root = parent[i] & GUF_MASK;
    const uint32_t have = parent[root] >> GUF_SHIFT;
    if (f & ~have)
        __hip_atomic_fetch_or(parent + root, f << GUF_SHIFT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(GUF_BX* GUF_BY) void k_guf_materialise(const uint32_t* parent, GufImg im) {
    const int x = blockIdx.x * GUF_BX + (int)(threadIdx.x % GUF_BX);
    const int y = blockIdx.y * GUF_BY + (int)(threadIdx.x / GUF_BX);
    if (x >= im.w || y >= im.h) return;
    const uint32_t i = (uint32_t)y * (uint32_t)im.w + (uint32_t)x;
    const uint32_t f = parent[parent[i] & GUF_MASK] >> GUF_SHIFT;
    const int64_t o = (int64_t)y * im.opitch + x;
    im.mx[o] = (f & 1u) ? 0 : 255;
    im.mn[o] = (f & 2u) ? 0 : 255;
}

size_t guf_bytes(int w, int h) { return align256((size_t)w * (size_t)h * 4); }

}  // namespace

}  // namespace ssx

using namespace ssx;

extern "C" {

ss_status ss_global_uf_workspace_size(int width, int height, size_t* bytes) {
    if (!bytes || width <= 0 || height <= 0) return SS_ERR_ARG;
    if ((uint64_t)width * (uint64_t)height > GUF_MASK) return SS_ERR_ARG;
    *bytes = guf_bytes(width, height);
    return SS_OK;
}

ss_status ss_global_uf_u8(const uint8_t* d_img, size_t pitch, int width, int height, uint8_t* d_max,
                          uint8_t* d_min, size_t out_pitch, void* d_ws, size_t ws_bytes, hipStream_t stream) {
    if (!d_img || !d_max || !d_min || width <= 0 || height <= 0) return SS_ERR_ARG;
    if (pitch < (size_t)width || out_pitch < (size_t)width) return SS_ERR_ARG;
    if ((uint64_t)width * (uint64_t)height > GUF_MASK) return SS_ERR_ARG;
    if (!d_ws || ws_bytes < guf_bytes(width, height) || (reinterpret_cast<uintptr_t>(d_ws) & 255u))
        return SS_ERR_WORKSPACE;
    uint32_t* parent = static_cast<uint32_t*>(d_ws);
    const GufImg im{d_img, d_max, d_min, (int64_t)pitch, (int64_t)out_pitch, width, height};
    const dim3 grid((unsigned)((width + GUF_BX - 1) / GUF_BX), (unsigned)((height + GUF_BY - 1) / GUF_BY));
    const uint32_t count = (uint32_t)width * (uint32_t)height;
    k_guf_init<<<grid, GUF_BX * GUF_BY, 0, stream>>>(parent, im);
    k_guf_union<<<grid, GUF_BX * GUF_BY, 0, stream>>>(parent, im);
    k_guf_compress<<<(count + 255) / 256, 256, 0, stream>>>(parent, count);
    k_guf_flags<<<grid, GUF_BX * GUF_BY, 0, stream>>>(parent, im);
    k_guf_materialise<<<grid, GUF_BX * GUF_BY, 0, stream>>>(parent, im);
    return hipGetLastError() == hipSuccess ? SS_OK : SS_ERR_HIP;
}

}  // extern "C"
