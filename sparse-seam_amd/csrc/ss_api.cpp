// ss_api.cpp -- C ABI of the single-image and batched operator (see
// include/ss_extrema.h for the contract and the reference functions replaced).
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "ss_internal.hpp"

namespace ssx {

static std::mutex g_cfg_mu;
static ss_config g_default_cfg = {64, 128};

ss_config resolve_config(const ss_config* cfg) {
    if (cfg) return *cfg;
    std::lock_guard<std::mutex> lk(g_cfg_mu);
    return g_default_cfg;
}

int device_cus() {
    // a fixed table: no host allocation on the launch path
    static std::atomic<int> cache[MAX_DEVICES];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (dev < 0 || dev >= MAX_DEVICES) return 256;
    int cus = cache[dev].load(std::memory_order_relaxed);
    if (!cus) {
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        cache[dev].store(cus, std::memory_order_relaxed);
    }
    return cus;
}

ProfileHook& profile_hook() {
    static thread_local ProfileHook hook;
    return hook;
}

static bool aligned256(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 255u) == 0; }

ss_status extrema_single(const uint8_t* d_img, size_t pitch, int width, int height, uint8_t* d_max,
                         uint8_t* d_min, size_t out_pitch, void* d_ws, size_t ws_bytes, const ss_config* cfg,
                         hipStream_t stream, size_t* need) {
    if (!d_img || !d_max || !d_min || width <= 0 || height <= 0) return SS_ERR_ARG;
    if (pitch < (size_t)width || out_pitch < (size_t)width) return SS_ERR_ARG;
    const ss_config c = resolve_config(cfg);
    if (!config_valid(c.tile, c.threads)) return SS_ERR_ARG;
    const int64_t nt = tiles_of(width, height, c.tile);
    if (nt * (4 * c.tile - 4) > (int64_t)0xFFFFFFF0u) return SS_ERR_ARG;
    Images ims{};
    ims.one = make_imgdev(d_img, d_max, d_min, (int64_t)pitch, (int64_t)out_pitch, width, height, c.tile, 0, 0, 0);
    const Layout L = make_layout(nt, blocks_of(ims.one), seams_of(ims.one), c.tile, 1);
    if (need) *need = L.bytes;
    if (!d_ws || ws_bytes < L.bytes || !aligned256(d_ws)) return SS_ERR_WORKSPACE;
    ims.table = nullptr;
    ims.count = 1;
    const Ws ws = make_ws(d_ws, L, c.threads);
    const hipError_t e = launch_pipeline(ims, ws, c.tile, c.threads, device_cus(), stream);
    return e == hipSuccess ? SS_OK : SS_ERR_HIP;
}

}  // namespace ssx

using namespace ssx;

struct ss_batch_plan {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" {

const char* ss_status_str(ss_status s) {
    switch (s) {
    case SS_OK: return "ok";
    case SS_ERR_ARG: return "invalid argument";
    case SS_ERR_WORKSPACE: return "workspace too small or misaligned";
    case SS_ERR_HIP: return "HIP runtime error";
    case SS_ERR_INVARIANT: return "device invariant failed";
    case SS_ERR_CAPACITY: return "capacity exceeded";
    case SS_ERR_NO_DEVICE: return "no usable device";
    }
    return "unknown status";
}

ss_status ss_profile_events(hipEvent_t tile_begin, hipEvent_t tile_end) {
    if ((tile_begin == nullptr) != (tile_end == nullptr)) return SS_ERR_ARG;
    profile_hook().tile_begin = tile_begin;
    profile_hook().tile_end = tile_end;
    return SS_OK;
}

ss_status ss_set_default_config(const ss_config* cfg) {
    if (!cfg || !config_valid(cfg->tile, cfg->threads)) return SS_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_cfg_mu);
    g_default_cfg = *cfg;
    return SS_OK;
}

ss_status ss_get_default_config(ss_config* cfg) {
    if (!cfg) return SS_ERR_ARG;
    *cfg = resolve_config(nullptr);
    return SS_OK;
}

ss_status ss_workspace_size(int width, int height, const ss_config* cfg, size_t* bytes) {
    if (!bytes || width <= 0 || height <= 0) return SS_ERR_ARG;
    const ss_config c = resolve_config(cfg);
    if (!config_valid(c.tile, c.threads)) return SS_ERR_ARG;
    const int64_t nt = tiles_of(width, height, c.tile);
    if (nt * (4 * c.tile - 4) > (int64_t)0xFFFFFFF0u) return SS_ERR_ARG;
    const ImgDev d = make_imgdev(nullptr, nullptr, nullptr, width, width, width, height, c.tile, 0, 0, 0);
    *bytes = make_layout(nt, blocks_of(d), seams_of(d), c.tile, 1).bytes;
    return SS_OK;
}

ss_status ss_extrema_u8(const uint8_t* d_img, size_t pitch, int width, int height, uint8_t* d_max,
                        uint8_t* d_min, size_t out_pitch, void* d_ws, size_t ws_bytes,
                        const ss_config* cfg, hipStream_t stream) {
    return extrema_single(d_img, pitch, width, height, d_max, d_min, out_pitch, d_ws, ws_bytes, cfg, stream,
                          nullptr);
}

ss_status ss_read_stats(const void* d_ws, size_t ws_bytes, ss_stats* out, hipStream_t stream) {
    if (!d_ws || !out || ws_bytes < sizeof(WsHeader)) return SS_ERR_ARG;
    WsHeader h;
    if (hipMemcpyAsync(&h, d_ws, sizeof(h), hipMemcpyDeviceToHost, stream) != hipSuccess) return SS_ERR_HIP;
    if (hipStreamSynchronize(stream) != hipSuccess) return SS_ERR_HIP;
    if (h.magic != WS_MAGIC) return SS_ERR_WORKSPACE;
    out->invariant = h.error;
    out->dirty_tiles = h.dirty_count;
    out->fragments = h.total;
    out->dsu_retries = h.retries;
    out->tiles = h.tiles;
    return h.error ? SS_ERR_INVARIANT : SS_OK;
}

ss_status ss_read_tile_counts(const void* d_ws, size_t ws_bytes, uint32_t* out, size_t count,
                              hipStream_t stream) {
    if (!d_ws || !out || ws_bytes < sizeof(WsHeader) + count * 4) return SS_ERR_ARG;
    const char* src = static_cast<const char*>(d_ws) + sizeof(WsHeader);
    if (hipMemcpyAsync(out, src, count * 4, hipMemcpyDeviceToHost, stream) != hipSuccess) return SS_ERR_HIP;
    if (hipStreamSynchronize(stream) != hipSuccess) return SS_ERR_HIP;
    for (size_t i = 0; i < count; ++i) out[i] &= COUNT_MASK;
    return SS_OK;
}

// ---- batch ----------------------------------------------------------------

ss_status ss_seam_profile_u8(const uint8_t* d_img, size_t pitch, int width, int height,
                             const ss_config* cfg, void* d_ws, size_t ws_bytes, hipStream_t stream,
                             ss_seam_profile* out) {
    if (!d_img || !out || width <= 0 || height <= 0 || pitch < (size_t)width) return SS_ERR_ARG;
    ss_config c = resolve_config(cfg);
    c.threads = 256;
    // the profiler also takes T = 128, the reference profiler's largest tile
    if (c.tile != 128 && !config_valid(c.tile, c.threads)) return SS_ERR_ARG;
    const int64_t nt = tiles_of(width, height, c.tile);
    if (nt * (4 * c.tile - 4) > (int64_t)0xFFFFFFF0u) return SS_ERR_ARG;
    Images ims{};
    ims.one = make_imgdev(d_img, nullptr, nullptr, (int64_t)pitch, (int64_t)width, width, height, c.tile, 0, 0, 0);
    ims.count = 1;
    const Layout L = make_layout(nt, blocks_of(ims.one), seams_of(ims.one), c.tile, 1);
    // the counters live in the workspace header; nothing else is used
    if (!d_ws || ws_bytes < sizeof(WsHeader) || !aligned256(d_ws)) return SS_ERR_WORKSPACE;
    const Ws ws = make_ws(d_ws, L, c.threads);
    if (launch_profile(ims, ws, c.tile, stream) != hipSuccess) return SS_ERR_HIP;
    uint64_t v[5];
    if (hipMemcpyAsync(v, &ws.hdr->profile[0], sizeof(v), hipMemcpyDeviceToHost, stream) != hipSuccess)
        return SS_ERR_HIP;
    if (hipStreamSynchronize(stream) != hipSuccess) return SS_ERR_HIP;
    out->fragments = v[0];
    out->seam_fragments = v[1];
    out->seam_pixels = v[2];
    out->equal_edges = v[3];
    out->seam_edges = v[4];
    return SS_OK;
}

static ss_status batch_prepare(const ss_image_desc* images, int n, const ss_config& c,
                               std::vector<ImgDev>& table, Layout& L) {
    if (!images || n <= 0) return SS_ERR_ARG;
    if (!config_valid(c.tile, c.threads)) return SS_ERR_ARG;
    table.resize((size_t)n);
    int64_t tiles = 0, blocks = 0, seams = 0, groups = 0;
    for (int i = 0; i < n; ++i) {
        const ss_image_desc& d = images[i];
        if (d.width <= 0 || d.height <= 0 || d.pitch < (uint64_t)d.width || d.out_pitch < (uint64_t)d.width)
            return SS_ERR_ARG;
        ImgDev& t = table[(size_t)i];
        t = make_imgdev(d.img, d.max, d.min, (int64_t)d.pitch, (int64_t)d.out_pitch, d.width, d.height, c.tile,
                        tiles, blocks, seams);
        t.group_begin = groups;
        groups += groups_of(t);
        tiles += (int64_t)t.tiles_x * t.tiles_y;
        blocks += blocks_of(t);
        seams += seams_of(t);
    }
    if (tiles * (4 * c.tile - 4) > (int64_t)0xFFFFFFF0u) return SS_ERR_CAPACITY;
    L = make_layout(tiles, blocks, seams, c.tile, n, groups);
    return SS_OK;
}

ss_status ss_batch_workspace_size(const ss_image_desc* images, int n, const ss_config* cfg,
                                  size_t* bytes) {
    if (!bytes) return SS_ERR_ARG;
    std::vector<ImgDev> table;
    Layout L;
    const ss_status st = batch_prepare(images, n, resolve_config(cfg), table, L);
    if (st != SS_OK) return st;
    *bytes = L.bytes;
    return SS_OK;
}

static ss_status batch_enqueue(const ss_image_desc* images, int n, void* d_ws, size_t ws_bytes,
                               const ss_config& c, hipStream_t stream, bool stage_descs) {
    std::vector<ImgDev> table;
    Layout L;
    ss_status st = batch_prepare(images, n, c, table, L);
    if (st != SS_OK) return st;
    for (int i = 0; i < n; ++i)
        if (!images[i].img || !images[i].max || !images[i].min) return SS_ERR_ARG;
    if (!d_ws || ws_bytes < L.bytes || !aligned256(d_ws)) return SS_ERR_WORKSPACE;
    Images ims{};
    ims.count = n;
    if (n == 1) {
        ims.one = table[0];
        ims.table = nullptr;
    } else {
        ImgDev* dst = reinterpret_cast<ImgDev*>(static_cast<char*>(d_ws) + L.off_descs);
        if (stage_descs) {
            if (hipMemcpyAsync(dst, table.data(), table.size() * sizeof(ImgDev), hipMemcpyHostToDevice,
                               stream) != hipSuccess)
                return SS_ERR_HIP;
        }
        ims.table = dst;
    }
    const Ws ws = make_ws(d_ws, L, c.threads);
    const hipError_t e = launch_pipeline(ims, ws, c.tile, c.threads, device_cus(), stream);
    return e == hipSuccess ? SS_OK : SS_ERR_HIP;
}

ss_status ss_extrema_batch_u8(const ss_image_desc* images, int n, void* d_ws, size_t ws_bytes,
                              const ss_config* cfg, hipStream_t stream) {
    return batch_enqueue(images, n, d_ws, ws_bytes, resolve_config(cfg), stream, true);
}

ss_status ss_batch_plan_create(const ss_image_desc* images, int n, void* d_ws, size_t ws_bytes,
                               const ss_config* cfg, hipStream_t stream, ss_batch_plan** plan) {
    if (!plan) return SS_ERR_ARG;
    *plan = nullptr;
    const ss_config c = resolve_config(cfg);
    std::vector<ImgDev> table;
    Layout L;
    ss_status st = batch_prepare(images, n, c, table, L);
    if (st != SS_OK) return st;
    if (!d_ws || ws_bytes < L.bytes || !aligned256(d_ws)) return SS_ERR_WORKSPACE;
    // Stage the descriptors once, outside the graph.
    if (n > 1) {
        ImgDev* dst = reinterpret_cast<ImgDev*>(static_cast<char*>(d_ws) + L.off_descs);
        if (hipMemcpyAsync(dst, table.data(), table.size() * sizeof(ImgDev), hipMemcpyHostToDevice,
                           stream) != hipSuccess)
            return SS_ERR_HIP;
        if (hipStreamSynchronize(stream) != hipSuccess) return SS_ERR_HIP;
    }
    ss_batch_plan* p = new (std::nothrow) ss_batch_plan();
    if (!p) return SS_ERR_CAPACITY;
    (void)device_cus();  // cache the device query outside the capture
    // Capture on a private stream: the caller's stream may be the null stream,
    // which cannot be captured.  The graph is later launched on any stream.
    hipStream_t cap = nullptr;
    if (hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess) {
        delete p;
        return SS_ERR_HIP;
    }
    if (hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipStreamDestroy(cap);
        delete p;
        return SS_ERR_HIP;
    }
    st = batch_enqueue(images, n, d_ws, ws_bytes, c, cap, false);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(cap, &g);
    (void)hipStreamDestroy(cap);
    if (st != SS_OK || ec != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        delete p;
        return st != SS_OK ? st : SS_ERR_HIP;
    }
    p->graph = g;
    if (hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(g);
        delete p;
        return SS_ERR_HIP;
    }
    *plan = p;
    return SS_OK;
}

ss_status ss_batch_plan_run(ss_batch_plan* plan, hipStream_t stream) {
    if (!plan || !plan->exec) return SS_ERR_ARG;
    return hipGraphLaunch(plan->exec, stream) == hipSuccess ? SS_OK : SS_ERR_HIP;
}

ss_status ss_batch_plan_destroy(ss_batch_plan* plan) {
    if (!plan) return SS_OK;
    if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
    if (plan->graph) (void)hipGraphDestroy(plan->graph);
    delete plan;
    return SS_OK;
}

}  // extern "C"
