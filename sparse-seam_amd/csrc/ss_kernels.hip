// ss_kernels.hip -- gfx950 kernels of the exact dual regional-extrema path and
// the host-side sequencer that enqueues them.
//
// Reference behaviour followed (harness/primary/snapshot_sparse_seam_cuda_v1_0.cu):
//   tile-local Komura labels        :194-289  -> tile_label()
//   classify / number / emit        :291-475  -> tile_stage()
//   seam merge (3 pairs per seam px):508-581  -> k_seam_merge
//   compress + reduce               :583-619  -> k_resolve
//   materialise seam records        :477-506  -> k_mark + k_writeback
// The reference writes a 1-byte record for every pixel and re-reads it for
// every pixel; here masks are written once (provisionally for seam zones), the
// record plane is written only by tiles holding a fragment that can still
// change, and only the tiles whose fragments did change are streamed again.
#include "ss_device.hpp"
#include "ss_internal.hpp"
#include "ss_tile.hpp"

namespace ssx {

// Plain form: one workgroup per tile, XCD-aware tile order.
template <int T, int NT, bool MULTI>
__global__ __launch_bounds__(NT) void k_tile_stage(Images ims, Ws ws) {
    __shared__ TileSmem<T> sm;
    const int64_t gtile = xcd_block(blockIdx.x, gridDim.x);
    if (gtile >= ws.ntiles) return;
    const ImgDev im = locate<MULTI>(ims, ws.tile_img, gtile);
    tile_stage<T, NT>(sm, im, gtile, ws);
}

// Persistent form: a fixed grid drains a device-side tile work queue.  The
// tile range is split into one contiguous queue per XCD; a workgroup drains
// its own XCD's queue first and then steals from the others, so neighbouring
// tiles share an L2 while the load still balances across the chip.
template <int T, int NT, bool MULTI>
__global__ __launch_bounds__(NT) void k_tile_stage_persistent(Images ims, Ws ws, int chunk) {
    __shared__ TileSmem<T> sm;
    const int home = blockIdx.x % NUM_XCD;
#pragma unroll 1
    for (int q = 0; q < NUM_XCD; ++q) {
        const int xq = (home + q) % NUM_XCD;
        for (;;) {
            __syncthreads();
            if (threadIdx.x == 0) sm.next = atomicAdd(&ws.hdr->queue[xq], (uint32_t)chunk);
            __syncthreads();
            // values read from LDS are not known to be uniform; make them scalar
            const uint32_t start = __builtin_amdgcn_readfirstlane(sm.next);
            const uint32_t end_q = ws.hdr->queue_end[xq];
            if (start >= end_q) break;
            const uint32_t stop = min(start + (uint32_t)chunk, end_q);
#pragma unroll 1
            for (uint32_t t = start; t < stop; ++t) {
                const ImgDev im = locate<MULTI>(ims, ws.tile_img, t);
                tile_stage<T, NT>(sm, im, t, ws);
                __syncthreads();
            }
        }
    }
}

// Per-tile image index of a batch: one workgroup per image.
__global__ __launch_bounds__(256) void k_tile_map(Images ims, Ws ws) {
    const ImgDev im = ims.table[blockIdx.x];
    const int64_t n = (int64_t)im.tiles_x * im.tiles_y;
    for (int64_t t = threadIdx.x; t < n; t += blockDim.x) ws.tile_img[im.tile_begin + t] = blockIdx.x;
}

// Resets the work queues of the persistent tile stage (graph-safe: a kernel,
// not a memset node).
__global__ void k_queue_reset(Ws ws) {
    const int i = threadIdx.x;
    if (i < NUM_XCD) {
        const int64_t n = ws.ntiles;
        const int64_t b = n * i / NUM_XCD, e = n * (i + 1) / NUM_XCD;
        ws.hdr->queue[i] = (uint32_t)b;
        ws.hdr->queue_end[i] = (uint32_t)e;
    }
}

// ---- compaction: exclusive scan of the per-tile counts --------------------

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_wave, uint32_t& total) {
    constexpr int W = SCAN_BLOCK / 64;
    const uint32_t inc = wave_incl_scan(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 63) s_wave[wid] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int i = 0; i < W; ++i) {
            const uint32_t t = s_wave[i];
            s_wave[i] = acc;
            acc += t;
        }
        s_wave[W] = acc;
    }
    __syncthreads();
    total = s_wave[W];
    const uint32_t r = s_wave[wid] + inc - v;
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_reduce(Ws ws) {
    __shared__ uint32_t s_wave[SCAN_BLOCK / 64 + 1];
    __shared__ uint32_t s_err;
    if (threadIdx.x == 0) s_err = 0;
    const int64_t t0 = (int64_t)blockIdx.x * SCAN_CHUNK + threadIdx.x * SCAN_ITEMS;
    uint32_t s = 0, e = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        if (t0 + i < ws.ntiles) {
            const uint32_t c = ws.counts[t0 + i];
            s += c & COUNT_MASK;
            e |= c >> COUNT_ERR_SHIFT;
        }
    }
    uint32_t total;
    block_excl_scan(s, s_wave, total);
    if (e) atomicOr(&s_err, e);
    __syncthreads();
    if (threadIdx.x == 0) {
        ws.bsum[blockIdx.x] = total;
        ws.berr[blockIdx.x] = s_err;
    }
}

// Single workgroup: scan of the block sums; also folds the tile-stage
// invariant word into the run's word and clears the per-run counters.
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_top(Ws ws) {
    __shared__ uint32_t s_wave[SCAN_BLOCK / 64 + 1];
    __shared__ uint32_t s_err;
    if (threadIdx.x == 0) s_err = 0;
    const uint32_t nb = ws.nblocks;
    uint32_t carry = 0, err = 0;
    for (uint32_t b0 = 0; b0 < nb; b0 += SCAN_BLOCK) {
        const uint32_t i = b0 + threadIdx.x;
        const uint32_t v = i < nb ? ws.bsum[i] : 0u;
        if (i < nb) err |= ws.berr[i];
        uint32_t total;
        const uint32_t ex = block_excl_scan(v, s_wave, total);
        if (i < nb) ws.bsum[i] = carry + ex;
        carry += total;
    }
    if (err) atomicOr(&s_err, err);
    __syncthreads();
    if (threadIdx.x == 0) {
        ws.bsum[nb] = carry;
        WsHeader* h = ws.hdr;
        h->magic = WS_MAGIC;
        h->total = carry;
        h->error = s_err;
        h->dirty_count = 0;
        h->retries = 0;
        h->tiles = (uint64_t)ws.ntiles;
        h->tile = ws.tile;
        h->threads = ws.threads;
    }
}

// Per scan block and part: the block's scan is recomputed by each of its
// APPLY_PARTS workgroups (1024 counts, cheap); part 0 writes the tile bases and
// part k initialises the dense union-find arrays of the k-th slice of the
// block's tiles (one wave per tile, lanes over the tile's fragments), so the
// fragment copy is spread over many CUs instead of one workgroup per block.
constexpr int APPLY_PARTS = 16;
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_apply(Ws ws) {
    __shared__ uint32_t s_wave[SCAN_BLOCK / 64 + 1];
    __shared__ uint32_t s_base[SCAN_CHUNK];
    __shared__ uint32_t s_cnt[SCAN_CHUNK];
    const uint32_t blk = blockIdx.x / APPLY_PARTS, part = blockIdx.x % APPLY_PARTS;
    const int64_t tb = (int64_t)blk * SCAN_CHUNK;
    const int64_t t0 = tb + threadIdx.x * SCAN_ITEMS;
    uint32_t c[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        c[i] = (t0 + i < ws.ntiles) ? (ws.counts[t0 + i] & COUNT_MASK) : 0u;
        s += c[i];
    }
    uint32_t total;
    uint32_t run = block_excl_scan(s, s_wave, total) + ws.bsum[blk];
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        const int li = threadIdx.x * SCAN_ITEMS + i;
        s_base[li] = run;
        s_cnt[li] = c[i];
        if (part == 0 && t0 + i < ws.ntiles) ws.base[t0 + i] = run;
        run += c[i];
    }
    __syncthreads();
    const int S = ws.slots;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    constexpr int SLICE = SCAN_CHUNK / APPLY_PARTS;
    for (int li = (int)part * SLICE + wid; li < (int)(part + 1) * SLICE; li += SCAN_BLOCK / 64) {
        const uint32_t cnt = s_cnt[li];
        if (!cnt) continue;
        const uint32_t b = s_base[li];
        const int64_t slot0 = (tb + li) * S;
        for (uint32_t j = lane; j < cnt; j += 64) {
            ws.dparent[b + j] = b + j;
            ws.dbits[b + j] = ws.sbits[slot0 + j];
        }
    }
}

// ---- seam merge -----------------------------------------------------------

// One thread per (tile, k): the tile's right-column pixel k against the three
// pixels right of it, and its bottom-row pixel k against the three below it.
// Together these visit every equal pixel pair that straddles a tile seam.
// A plateau crossing a seam yields the same fragment pair at many consecutive
// k, so a pair already present in the thread or in the neighbouring lane (the
// previous k of the same tile edge) is not united again.
template <int T, bool MULTI>
__global__ __launch_bounds__(256) void k_seam_merge(Images ims, Ws ws) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t gtile = g / T;
    const int k = (int)(g % T);
    const int lane = threadIdx.x & 63;
    uint32_t retries = 0;
    bool bad = false;
    constexpr uint32_t NONE = 0xFFFFFFFFu;
    uint32_t ra = NONE, rb[3] = {NONE, NONE, NONE};   // right edge: fa and up to 3 fb
    uint32_t ba = NONE, bb[3] = {NONE, NONE, NONE};   // bottom edge
    const uint32_t cnt = gtile < ws.ntiles ? (ws.counts[gtile] & COUNT_MASK) : 0u;
    if (cnt != 0) {
        const ImgDev im = locate<MULTI>(ims, ws.tile_img, gtile);
        const int64_t lt = gtile - im.tile_begin;
        const int tx = (int)(lt % im.tiles_x), ty = (int)(lt / im.tiles_x);
        const uint8_t* eq = ws.eq + gtile * 2 * T;
        const uint8_t* rec = ws.rec + gtile * 4 * T;
        const uint32_t base = ws.base[gtile];
        if (tx + 1 < im.tiles_x) {
            const uint32_t e = eq[k];
            if (e) {
                const uint32_t ja = rec[3 * T + k];
                if (ja < cnt) ra = base + ja; else bad = true;
                const int y = ty * T + k;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    if (!(e & (1u << d)) || ra == NONE) continue;
                    const int ny = y + d - 1;
                    const int nty = ny / T;
                    const int64_t nt = im.tile_begin + (int64_t)nty * im.tiles_x + tx + 1;
                    const uint32_t jb = ws.rec[nt * 4 * T + 2 * T + (ny - nty * T)];
                    if (jb >= (ws.counts[nt] & COUNT_MASK)) { bad = true; continue; }
                    rb[d] = ws.base[nt] + jb;
                }
            }
        }
        if (ty + 1 < im.tiles_y) {
            const uint32_t e = eq[T + k];
            if (e) {
                const uint32_t ja = rec[T + k];
                if (ja < cnt) ba = base + ja; else bad = true;
                const int x = tx * T + k;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    if (!(e & (1u << d)) || ba == NONE) continue;
                    const int nx = x + d - 1;
                    const int ntx = nx / T;
                    const int64_t nt = im.tile_begin + (int64_t)(ty + 1) * im.tiles_x + ntx;
                    const uint32_t jb = ws.rec[nt * 4 * T + (nx - ntx * T)];
                    if (jb >= (ws.counts[nt] & COUNT_MASK)) { bad = true; continue; }
                    bb[d] = ws.base[nt] + jb;
                }
            }
        }
    }
    // pairs of the previous k (lane - 1) on the same kind of edge
    const uint32_t pra = __shfl_up(ra, 1, 64), pba = __shfl_up(ba, 1, 64);
    uint32_t prb[3], pbb[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        prb[d] = __shfl_up(rb[d], 1, 64);
        pbb[d] = __shfl_up(bb[d], 1, 64);
    }
    const bool has_prev = lane > 0 && k > 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (rb[d] != NONE) {
            bool dup = (d > 0 && rb[d] == rb[d - 1]) || (d > 1 && rb[d] == rb[d - 2]);
            if (has_prev && pra == ra) dup = dup || rb[d] == prb[0] || rb[d] == prb[1] || rb[d] == prb[2];
            if (!dup) gunion(ws.dparent, ra, rb[d], retries);
        }
        if (bb[d] != NONE) {
            bool dup = (d > 0 && bb[d] == bb[d - 1]) || (d > 1 && bb[d] == bb[d - 2]);
            if (has_prev && pba == ba) dup = dup || bb[d] == pbb[0] || bb[d] == pbb[1] || bb[d] == pbb[2];
            if (!dup) gunion(ws.dparent, ba, bb[d], retries);
        }
    }
    const uint32_t r = wave_sum(retries);
    if (lane == 0 && r) atomicAdd((unsigned long long*)&ws.hdr->retries, (unsigned long long)r);
    // a perimeter record out of range would mean a broken tile stage: report
    // it instead of following it
    if (bad) atomicOr(&ws.hdr->error, 2u);
}

// ---- resolve: compress + OR local evidence into the roots ----------------

// A component can hold millions of fragments (a near-percolating plateau), so
// OR-ing every fragment's evidence into the one root word would serialise on
// that address.  Each wave therefore issues one atomic per distinct root and
// only for bits the (possibly stale, never wrong) cached root word lacks.
__global__ __launch_bounds__(256) void k_resolve(Ws ws) {
    const uint32_t total = ws.hdr->total;
    const int lane = threadIdx.x & 63;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t wbase = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); wbase < total; wbase += stride) {
        const uint32_t i = wbase + lane;
        uint32_t r = 0, lb = 0;
        if (i < total) {
            // halving walk: compresses the trees for the other threads; a later
            // halving store may replace this entry with an ancestor, so
            // consumers finish the walk with gfind_ro
            r = gfind(ws.dparent, i);
            if (r != i) ws.dparent[i] = r;
            lb = ws.dbits[i] & 3u;
            if (lb && ((ws.dbits[r] >> 2) & lb) == lb) lb = 0;
        }
        uint64_t pending = __ballot(lb != 0);
        while (pending) {
            const int leader = __ffsll((unsigned long long)pending) - 1;
            const uint32_t rl = __shfl(r, leader, 64);
            const bool mine = lb != 0 && r == rl;
            const uint64_t m = __ballot(mine);
            const uint32_t bits = (__ballot(mine && (lb & 1u)) ? 1u : 0u) | (__ballot(mine && (lb & 2u)) ? 2u : 0u);
            if (lane == leader)
                __hip_atomic_fetch_or(ws.dbits + rl, bits << 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (mine) lb = 0;
            pending &= ~m;
        }
    }
}

// ---- mark: fragments whose output changes, dirty-tile list ---------------

__global__ __launch_bounds__(256) void k_mark(Ws ws) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool dirty = false;
    if (t < ws.ntiles) {
        const uint32_t cnt = ws.counts[t] & COUNT_MASK;
        const uint32_t b = ws.base[t];
        for (uint32_t j = 0; j < cnt; ++j) {
            const uint32_t i = b + j;
            const uint32_t own = ws.dbits[i];
            const uint32_t res = (ws.dbits[gfind_ro(ws.dparent, i)] >> 2) & 3u;
            if (res & ~own & 3u) {
                ws.dbits[i] = own | DB_CHANGED;
                dirty = true;
            }
        }
    }
    const uint64_t m = __ballot(dirty);
    if (!m) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t slot0 = 0;
    if (lane == leader) slot0 = atomicAdd(&ws.hdr->dirty_count, (uint32_t)__popcll(m));
    slot0 = __shfl(slot0, leader, 64);
    if (dirty) {
        const uint32_t rank = __popcll(m & ((1ull << lane) - 1ull));
        ws.dirty[slot0 + rank] = (uint32_t)t;
    }
}

// ---- write-back of changed seam zones ------------------------------------

template <int T, int NT, bool MULTI>
__global__ __launch_bounds__(NT) void k_writeback(Images ims, Ws ws) {
    __shared__ TileSmem<T> sm;
    const uint32_t nd = ws.hdr->dirty_count;
#pragma unroll 1
    for (uint32_t d = blockIdx.x; d < nd; d += gridDim.x) {
        const int64_t gtile = ws.dirty[d];
        const ImgDev im = locate<MULTI>(ims, ws.tile_img, gtile);
        tile_writeback<T, NT>(sm, im, gtile, ws);
    }
}

// ---- host sequencer -------------------------------------------------------

template <int T, int NT, bool MULTI>
static void launch_tile_kernels(const Images& ims, const Ws& ws, bool persistent, int grid_cap,
                                hipStream_t s) {
    if (persistent) {
        k_queue_reset<<<1, 64, 0, s>>>(ws);
        const int64_t grid = std::min<int64_t>(ws.ntiles, grid_cap);
        // claims of several tiles keep the queue atomics off the critical path
        const int chunk = (int)std::max<int64_t>(1, std::min<int64_t>(32, ws.ntiles / (grid * 16)));
        k_tile_stage_persistent<T, NT, MULTI><<<(unsigned)grid, NT, 0, s>>>(ims, ws, chunk);
    } else {
        k_tile_stage<T, NT, MULTI><<<(unsigned)ws.ntiles, NT, 0, s>>>(ims, ws);
    }
}

template <int T, int NT, bool MULTI>
static void launch_writeback(const Images& ims, const Ws& ws, int grid_cap, hipStream_t s) {
    const int64_t grid = std::min<int64_t>(ws.ntiles, grid_cap);
    k_writeback<T, NT, MULTI><<<(unsigned)grid, NT, 0, s>>>(ims, ws);
}

template <int T, bool MULTI>
static void launch_merge(const Images& ims, const Ws& ws, hipStream_t s) {
    const int64_t merge_threads = ws.ntiles * T;
    const unsigned merge_grid = (unsigned)((merge_threads + 255) / 256);
    k_seam_merge<T, MULTI><<<merge_grid, 256, 0, s>>>(ims, ws);
}

#define SSX_DISPATCH(T_, NT_, CALL)                                        \
    switch ((T_) * 1000 + (NT_)) {                                         \
    case 16064: CALL(16, 64); break;                                       \
    case 16128: CALL(16, 128); break;                                      \
    case 16256: CALL(16, 256); break;                                      \
    case 32064: CALL(32, 64); break;                                       \
    case 32128: CALL(32, 128); break;                                      \
    case 32256: CALL(32, 256); break;                                      \
    case 32512: CALL(32, 512); break;                                      \
    case 64064: CALL(64, 64); break;                                       \
    case 64128: CALL(64, 128); break;                                      \
    case 64256: CALL(64, 256); break;                                      \
    case 64512: CALL(64, 512); break;                                      \
    default: return hipErrorInvalidValue;                                  \
    }

template <bool MULTI>
static hipError_t launch_all(const Images& ims, const Ws& ws, int T, int NT, bool persistent,
                             int cus, hipStream_t s) {
    const int grid_cap = std::max(1, cus) * 8;
    if (MULTI) k_tile_map<<<(unsigned)ims.count, 256, 0, s>>>(ims, ws);
    const ProfileHook& hook = profile_hook();
    if (hook.tile_begin) (void)hipEventRecord(hook.tile_begin, s);
#define CALL_TILE(a, b) launch_tile_kernels<a, b, MULTI>(ims, ws, persistent, grid_cap, s)
    SSX_DISPATCH(T, NT, CALL_TILE)
#undef CALL_TILE
    if (hook.tile_end) (void)hipEventRecord(hook.tile_end, s);
    const uint32_t nb = ws.nblocks;
    k_scan_reduce<<<nb, SCAN_BLOCK, 0, s>>>(ws);
    k_scan_top<<<1, SCAN_BLOCK, 0, s>>>(ws);
    k_scan_apply<<<nb * APPLY_PARTS, SCAN_BLOCK, 0, s>>>(ws);
    switch (T) {
    case 16: launch_merge<16, MULTI>(ims, ws, s); break;
    case 32: launch_merge<32, MULTI>(ims, ws, s); break;
    case 64: launch_merge<64, MULTI>(ims, ws, s); break;
    default: return hipErrorInvalidValue;
    }
    const int64_t max_frag = ws.ntiles * (int64_t)ws.slots;
    const unsigned resolve_grid =
        (unsigned)std::max<int64_t>(1, std::min<int64_t>((max_frag + 255) / 256, (int64_t)grid_cap * 2));
    k_resolve<<<resolve_grid, 256, 0, s>>>(ws);
    k_mark<<<(unsigned)((ws.ntiles + 255) / 256), 256, 0, s>>>(ws);
#define CALL_WB(a, b) launch_writeback<a, b, MULTI>(ims, ws, grid_cap, s)
    SSX_DISPATCH(T, NT, CALL_WB)
#undef CALL_WB
    return hipGetLastError();
}

hipError_t launch_pipeline(const Images& ims, const Ws& ws, int T, int NT, bool persistent,
                           int cus, hipStream_t s) {
    if (ims.count > 1) return launch_all<true>(ims, ws, T, NT, persistent, cus, s);
    return launch_all<false>(ims, ws, T, NT, persistent, cus, s);
}

}  // namespace ssx
