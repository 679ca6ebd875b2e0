// ss_device.hpp -- device-side building blocks of the dual regional-extrema
// pipeline for gfx950 (wave64, 160 KB LDS per CU, 8 XCDs).
//
// Pipeline for one launch set (single image or a whole batch):
//   1. tile stage   one workgroup per T x T tile: stage the tile plus a
//                   one-pixel halo in LDS, label tile-local flat zones
//                   (Komura init + min-root union-find in LDS), OR the
//                   greater / smaller / touches-seam evidence into each zone
//                   root, write both masks for every pixel (final for
//                   tile-contained zones, provisional for seam zones), number
//                   the seam zones ("fragments") and publish per perimeter
//                   pixel its fragment and its cross-seam equality bits.
//   2. compaction   exclusive scan of the per-tile fragment counts gives every
//                   fragment a dense global id.
//   3. seam merge   a) per block of 8 x 8 tiles: union-find in LDS over the
//                   seams inside the block; every fragment's dense parent
//                   becomes its block-local root (dense arrays initialised);
//                   b) per group of 2 x 2 blocks: union-find in LDS over the
//                   block roots that meet across the two block seams inside
//                   the group, linked with plain stores (one workgroup owns
//                   the group's roots);
//                   c) lock-free min-root CAS union with path halving over the
//                   dense ids for the seams between groups, one thread per
//                   pixel of a tile edge on a block boundary; the edges of a
//                   block seam are consecutive and a workgroup unites each
//                   distinct pair of block roots once.
//   4. resolve      full compression and OR of the local evidence into roots;
//                   then one thread per fragment turns its zone's resolved
//                   evidence into the fragment's final masks and change bits
//                   (one byte per dense id).
//   5. mark and     one wave per tile holding a fragment that may still change:
//      write-back   the tile's change bytes are read; if any fragment changed,
//                   the tile streams its 1-byte per-pixel record plane
//                   (written in step 1 only by such tiles: fragment ids, and
//                   the final masks of the other pixels) and rewrites every
//                   tile row holding a changed pixel.
//
// Everything is exact integer work; no MFMA (nothing here is a contraction).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ss_seam.hpp"

namespace ssx {

constexpr uint32_t F_G = 1u;      // a strictly greater 8-neighbour exists
constexpr uint32_t F_S = 2u;      // a strictly smaller 8-neighbour exists
constexpr uint32_t F_SEAM = 4u;   // an equal 8-neighbour lies in another tile
constexpr uint32_t LMASK = 0x3FFFu;     // label bits (T*T <= 16384: the profiler also runs T = 128)
constexpr int FLAG_SHIFT = 16;          // zone evidence OR'ed at the root
constexpr int FRAG_SHIFT = 20;          // fragment id + 1 stored at seam roots
constexpr uint32_t DB_BROOT = 16u;      // dense bits: root of a block-local component
constexpr int SCAN_BLOCK = 1024;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_CHUNK = SCAN_BLOCK * SCAN_ITEMS;  // tiles per scan block
constexpr int NUM_XCD = 8;
constexpr int MERGE_THREADS = 1024;

// One image of a launch set.  tile_begin is the global id of its first tile,
// block_begin that of its first merge block (MERGE_BT x MERGE_BT tiles).
struct ImgDev {
    const uint8_t* img;
    uint8_t* mx;
    uint8_t* mn;
    int64_t pitch;
    int64_t opitch;
    int32_t w, h, tiles_x, tiles_y;
    int64_t tile_begin;
    int64_t block_begin;
    int32_t blocks_x, blocks_y;
    // seam-merge units: the right edges of the last tile column of every
    // merge-block column but the last (seam_cols of them, tiles_y tiles
    // each), then the bottom edges of the last tile row of every merge-block
    // row but the last (seam_rows x tiles_x)
    int64_t seam_begin;
    int32_t seam_cols, seam_rows;
    // global index of its first group of 2 x 2 merge blocks (group merge;
    // groups row-major, ceil(blocks_x / 2) per row)
    int64_t group_begin;
};

struct Images {
    ImgDev one;           // count == 1: the image itself (table may be null)
    const ImgDev* table;  // count > 1: device table sorted by tile_begin
    int count;
};

// Workspace header (first 256 bytes of every workspace).
struct WsHeader {
    uint32_t magic;
    uint32_t error;        // invariant word of the last run
    uint32_t reserved;
    uint32_t dirty_count;  // tiles rewritten by the write-back
    uint32_t total;        // fragments in the dense array
    int32_t tile;
    int32_t threads;
    uint32_t nblocks;      // scan blocks
    uint64_t retries;      // failed CAS in the seam union-find
    uint64_t tiles;
    uint64_t profile[5];   // seam-sparsity counters (ss_seam_profile_u8 only)
    uint32_t pad[64 - 10 - 2 - 10];
};
static_assert(sizeof(WsHeader) == 256, "header must be 256 bytes");

// counts[t] holds the tile's seam-fragment count in bits 0-23, two tile flags,
// and the tile-stage invariant code in bits 28-31 (folded into
// WsHeader::error by the scan, so no per-run clearing is needed).
constexpr uint32_t COUNT_MASK = 0x00FFFFFFu;
constexpr uint32_t CNT_UNIFORM = 1u << 24;   // the tile is one zone (fragment 0 if seam)
constexpr uint32_t CNT_OPEN = 1u << 25;      // the tile wrote its per-pixel record plane
constexpr uint32_t CNT_ROWS = 1u << 26;      // only the plane rows in orows[t] were written
constexpr int COUNT_ERR_SHIFT = 28;
// Record-plane byte of a pixel outside every open fragment: REC_CLOSED plus
// its final mask bits (bit 0 max, bit 1 min).  Fragment ids stay below it
// (at most 4T - 4 = 252 slots, ids 0 .. 251).
constexpr uint32_t REC_CLOSED = 0xFCu;

struct Ws {
    WsHeader* hdr;
    uint32_t* counts;    // [tiles] seam fragments per tile (+ invariant code)
    uint32_t* base;      // [tiles] dense id of the tile's first fragment
    uint64_t* bstate;    // [nblocks] look-back state of each scan block: flag (bits
                         //   32-33: 0 none yet, 1 its own sum, 2 inclusive prefix) and
                         //   value (bits 0-31); cleared by the tile stage
    uint8_t* sbits;      // [tiles * S] local G/S evidence per fragment slot (read by the
                         //   block merge only)
    uint8_t* flut;       // [tiles * S] per dense id, once resolved: final max / min mask
                         //   (bits 0 / 1) and whether each changed (bits 6 / 7); the
                         //   sbits memory, dead by then
    uint8_t* rec8;       // [tiles * T*T] per-pixel record (open tiles only): the
                         //   pixel's open fragment, or REC_CLOSED | its final masks
    uint64_t* orows;     // [tiles] with CNT_ROWS: the record-plane rows that hold an
                         //   open pixel (bit r = tile row r), the only rows written
    uint8_t* rec;        // [tiles * 4T] fragment id of seam perimeter pixels
                         //   [0,T) top row, [T,2T) bottom row, [2T,3T) left col, [3T,4T) right col
    uint8_t* eq;         // [tiles * 2T] cross-seam equality bits: [0,T) right col, [T,2T) bottom row
    uint32_t* dparent;   // [tiles * S] dense union-find parents
    uint32_t* dbits;     // [tiles * S] bits 0-1 local, 2-3 component OR (roots),
                         //   4 block root
    uint32_t* tile_img;  // [tiles] image index of each tile (batches only)
    uint32_t* block_img; // [nblk] image index of each merge block (batches only)
    uint32_t* seam_img;  // [nseam] image index of each seam-merge unit (batches only)
    uint32_t* group_img; // [ngroup] image index of each merge-block group (batches only)
    int64_t ntiles;
    int64_t nblk;        // merge blocks
    int64_t nseam;       // seam-merge units (one tile edge each)
    int64_t ngroup;      // groups of 2 x 2 merge blocks
    uint32_t nblocks;    // scan blocks = ceil(ntiles / SCAN_CHUNK)
    int tile;            // T
    int threads;         // threads per tile
    int slots;           // S = 4T - 4
};

constexpr uint32_t WS_MAGIC = 0x53535831u;  // "SSX1"

// Image of a global tile id.  MULTI is a template parameter so that the
// single-image kernels read the descriptor straight from the kernel
// arguments; batched kernels look it up through the per-tile image index.
template <bool MULTI>
__device__ __forceinline__ ImgDev locate(const Images& ims, const uint32_t* tile_img, int64_t tile) {
    if constexpr (!MULTI) {
        return ims.one;
    } else {
        return ims.table[tile_img[tile]];
    }
}

// XCD-aware logical block id: workgroups are dispatched round-robin over the
// 8 XCDs, so block b runs on XCD b % 8.  Give each XCD a contiguous run of
// logical ids so that horizontally adjacent tiles (which share image and mask
// cache lines) are processed by the same L2.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t g) {
    const int64_t per = g / NUM_XCD, rem = g % NUM_XCD;
    const int64_t k = b % NUM_XCD, i = b / NUM_XCD;
    return (k < rem ? k * (per + 1) : rem * (per + 1) + (k - rem) * per) + i;
}

// ---- LDS union-find (workgroup scope) ------------------------------------

__device__ __forceinline__ uint32_t lds_load(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Read-only LDS find (plain labels).
__device__ __forceinline__ uint32_t lfind(const uint32_t* lab, uint32_t x) {
    uint32_t r = lds_load(lab + x);
    for (;;) {
        const uint32_t n = lds_load(lab + r);
        if (n == r) return r;
        r = n;
    }
}

// Min-root LDS union (plain labels, no flag bits).  The two walks advance
// together, one path-halving step each per round, so a union waits for the
// deeper walk rather than for both in turn; a failed CAS resumes from the
// parent it returns.
__device__ __forceinline__ void lunion_pair(uint32_t* lab, uint32_t a, uint32_t b) {
    for (;;) {
        for (;;) {
            if (a == b) return;
            const uint32_t pa = lds_load(lab + a), pb = lds_load(lab + b);
            if (pa == a && pb == b) break;
            // at a root, lab[pa] is pa itself and nothing is stored
            const uint32_t ga = lds_load(lab + pa), gb = lds_load(lab + pb);
            if (ga != pa) __hip_atomic_store(lab + a, ga, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (gb != pb) __hip_atomic_store(lab + b, gb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            a = ga;
            b = gb;
        }
        const uint32_t hi = a > b ? a : b, lo = a > b ? b : a;
        uint32_t expected = hi;
        if (__hip_atomic_compare_exchange_strong(lab + hi, &expected, lo, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))
            return;
        a = expected;
        b = lo;
    }
}

// ---- global union-find over dense fragment ids ----------------------------
//
// Walks use ordinary (cached) loads and plain halving stores.  A cached parent
// may be stale, but it is always an ancestor that was current at some point:
// a link only ever turns a root into a child of a smaller id, and halving only
// rewrites a non-root (which never becomes a root again) to one of its
// ancestors.  Links are agent-scope CAS on a node believed to be a root.  A
// failed CAS returns that node's current parent, a strictly smaller id, and
// the next walk resumes from it, so every retry lowers max(a, b) and the
// union terminates.

__device__ __forceinline__ uint32_t gfind(uint32_t* parent, uint32_t x) {
    for (;;) {
        const uint32_t p = parent[x];
        if (p == x) return x;
        const uint32_t gp = parent[p];
        if (gp == p) return p;
        parent[x] = gp;
        x = gp;
    }
}

// Read-only walk to the root (for consumers after all links are done).
__device__ __forceinline__ uint32_t gfind_ro(const uint32_t* parent, uint32_t x) {
    for (;;) {
        const uint32_t p = parent[x];
        if (p == x) return x;
        x = p;
    }
}

// Min-root union: the larger root is linked under the smaller one by CAS.
__device__ __forceinline__ void gunion(uint32_t* parent, uint32_t a, uint32_t b, uint32_t& retries) {
    for (;;) {
        a = gfind(parent, a);
        b = gfind(parent, b);
        if (a == b) return;
        const uint32_t hi = a > b ? a : b, lo = a > b ? b : a;
        uint32_t expected = hi;
        if (__hip_atomic_compare_exchange_strong(parent + hi, &expected, lo, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            return;
        ++retries;
        a = expected;
        b = lo;
    }
}

// ---- wave / block scans ---------------------------------------------------

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t n = __shfl_up(v, d, 64);
        if (lane >= d) v += n;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

}  // namespace ssx
