// ss_rowtile.hpp -- row tile stage: the tile-local step of the pipeline for
// T = 64 with 128 threads, formulated on 32-bit row masks.
//
// Mapping: one workgroup of two waves per 64 x 64 tile; wave h holds columns
// 32h .. 32h + 31, lane y holds tile row y.  Bit b of a lane's masks is
// column 32h + b of its row.
//
//   load      the tile plus its one-pixel halo into LDS (16-byte loads);
//   classify  each lane reads its row and the rows above and below from LDS
//             and computes, four pixels per dword, the greater / smaller
//             evidence and its equal right / down / down-left / down-right
//             neighbours; the up masks are the down masks of the lane above
//             (DPP wave shift).  Seam bits: equal neighbours outside the tile
//             and inside the image.
//   runs      non-singleton pixels joined by equal right neighbours form
//             runs; only a run's first pixel is a union-find node.  Unions
//             (16-bit LDS parents, min-root CAS) follow row_edges: one per
//             change of run above along a run, plus the three cross-half
//             relations at columns 31 | 32 (ss_rows.hpp, checked on the host
//             against a breadth-first labelling).
//   evidence  a run's greater / smaller / seam bits reach its start by a
//             backward fill along the run and are OR'ed into its root.
//   output    seam roots are numbered (fragments); the masks are written
//             (final for tile-contained zones, provisional for seam zones);
//             if the tile holds an open fragment, its per-pixel record plane
//             is built in LDS and streamed out; perimeter records and
//             cross-seam equality bits as in the quad tile stage.
//
// The outputs (counts word, fragment evidence, perimeter records, cross-seam
// bits, record plane, masks) are those of tile_stage in ss_tile.hpp, so every
// later kernel is shared.
#pragma once
#include "ss_device.hpp"
#include "ss_rows.hpp"
#include "ss_tile.hpp"

namespace ssx {

constexpr int RT = 64;            // tile side
constexpr int RT_THREADS = 128;   // two waves per tile
constexpr int RT_NTW = 1;         // tiles per workgroup
constexpr int RT_P = 72;          // staged row pitch: tile column c of staged row r at r * RT_P + 16 + c
                                  // (a multiple of 8: the LDS footprint lets 12 workgroups share a CU)
constexpr int RT_ROWS = 66;       // staged rows: tile rows -1 .. 64
// Union queues (two waves) in the staged rows' LDS after the GS words
// (the rows are no longer read when the unions run).
constexpr int RT_UQ_OFF = 512;
constexpr int RT_UQ = 528;        // node pairs per wave

// Union-find nodes are indexed column-major, index 64 x + y for pixel (y, x),
// so the entries the 64 lanes (rows) touch at one column are consecutive
// halfwords, not one per 128-byte row (which all fall in the same LDS bank).
// A node is named by its entry's byte offset in the label array (twice its
// index), and parents are stored in that form: a loaded parent is the next
// address as it is.
__device__ __forceinline__ uint32_t rnode(int y, int x) { return (uint32_t)((x << 7) | (y << 1)); }
constexpr int RL_COL = 7;       // rnode(y, x + 1) = rnode(y, x) + (1 << RL_COL)
constexpr uint32_t RL_ROW = 2;  // rnode(y + 1, x) = rnode(y, x) + RL_ROW

// Label entries (16 bits): the parent's offset (bits 1-12; bit 0 clear) while
// the unions run; evidence bits 13-15 OR'ed at roots; once numbered, a root
// holds RL_ROOT | fragment id << 1 | evidence (id RL_NOID if it is not a seam
// zone).
constexpr uint32_t RL_IDX = 0x1FFEu;
constexpr int RL_FSHIFT = 13;
constexpr uint32_t RL_ROOT = 1u;
constexpr uint32_t RL_NOID = 0x0FFFu;
__device__ __forceinline__ uint32_t rl_id(uint32_t e) { return (e >> 1) & 0xFFFu; }

struct RowTileSmem {
    union {
        alignas(16) uint8_t img[RT_ROWS * RT_P + 16];   // staged rows (+ the last row's right halo dword)
        alignas(16) uint8_t plane[RT * RT_P];           // record plane (row pitch RT_P), built after the staged rows are read
    };
static_assert(RT_UQ_OFF + 2 * RT_UQ * 4 <= RT_ROWS * RT_P + 16, "union queues must fit in the staged rows");
    alignas(16) uint16_t lab[RT * RT];
    // edge pixels at 255 / at 0: bit i = pixel i of tile row 0, row th - 1,
    // column 0, column tw - 1
    uint64_t e255[4], e0[4];
    uint32_t count, open, err, zflags, nonuni, anygs, orow[2], qn[2];
};

// Entry loads are plain (an atomic 16-bit load costs a masking instruction
// after every ds_read_u16): every walk reads a different entry each step, and
// a stale read is harmless where the entries may change meanwhile (a find
// that stops at a node that is no longer a root has its CAS fail and walks
// again; an evidence test that misses bits only repeats an atomic OR).
__device__ __forceinline__ uint32_t rl_ld(const uint16_t* lab, uint32_t o) {
    return *reinterpret_cast<const uint16_t*>(reinterpret_cast<const uint8_t*>(lab) + o);
}
__device__ __forceinline__ void rl_st(uint16_t* lab, uint32_t o, uint32_t v) {
    __hip_atomic_store(reinterpret_cast<uint16_t*>(reinterpret_cast<uint8_t*>(lab) + o), (uint16_t)v,
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// The 32-bit word holding entry o, and the entry's bit position in it.
__device__ __forceinline__ uint32_t* rl_word(uint16_t* lab, uint32_t o) {
    return reinterpret_cast<uint32_t*>(reinterpret_cast<uint8_t*>(lab) + (o & ~3u));
}
__device__ __forceinline__ int rl_shift(uint32_t o) { return (int)(o & 2u) * 8; }

// Find with path halving (unions only: entries hold plain parents).
__device__ __forceinline__ uint32_t rl_find_h(uint16_t* lab, uint32_t x) {
    uint32_t p = rl_ld(lab, x);
    if (p == x) return x;
    for (;;) {
        const uint32_t gp = rl_ld(lab, p);
        if (gp == p) return p;
        rl_st(lab, x, gp);
        x = gp;
        p = rl_ld(lab, x);
    }
}

// Run starts whose run holds both a strictly greater and a strictly smaller
// neighbour ("GS runs") are marked per row and half in a 32-bit word of an
// LDS array (owngs[2 y + h], bit b = column 32 h + b).  A node's key orders
// GS runs before the others, then by row, then by column.
// Every link goes from a larger key to a smaller one, so the root of a tree
// holds the smallest key in it: a zone's root is a GS run whenever the zone
// has one, and such a zone has both kinds of evidence.
__device__ __forceinline__ uint32_t rl_gs(const uint32_t* owngs, uint32_t o) {
    const uint32_t y = (o >> 1) & 63u, x = o >> 7;
    return (owngs[2 * y + (x >> 5)] >> (x & 31u)) & 1u;
}
// Row before column: the offset rotated right by 7 bits is y << 26 | x.
__device__ __forceinline__ uint32_t rl_rowmajor(uint32_t o) { return __builtin_amdgcn_alignbit(o, o, 7); }
// Whether a's key is larger than b's.
__device__ __forceinline__ bool rl_key_gt(const uint32_t* owngs, bool anygs, uint32_t a, uint32_t b) {
    const uint32_t ra = rl_rowmajor(a), rb = rl_rowmajor(b);
    if (!anygs) return ra > rb;   // without GS runs in the tile every key has the same top bit
    const uint32_t ga = rl_gs(owngs, a), gb = rl_gs(owngs, b);
    return ga != gb ? ga < gb : ra > rb;
}

// Union of two trees: the root with the larger key is linked under the
// other by a CAS on the 32-bit word holding its entry (the other half may
// change meanwhile: then the CAS is retried as long as the entry is still a
// root); a root that got a parent meanwhile is walked again from that parent.
// Every link joins two trees, so the parents stay a forest.
__device__ __forceinline__ void rl_unite(uint16_t* lab, const uint32_t* owngs, bool anygs, uint32_t a, uint32_t b) {
    for (;;) {
        a = rl_find_h(lab, a);
        b = rl_find_h(lab, b);
        if (a == b) return;
        const bool ahi = rl_key_gt(owngs, anygs, a, b);
        const uint32_t hi = ahi ? a : b, lo = ahi ? b : a;
        uint32_t* w = rl_word(lab, hi);
        const int sh = rl_shift(hi);
        uint32_t old = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        for (;;) {
            const uint32_t cur = (old >> sh) & 0xFFFFu;
            if (cur != hi) break;
            const uint32_t nw = (old & ~(0xFFFFu << sh)) | (lo << sh);
            if (__hip_atomic_compare_exchange_strong(w, &old, nw, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_WORKGROUP))
                return;
        }
        a = (old >> sh) & RL_IDX;
        b = lo;
    }
}

// Zone entry of a run start whose path is compressed (its entry is its root
// or names it), and the fragment id it carries: a root that was never
// numbered is not a seam zone.
__device__ __forceinline__ uint32_t rl_zone(const uint16_t* lab, uint32_t p) {
    const uint32_t e = rl_ld(lab, p);
    return ((e & RL_ROOT) || (e & RL_IDX) == p) ? e : rl_ld(lab, e & RL_IDX);
}
__device__ __forceinline__ uint32_t rl_zone_id(uint32_t z) { return (z & RL_ROOT) ? rl_id(z) : RL_NOID; }

// Lane y receives lane y - 1's value (lane 0: 0).
__device__ __forceinline__ uint32_t from_lane_above(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xF, 0xF, false);   // wave_shr:1
}

// Top bits of a 0x80-per-byte flag word -> 4-bit mask (bit k = byte k).
__device__ __forceinline__ uint32_t nib80(uint32_t f) { return (f * 0x00204081u) >> 28; }
// Flags at bits 3 + 8k (bits 0-3 out) and 7 + 8k (bits 4-7 out) -> 8-bit mask.
// The multiplier's four partial products of the eight flags fall on distinct
// bits (no carries), bits 24-31 in order.
__device__ __forceinline__ uint32_t nib80x2(uint32_t f) { return (f * 0x00204081u) >> 24; }
// Bits 4d .. 4d + 3 of m -> 0x00 / 0xFF bytes.
__device__ __forceinline__ uint32_t mask_bytes(uint32_t m, int d) {
    return __builtin_amdgcn_perm(0u, 0u, 0x0C0C0C0Cu | spread4((m >> (4 * d)) & 0xFu));
}
__device__ __forceinline__ uint32_t eq80(uint32_t a, uint32_t b) { return ~ne_top(a, b) & 0x80808080u; }
__device__ __forceinline__ uint32_t bit(uint32_t m, int b) { return (m >> b) & 1u; }

// Three-way packed maximum / minimum of 16-bit lanes that hold values 0 ..
// 255.  Taken as half-precision bit patterns these are +0 and positive
// subnormals, whose order is that of the integers; v_pk_maximum3_f16 /
// v_pk_minimum3_f16 return one of their inputs (half-precision subnormals are
// kept, not flushed, in the kernels' default floating-point mode), so the
// result is the integer maximum / minimum in one instruction instead of two.
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pk_max3(uint32_t a, uint32_t b, uint32_t c) {
    const f16x2 x = __builtin_bit_cast(f16x2, a), y = __builtin_bit_cast(f16x2, b), z = __builtin_bit_cast(f16x2, c);
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_maximum(__builtin_elementwise_maximum(x, y), z));
}
__device__ __forceinline__ uint32_t pk_min3(uint32_t a, uint32_t b, uint32_t c) {
    const f16x2 x = __builtin_bit_cast(f16x2, a), y = __builtin_bit_cast(f16x2, b), z = __builtin_bit_cast(f16x2, c);
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_minimum(__builtin_elementwise_minimum(x, y), z));
}

// 16 bytes of LDS at an 8-byte aligned address (the staged rows' pitch is a
// multiple of 8, not of 16): two 8-byte accesses, which the compiler pairs.
__device__ __forceinline__ uint4 lds_ld16(const uint8_t* p) {
    const uint2 a = reinterpret_cast<const uint2*>(p)[0], b = reinterpret_cast<const uint2*>(p)[1];
    return make_uint4(a.x, a.y, b.x, b.y);
}
__device__ __forceinline__ void lds_st16(uint8_t* p, const uint4& v) {
    reinterpret_cast<uint2*>(p)[0] = make_uint2(v.x, v.y);
    reinterpret_cast<uint2*>(p)[1] = make_uint2(v.z, v.w);
}

// Staged rows: tile rows -1 .. th (clamped to the image, so an out-of-image
// halo pixel replicates the nearest image pixel; see tile_load), columns
// 0 .. tw - 1 and the halo columns -1 and tw.
__device__ __forceinline__ void row_tile_load(RowTileSmem& sm, const ImgDev& im, int x0, int y0, int tw, int th,
                                              int t) {
    constexpr int NV = RT_ROWS * 4;                                  // 16-byte items
    constexpr int ITEMS = (NV + RT_THREADS - 1) / RT_THREADS;
    constexpr int HALO = (2 * RT_ROWS + RT_THREADS - 1) / RT_THREADS;
    if (x0 > 0 && y0 > 0 && x0 + RT < im.w && y0 + RT < im.h) {
        const uint8_t* base = im.img + (int64_t)(y0 - 1) * im.pitch + x0;
        uint4 v[ITEMS];
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const int it = t + i * RT_THREADS;
            v[i] = make_uint4(0, 0, 0, 0);
            if (it < NV) __builtin_memcpy(&v[i], base + (int64_t)(it >> 2) * im.pitch + 16 * (it & 3), 16);
        }
        uint8_t hv[HALO];
#pragma unroll
        for (int i = 0; i < HALO; ++i) {
            const int it = t + i * RT_THREADS;
            hv[i] = 0;
            if (it < 2 * RT_ROWS) hv[i] = base[(int64_t)(it >> 1) * im.pitch + ((it & 1) ? RT : -1)];
        }
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const int it = t + i * RT_THREADS;
            if (it < NV) lds_st16(&sm.img[(it >> 2) * RT_P + 16 + 16 * (it & 3)], v[i]);
        }
#pragma unroll
        for (int i = 0; i < HALO; ++i) {
            const int it = t + i * RT_THREADS;
            if (it < 2 * RT_ROWS) sm.img[(it >> 1) * RT_P + 16 + ((it & 1) ? RT : -1)] = hv[i];
        }
        return;
    }
    auto row_of = [&](int r) -> const uint8_t* {   // staged row r = tile row r - 1
        int gy = y0 + r - 1;
        gy = gy < 0 ? 0 : (gy >= im.h ? im.h - 1 : gy);
        return im.img + (int64_t)gy * im.pitch + x0;
    };
    const int nrows = th + 2;
    uint4 v[ITEMS];
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const int it = t + i * RT_THREADS;
        const int r = it >> 2, c = 16 * (it & 3);
        v[i] = make_uint4(0, 0, 0, 0);
        if (r >= nrows || c >= tw) continue;
        const uint8_t* src = row_of(r) + c;
        if (c + 16 <= tw) {
            __builtin_memcpy(&v[i], src, 16);
        } else {
            uint32_t q[4] = {0, 0, 0, 0};
#pragma unroll
            for (int b = 0; b < 16; ++b)
                if (c + b < tw) q[b >> 2] |= (uint32_t)src[b] << (8 * (b & 3));
            v[i] = make_uint4(q[0], q[1], q[2], q[3]);
        }
    }
    uint8_t hv[HALO];
#pragma unroll
    for (int i = 0; i < HALO; ++i) {
        const int it = t + i * RT_THREADS;
        hv[i] = 0;
        if ((it >> 1) >= nrows) continue;
        int gx = x0 + ((it & 1) ? tw : -1);
        gx = gx < 0 ? 0 : (gx >= im.w ? im.w - 1 : gx);
        hv[i] = row_of(it >> 1)[gx - x0];
    }
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const int it = t + i * RT_THREADS;
        const int r = it >> 2, c = 16 * (it & 3);
        if (r >= nrows || c >= tw) continue;
        uint8_t* dst = &sm.img[r * RT_P + 16 + c];
        if (c + 16 <= tw) {
            lds_st16(dst, v[i]);
        } else {
            // bytes past tw stay untouched: column tw is the halo column
            const uint32_t q[4] = {v[i].x, v[i].y, v[i].z, v[i].w};
#pragma unroll
            for (int b = 0; b < 16; ++b)
                if (c + b < tw) dst[b] = (uint8_t)(q[b >> 2] >> (8 * (b & 3)));
        }
    }
#pragma unroll
    for (int i = 0; i < HALO; ++i) {
        const int it = t + i * RT_THREADS;
        if ((it >> 1) >= nrows) continue;
        sm.img[(it >> 1) * RT_P + 16 + ((it & 1) ? tw : -1)] = hv[i];
    }
}

// A lane's view of one staged row: w[0] columns 32h - 4 .. 32h - 1, w[1..8]
// columns 32h .. 32h + 31, w[9] columns 32h + 32 .. 32h + 35.
struct HalfRow {
    uint32_t w[10];
};
__device__ __forceinline__ HalfRow read_half_row(const RowTileSmem& sm, int r, int h) {
    const uint8_t* p = sm.img + (r + 1) * RT_P + 16 + 32 * h;
    HalfRow o;
    o.w[0] = *reinterpret_cast<const uint32_t*>(p - 4);
    const uint4 q0 = lds_ld16(p), q1 = lds_ld16(p + 16);
    o.w[1] = q0.x; o.w[2] = q0.y; o.w[3] = q0.z; o.w[4] = q0.w;
    o.w[5] = q1.x; o.w[6] = q1.y; o.w[7] = q1.z; o.w[8] = q1.w;
    o.w[9] = *reinterpret_cast<const uint32_t*>(p + 32);
    return o;
}

template <bool MULTI>
__device__ __forceinline__ void row_tile_stage(RowTileSmem& sm, const ImgDev& im, int64_t gtile, const Ws& ws,
                                               bool live, int tid) {
    constexpr int S = 4 * RT - 4;
    const int y = tid & 63, h = tid >> 6;
    int x0 = 0, y0 = 0, tw = 0, th = 0;   // a workgroup's missing last tile has no pixels
    if (live) {
        const uint32_t lt = (uint32_t)(gtile - im.tile_begin);
        x0 = (int)(lt % (uint32_t)im.tiles_x) * RT;
        y0 = (int)(lt / (uint32_t)im.tiles_x) * RT;
        tw = min(RT, im.w - x0);
        th = min(RT, im.h - y0);
    }
    const bool hasL = x0 > 0, hasR = x0 + tw < im.w, hasU = y0 > 0, hasD = y0 + th < im.h;
    if (tid == 0) {
        sm.count = 0;
        sm.open = 0;
        sm.err = 0;
        sm.zflags = 0;
        sm.nonuni = 0;
        sm.orow[0] = 0;
        sm.orow[1] = 0;
        sm.qn[0] = 0;
        sm.qn[1] = 0;
        sm.anygs = 0;
    }
    if (live) row_tile_load(sm, im, x0, y0, tw, th, tid);
    __syncthreads();
    // Which edge pixels are at 255 / at 0, for the numbering of the seam
    // zones after the staged rows have been reused (bits beyond tw / th are
    // not read): wave 0 takes columns 0 and tw - 1, wave 1 rows 0 and th - 1.
    if (live) {
        const uint8_t* st = sm.img + 16;
        const uint32_t p = h == 0 ? st[(y + 1) * RT_P] : st[RT_P + y];
        const uint32_t q = h == 0 ? st[(y + 1) * RT_P + tw - 1] : st[th * RT_P + y];
        const uint64_t p255 = __ballot(p == 255u), p0 = __ballot(p == 0u);
        const uint64_t q255 = __ballot(q == 255u), q0 = __ballot(q == 0u);
        if (y == 0) {
            sm.e255[h == 0 ? 2 : 0] = p255;
            sm.e0[h == 0 ? 2 : 0] = p0;
            sm.e255[h == 0 ? 3 : 1] = q255;
            sm.e0[h == 0 ? 3 : 1] = q0;
        }
    }

    // ---- classification ---------------------------------------------------
    const int nb = min(max(tw - 32 * h, 0), 32);   // valid columns of this half
    const uint32_t VM = (y < th) ? (nb >= 32 ? 0xFFFFFFFFu : ((1u << nb) - 1u)) : 0u;
    const HalfRow a = read_half_row(sm, y - 1, h), b = read_half_row(sm, y, h), c = read_half_row(sm, y + 1, h);
    // The 3 x 3 maximum and minimum are taken down the columns first (the
    // three rows' 16-bit lanes of even pixels, bytes 0 / 2, and of odd
    // pixels, bytes 1 / 3, of each dword), then across: pixel 0 of a dword
    // needs column -1 (the previous dword's odd high lane), pixel 3 column 4
    // (the next dword's even low lane).  Dwords 0 and 9 hold the columns left
    // and right of the half: of these only the odd / even lanes are needed.
    uint32_t pxo, pno;   // previous dword: odd max / min
    {
        const uint32_t sel = 0x0C030C01u;
        const uint32_t ao = __builtin_amdgcn_perm(0u, a.w[0], sel), bo = __builtin_amdgcn_perm(0u, b.w[0], sel),
                       co = __builtin_amdgcn_perm(0u, c.w[0], sel);
        pxo = pk_max3(ao, bo, co);
        pno = pk_min3(ao, bo, co);
    }
    auto vert = [&](int i, uint32_t& xe, uint32_t& xo, uint32_t& ne, uint32_t& no) {
        const uint32_t ae = __builtin_amdgcn_perm(0u, a.w[i], 0x0C020C00u), ao = __builtin_amdgcn_perm(0u, a.w[i], 0x0C030C01u);
        const uint32_t be = __builtin_amdgcn_perm(0u, b.w[i], 0x0C020C00u), bo = __builtin_amdgcn_perm(0u, b.w[i], 0x0C030C01u);
        const uint32_t ce = __builtin_amdgcn_perm(0u, c.w[i], 0x0C020C00u), co = __builtin_amdgcn_perm(0u, c.w[i], 0x0C030C01u);
        xe = pk_max3(ae, be, ce);
        xo = pk_max3(ao, bo, co);
        ne = pk_min3(ae, be, ce);
        no = pk_min3(ao, bo, co);
    };
    uint32_t cxe, cxo, cne, cno;   // current dword: even / odd max / min
    vert(1, cxe, cxo, cne, cno);
    uint32_t G = 0, Sm = 0, ER = 0, ED = 0, EDL = 0, EDR = 0;
    uint32_t pG = 0, pS = 0, pR = 0, pD = 0, pDL = 0, pDR = 0;   // flags of the even dword of a pair
#pragma unroll
    for (int d = 0; d < 8; ++d) {
        uint32_t nxe, nxo = 0, nne, nno = 0;   // next dword
        if (d < 7) {
            vert(d + 2, nxe, nxo, nne, nno);
        } else {
            const uint32_t sel = 0x0C020C00u;
            const uint32_t ae = __builtin_amdgcn_perm(0u, a.w[9], sel), be = __builtin_amdgcn_perm(0u, b.w[9], sel),
                           ce = __builtin_amdgcn_perm(0u, c.w[9], sel);
            nxe = pk_max3(ae, be, ce);
            nne = pk_min3(ae, be, ce);
        }
        // lanes (pixel 0, pixel 2) take (column -1, column 1), lanes (pixel 1,
        // pixel 3) take (column 2, column 4); back in byte form a strictly
        // greater neighbour exists iff the maximum differs from the pixel, a
        // strictly smaller one iff the minimum does
        const uint32_t mx3 = pk_max3(cxe, cxo, __builtin_amdgcn_alignbit(cxo, pxo, 16)) |
                             pk_max3(cxe, cxo, __builtin_amdgcn_alignbit(nxe, cxe, 16)) << 8;
        const uint32_t mn3 = pk_min3(cne, cno, __builtin_amdgcn_alignbit(cno, pno, 16)) |
                             pk_min3(cne, cno, __builtin_amdgcn_alignbit(nne, cne, 16)) << 8;
        const uint32_t V = b.w[d + 1];
        const uint32_t eR = eq80(V, __builtin_amdgcn_alignbyte(b.w[d + 2], V, 1));
        const uint32_t eD = eq80(V, c.w[d + 1]);
        const uint32_t eDL = eq80(V, __builtin_amdgcn_alignbyte(c.w[d + 1], c.w[d], 3));
        const uint32_t eDR = eq80(V, __builtin_amdgcn_alignbyte(c.w[d + 2], c.w[d + 1], 1));
        const uint32_t fG = ne_top(mx3, V), fS = ne_top(mn3, V);
        if (d & 1) {
            // two dwords' flags (the earlier one's moved to bit 3 of each
            // byte) in one multiply: eight mask bits
            G |= nib80x2((pG >> 4) | fG) << (8 * (d >> 1));
            Sm |= nib80x2((pS >> 4) | fS) << (8 * (d >> 1));
            ER |= nib80x2((pR >> 4) | eR) << (8 * (d >> 1));
            ED |= nib80x2((pD >> 4) | eD) << (8 * (d >> 1));
            EDL |= nib80x2((pDL >> 4) | eDL) << (8 * (d >> 1));
            EDR |= nib80x2((pDR >> 4) | eDR) << (8 * (d >> 1));
        } else {
            pG = fG;
            pS = fS;
            pR = eR;
            pD = eD;
            pDL = eDL;
            pDR = eDR;
        }
        pxo = cxo;
        pno = cno;
        cxe = nxe;
        cxo = nxo;
        cne = nne;
        cno = nno;
    }
    G &= VM;
    Sm &= VM;
    // links inside the tile and the wave's half
    const bool dv = y + 1 < th;
    const uint32_t H = ER & VM & (VM >> 1);
    const uint32_t D = dv ? ED & VM : 0u;
    const uint32_t DL = dv ? EDL & VM & (VM << 1) : 0u;
    const uint32_t DR = dv ? EDR & VM & (VM >> 1) : 0u;
    const uint32_t U = from_lane_above(D), UL = from_lane_above(DR) << 1, UR = from_lane_above(DL) >> 1;
    // relations across columns 31 | 32 (both pixels inside the tile)
    const bool v31 = (VM >> 31) & 1u && tw > 32;   // wave 0: (y, 31) and (y, 32) valid
    const bool v32 = h == 1 && (VM & 1u);          // wave 1: (y, 32) valid
    const uint32_t XH = (h == 0 && v31) ? bit(ER, 31) : 0u;
    const uint32_t XDR = (h == 0 && v31 && dv) ? bit(EDR, 31) : 0u;
    const uint32_t XUR = (h == 0 && v31 && y > 0) ? (uint32_t)((b.w[8] >> 24) == (a.w[9] & 0xFFu)) : 0u;
    uint32_t cross = (XH | XDR | XUR) << 31;
    if (v32) {
        const uint32_t p32 = b.w[1] & 0xFFu;
        if (p32 == (b.w[0] >> 24) || (dv && bit(EDL, 0)) || (y > 0 && p32 == (a.w[0] >> 24))) cross |= 1u;
    }
    // seam bits: an equal 8-neighbour outside the tile and inside the image
    uint32_t SEAM = 0;
    {
        // top row against the halo row: lane b of each wave takes column 32h + b
        const int bl = y & 31, x = 32 * h + bl;
        const uint8_t* r0 = sm.img + RT_P + 16;   // tile row 0
        const uint8_t* rm = sm.img + 16;          // tile row -1
        const uint32_t v = r0[x];
        const bool in = y < 32 && x < tw && hasU;
        const bool tu = rm[x] == v, tul = rm[x - 1] == v && (x > 0 || hasL), tur = rm[x + 1] == v && (x + 1 < tw || hasR);
        const uint32_t top = (uint32_t)__ballot(in && (tu || tul || tur));
        if (y == 0) SEAM |= top;
    }
    if (y == th - 1 && hasD) {
        // bottom row against the halo row (the row below): the raw down masks
        // corners: column -1 needs a tile on the left, column tw one on the right
        const int br = tw - 1 - 32 * h;
        const uint32_t rcorner = (!hasR && br >= 0 && br < 32) ? 1u << br : 0u;
        SEAM |= (ED | (EDL & ~(h == 0 && !hasL ? 1u : 0u)) | (EDR & ~rcorner)) & VM;
    }
    if (h == 0 && hasL && (VM & 1u)) {
        const uint32_t p0 = b.w[1] & 0xFFu;
        if (p0 == (b.w[0] >> 24) || (p0 == (a.w[0] >> 24) && (y > 0 || hasU)) ||
            (p0 == (c.w[0] >> 24) && (dv || hasD)))
            SEAM |= 1u;
    }
    if (h == 1 && tw == 64 && hasR && (VM >> 31)) {
        const uint32_t p63 = b.w[8] >> 24;
        if (p63 == (b.w[9] & 0xFFu) || (p63 == (a.w[9] & 0xFFu) && (y > 0 || hasU)) ||
            (p63 == (c.w[9] & 0xFFu) && (dv || hasD)))
            SEAM |= 0x80000000u;
    }
    const uint32_t NS = VM & (H | (H << 1) | U | UL | UR | D | DL | DR | cross | SEAM);
    const uint32_t RS = NS & ~(H << 1);
    const uint32_t SGL = VM & ~NS;
    const uint32_t nrow = rnode(y, 32 * h);   // node of column 32h of this row; column 32h + b: nrow + (b << RL_COL)
    // cross-seam equality bits (right column / bottom row)
    uint32_t eqr = 0;
    if (h == 1 && tw == 64 && hasR && y < th) {
        const uint32_t p63 = b.w[8] >> 24;
        eqr = (uint32_t)(p63 == (a.w[9] & 0xFFu) && (y > 0 || hasU)) | (uint32_t)(p63 == (b.w[9] & 0xFFu)) << 1 |
              (uint32_t)(p63 == (c.w[9] & 0xFFu) && (dv || hasD)) << 2;
    }
    uint32_t eqb = 0;
    if (th == 64 && hasD) {
        const int bl = y & 31, x = 32 * h + bl;
        const uint32_t ed = __builtin_amdgcn_readlane(ED, 63), edl = __builtin_amdgcn_readlane(EDL, 63),
                       edr = __builtin_amdgcn_readlane(EDR, 63);
        if (x < tw)
            eqb = (bit(edl, bl) & (uint32_t)(x > 0 || hasL)) | bit(ed, bl) << 1 |
                  (bit(edr, bl) & (uint32_t)(x + 1 < tw || hasR)) << 2;
    }

    // a uniform tile (one value): every valid pixel equals its right and lower
    // neighbours inside the tile, and the halves meet equal
    const bool all_eq = H == (VM & (VM >> 1)) && (!dv || D == VM) && (h == 1 || !v31 || XH);
    if (!__all(all_eq)) sm.nonuni = 1u;
    {
        // a uniform tile is one zone: its evidence is the OR over the tile
        const uint32_t f = (G ? F_G : 0u) | (Sm ? F_S : 0u) | (SEAM ? F_SEAM : 0u);
        const uint32_t wf = (__ballot(f & F_G) ? F_G : 0u) | (__ballot(f & F_S) ? F_S : 0u) |
                            (__ballot(f & F_SEAM) ? F_SEAM : 0u);
        if (y == 0 && wf) atomicOr(&sm.zflags, wf);
    }
    {
        // every label entry its own parent, 16 bytes (8 entries) per thread
        // and store, consecutive threads on consecutive bytes
        uint4* dst = reinterpret_cast<uint4*>(sm.lab);
#pragma unroll
        for (int q = 0; q < RT * RT / 8 / RT_THREADS; ++q) {
            const uint32_t c = tid + q * RT_THREADS;
            // entry i holds its own offset 2i
            const uint32_t k = (16 * c) * 0x00010001u + 0x00020000u;
            dst[c] = make_uint4(k, k + 0x00040004u, k + 0x00080008u, k + 0x000C000Cu);
        }
    }
    // Every barrier below is passed by all threads of the workgroup, which
    // may hold tiles that take different branches.
    __syncthreads();
    const bool uniform = sm.nonuni == 0;
    uint8_t* rec = ws.rec + gtile * 4 * RT;

    // ---- runs: union-find over run starts ----------------------------------
    const uint32_t RSup = from_lane_above(RS);
    const uint32_t xkeep = cross_keep(XH, XDR, XUR, from_lane_above(XH), bit(D, 31), bit(U, 31));
    const RowEdges e = row_edges(U, UL, UR, H);
    const uint32_t nup = nrow - RL_ROW;   // the row above
    // each run's evidence at its start: greater, smaller and seam bits
    const uint32_t rfg = fill_bwd(G & NS, H) & RS, rfs = fill_bwd(Sm & NS, H) & RS, rfx = fill_bwd(SEAM, H) & RS;
    const uint32_t GSR = rfg & rfs;   // GS runs (both kinds of evidence)
    // A run's first edge is its initial link (Komura): a plain store of the
    // run above as its parent.  Links always point one row up, so they form a
    // forest; the other edges are unions on it.  A plain link goes to a
    // smaller key (rl_key) unless a GS run would link under a run that is not
    // one: such a first edge is a union instead.
    const uint32_t first1 = e.first & ~((fill_fwd(e.first, H) & H) << 1);
    const uint32_t NGSup = from_lane_above(fill_fwd(RS & ~GSR, H));   // pixels of non-GS runs above
    const uint32_t viol = first1 & fill_fwd(GSR, H) &
                          ((NGSup & ~e.kl & ~e.kr) | ((NGSup << 1) & e.kl) | ((NGSup >> 1) & e.kr));
    const uint32_t link1 = first1 & ~viol;
    // the staged rows are no longer read: their first 512 bytes hold the GS run starts
    uint32_t* owngs = reinterpret_cast<uint32_t*>(sm.img);
    if (!uniform) {
        owngs[2 * y + h] = GSR;
        if (__ballot(GSR != 0u) && y == 0) sm.anygs = 1u;
        uint32_t todo = link1;
        while (todo) {
            const int bb = __builtin_ctz(todo);
            todo &= todo - 1u;
            const int k = bb - (int)bit(e.kl, bb) + (int)bit(e.kr, bb);
            rl_st(sm.lab, nrow + (high_at_or_below(RS, bb) << RL_COL), nup + (high_at_or_below(RSup, k) << RL_COL));
        }
    }
    __syncthreads();
    const bool anygs = sm.anygs != 0u;
    if (!uniform) {
        // A lane's unions: the first edges without a plain link, the second
        // edges, then (wave 0) the cross-half relations that cross_keep
        // keeps, at the run through column 31.  Their number varies a lot
        // across the wave and a union costs over a hundred instructions, so
        // each lane appends its node pairs to a per-wave LDS queue and the
        // lanes then take the queue 64 at a time (a lane whose pairs do not
        // fit unites them itself).  A wave whose lanes have at most two each
        // skips the queue.
        uint32_t tf = e.first & ~link1, ts = e.second, tx = xkeep;
        auto next_pair = [&](uint32_t& na, uint32_t& nb) {
            const bool isf = tf != 0u, iss = !isf && ts != 0u, isx = !isf && !iss;
            const uint32_t cur = isf ? tf : ts;
            const int bb = isx ? 31 : __builtin_ctz(cur);
            const int k = isf ? bb - (int)bit(e.kl, bb) + (int)bit(e.kr, bb) : bb + 1;
            const int j = __builtin_ctz(tx | 8u);   // 0: XH, 1: XDR, 2: XUR
            nb = isx ? rnode(y, 32) + (j == 1 ? RL_ROW : 0u) - (j == 2 ? RL_ROW : 0u)
                     : nup + (high_at_or_below(RSup, k) << RL_COL);
            na = nrow + (high_at_or_below(RS, bb) << RL_COL);
            tf = isf ? tf & (tf - 1u) : tf;
            ts = iss ? ts & (ts - 1u) : ts;
            tx = isx ? tx & (tx - 1u) : tx;
        };
        const uint32_t n = __builtin_popcount(tf) + __builtin_popcount(ts) + __builtin_popcount(tx);
        if (__ballot(n >= 3u)) {
            uint32_t* uq = reinterpret_cast<uint32_t*>(sm.img + RT_UQ_OFF) + h * RT_UQ;
            uint32_t slot = atomicAdd(&sm.qn[h], n);
            while (tf | ts | tx) {
                uint32_t na, nb;
                next_pair(na, nb);
                if (slot < (uint32_t)RT_UQ)
                    uq[slot] = na | nb << 16;
                else
                    rl_unite(sm.lab, owngs, anygs, na, nb);
                ++slot;
            }
            // the wave's appends are done (LDS operations of one wave complete
            // in order; the fences keep the compiler from moving the reads up)
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const uint32_t tot = min(sm.qn[h], (uint32_t)RT_UQ);
            for (uint32_t i = (uint32_t)y; i < tot; i += 64) {
                const uint32_t w = uq[i];
                rl_unite(sm.lab, owngs, anygs, w & 0xFFFFu, w >> 16);
            }
        } else {
            // at most two per lane: each lane unites its own
            while (tf | ts | tx) {
                uint32_t na, nb;
                next_pair(na, nb);
                rl_unite(sm.lab, owngs, anygs, na, nb);
            }
        }
    }
    __syncthreads();
    uint32_t gsr = 0;   // runs (not GS runs) whose root is a GS run: zone with both kinds
    if (!uniform) {
        // ---- evidence of the other runs at their roots; compression ---------
        // A GS run's zone is rooted at a GS run, which tells its evidence, so
        // only the runs that are not GS runs report theirs, and one whose
        // root turns out to be a GS run need not (its masks are 0, its zone
        // closed).  Two runs per iteration (the second repeats the first when
        // the row has none left), their two finds walked together: the
        // dependent LDS loads of one overlap those of the other.
        // the index mask in a register the compiler cannot see through: with
        // the constant it narrows the walk to 16-bit compares and then widens
        // every stored parent again
        uint32_t idx = RL_IDX;
        asm volatile("" : "+v"(idx));
        uint32_t m = RS & ~GSR;
        while (m) {
            const int b1 = __builtin_ctz(m);
            m &= m - 1u;
            const int b2 = m ? __builtin_ctz(m) : b1;
            m &= m - 1u;
            const uint32_t p1 = nrow + (b1 << RL_COL), p2 = nrow + (b2 << RL_COL);
            // find with path halving, one exit: at a root the grandparent
            // is the parent itself
            uint32_t x1 = p1, x2 = p2;
            for (;;) {
                const uint32_t q1 = rl_ld(sm.lab, x1) & idx, q2 = rl_ld(sm.lab, x2) & idx;
                const uint32_t g1 = rl_ld(sm.lab, q1) & idx, g2 = rl_ld(sm.lab, q2) & idx;
                const bool d1 = g1 == q1, d2 = g2 == q2;
                if (d1 && d2) {
                    x1 = q1;
                    x2 = q2;
                    break;
                }
                // a node that is not a root has no evidence bits
                if (!d1) { rl_st(sm.lab, x1, g1); x1 = g1; }
                if (!d2) { rl_st(sm.lab, x2, g2); x2 = g2; }
            }
            if (x1 != p1) rl_st(sm.lab, p1, x1);
            if (x2 != p2) rl_st(sm.lab, p2, x2);
            const uint32_t f1 = bit(rfg, b1) | bit(rfs, b1) << 1 | bit(rfx, b1) << 2;
            const uint32_t f2 = bit(rfg, b2) | bit(rfs, b2) << 1 | bit(rfx, b2) << 2;
            uint32_t g1 = 0, g2 = 0;
            if (anygs) {
                g1 = rl_gs(owngs, x1);
                g2 = rl_gs(owngs, x2);
                gsr |= g1 << b1 | g2 << b2;
            }
            const uint32_t h1 = rl_ld(sm.lab, x1) >> RL_FSHIFT, h2 = rl_ld(sm.lab, x2) >> RL_FSHIFT;
            if (!g1 && (f1 & (h1 ^ 7u)))   // h1 is at most 7: the bits its root lacks
                __hip_atomic_fetch_or(rl_word(sm.lab, x1), f1 << (RL_FSHIFT + rl_shift(x1)), __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_WORKGROUP);
            if (!g2 && (f2 & (h2 ^ 7u)))
                __hip_atomic_fetch_or(rl_word(sm.lab, x2), f2 << (RL_FSHIFT + rl_shift(x2)), __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
    __syncthreads();
    uint32_t outMX = 0, outMN = 0;   // this half-row's final (or provisional) masks
    uint32_t opens = 0;              // run starts of open fragments
    if (!uniform) {
        // ---- number the seam zones (fragments) -----------------------------
        // A run holding a seam pixel starts in the top or bottom row, or holds
        // column 0 or column 63.  Lane b of each half takes the run start at
        // column 32 h + b of the top row (lanes 0-31) or the bottom row (lanes
        // 32-63); a lane of any other row takes its one seam run (its left or
        // right column).  Each finds its root (compressing its own entry to
        // it) and claims it: the first claim sets RL_ROOT and numbers the
        // root; its entry becomes RL_ROOT | id << 1 | evidence.  Other lanes
        // meanwhile read only the evidence bits, which stay, except the seam
        // bit: the numbering keeps it only if a mask of the zone can still
        // change (a zone at value 255 never gains greater evidence and one at
        // 0 never smaller, so a zone with either kind of evidence or its value
        // at that end of the range is final for that mask).  A run whose mask
        // pass reads the bit before it is cleared treats its zone as open,
        // which only costs record-plane bytes: the write-back finds nothing
        // changed in it and rewrites the same mask values.
        {
            const int bl = y & 31;
            const int ry = y < 32 ? 0 : th - 1;
            const uint32_t xr = __builtin_amdgcn_readlane(rfx, 0), xb = __builtin_amdgcn_readlane(rfx, th - 1);
            const uint32_t ctb = bit(y < 32 ? xr : xb, bl) ? rnode(ry, 32 * h + bl) : 1u;
            const uint32_t cmid = (y > 0 && y < th - 1 && rfx) ? nrow + (__builtin_ctz(rfx) << RL_COL) : 1u;
            // each candidate run's value: its pixel in the top / bottom row, or
            // the left / right edge pixel of its row (a middle row's seam run
            // holds column 0 or column tw - 1)
            const int ee[2] = {y < 32 ? 0 : 1, (cmid >> RL_COL) == 0 ? 2 : 3};
            const int eb[2] = {32 * h + bl, y};
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const uint32_t pc = c == 0 ? ctb : cmid;
                if (pc & 1u) continue;
                uint32_t x = pc;
                for (;;) {
                    const uint32_t ex = rl_ld(sm.lab, x);
                    const uint32_t q = ex & RL_IDX;
                    if ((ex & RL_ROOT) || q == x) break;
                    const uint32_t eq_ = rl_ld(sm.lab, q);
                    const uint32_t g = eq_ & RL_IDX;
                    if ((eq_ & RL_ROOT) || g == q) {
                        x = q;
                        break;
                    }
                    rl_st(sm.lab, x, g);
                    x = g;
                }
                if (x != pc) rl_st(sm.lab, pc, x);
                const int sh = rl_shift(x);
                const uint32_t old = __hip_atomic_fetch_or(rl_word(sm.lab, x), RL_ROOT << sh, __ATOMIC_RELAXED,
                                                           __HIP_MEMORY_SCOPE_WORKGROUP) >> sh;
                if (!(old & RL_ROOT)) {
                    uint32_t f = ((old >> RL_FSHIFT) & 7u) | (anygs && rl_gs(owngs, x) ? (F_G | F_S) : 0u) | F_SEAM;
                    const bool at255 = (sm.e255[ee[c]] >> eb[c]) & 1u, at0 = (sm.e0[ee[c]] >> eb[c]) & 1u;
                    if (((f & F_G) || at255) && ((f & F_S) || at0)) f &= ~F_SEAM;
                    uint32_t id = RL_NOID;
                    const uint32_t j = atomicAdd(&sm.count, 1u);
                    if (j >= (uint32_t)S) {
                        sm.err = 1u;
                    } else {
                        id = j;
                        ws.sbits[gtile * S + j] = (uint8_t)(f & 3u);
                    }
                    rl_st(sm.lab, x, RL_ROOT | id << 1 | f << RL_FSHIFT);
                }
            }
        }
        // ---- masks: each run takes its zone's evidence ----------------------
        // A zone rooted at a GS run has both kinds (masks 0, closed): the GS
        // runs and gsr.  Every other run's entry is its root or names it
        // (compressed above), a root that is not a GS run, whose entry holds
        // the zone's evidence.
        uint32_t zg = GSR | gsr, zs = GSR | gsr, zx = 0;   // at run starts: the zone's G / S / seam bits
        {
            uint32_t m = RS & ~(GSR | gsr);
            while (m) {
                const int b1 = __builtin_ctz(m);
                m &= m - 1u;
                const int b2 = m ? __builtin_ctz(m) : b1;
                m &= m - 1u;
                const uint32_t p1 = nrow + (b1 << RL_COL), p2 = nrow + (b2 << RL_COL);
                const uint32_t e1 = rl_ld(sm.lab, p1), e2 = rl_ld(sm.lab, p2);
                const uint32_t r1 = ((e1 & RL_ROOT) || (e1 & RL_IDX) == p1) ? p1 : (e1 & RL_IDX);
                const uint32_t r2 = ((e2 & RL_ROOT) || (e2 & RL_IDX) == p2) ? p2 : (e2 & RL_IDX);
                const uint32_t z1 = rl_ld(sm.lab, r1) >> RL_FSHIFT, z2 = rl_ld(sm.lab, r2) >> RL_FSHIFT;
                zg |= (z1 & 1u) << b1 | (z2 & 1u) << b2;
                zs |= ((z1 >> 1) & 1u) << b1 | ((z2 >> 1) & 1u) << b2;
                zx |= ((z1 >> 2) & 1u) << b1 | ((z2 >> 2) & 1u) << b2;
            }
        }
        const uint32_t mxs = RS & ~zg, mns = RS & ~zs;
        // a fragment that already has both kinds of evidence is final
        opens = zx & ~(zg & zs);
        if (opens) sm.open = 1u;
        {
            // the tile rows holding an open pixel (either half)
            const uint64_t orw = __ballot(opens != 0u);
            if (y == 0 && orw) {
                atomicOr(&sm.orow[0], (uint32_t)orw);
                atomicOr(&sm.orow[1], (uint32_t)(orw >> 32));
            }
        }
        const uint32_t MX = ((SGL & ~G) | fill_fwd(mxs, H)) & VM;
        const uint32_t MN = ((SGL & ~Sm) | fill_fwd(mns, H)) & VM;
        outMX = MX;
        outMN = MN;
    } else {
        const uint32_t zf = sm.zflags;
        if (tid == 0 && (zf & F_SEAM)) {
            sm.count = 1;
            ws.sbits[gtile * S] = (uint8_t)(zf & 3u);
        }
        outMX = (zf & F_G) ? 0u : VM;
        outMN = (zf & F_S) ? 0u : VM;
    }
    __syncthreads();   // every root numbered
    if (!uniform) {
        // ---- perimeter records -------------------------------------------------
        // Records are read only at seam pixels (the merges follow equal pixel
        // pairs across tile edges), whose runs were compressed by their
        // claims.  A seam pixel's zone must have a fragment id (invariant
        // code 2).
        {
            // top row (lanes 0-31) and bottom row (lanes 32-63), column 32h + (y & 31)
            const int bl = y & 31, x = 32 * h + bl;
            const int ry = y < 32 ? 0 : th - 1;
            const uint32_t rs_r = __builtin_amdgcn_readlane(RS, 0), rs_b = __builtin_amdgcn_readlane(RS, th - 1);
            const uint32_t sx_r = __builtin_amdgcn_readlane(SEAM, 0), sx_b = __builtin_amdgcn_readlane(SEAM, th - 1);
            const uint32_t rsm = y < 32 ? rs_r : rs_b, sxm = y < 32 ? sx_r : sx_b;
            if (bit(sxm, bl)) {
                const uint32_t id = rl_zone_id(rl_zone(sm.lab, rnode(ry, 32 * h + high_at_or_below(rsm, bl))));
                rec[(y < 32 ? 0 : RT) + x] = (uint8_t)id;
                if (id == RL_NOID) sm.err = 2u;
            }
            // left column (wave 0) and right column (the wave holding column tw - 1)
            if (h == 0 && (SEAM & 1u)) {
                const uint32_t id = rl_zone_id(rl_zone(sm.lab, nrow));
                rec[2 * RT + y] = (uint8_t)id;
                if (id == RL_NOID) sm.err = 2u;
            }
            const int br = tw - 1 - 32 * h;
            if (br >= 0 && br < 32 && bit(SEAM, br)) {
                const uint32_t id = rl_zone_id(rl_zone(sm.lab, nrow + (high_at_or_below(RS, br) << RL_COL)));
                rec[3 * RT + y] = (uint8_t)id;
                if (id == RL_NOID) sm.err = 2u;
            }
        }
        // ---- record plane (tiles with an open fragment) ----------------------
        if (sm.open) {
            uint8_t* row = sm.plane + RT_P * y + 32 * h;
            uint32_t cw[8];
#pragma unroll
            for (int d = 0; d < 8; ++d)
                cw[d] = REC_CLOSED * 0x01010101u | spread4((outMX >> (4 * d)) & 0xFu) |
                        spread4((outMN >> (4 * d)) & 0xFu) << 1;
            lds_st16(row, make_uint4(cw[0], cw[1], cw[2], cw[3]));
            lds_st16(row + 16, make_uint4(cw[4], cw[5], cw[6], cw[7]));
            // the bytes of open runs after their first: filled from it
            const uint32_t OF = fill_fwd(opens, H) & ~opens;
            // A wave with long open runs writes each run's id at its first
            // byte and fills the rest by doubling over the row's bytes (byte
            // i takes byte i - s while both are fill bytes, s = 1, 2, 4, 8,
            // 16); otherwise each run is written byte / dword-wise.
            const bool dbl = __ballot(__builtin_popcount(OF) >= 12) != 0;
            uint32_t m = opens;
            while (m) {
                const int bb = __builtin_ctz(m);
                m &= m - 1u;
                const uint32_t id = rl_zone_id(rl_zone(sm.lab, nrow + (bb << RL_COL))) & 0xFFu;
                // the run: bb .. the first column without an equal right link
                const int be = dbl ? bb : __builtin_ctz(~H & (0xFFFFFFFFu << bb));
                int q = bb;
                while (q <= be) {
                    if ((q & 3) == 0 && q + 3 <= be) {
                        *reinterpret_cast<uint32_t*>(row + q) = id * 0x01010101u;
                        q += 4;
                    } else {
                        row[q] = (uint8_t)id;
                        q += 1;
                    }
                }
            }
            if (dbl) {
                const uint4 r0 = lds_ld16(row), r1 = lds_ld16(row + 16);
                uint32_t v[8] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
                uint32_t f[8];
#pragma unroll
                for (int d = 0; d < 8; ++d) f[d] = mask_bytes(OF, d);
#pragma unroll
                for (int sb = 1; sb <= 16; sb *= 2) {
#pragma unroll
                    for (int d = 7; d >= 0; --d) {
                        uint32_t vs, fs;   // bytes 4d + j - sb
                        if (sb < 4) {
                            const uint32_t vp = d > 0 ? v[d - 1] : 0u, fp = d > 0 ? f[d - 1] : 0u;
                            vs = __builtin_amdgcn_alignbyte(v[d], vp, 4 - sb);
                            fs = __builtin_amdgcn_alignbyte(f[d], fp, 4 - sb);
                        } else {
                            vs = d >= sb / 4 ? v[d - sb / 4] : 0u;
                            fs = d >= sb / 4 ? f[d - sb / 4] : 0u;
                        }
                        v[d] = (f[d] & vs) | (~f[d] & v[d]);
                        f[d] &= fs;
                    }
                }
                lds_st16(row, make_uint4(v[0], v[1], v[2], v[3]));
                lds_st16(row + 16, make_uint4(v[4], v[5], v[6], v[7]));
            }
        }
    } else if (live) {
        // records: fragment 0 everywhere on the perimeter
        if (y < 32) {
            rec[32 * h + y] = 0;
            rec[RT + 32 * h + y] = 0;
        }
        if (h == 0) rec[2 * RT + y] = 0;
        if (h == 1) rec[3 * RT + y] = 0;
    }
    __syncthreads();
    // The masks leave through LDS as bits (the labels are no longer read), so
    // that each store covers whole 64-byte tile rows: a lane's own row would
    // make every store instruction touch 64 rows.
    reinterpret_cast<uint2*>(sm.lab)[tid] = make_uint2(outMX, outMN);
    // byte -> eight 0x00 / 0xFF bytes, for the mask stores (entries 2 tid, 2 tid + 1)
    uint2* const bytes8 = reinterpret_cast<uint2*>(sm.lab) + RT_THREADS;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const uint32_t v = 2 * tid + k;
        bytes8[v] = make_uint2(((((v & 0xFu) * 0x00204081u) & 0x01010101u) * 0xFFu),
                               ((((v >> 4) * 0x00204081u) & 0x01010101u) * 0xFFu));
    }
    const uint32_t count = sm.count;
    const bool open = !uniform && sm.open != 0;
    const uint64_t orows = (uint64_t)sm.orow[1] << 32 | sm.orow[0];
    if (open) {
        // only the rows with an open pixel: the others hold closed pixels,
        // which the write-back never changes
        uint8_t* plane = ws.rec8 + gtile * (int64_t)(RT * RT);
        uint4* dst = reinterpret_cast<uint4*>(plane);
#pragma unroll
        for (int i = 0; i < RT * RT / 16 / RT_THREADS; ++i) {
            const int c = tid + i * RT_THREADS;   // 16-byte chunk: row c / 4, part c % 4
            if ((orows >> (c >> 2)) & 1u)
                dst[c] = lds_ld16(sm.plane + (c >> 2) * RT_P + 16 * (c & 3));
        }
        if (tid == 0) ws.orows[gtile] = orows;
    }
    if (count) {
        uint8_t* eq = ws.eq + gtile * 2 * RT;
        if (h == 1) eq[y] = (uint8_t)eqr;
        if (y < 32) eq[RT + 32 * h + y] = (uint8_t)eqb;
    }
    if (tid == 0 && live)
        ws.counts[gtile] = min(count, (uint32_t)S) | (uniform ? CNT_UNIFORM : 0u) | (open ? CNT_OPEN | CNT_ROWS : 0u) |
                           (sm.err << COUNT_ERR_SHIFT);
    __syncthreads();
    // 16-pixel chunk c: tile row c / 4, columns 16 (c % 4) .. + 15
#pragma unroll
    for (int i = 0; i < RT * RT / 16 / RT_THREADS; ++i) {
        const int c = tid + i * RT_THREADS, r = c >> 2, q = c & 3;
        if (r >= th || 16 * q >= tw) continue;
        const uint2 bm = reinterpret_cast<const uint2*>(sm.lab)[(q >> 1) * 64 + r];
        const uint32_t mx16 = bm.x >> (16 * (q & 1)), mn16 = bm.y >> (16 * (q & 1));
        const int64_t o = (int64_t)(y0 + r) * im.opitch + x0 + 16 * q;
        if (16 * q + 16 <= tw) {
            const uint2 x0b = bytes8[mx16 & 0xFFu], x1b = bytes8[(mx16 >> 8) & 0xFFu];
            const uint2 n0b = bytes8[mn16 & 0xFFu], n1b = bytes8[(mn16 >> 8) & 0xFFu];
            store_u128(im.mx + o, make_uint4(x0b.x, x0b.y, x1b.x, x1b.y));
            store_u128(im.mn + o, make_uint4(n0b.x, n0b.y, n1b.x, n1b.y));
        } else {
            for (int k = 0; 16 * q + k < tw; ++k) {
                im.mx[o + k] = (uint8_t)(bit(mx16, k) * 0xFFu);
                im.mn[o + k] = (uint8_t)(bit(mn16, k) * 0xFFu);
            }
        }
    }
}

}  // namespace ssx
