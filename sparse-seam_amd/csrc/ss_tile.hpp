// ss_tile.hpp -- tile-local stage of the dual regional-extrema pipeline:
// LDS staging, Komura flat-zone labelling, evidence, fragments, masks, and the
// record-plane write-back.
//
// Work mapping: a tile of T x T pixels is cut into quads of 4 horizontally
// adjacent pixels; each thread owns the same quads through every phase and
// keeps their per-pixel evidence and edge bits in one register word per quad,
// so LDS holds only the image (for neighbour reads) and the shared labels.
// LDS image rows are dword aligned (pixel (lx, ly) at (ly + 1) * P + 4 + lx),
// so a quad reads its 3 x 6 neighbourhood with nine 32-bit LDS loads.
// The exact seam tests (which need the image-border rules) run in a separate
// pass over the 4T-4 perimeter pixels instead of inside every quad.
//
// Two shortcuts keep the common cases off the union-find:
//   singleton pixels  no equal 8-neighbour at all: the zone is the pixel, its
//                     evidence is final, and it takes no label or atomic;
//   uniform tiles     every pixel has one value: one zone, no labelling.
// Both produce the same zones (roots are minimum raster indices) as the
// general path.
#pragma once
#include "ss_device.hpp"
#include "ss_quad.hpp"

namespace ssx {

// 4-bit mask -> 0x00 / 0xFF bytes.
__device__ __forceinline__ uint32_t bytes4(uint32_t m) { return ((m * 0x00204081u) & 0x01010101u) * 0xFFu; }
// Top bit set in every byte where x and v differ (exact, no carry between bytes).
__device__ __forceinline__ uint32_t ne_top(uint32_t x, uint32_t v) {
    const uint32_t z = x ^ v;
    return (((z & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | z) & 0x80808080u;
}

// Two 16-bit lanes (v_pk_*_u16).
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pk_max(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(u16x2, a),
                                                                  __builtin_bit_cast(u16x2, b)));
}
__device__ __forceinline__ uint32_t pk_min(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(u16x2, a),
                                                                  __builtin_bit_cast(u16x2, b)));
}
__device__ __forceinline__ uint32_t pk_subsat(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_sub_sat(__builtin_bit_cast(u16x2, a),
                                                                      __builtin_bit_cast(u16x2, b)));
}
__device__ __forceinline__ uint32_t pk_add(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, a) + __builtin_bit_cast(u16x2, b));
}
// 16-bit lanes holding 0..255: top bit of the lane set iff the lane is non-zero.
__device__ __forceinline__ uint32_t nz_top16(uint32_t a) { return pk_add(a, 0x7FFF7FFFu) & 0x80008000u; }

// Horizontal 3-wide max and min of one image row around a quad, as two 16-bit
// lane pairs: even pixels (0, 2) and odd pixels (1, 3).  x0, x1, x2 are the
// dwords left of, at and right of the quad.  Also returns the quad's own even
// and odd bytes.
struct RowMinMax {
    uint32_t maxe, maxo, mine, mino, mid_e, mid_o;
};
__device__ __forceinline__ RowMinMax row_minmax(uint32_t x0, uint32_t x1, uint32_t x2) {
    RowMinMax r;
    r.mid_e = __builtin_amdgcn_perm(0u, x1, 0x0C020C00u);       // x1.0, x1.2
    r.mid_o = __builtin_amdgcn_perm(0u, x1, 0x0C030C01u);       // x1.1, x1.3
    const uint32_t le = __builtin_amdgcn_perm(x1, x0, 0x0C050C03u);  // x0.3, x1.1
    const uint32_t ro = __builtin_amdgcn_perm(x2, x1, 0x0C040C02u);  // x1.2, x2.0
    const uint32_t mx = pk_max(r.mid_e, r.mid_o), mn = pk_min(r.mid_e, r.mid_o);
    r.maxe = pk_max(le, mx);
    r.maxo = pk_max(mx, ro);
    r.mine = pk_min(le, mn);
    r.mino = pk_min(mn, ro);
    return r;
}

template <int T>
struct TileSmem {
    static constexpr int P = T + 16;  // LDS row pitch in bytes (strips of 4 rows hit disjoint banks)
    // The staged rows are read only by the classification and the perimeter
    // pass; the labels are stored after the barrier that follows them, so the
    // two share their LDS (one more workgroup per CU at T = 64).
    union {
        alignas(16) uint8_t img[(T + 2) * P];
        alignas(16) uint32_t lab[T * T];  // labels; at roots also evidence and fragment id
    };
    uint8_t eqr[T];                   // cross-seam equality bits, right column
    uint8_t eqb[T];                   // cross-seam equality bits, bottom row
    uint32_t count;
    uint32_t err;
    uint32_t zflags;                  // uniform tile: OR of the evidence
    uint32_t open;                    // the tile holds a fragment that may still change
    // Work list of the union pass (quad_label): the unions are spread over all
    // threads of the workgroup, so the pass takes as many rounds as the tile
    // has unions per thread rather than as many as its busiest thread has.
    // Unions beyond the capacity are done at once by the thread that has them.
    static constexpr int CAP_PAIRS = T * T / 8;
    // Staging of store_labels (earlier in the tile, so the two share LDS):
    // one slot per thread of a workgroup of up to STAGE threads.
    static constexpr int STAGE = 256;
    union {
        uint32_t pairs[CAP_PAIRS];    // pixel index | pixel index << 16
        uint2 stage[STAGE];           // quad word, pixel index of its pixel 0
    };
    uint32_t npairs;
};

// Thread -> quad mapping.  With fewer threads than quads a thread owns a
// vertical strip of PER quads (rows r0 .. r0 + PER - 1 of one quad column), so
// the strip's rows are loaded and reduced once and shared by the quads above
// and below them.
template <int T, int NT>
struct Quads {
    static constexpr int D = T / 4;                  // quads per row
    static constexpr int Q = T * T / 4;              // quads per tile
    static constexpr int PER = (Q + NT - 1) / NT;    // quads per thread
    static constexpr int PPER = (4 * T + NT - 1) / NT;  // perimeter slots per thread
    static_assert(PER == 1 || Q % NT == 0, "strips need NT | Q");
    // Row and first column of the calling thread's quad k; false if the
    // thread has no quad k.
    __device__ static __forceinline__ bool at(int k, int& ly, int& lx0) {
        if constexpr (PER == 1) {
            const int q = threadIdx.x;
            ly = q / D;
            lx0 = 4 * (q % D);
            return Q % NT == 0 || q < Q;
        } else {
            ly = (int)(threadIdx.x / D) * PER + k;
            lx0 = 4 * (int)(threadIdx.x % D);
            return true;
        }
    }
};

// The calling thread's quads as one bitmask (bit 4k + i = pixel i of quad k),
// so that a pass over the heads of all its quads runs as many iterations as
// the thread has heads, not as the sum over k of the wave's busiest lane.
template <int PER>
struct QuadMask {
    using type = uint32_t;
};
template <>
struct QuadMask<16> {
    using type = uint64_t;
};
template <int T, int NT>
struct QuadSet {
    static constexpr int PER = Quads<T, NT>::PER;
    using M = typename QuadMask<PER>::type;
    // Raster index of pixel 0 of quad 0; quad k starts T * k further down.
    __device__ static __forceinline__ uint32_t base() {
        int ly, lx0;
        Quads<T, NT>::at(0, ly, lx0);
        return (uint32_t)(ly * T + lx0);
    }
    template <typename G>
    __device__ static __forceinline__ M gather(G&& nibble) {
        M m = 0;
#pragma unroll
        for (int k = 0; k < PER; ++k) m |= (M)nibble(k) << (4 * k);
        return m;
    }
    // Calls f(k, i) for every set bit.
    template <typename F>
    __device__ static __forceinline__ void each(M m, F&& f) {
        while (m) {
            const int b = sizeof(M) == 8 ? __builtin_ctzll((uint64_t)m) : __builtin_ctz((uint32_t)m);
            m &= m - 1;
            f(b >> 2, b & 3);
        }
    }
    // v[k] for a k known only at run time, as a chain of selects between
    // registers.  The empty asm makes each operand an opaque register value:
    // otherwise the compiler turns the chain into an indexed load and keeps
    // the whole array in private memory.
    __device__ static __forceinline__ uint32_t pick(const uint32_t (&v)[PER], int k) {
        uint32_t r = v[0];
        __asm__("" : "+v"(r));
#pragma unroll
        for (int j = 1; j < PER; ++j) {
            uint32_t x = v[j];
            __asm__("" : "+v"(x));
            r = k == j ? x : r;
        }
        return r;
    }
};

// Per-quad evidence in byte form (bit b of byte i = flag of pixel i), the
// form the classification produces and the mask stores consume.
constexpr uint32_t B_G = 0x80808080u;     // a strictly greater 8-neighbour exists
constexpr uint32_t B_S = 0x40404040u;     // a strictly smaller 8-neighbour exists
constexpr uint32_t B_SGL = 0x20202020u;   // singleton
constexpr uint32_t B_VAL = 0x10101010u;   // pixel inside the image part of the tile
// Bits 8-11 (free in the byte form, set by store_labels): the run heads whose
// Komura link is the head itself, the only heads that can end as roots (a
// union only ever turns a root into a non-root).
constexpr int B_RT_SHIFT = 8;
__device__ __forceinline__ uint32_t b_roots(uint32_t sb) { return (sb >> B_RT_SHIFT) & 0xFu; }
// 0x80-per-byte flags -> 0xFF bytes.
__device__ __forceinline__ uint32_t bytes80(uint32_t x) { return (x - (x >> 7)) | x; }
// Two 0x80-per-byte flag words -> one byte: a's pixels in bits 0-3, b's in 4-7.
__device__ __forceinline__ uint32_t movemask_pair(uint32_t a, uint32_t b) {
    return ((((a >> 7) | (b >> 3)) & 0x11111111u) * 0x01020408u) >> 24;
}

__device__ __forceinline__ uint32_t lds32(const uint8_t* p) { return *reinterpret_cast<const uint32_t*>(p); }

// LDS find with path halving (label words of non-roots carry no flag bits).
// Not for the final compression: a halving store could replace a finished
// root entry with a mere ancestor.
__device__ __forceinline__ uint32_t lfind_h(uint32_t* lab, uint32_t x) {
    for (;;) {
        // at a root x the grandparent is x itself, so one test covers both
        const uint32_t p = lds_load(lab + x) & LMASK;
        const uint32_t gp = lds_load(lab + p) & LMASK;
        if (gp == p) return p;
        __hip_atomic_store(lab + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        x = gp;
    }
}

// Read-only LDS find; masks the evidence bits that roots accumulate.
__device__ __forceinline__ uint32_t lfindm(const uint32_t* lab, uint32_t x) {
    uint32_t r = lds_load(lab + x) & LMASK;
    for (;;) {
        const uint32_t n = lds_load(lab + r) & LMASK;
        if (n == r) return r;
        r = n;
    }
}

__device__ __forceinline__ void lunion_h(uint32_t* lab, uint32_t a, uint32_t b) {
    for (;;) {
        a = lfind_h(lab, a);
        b = lfind_h(lab, b);
        if (a == b) return;
        const uint32_t hi = a > b ? a : b, lo = a > b ? b : a;
        uint32_t expected = hi;
        if (__hip_atomic_compare_exchange_strong(lab + hi, &expected, lo, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))
            return;
        a = expected;
        b = lo;
    }
}

// Image and mask rows may start at any byte (pitches need not be multiples of
// 4): gfx950 performs unaligned global loads and stores, and a memcpy of 4 or
// 16 bytes compiles to one dword / dwordx4 access whatever the address.
__device__ __forceinline__ uint32_t load_u32(const uint8_t* p) {
    uint32_t v;
    __builtin_memcpy(&v, p, 4);
    return v;
}
__device__ __forceinline__ void store_u32(uint8_t* p, uint32_t v) { __builtin_memcpy(p, &v, 4); }
__device__ __forceinline__ void store_u128(uint8_t* p, const uint4& v) { __builtin_memcpy(p, &v, 16); }

// Loads the tile and its halo rows -1, th and halo columns -1, tw.  Halo
// pixels outside the image replicate the nearest image pixel: such a value
// belongs to the pixel itself or to one of its genuine 8-neighbours, so the
// 3x3 greater / smaller tests equal the "absent neighbour" rule exactly
// (equality across the image border is excluded by the perimeter pass).
// All of a thread's global loads are issued before the first LDS store, so a
// tile costs one memory round trip rather than one per loop iteration.
template <int T, int NT>
__device__ __forceinline__ void tile_load(TileSmem<T>& sm, const ImgDev& im, int x0, int y0, int tw, int th) {
    constexpr int P = TileSmem<T>::P, D = T / 4;
    constexpr int ITEMS = ((T + 2) * D + NT - 1) / NT;   // dwords per thread
    constexpr int HALO = (2 * (T + 2) + NT - 1) / NT;    // halo-column bytes per thread
    if (x0 > 0 && y0 > 0 && x0 + T < im.w && y0 + T < im.h) {
        // interior tile: every staged row and column lies in the image, so
        // the addresses are one uniform base plus per-thread offsets
        const uint8_t* base = im.img + (int64_t)(y0 - 1) * im.pitch + x0;
        uint32_t v[ITEMS];
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const int it = threadIdx.x + i * NT;
            v[i] = 0;
            if (it < (T + 2) * D)
                v[i] = load_u32(base + (int64_t)(it / D) * im.pitch + 4 * (it % D));
        }
        uint8_t hv[HALO];
#pragma unroll
        for (int i = 0; i < HALO; ++i) {
            const int it = threadIdx.x + i * NT;
            hv[i] = 0;
            if (it < 2 * (T + 2)) hv[i] = base[(int64_t)(it / 2) * im.pitch + ((it & 1) ? T : -1)];
        }
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const int it = threadIdx.x + i * NT;
            if (it < (T + 2) * D) *reinterpret_cast<uint32_t*>(&sm.img[(it / D) * P + 4 + 4 * (it % D)]) = v[i];
        }
#pragma unroll
        for (int i = 0; i < HALO; ++i) {
            const int it = threadIdx.x + i * NT;
            if (it < 2 * (T + 2)) sm.img[(it / 2) * P + 4 + ((it & 1) ? T : -1)] = hv[i];
        }
        return;
    }
    auto row_of = [&](int r) -> const uint8_t* {
        int gy = y0 + r;
        gy = gy < 0 ? 0 : (gy >= im.h ? im.h - 1 : gy);
        return im.img + (int64_t)gy * im.pitch + x0;
    };
    uint32_t v[ITEMS];
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const int it = threadIdx.x + i * NT;
        const int r = it / D - 1, c = 4 * (it % D);
        v[i] = 0;
        if (it >= (th + 2) * D || c >= tw) continue;
        const uint8_t* src = row_of(r) + c;
        if (c + 3 < tw) {
            v[i] = load_u32(src);
        } else {
            for (int b = 0; b < 4 && c + b < tw; ++b) v[i] |= (uint32_t)src[b] << (8 * b);
        }
    }
    uint8_t hv[HALO];
#pragma unroll
    for (int i = 0; i < HALO; ++i) {
        const int it = threadIdx.x + i * NT;
        hv[i] = 0;
        if (it >= 2 * (th + 2)) continue;
        const int r = it / 2 - 1, col = (it & 1) ? tw : -1;
        int gx = x0 + col;
        gx = gx < 0 ? 0 : (gx >= im.w ? im.w - 1 : gx);
        hv[i] = row_of(r)[gx - x0];
    }
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const int it = threadIdx.x + i * NT;
        const int r = it / D - 1, c = 4 * (it % D);
        if (it >= (th + 2) * D || c >= tw) continue;
        uint8_t* dst = &sm.img[(r + 1) * P + 4 + c];
        if (c + 3 < tw) {
            *reinterpret_cast<uint32_t*>(dst) = v[i];
        } else {
            // bytes past tw stay untouched: column tw is the halo column
            for (int b = 0; c + b < tw; ++b) dst[b] = (uint8_t)(v[i] >> (8 * b));
        }
    }
#pragma unroll
    for (int i = 0; i < HALO; ++i) {
        const int it = threadIdx.x + i * NT;
        if (it >= 2 * (th + 2)) continue;
        const int r = it / 2 - 1, col = (it & 1) ? tw : -1;
        sm.img[(r + 1) * P + 4 + col] = hv[i];
    }
}

// One image row around a quad: the quad's dword, its left- and right-shifted
// neighbour words (byte i = neighbour of pixel i) and the row's 3-wide
// extrema.  Row ly = -1 and ly = T are the halo rows.
struct RowW {
    uint32_t w, l, r;
    RowMinMax m;
};
template <int T>
__device__ __forceinline__ RowW row_at(const TileSmem<T>& sm, int ly, int lx0) {
    constexpr int P = TileSmem<T>::P;
    const int base = (ly + 1) * P + 4 + lx0;
    const uint32_t x0 = lds32(&sm.img[base - 4]), x1 = lds32(&sm.img[base]), x2 = lds32(&sm.img[base + 4]);
    RowW r;
    r.w = x1;
    r.l = __builtin_amdgcn_alignbyte(x1, x0, 3);
    r.r = __builtin_amdgcn_alignbyte(x2, x1, 1);
    r.m = row_minmax(x0, x1, x2);
    return r;
}

// 0x80 in bytes 0 .. n-1 (n >= 4: all four, n <= 0: none).
__device__ __forceinline__ uint32_t first_bytes80(int n) {
    return n >= 4 ? 0x80808080u : (n <= 0 ? 0u : (0x80808080u >> (8 * (4 - n))));
}

// Classifies one quad from its three rows: byte-form evidence (returned), and
// for a quad holding a non-singleton pixel also the edge word of the labelling
// passes (else 0), from which store_labels later derives its Komura labels.
// Singletons (no equal value among
// the eight LDS neighbours, halo included -- conservative at the image border,
// where a replicated halo value may only make a pixel non-singleton) take no
// label.  Pixels past tw or th are masked out of every flag.
template <int T>
__device__ __forceinline__ uint32_t classify_quad(const RowW& a, const RowW& b, const RowW& c,
                                                  int ly, int lx0, int tw, int th, uint32_t v4, bool& all_eq,
                                                  uint32_t& st) {
    const uint32_t V = b.w;
    const uint32_t nU = ne_top(a.w, V), nUL = ne_top(a.l, V), nUR = ne_top(a.r, V);
    const uint32_t nL = ne_top(b.l, V), nR = ne_top(b.r, V);
    const uint32_t nD = ne_top(c.w, V), nDL = ne_top(c.l, V), nDR = ne_top(c.r, V);
    // 3x3 max / min (self included: it never makes a neighbour greater or
    // smaller) on even / odd 16-bit lanes
    const uint32_t ge = pk_subsat(pk_max(pk_max(a.m.maxe, b.m.maxe), c.m.maxe), b.m.mid_e);
    const uint32_t go = pk_subsat(pk_max(pk_max(a.m.maxo, b.m.maxo), c.m.maxo), b.m.mid_o);
    const uint32_t se = pk_subsat(b.m.mid_e, pk_min(pk_min(a.m.mine, b.m.mine), c.m.mine));
    const uint32_t so = pk_subsat(b.m.mid_o, pk_min(pk_min(a.m.mino, b.m.mino), c.m.mino));
    const uint32_t g80 = (nz_top16(ge) >> 8) | nz_top16(go);
    const uint32_t s80 = (nz_top16(se) >> 8) | nz_top16(so);
    const uint32_t val80 = (ly < th) ? first_bytes80(tw - lx0) : 0u;
    const uint32_t neq = nU & nUL & nUR & nL & nR & nD & nDL & nDR;   // 0x80: no equal neighbour
    const uint32_t sgl80 = neq & val80;
    all_eq = all_eq && (ne_top(V, v4) & val80) == 0;
    st = 0;
    if (val80 & ~neq) {
        const uint32_t lv80 = lx0 > 0 ? 0x80808080u : 0x80808000u;   // left neighbour in the tile
        const uint32_t rv80 = first_bytes80(tw - lx0 - 1);              // right neighbour in the tile
        const uint32_t up80 = ly > 0 ? val80 : 0u;
        const uint32_t eu80 = ~nU & up80, eul80 = ~nUL & up80 & lv80, eur80 = ~nUR & up80 & rv80;
        const uint32_t el80 = ~nL & lv80 & val80;
        uint32_t s = movemask_pair(g80, s80) << Q_G | movemask_pair(el80, eu80) << Q_EL |
                     movemask_pair(eul80, eur80) << Q_EUL;
        const uint32_t nsg4 = movemask_pair(val80 & ~neq, 0u);
        // the links, union heads and root candidates follow from s alone:
        // store_labels derives them once the staged rows are no longer read
        st = s | (nsg4 & ~((s >> Q_EL) & 0xEu)) << Q_HD;
    }
    return g80 | (s80 >> 1) | (sgl80 >> 2) | (val80 >> 3);
}

// First pass over the thread's quads (a vertical strip): evidence, singleton
// detection and the edge bits of the union passes and of the Komura links
// (each head links to an equal pixel above, else left; stored later by
// store_labels).  sb[k] gets the byte-form evidence, st[k] the edge word (0
// for a quad of singletons).  Returns whether every pixel equals v0.
template <int T, int NT>
__device__ __forceinline__ bool quad_init(TileSmem<T>& sm, int tw, int th, uint32_t (&st)[Quads<T, NT>::PER],
                                          uint32_t (&sb)[Quads<T, NT>::PER], uint32_t v0) {
    constexpr int PER = Quads<T, NT>::PER;
    const uint32_t v4 = v0 * 0x01010101u;
    bool all_eq = true;
    int ly, lx0;
    const bool has = Quads<T, NT>::at(0, ly, lx0);
    if constexpr (PER == 1) {
        st[0] = 0;
        sb[0] = 0;
        if (!has || ly >= th || lx0 >= tw) return true;
        const RowW a = row_at<T>(sm, ly - 1, lx0), b = row_at<T>(sm, ly, lx0), c = row_at<T>(sm, ly + 1, lx0);
        sb[0] = classify_quad<T>(a, b, c, ly, lx0, tw, th, v4, all_eq, st[0]);
    } else {
        const int r0 = ly;
        const bool col_ok = lx0 < tw;
        RowW a = row_at<T>(sm, r0 - 1, lx0), b = row_at<T>(sm, r0, lx0);
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const RowW c = row_at<T>(sm, r0 + k + 1, lx0);
            st[k] = 0;
            sb[k] = 0;
            if (col_ok && r0 + k < th)
                sb[k] = classify_quad<T>(a, b, c, r0 + k, lx0, tw, th, v4, all_eq, st[k]);
            a = b;
            b = c;
        }
    }
    return all_eq;
}

// Komura labels of the thread's non-singleton quads, derived from their edge
// words, one 16-byte store per quad: a head links to its first contact above
// (else to the pixel left of the quad, else to itself), every other run pixel
// to its head; a singleton in such a quad gets a dummy entry that nothing
// reads.  Every link points at a smaller index.  Also completes the quad
// words: the union heads into st (Q_UN), the root candidates into sb (bits
// B_RT_SHIFT ..).  Called once nothing reads the staged rows any more (the
// labels overwrite them) and followed by a barrier.
//
// A quad that is one run (a plateau's interior) takes the short form
// quad_links_one_run.  A wave whose other non-singleton quads are few (noise-
// like content: most quads hold singletons only; plateaus: only their rims)
// spreads those over its lanes: each is staged in LDS by its owner, taken by
// the lane of its rank, and its union heads and root candidates are handed
// back; otherwise every lane would step through all its quads for the sake of
// the few lanes that have work in each.
template <int T, int NT>
__device__ __forceinline__ void store_labels(TileSmem<T>& sm, uint32_t (&st)[Quads<T, NT>::PER],
                                             uint32_t (&sb)[Quads<T, NT>::PER]) {
    constexpr int PER = Quads<T, NT>::PER;
    auto links = [&](uint32_t s, uint32_t p0) {
        uint32_t lab4[4], un, rt;
        quad_links(s, p0, (uint32_t)T, lab4, un, rt);
        *reinterpret_cast<uint4*>(&sm.lab[p0]) = make_uint4(lab4[0], lab4[1], lab4[2], lab4[3]);
        return un | rt << 4;
    };
    if constexpr (PER > 1 && NT <= TileSmem<T>::STAGE) {
        uint64_t bal[PER];
        uint32_t rank[PER], n = 0;
        bool gen[PER];   // needs the general form
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            gen[k] = st[k] != 0 && !q_one_run(st[k]);
            bal[k] = __ballot(gen[k]);
            rank[k] = n + __builtin_amdgcn_mbcnt_hi((uint32_t)(bal[k] >> 32),
                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)bal[k], 0u));
            n += (uint32_t)__builtin_popcountll(bal[k]);
        }
        if (n <= 64) {   // wave-uniform: one round
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                if (!st[k] || gen[k]) continue;
                int ly, lx0;
                Quads<T, NT>::at(k, ly, lx0);
                const uint32_t p0 = ly * T + lx0;
                uint32_t lab4[4], un, rt;
                quad_links_one_run(st[k], p0, (uint32_t)T, lab4, un, rt);
                *reinterpret_cast<uint4*>(&sm.lab[p0]) = make_uint4(lab4[0], lab4[1], lab4[2], lab4[3]);
                st[k] |= un << Q_UN;
                sb[k] |= rt << B_RT_SHIFT;
            }
            uint2* stage = sm.stage + (threadIdx.x & ~63u);
            const uint32_t lane = threadIdx.x & 63u;
            auto wave_sync = [] {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            };
            for (uint32_t r0 = 0; r0 < n; r0 += 64) {
#pragma unroll
                for (int k = 0; k < PER; ++k) {
                    if (!gen[k] || rank[k] - r0 >= 64u) continue;
                    int ly, lx0;
                    Quads<T, NT>::at(k, ly, lx0);
                    stage[rank[k] - r0] = make_uint2(st[k], (uint32_t)(ly * T + lx0));
                }
                wave_sync();
                if (lane < n - r0) {
                    const uint2 it = stage[lane];
                    stage[lane].x = links(it.x, it.y);
                }
                wave_sync();
#pragma unroll
                for (int k = 0; k < PER; ++k) {
                    if (!gen[k] || rank[k] - r0 >= 64u) continue;
                    const uint32_t res = stage[rank[k] - r0].x;
                    st[k] |= (res & 0xFu) << Q_UN;
                    sb[k] |= (res >> 4) << B_RT_SHIFT;
                }
                wave_sync();
            }
            return;
        }
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        if (!st[k]) continue;
        int ly, lx0;
        Quads<T, NT>::at(k, ly, lx0);
        const uint32_t res = links(st[k], (uint32_t)(ly * T + lx0));
        st[k] |= (res & 0xFu) << Q_UN;
        sb[k] |= (res >> 4) << B_RT_SHIFT;
    }
}

// Perimeter slot e in [0, 4T): top row, bottom row, left and right columns
// without their corners; every perimeter pixel appears exactly once.
template <int T>
__device__ __forceinline__ bool perim_pixel(int e, int tw, int th, int& lx, int& ly) {
    const int side = e / T, k = e % T;
    if (side == 0) { lx = k; ly = 0; return k < tw; }
    if (side == 1) { lx = k; ly = th - 1; return k < tw && th > 1; }
    if (side == 2) { lx = 0; ly = k; return k > 0 && k < th - 1; }
    lx = tw - 1; ly = k;
    return k > 0 && k < th - 1 && tw > 1;
}

// Exact seam tests of the thread's perimeter slots: an equal neighbour that is
// outside the tile but inside the image.  Writes the cross-seam equality bits
// of the right column and bottom row; returns one seam bit per slot.  The nine
// bytes around a perimeter pixel all lie in the staged rows -1 .. th and
// columns -1 .. tw, so they are read unconditionally (one LDS round trip) and
// the neighbour rules are applied as masks.
template <int T, int NT>
__device__ __forceinline__ uint32_t perimeter_pass(TileSmem<T>& sm, const ImgDev& im, int x0, int y0, int tw,
                                                   int th) {
    constexpr int P = TileSmem<T>::P;
    const bool hasL = x0 > 0, hasR = x0 + tw < im.w, hasU = y0 > 0, hasD = y0 + th < im.h;
    uint32_t seam_bits = 0;
#pragma unroll
    for (int k = 0; k < Quads<T, NT>::PPER; ++k) {
        const int e = threadIdx.x + k * NT;
        int lx, ly;
        if (e >= 4 * T || !perim_pixel<T>(e, tw, th, lx, ly)) continue;
        const uint8_t* c = &sm.img[(ly + 1) * P + 4 + lx];
        const uint32_t v = c[0];
        const uint32_t eul = c[-P - 1] == v, eu = c[-P] == v, eur = c[-P + 1] == v;
        const uint32_t el = c[-1] == v, er = c[1] == v;
        const uint32_t edl = c[P - 1] == v, ed = c[P] == v, edr = c[P + 1] == v;
        // a neighbour counts if it is outside the tile in some direction and
        // inside the image in both
        const bool oL = lx == 0, oR = lx == tw - 1, oU = ly == 0, oD = ly == th - 1;
        const bool iL = !oL || hasL, iR = !oR || hasR, iU = !oU || hasU, iD = !oD || hasD;
        const uint32_t sul = eul & ((oL || oU) && iL && iU), su = eu & (oU && iU), sur = eur & ((oR || oU) && iR && iU);
        const uint32_t sl = el & (oL && iL), sr = er & (oR && iR);
        const uint32_t sdl = edl & ((oL || oD) && iL && iD), sd = ed & (oD && iD), sdr = edr & ((oR || oD) && iR && iD);
        if (sul | su | sur | sl | sr | sdl | sd | sdr) seam_bits |= 1u << k;
        if (oR) sm.eqr[ly] = (uint8_t)(sur | (sr << 1) | (sdr << 2));
        if (oD) sm.eqb[lx] = (uint8_t)(sdl | (sd << 1) | (sdr << 2));
    }
    return seam_bits;
}

// Labels of the non-singleton pixels: compression of the run heads, unions of
// the edges the Komura links do not cover (further contact segments above a
// run: a segment is one zone, its pixels being equal and horizontally
// adjacent; and the left contact of a run that starts the quad), then a
// read-only final compression that also ORs each run's evidence, and the seam
// evidence of the perimeter slots, into the roots.  Assumes the labels stored
// (store_labels) and a barrier; ends with one.
template <int T, int NT>
__device__ __forceinline__ void quad_label(TileSmem<T>& sm, const uint32_t (&st)[Quads<T, NT>::PER],
                                           const uint32_t (&sb)[Quads<T, NT>::PER], int tw, int th,
                                           uint32_t seam_bits) {
    using QS = QuadSet<T, NT>;
    const uint32_t base = QS::base();
    if (threadIdx.x == 0) sm.npairs = 0;
    // a self-linked head is already compressed
    QS::each(QS::gather([&](int k) { return q_heads(st[k]) & ~b_roots(sb[k]); }), [&](int k, int i) {
        const uint32_t p = base + k * T + i;
        sm.lab[p] = lfind_h(sm.lab, p);
    });
    __syncthreads();
    auto unite = [&](uint32_t a, uint32_t b) {
        const uint32_t j = atomicAdd(&sm.npairs, 1u);
        if (j < (uint32_t)TileSmem<T>::CAP_PAIRS)
            sm.pairs[j] = a | b << 16;
        else
            lunion_h(sm.lab, a, b);
    };
    QS::each(QS::gather([&](int k) { return q_unions(st[k]); }), [&](int k, int i) {
        const uint32_t s = QS::pick(st, k), p0 = base + k * T;
        const uint32_t ab = q_above(s, q_run(s, i));
        uint32_t seg = ab & ~(ab << 1);
        seg &= seg - 1u;      // the first segment is the head's link
        while (seg) {
            const int j = __builtin_ctz(seg);
            seg &= seg - 1u;
            unite(p0 + i, p0 - T + j - 1);
        }
        // The left pixel needs a union only if the run touches neither of
        // its two upper neighbours: such a contact is one of the left
        // pixel's own contacts, to which both runs are already joined.
        if (i == 0 && ((s >> Q_EL) & 1u) && !(ab & 3u)) unite(p0, p0 - 1);
    });
    __syncthreads();
    {
        const uint32_t n = min(sm.npairs, (uint32_t)TileSmem<T>::CAP_PAIRS);
        for (uint32_t j = threadIdx.x; j < n; j += NT) {
            const uint32_t w = sm.pairs[j];
            lunion_h(sm.lab, w & 0xFFFFu, w >> 16);
        }
    }
    __syncthreads();
    QS::each(QS::gather([&](int k) { return q_heads(st[k]); }), [&](int k, int i) {
        const uint32_t s = QS::pick(st, k), p = base + k * T + i;
        const uint32_t r = lfindm(sm.lab, p);
        if (r != p) sm.lab[p] = r;
        const uint32_t run = q_run(s, i);
        const uint32_t f = (((s >> Q_G) & run) ? F_G : 0u) | (((s >> Q_S) & run) ? F_S : 0u);
        // skip the atomic when the root already holds the evidence: every
        // run of a plateau's rim would otherwise serialise on one word
        if (f & ~(lds_load(&sm.lab[r]) >> FLAG_SHIFT))
            __hip_atomic_fetch_or(&sm.lab[r], f << FLAG_SHIFT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    });
#pragma unroll
    for (int k = 0; k < Quads<T, NT>::PPER; ++k) {
        if (!((seam_bits >> k) & 1u)) continue;
        int lx, ly;
        perim_pixel<T>(threadIdx.x + k * NT, tw, th, lx, ly);
        const uint32_t r = lfindm(sm.lab, ly * T + lx);
        __hip_atomic_fetch_or(&sm.lab[r], F_SEAM << FLAG_SHIFT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __syncthreads();
}

// Stores four mask bytes of a quad: one dword when the quad is whole, else
// byte by byte (only the pixels in `sel`).
__device__ __forceinline__ void store_quad(uint8_t* dst, uint32_t packed, uint32_t sel) {
    if (sel == 0xFu) {
        store_u32(dst, packed);
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if ((sel >> i) & 1u) dst[i] = (uint8_t)(packed >> (8 * i));
    }
}

// Same with the selection as 0xFF bytes.
__device__ __forceinline__ void store_bytes(uint8_t* dst, uint32_t packed, uint32_t selb) {
    if (selb == 0xFFFFFFFFu) {
        store_u32(dst, packed);
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if ((selb >> (8 * i)) & 1u) dst[i] = (uint8_t)(packed >> (8 * i));
    }
}

// Perimeter records of the thread's seam slots: fragment j in every perimeter
// array the pixel belongs to.
template <int T, int NT>
__device__ __forceinline__ void write_perimeter_records(TileSmem<T>& sm, uint8_t* rec, int tw, int th,
                                                        uint32_t seam_bits, bool uniform) {
#pragma unroll
    for (int k = 0; k < Quads<T, NT>::PPER; ++k) {
        if (!((seam_bits >> k) & 1u)) continue;
        int lx, ly;
        perim_pixel<T>(threadIdx.x + k * NT, tw, th, lx, ly);
        uint8_t j = 0;
        if (!uniform) {
            const uint32_t jr = (sm.lab[lfindm(sm.lab, ly * T + lx)] >> FRAG_SHIFT) & 0xFFu;
            if (jr == 0u) {
                sm.err = 2u;
                continue;
            }
            j = (uint8_t)(jr - 1u);
        }
        if (ly == 0) rec[lx] = j;
        if (ly == th - 1) rec[T + lx] = j;
        if (lx == 0) rec[2 * T + ly] = j;
        if (lx == tw - 1) rec[3 * T + ly] = j;
    }
}

// Tile stage for one tile (ss_device.hpp, step 1).
template <int T, int NT>
__device__ __forceinline__ void tile_stage(TileSmem<T>& sm, const ImgDev& im, int64_t gtile, const Ws& ws) {
    constexpr int S = 4 * T - 4, PER = Quads<T, NT>::PER;
    // fragment slots are 32-bit indices, so an image has fewer than 2^32 tiles
    const uint32_t lt = (uint32_t)(gtile - im.tile_begin);
    const int tx = (int)(lt % (uint32_t)im.tiles_x), ty = (int)(lt / (uint32_t)im.tiles_x);
    const int x0 = tx * T, y0 = ty * T;
    const int tw = min(T, im.w - x0), th = min(T, im.h - y0);
    if (threadIdx.x == 0) {
        sm.count = 0;
        sm.err = 0;
        sm.zflags = 0;
        sm.open = 0;
    }
    tile_load<T, NT>(sm, im, x0, y0, tw, th);
    __syncthreads();
    const uint32_t v0 = sm.img[TileSmem<T>::P + 4];
    uint32_t st[PER], sb[PER];
    const bool mine_eq = quad_init<T, NT>(sm, tw, th, st, sb, v0);
    const uint32_t seam_bits = perimeter_pass<T, NT>(sm, im, x0, y0, tw, th);
    const bool uniform = __syncthreads_and(mine_eq) != 0;
    uint8_t* rec = ws.rec + gtile * 4 * T;

    if (uniform) {
        // One zone, root 0: OR the evidence, then everything is known.
        uint32_t f = seam_bits ? F_SEAM : 0u;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            f |= (sb[k] & (sb[k] << 3) & B_G) ? F_G : 0u;          // evidence of a valid pixel
            f |= ((sb[k] << 1) & (sb[k] << 3) & B_G) ? F_S : 0u;
        }
        if (f) atomicOr(&sm.zflags, f);
        __syncthreads();
        const uint32_t zf = sm.zflags;
        if (threadIdx.x == 0 && (zf & F_SEAM)) {
            sm.count = 1;
            ws.sbits[gtile * S] = (uint8_t)(zf & 3u);
        }
        const uint32_t mxb = (zf & F_G) ? 0u : 0xFFFFFFFFu, mnb = (zf & F_S) ? 0u : 0xFFFFFFFFu;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            if (!(sb[k] & B_VAL)) continue;
            int ly, lx0;
            Quads<T, NT>::at(k, ly, lx0);
            const uint32_t selb = bytes80((sb[k] << 3) & B_G);
            const int64_t o = (int64_t)(y0 + ly) * im.opitch + x0 + lx0;
            store_bytes(im.mx + o, mxb, selb);
            store_bytes(im.mn + o, mnb, selb);
        }
        write_perimeter_records<T, NT>(sm, rec, tw, th, seam_bits, true);
    } else {
        store_labels<T, NT>(sm, st, sb);
        __syncthreads();
        quad_label<T, NT>(sm, st, sb, tw, th, seam_bits);
        // Number the seam zones (their roots) and fill their fragment slots.
        {
            using QS = QuadSet<T, NT>;
            const uint32_t base = QS::base();
            QS::each(QS::gather([&](int k) { return b_roots(sb[k]); }), [&](int k, int i) {
                const uint32_t p = base + k * T + i;
                const uint32_t w = lds_load(&sm.lab[p]);
                if ((w & LMASK) != p || !((w >> FLAG_SHIFT) & F_SEAM)) return;
                const uint32_t j = atomicAdd(&sm.count, 1u);
                if (j >= (uint32_t)S) {
                    sm.err = 1u;
                    return;
                }
                sm.lab[p] = w | ((j + 1) << FRAG_SHIFT);
                ws.sbits[gtile * S + j] = (uint8_t)((w >> FLAG_SHIFT) & 3u);
                // a fragment that already has both kinds of evidence is final
                if (((w >> FLAG_SHIFT) & 3u) != 3u) sm.open = 1u;
            });
        }
        __syncthreads();
        // Masks (final for tile-contained zones, provisional for seam zones)
        // and, when the tile holds an open fragment, its per-pixel record
        // plane for the write-back.
        const bool open = sm.open != 0;
        uint8_t* plane = ws.rec8 + gtile * (int64_t)(T * T);
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const uint32_t b = sb[k];
            if (!(b & B_VAL)) continue;
            int ly, lx0;
            Quads<T, NT>::at(k, ly, lx0);
            const uint32_t selb = bytes80((b << 3) & B_G);
            // singletons straight from the evidence bits, the others from their root
            uint32_t mxw = bytes80((b << 2) & ~b & B_G), mnw = bytes80((b << 2) & ~(b << 1) & B_G);
            uint32_t om = 0, ids = 0;   // bytes of open fragments, and their ids
            uint32_t hd = q_heads(st[k]);
            while (hd) {
                const int i = __builtin_ctz(hd);
                hd &= hd - 1u;
                const uint32_t rb = bytes4(q_run(st[k], i));
                const uint32_t rw = sm.lab[sm.lab[ly * T + lx0 + i] & LMASK];
                const uint32_t zf = rw >> FLAG_SHIFT;
                const uint32_t jr = (rw >> FRAG_SHIFT) & 0xFFu;
                if (jr && (zf & 3u) != 3u) {
                    om |= rb;
                    ids |= (jr - 1u) * 0x01010101u & rb;
                }
                if (!(zf & F_G)) mxw |= rb;
                if (!(zf & F_S)) mnw |= rb;
            }
            const int64_t o = (int64_t)(y0 + ly) * im.opitch + x0 + lx0;
            store_bytes(im.mx + o, mxw, selb);
            store_bytes(im.mn + o, mnw, selb);
            if (open) {
                const uint32_t closed = REC_CLOSED * 0x01010101u | (mxw & 0x01010101u) | (mnw & 0x02020202u);
                *reinterpret_cast<uint32_t*>(plane + ly * T + lx0) = (closed & ~om) | ids;
            }
        }
        write_perimeter_records<T, NT>(sm, rec, tw, th, seam_bits, false);
    }
    __syncthreads();
    const uint32_t count = sm.count;
    const uint32_t tflags = (uniform ? CNT_UNIFORM : 0u) | (sm.open ? CNT_OPEN : 0u);
    if (count) {
        uint8_t* eq = ws.eq + gtile * 2 * T;
        for (int k = threadIdx.x; k < T; k += NT) {
            eq[k] = (k < th && tw == T) ? sm.eqr[k] : 0;
            eq[T + k] = (k < tw && th == T) ? sm.eqb[k] : 0;
        }
    }
    if (threadIdx.x == 0) ws.counts[gtile] = min(count, (uint32_t)S) | tflags | (sm.err << COUNT_ERR_SHIFT);
}

// Seam-sparsity counters of one tile (diagnostic, not on the extrema path),
// the quantities of the reference profiler
// (gpu_seam_sparsity_profiler_v0_1.cu:133-294): c[0] tile-local zones
// (singletons included), c[1] zones with an equal 8-neighbour in another tile,
// c[2] their pixels, c[3] equal 8-neighbour pairs with both pixels in the
// image, each counted once from its upper / left pixel (right, down-left,
// down, down-right), c[4] those pairs whose pixels lie in different tiles.
// Reuses the tile stage's staging, classification, seam tests and labelling.
template <int T, int NT>
__device__ __forceinline__ void tile_profile(TileSmem<T>& sm, const ImgDev& im, int64_t gtile, uint32_t* c) {
    constexpr int PER = Quads<T, NT>::PER;
    const uint32_t lt = (uint32_t)(gtile - im.tile_begin);
    const int tx = (int)(lt % (uint32_t)im.tiles_x), ty = (int)(lt / (uint32_t)im.tiles_x);
    const int x0 = tx * T, y0 = ty * T;
    const int tw = min(T, im.w - x0), th = min(T, im.h - y0);
    if (threadIdx.x == 0) sm.zflags = 0;
    tile_load<T, NT>(sm, im, x0, y0, tw, th);
    __syncthreads();
    const uint32_t v0 = sm.img[TileSmem<T>::P + 4];
    uint32_t st[PER], sb[PER];
    const bool mine_eq = quad_init<T, NT>(sm, tw, th, st, sb, v0);
    const uint32_t seam_bits = perimeter_pass<T, NT>(sm, im, x0, y0, tw, th);
    const bool uniform = __syncthreads_and(mine_eq) != 0;
    uint32_t zones = 0, seam_zones = 0, seam_px = 0, equal = 0, seam_edges = 0;
    // equal forward pairs from the staged rows (halo included; out-of-image
    // neighbours masked)
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        int ly, lx0;
        if (!Quads<T, NT>::at(k, ly, lx0) || ly >= th || lx0 >= tw) continue;
        const RowW b = row_at<T>(sm, ly, lx0), d = row_at<T>(sm, ly + 1, lx0);
        const uint32_t val = first_bytes80(tw - lx0);
        const int gx = x0 + lx0;
        const uint32_t right_img = first_bytes80(im.w - gx - 1), left_img = gx > 0 ? 0x80808080u : 0x80808000u;
        const uint32_t down_in_img = (y0 + ly + 1 < im.h) ? 0x80808080u : 0u;
        const uint32_t right_tile = first_bytes80(tw - lx0 - 1), left_tile = lx0 > 0 ? 0x80808080u : 0x80808000u;
        const uint32_t down_tile = (ly + 1 < th) ? 0x80808080u : 0u;
        const uint32_t eR = ~ne_top(b.r, b.w) & val & right_img;
        const uint32_t eD = ~ne_top(d.w, b.w) & val & down_in_img;
        const uint32_t eDL = ~ne_top(d.l, b.w) & val & left_img & down_in_img;
        const uint32_t eDR = ~ne_top(d.r, b.w) & val & right_img & down_in_img;
        equal += __builtin_popcount((eR | (eD >> 1) | (eDL >> 2) | (eDR >> 3)) & 0xF0F0F0F0u);
        seam_edges += __builtin_popcount(((eR & ~right_tile) | ((eD & ~down_tile) >> 1) |
                                          ((eDL & ~(left_tile & down_tile)) >> 2) |
                                          ((eDR & ~(right_tile & down_tile)) >> 3)) & 0xF0F0F0F0u);
    }
    if (uniform) {
        const bool any_seam = __syncthreads_or(seam_bits != 0) != 0;
        if (threadIdx.x == 0) {
            zones = 1;
            seam_zones = any_seam ? 1u : 0u;
            seam_px = any_seam ? (uint32_t)(tw * th) : 0u;
        }
    } else {
        __syncthreads();   // the equal-pair counts above read the staged rows
        store_labels<T, NT>(sm, st, sb);
        __syncthreads();
        quad_label<T, NT>(sm, st, sb, tw, th, seam_bits);
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            int ly, lx0;
            Quads<T, NT>::at(k, ly, lx0);
            zones += __builtin_popcount(sb[k] & B_SGL);   // singletons: never on a seam
            const uint32_t p0 = ly * T + lx0;
            uint32_t hd = q_heads(st[k]);
            while (hd) {
                const int i = __builtin_ctz(hd);
                hd &= hd - 1u;
                const uint32_t w = lds_load(&sm.lab[p0 + i]);
                const uint32_t rw = (w & LMASK) == p0 + i ? w : lds_load(&sm.lab[w & LMASK]);
                const bool seam = (rw >> FLAG_SHIFT) & F_SEAM;
                if ((w & LMASK) == p0 + i) {
                    ++zones;
                    seam_zones += seam ? 1u : 0u;
                }
                if (seam) seam_px += __builtin_popcount(q_run(st[k], i));
            }
        }
    }
    const uint32_t vals[5] = {zones, seam_zones, seam_px, equal, seam_edges};
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        const uint32_t v = wave_sum(vals[q]);
        if ((threadIdx.x & 63) == 0 && v) atomicAdd(&c[q], v);
    }
}

// Mark and write-back of one tile by one wave (ss_device.hpp, step 5).
// cw is the tile's counts word; only a tile with an open fragment, or a
// uniform tile (whose one fragment may be open), can change.  Lanes copy the
// tile's resolved fragment bytes (k_final) into the wave's 256-byte LDS table
// lut[wid] (bits 6 / 7: the fragment's max / min mask changed; bits 0-1: its
// final max / min masks; entries REC_CLOSED + c hold the closed code c, so a
// record byte of either kind is one lookup).  If any changed, the wave
// rewrites the tile: a uniform tile is
// one fragment; any other tile streams its record plane (fragment ids, or the
// final masks of closed pixels) and rewrites, per mask, every whole tile row
// that holds a pixel whose value in that mask changed, so HBM sees full row
// segments rather than scattered partial lines, and a mask no fragment of the
// row changed in is not written.  No image read, no relabelling.  b is the
// tile's base (dense id of its first fragment).  Returns whether the tile
// changed.
template <int T>
__device__ __forceinline__ bool tile_writeback(uint8_t* lut0, uint32_t wid, const ImgDev& im, int64_t gtile,
                                               uint32_t cw, uint32_t b, const Ws& ws) {
    static_assert(4 * T - 4 <= (int)REC_CLOSED, "fragment ids must stay below the closed codes");
    const int lane = threadIdx.x & 63;
    uint8_t* lut = lut0 + wid * 256;
    const uint32_t cnt = cw & COUNT_MASK;
    bool changed = false;
    for (uint32_t j = lane; j < cnt; j += 64) {
        const uint32_t v = ws.flut[b + j];
        changed = changed || (v & 0xC0u) != 0;
        lut[j] = (uint8_t)v;
    }
    if (!__ballot(changed)) return false;
    // the table was written by other lanes of this wave: LDS operations of one
    // wave complete in order; the fence keeps the compiler from reordering
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint32_t lt = (uint32_t)(gtile - im.tile_begin);
    const int tx = (int)(lt % (uint32_t)im.tiles_x), ty = (int)(lt / (uint32_t)im.tiles_x);
    const int x0 = tx * T, y0 = ty * T;
    const int tw = min(T, im.w - x0), th = min(T, im.h - y0);
    if (cw & CNT_UNIFORM) {
        const uint32_t v = lut[0];
        const uint32_t mxb = (v & 1u) ? 0xFFFFFFFFu : 0u, mnb = (v & 2u) ? 0xFFFFFFFFu : 0u;
        const bool wx = (v & 0x40u) != 0, wn = (v & 0x80u) != 0;
        for (int q = lane; q < T * T / 4; q += 64) {
            const int ly = q / (T / 4), lx0 = 4 * (q % (T / 4));
            if (ly >= th || lx0 >= tw) continue;
            const uint32_t sel = tw - lx0 >= 4 ? 0xFu : ((1u << (tw - lx0)) - 1u);
            const int64_t o = (int64_t)(y0 + ly) * im.opitch + x0 + lx0;
            if (wx) store_quad(im.mx + o, mxb, sel);
            if (wn) store_quad(im.mn + o, mnb, sel);
        }
        return true;
    }
    // record plane: 16 bytes per lane and step; the VR lanes of one tile row
    // are consecutive and aligned, so one ballot tells each lane whether its
    // row changed.  A changed row is rewritten whole: one full segment per
    // mask instead of partial lines that HBM must read back and merge.  Plane
    // bytes outside the image part of the tile were never written and count
    // as closed, unchanged pixels.
    const uint8_t* plane = ws.rec8 + gtile * (int64_t)(T * T);
    constexpr int VR = T / 16;   // 16-byte vectors per row
    constexpr uint32_t CL4 = REC_CLOSED * 0x01010101u;
    // a tile stage that wrote only the rows with an open pixel says which
    // (the others hold closed pixels, which cannot change)
    const uint64_t rows = (cw & CNT_ROWS) ? ws.orows[gtile] : ~0ull;
    for (int u = lane; u < T * T / 16; u += 64) {
        const int ly = u / VR, c0 = 16 * (u % VR);
        // a row without an open pixel holds closed pixels only: unchanged
        if (!((rows >> ly) & 1u) || ly >= th || c0 >= tw) continue;
        const uint4 rv = *reinterpret_cast<const uint4*>(plane + ly * T + c0);
        uint32_t rw[4] = {rv.x, rv.y, rv.z, rv.w};
        if (c0 + 16 > tw) {
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                const int nv = tw - c0 - 4 * d;   // valid bytes of this dword
                if (nv < 4) rw[d] = nv <= 0 ? CL4 : (rw[d] & (0xFFFFFFFFu >> (32 - 8 * nv))) | (CL4 << (8 * nv));
            }
        }
        // one table lookup per byte: the LDS address lut0 + wid * 256 + byte
        // is assembled by one byte permute (byte 0: the record byte, byte 1:
        // wid); the four results are packed back into a dword
        uint32_t mxw[4], mnw[4];
        uint32_t ch = 0;   // bits 6 / 7 of any byte: a max / min value changed
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            const uint32_t w = rw[d];
            const uint32_t v = (uint32_t)lut0[__builtin_amdgcn_perm(wid, w, 0x0C0C0400u)] |
                               (uint32_t)lut0[__builtin_amdgcn_perm(wid, w, 0x0C0C0401u)] << 8 |
                               (uint32_t)lut0[__builtin_amdgcn_perm(wid, w, 0x0C0C0402u)] << 16 |
                               (uint32_t)lut0[__builtin_amdgcn_perm(wid, w, 0x0C0C0403u)] << 24;
            ch |= v;
            const uint32_t x = v & 0x01010101u, n = v & 0x02020202u;
            mxw[d] = (x << 8) - x;          // bit 0 of each byte -> 0x00 / 0xFF
            mnw[d] = (n << 7) - (n >> 1);   // bit 1 of each byte -> 0x00 / 0xFF
        }
        const uint32_t rowbits = (1u << VR) - 1u, first = (uint32_t)lane & ~(VR - 1u);
        const bool wx = (__ballot((ch & 0x40404040u) != 0) >> first) & rowbits;
        const bool wn = (__ballot((ch & 0x80808080u) != 0) >> first) & rowbits;
        if (!(wx || wn)) continue;
        const int64_t o = (int64_t)(y0 + ly) * im.opitch + x0 + c0;
        if (c0 + 16 <= tw) {
            if (wx) store_u128(im.mx + o, make_uint4(mxw[0], mxw[1], mxw[2], mxw[3]));
            if (wn) store_u128(im.mn + o, make_uint4(mnw[0], mnw[1], mnw[2], mnw[3]));
        } else {
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                const int nv = tw - c0 - 4 * d;
                if (nv <= 0) break;
                const uint32_t sel = nv >= 4 ? 0xFu : ((1u << nv) - 1u);
                if (wx) store_quad(im.mx + o + 4 * d, mxw[d], sel);
                if (wn) store_quad(im.mn + o + 4 * d, mnw[d], sel);
            }
        }
    }
    return true;
}

}  // namespace ssx
