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

namespace ssx {

// Per-quad register word: eight 4-bit fields, bit i = pixel i of the quad.
constexpr int Q_G = 0;        // a strictly greater 8-neighbour exists
constexpr int Q_S = 4;        // a strictly smaller 8-neighbour exists
constexpr int Q_EL = 8;       // equal left neighbour in the tile
constexpr int Q_EU = 12;      // equal up neighbour in the tile
constexpr int Q_EUL = 16;     // equal up-left neighbour in the tile
constexpr int Q_EUR = 20;     // equal up-right neighbour in the tile
constexpr int Q_SGL = 24;     // singleton (no equal 8-neighbour)
constexpr int Q_VAL = 28;     // pixel lies inside the image part of the tile

// Runs: the non-singleton pixels of a quad split into maximal runs of equal
// horizontal neighbours.  Only a run's first pixel (its head) is a union-find
// node; the other pixels of the run point at the head and never change.  A
// plateau quad is one node instead of four.
__device__ __forceinline__ uint32_t q_nsg(uint32_t st) { return (st >> Q_VAL) & ~(st >> Q_SGL) & 0xFu; }
__device__ __forceinline__ uint32_t q_heads(uint32_t st) { return q_nsg(st) & ~((st >> Q_EL) & 0xEu); }
// Pixels of the run that starts at head h.
__device__ __forceinline__ uint32_t q_run(uint32_t st, int h) {
    const uint32_t c = (st >> Q_EL) & 0xEu;
    uint32_t r = 1u << h;
    r |= (r << 1) & c;
    r |= (r << 1) & c;
    r |= (r << 1) & c;
    return r;
}
// Row-above contacts of run r: bit j set iff the pixel above at quad column
// j - 1 (j = 0..5) equals the run's value.
__device__ __forceinline__ uint32_t q_above(uint32_t st, uint32_t r) {
    return (((st >> Q_EUL) & r) | (((st >> Q_EU) & r) << 1) | (((st >> Q_EUR) & r) << 2)) & 0x3Fu;
}

// SWAR helpers on four bytes (one per pixel of a quad).
// Top bit of every byte -> 4-bit mask.
__device__ __forceinline__ uint32_t movemask4(uint32_t w) {
    return (((w >> 7) & 0x01010101u) * 0x01020408u) >> 24;
}
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
    static constexpr int P = T + 8;   // LDS row pitch in bytes
    alignas(16) uint8_t img[(T + 2) * P];
    alignas(16) uint32_t lab[T * T];  // labels; at roots also evidence and fragment id
    uint8_t lut[4 * T];               // write-back: fragment -> 0x80 | new mask bits, or 0
    uint8_t eqr[T];                   // cross-seam equality bits, right column
    uint8_t eqb[T];                   // cross-seam equality bits, bottom row
    uint32_t count;
    uint32_t err;
    uint32_t next;                    // work-queue broadcast
    uint32_t zflags;                  // uniform tile: OR of the evidence
    uint32_t open;                    // the tile holds a fragment that may still change
};

template <int T, int NT>
struct Quads {
    static constexpr int D = T / 4;                  // quads per row
    static constexpr int Q = T * T / 4;              // quads per tile
    static constexpr int PER = (Q + NT - 1) / NT;    // quads per thread
    static constexpr int PPER = (4 * T + NT - 1) / NT;  // perimeter slots per thread
};

__device__ __forceinline__ uint32_t lds32(const uint8_t* p) { return *reinterpret_cast<const uint32_t*>(p); }

// LDS find with path halving (label words of non-roots carry no flag bits).
// Not for the final compression: a halving store could replace a finished
// root entry with a mere ancestor.
__device__ __forceinline__ uint32_t lfind_h(uint32_t* lab, uint32_t x) {
    for (;;) {
        const uint32_t p = lds_load(lab + x) & LMASK;
        if (p == x) return x;
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

// Loads the tile and its halo rows -1, th and halo columns -1, tw.  Halo
// pixels outside the image replicate the nearest image pixel: such a value
// belongs to the pixel itself or to one of its genuine 8-neighbours, so the
// 3x3 greater / smaller tests equal the "absent neighbour" rule exactly
// (equality across the image border is excluded by the perimeter pass).
template <int T, int NT>
__device__ __forceinline__ void tile_load(TileSmem<T>& sm, const ImgDev& im, int x0, int y0, int tw, int th) {
    constexpr int P = TileSmem<T>::P, D = T / 4;
    const bool al4 = ((reinterpret_cast<uintptr_t>(im.img) | (uintptr_t)im.pitch) & 3u) == 0;
    for (int it = threadIdx.x; it < (th + 2) * D; it += NT) {
        const int r = it / D - 1, c = 4 * (it % D);
        if (c >= tw) continue;
        int gy = y0 + r;
        gy = gy < 0 ? 0 : (gy >= im.h ? im.h - 1 : gy);
        const uint8_t* src = im.img + (int64_t)gy * im.pitch + x0 + c;
        uint8_t* dst = &sm.img[(r + 1) * P + 4 + c];
        if (al4 && c + 3 < tw) {
            *reinterpret_cast<uint32_t*>(dst) = *reinterpret_cast<const uint32_t*>(src);
        } else {
            for (int b = 0; b < 4 && c + b < tw; ++b) dst[b] = src[b];
        }
    }
    for (int it = threadIdx.x; it < 2 * (th + 2); it += NT) {
        const int r = it / 2 - 1, col = (it & 1) ? tw : -1;
        int gy = y0 + r, gx = x0 + col;
        gy = gy < 0 ? 0 : (gy >= im.h ? im.h - 1 : gy);
        gx = gx < 0 ? 0 : (gx >= im.w ? im.w - 1 : gx);
        sm.img[(r + 1) * P + 4 + col] = im.img[(int64_t)gy * im.pitch + gx];
    }
}

// First pass over the thread's quads: pixel evidence, Komura init (link to the
// first equal predecessor in the order up, up-left, up-right, left), the edge
// bits of the union pass and singleton detection (no equal value among the
// eight LDS neighbours, halo included -- conservative at the image border,
// where a replicated halo value may only make a pixel non-singleton).
// Returns whether every pixel equals v0.
template <int T, int NT>
__device__ __forceinline__ bool quad_init(TileSmem<T>& sm, int tw, int th, uint32_t (&st)[Quads<T, NT>::PER],
                                          uint32_t v0) {
    constexpr int P = TileSmem<T>::P, D = Quads<T, NT>::D, Q = Quads<T, NT>::Q;
    bool all_eq = true;
#pragma unroll
    for (int k = 0; k < Quads<T, NT>::PER; ++k) {
        st[k] = 0;
        const int q = threadIdx.x + k * NT;
        if (Q % NT != 0 && q >= Q) continue;
        const int ly = q / D, lx0 = 4 * (q % D);
        if (ly >= th || lx0 >= tw) continue;
        const int base = (ly + 1) * P + 4 + lx0;
        const uint32_t a0 = lds32(&sm.img[base - P - 4]), a1 = lds32(&sm.img[base - P]),
                       a2 = lds32(&sm.img[base - P + 4]);
        const uint32_t b0 = lds32(&sm.img[base - 4]), b1 = lds32(&sm.img[base]), b2 = lds32(&sm.img[base + 4]);
        const uint32_t c0 = lds32(&sm.img[base + P - 4]), c1 = lds32(&sm.img[base + P]),
                       c2 = lds32(&sm.img[base + P + 4]);
        // Neighbour words: byte i is the neighbour of pixel i in that direction.
        const uint32_t V = b1;
        const uint32_t nU = ne_top(a1, V), nUL = ne_top(__builtin_amdgcn_alignbyte(a1, a0, 3), V),
                       nUR = ne_top(__builtin_amdgcn_alignbyte(a2, a1, 1), V);
        const uint32_t nL = ne_top(__builtin_amdgcn_alignbyte(b1, b0, 3), V),
                       nR = ne_top(__builtin_amdgcn_alignbyte(b2, b1, 1), V);
        const uint32_t nD = ne_top(c1, V), nDL = ne_top(__builtin_amdgcn_alignbyte(c1, c0, 3), V),
                       nDR = ne_top(__builtin_amdgcn_alignbyte(c2, c1, 1), V);
        // 3x3 max / min (self included: it never makes a neighbour greater or
        // smaller) on even / odd 16-bit lanes.
        const RowMinMax ru = row_minmax(a0, a1, a2), rc = row_minmax(b0, b1, b2), rd = row_minmax(c0, c1, c2);
        const uint32_t ge = pk_subsat(pk_max(pk_max(ru.maxe, rc.maxe), rd.maxe), rc.mid_e);
        const uint32_t go = pk_subsat(pk_max(pk_max(ru.maxo, rc.maxo), rd.maxo), rc.mid_o);
        const uint32_t se = pk_subsat(rc.mid_e, pk_min(pk_min(ru.mine, rc.mine), rd.mine));
        const uint32_t so = pk_subsat(rc.mid_o, pk_min(pk_min(ru.mino, rc.mino), rd.mino));
        const uint32_t g4 = movemask4((nz_top16(ge) >> 8) | nz_top16(go));
        const uint32_t s4 = movemask4((nz_top16(se) >> 8) | nz_top16(so));
        const uint32_t val4 = tw - lx0 >= 4 ? 0xFu : ((1u << (tw - lx0)) - 1u);
        const uint32_t lv4 = lx0 > 0 ? 0xFu : 0xEu;   // pixel has a left neighbour in the tile
        // pixel i has a right neighbour in the tile iff lx0 + i + 1 < tw
        const uint32_t rv4 = tw - lx0 - 1 >= 4 ? 0xFu : ((1u << (tw - lx0 - 1)) - 1u);
        // pixels past tw hold no image data: every edge bit is masked to valid pixels
        const uint32_t up4 = ly > 0 ? val4 : 0u;
        const uint32_t any_eq4 = movemask4(~(nU & nUL & nUR & nL & nR & nD & nDL & nDR) & 0x80808080u);
        const uint32_t eu4 = movemask4(nU ^ 0x80808080u) & up4;
        const uint32_t eul4 = movemask4(nUL ^ 0x80808080u) & up4 & lv4;
        const uint32_t eur4 = movemask4(nUR ^ 0x80808080u) & up4 & rv4;
        const uint32_t el4 = movemask4(nL ^ 0x80808080u) & lv4 & val4;
        all_eq = all_eq && (movemask4(ne_top(V, v0 * 0x01010101u)) & val4) == 0;
        const uint32_t sgl4 = ~any_eq4 & val4;
        const uint32_t s = g4 << Q_G | s4 << Q_S | el4 << Q_EL | eu4 << Q_EU | eul4 << Q_EUL | eur4 << Q_EUR |
                           sgl4 << Q_SGL | val4 << Q_VAL;
        st[k] = s;
        const uint32_t nsg = val4 & ~sgl4;
        if (!nsg) continue;
        // Labels of the quad in one 16-byte store: a head links to its first
        // contact above (else to the pixel left of the quad, else to itself),
        // every other run pixel to its head.  Singletons get a dummy entry
        // that nothing reads.  Every link points at a smaller index.
        const uint32_t p0 = ly * T + lx0;
        const uint32_t hd = nsg & ~(el4 & 0xEu);
        uint32_t lab4[4];
        uint32_t head = p0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if ((hd >> i) & 1u) {
                head = p0 + i;
                const uint32_t a = q_above(s, q_run(s, i));
                lab4[i] = a ? (p0 - T + __builtin_ctz(a) - 1u) : ((i == 0 && (el4 & 1u)) ? p0 - 1u : p0 + i);
            } else {
                lab4[i] = head;   // run pixel (or an unread dummy)
            }
        }
        *reinterpret_cast<uint4*>(&sm.lab[p0]) = make_uint4(lab4[0], lab4[1], lab4[2], lab4[3]);
    }
    return all_eq;
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
// of the right column and bottom row; returns one seam bit per slot.
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
        auto sq = [&](int dx, int dy) -> uint32_t {
            const int nx = lx + dx, ny = ly + dy;
            const bool out = nx < 0 || nx >= tw || ny < 0 || ny >= th;
            const bool in_img = (nx >= 0 || hasL) && (nx < tw || hasR) && (ny >= 0 || hasU) && (ny < th || hasD);
            return (out && in_img && c[dy * P + dx] == v) ? 1u : 0u;
        };
        const uint32_t sul = sq(-1, -1), su = sq(0, -1), sur = sq(1, -1);
        const uint32_t sl = sq(-1, 0), sr = sq(1, 0);
        const uint32_t sdl = sq(-1, 1), sd = sq(0, 1), sdr = sq(1, 1);
        if (sul | su | sur | sl | sr | sdl | sd | sdr) seam_bits |= 1u << k;
        if (lx == tw - 1) sm.eqr[ly] = (uint8_t)(sur | (sr << 1) | (sdr << 2));
        if (ly == th - 1) sm.eqb[lx] = (uint8_t)(sdl | (sd << 1) | (sdr << 2));
    }
    return seam_bits;
}

// Labels of the non-singleton pixels: compress, unite the left and up-right
// edges (min-root CAS), then a read-only final compression that also ORs each
// pixel's evidence, and the seam evidence of the perimeter slots, into the
// roots.  Assumes a barrier after quad_init; ends with a barrier.
template <int T, int NT>
__device__ __forceinline__ void quad_label(TileSmem<T>& sm, const uint32_t (&st)[Quads<T, NT>::PER], int tw,
                                           int th, uint32_t seam_bits) {
    constexpr int D = Quads<T, NT>::D;
#pragma unroll
    for (int k = 0; k < Quads<T, NT>::PER; ++k) {
        const int q = threadIdx.x + k * NT, p0 = (q / D) * T + 4 * (q % D);
        uint32_t hd = q_heads(st[k]);
        while (hd) {
            const int i = __builtin_ctz(hd);
            hd &= hd - 1u;
            sm.lab[p0 + i] = lfind_h(sm.lab, p0 + i);
        }
    }
    __syncthreads();
    // Unions of the edges the head links do not cover: the further contact
    // segments above a run (a segment is one zone: its pixels are equal and
    // horizontally adjacent) and the left contact of a run that starts the quad.
#pragma unroll
    for (int k = 0; k < Quads<T, NT>::PER; ++k) {
        const int q = threadIdx.x + k * NT, ly = q / D, p0 = ly * T + 4 * (q % D);
        uint32_t hd = q_heads(st[k]);
        while (hd) {
            const int i = __builtin_ctz(hd);
            hd &= hd - 1u;
            const uint32_t a = q_above(st[k], q_run(st[k], i));
            uint32_t seg = a & ~(a << 1);
            if (!seg) continue;   // linked left or alone: nothing more to unite
            seg &= seg - 1u;      // the first segment is the head's link
            while (seg) {
                const int j = __builtin_ctz(seg);
                seg &= seg - 1u;
                lunion_h(sm.lab, p0 + i, p0 - T + j - 1);
            }
            if (i == 0 && ((st[k] >> Q_EL) & 1u)) lunion_h(sm.lab, p0, p0 - 1);
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < Quads<T, NT>::PER; ++k) {
        const int q = threadIdx.x + k * NT, p0 = (q / D) * T + 4 * (q % D);
        uint32_t hd = q_heads(st[k]);
        while (hd) {
            const int i = __builtin_ctz(hd);
            hd &= hd - 1u;
            const uint32_t p = p0 + i;
            const uint32_t r = lfindm(sm.lab, p);
            if (r != p) sm.lab[p] = r;
            const uint32_t run = q_run(st[k], i);
            const uint32_t f = (((st[k] >> Q_G) & run) ? F_G : 0u) | (((st[k] >> Q_S) & run) ? F_S : 0u);
            // skip the atomic when the root already holds the evidence: every
            // run of a plateau's rim would otherwise serialise on one word
            if (f & ~(lds_load(&sm.lab[r]) >> FLAG_SHIFT))
                __hip_atomic_fetch_or(&sm.lab[r], f << FLAG_SHIFT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
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

// Stores four mask bytes of a quad: one dword when the quad is whole and the
// destination aligned, else byte by byte (only the pixels in `sel`).
__device__ __forceinline__ void store_quad(uint8_t* dst, uint32_t packed, uint32_t sel, bool aligned) {
    if (sel == 0xFu && aligned) {
        *reinterpret_cast<uint32_t*>(dst) = packed;
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if ((sel >> i) & 1u) dst[i] = (uint8_t)(packed >> (8 * i));
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
    constexpr int S = 4 * T - 4, D = Quads<T, NT>::D, PER = Quads<T, NT>::PER;
    const int64_t lt = gtile - im.tile_begin;
    const int tx = (int)(lt % im.tiles_x), ty = (int)(lt / im.tiles_x);
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
    uint32_t st[PER];
    const bool mine_eq = quad_init<T, NT>(sm, tw, th, st, v0);
    const uint32_t seam_bits = perimeter_pass<T, NT>(sm, im, x0, y0, tw, th);
    const bool uniform = __syncthreads_and(mine_eq) != 0;
    const bool al_out = (((uintptr_t)im.mx | (uintptr_t)im.mn | (uintptr_t)im.opitch) & 3u) == 0;
    uint8_t* rec = ws.rec + gtile * 4 * T;

    if (uniform) {
        // One zone, root 0: OR the evidence, then everything is known.
        uint32_t f = seam_bits ? F_SEAM : 0u;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const uint32_t val4 = (st[k] >> Q_VAL) & 0xFu;
            f |= ((st[k] >> Q_G) & val4) ? F_G : 0u;
            f |= ((st[k] >> Q_S) & val4) ? F_S : 0u;
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
            const uint32_t sel = (st[k] >> Q_VAL) & 0xFu;
            if (!sel) continue;
            const int q = threadIdx.x + k * NT, ly = q / D, lx0 = 4 * (q % D);
            const int64_t o = (int64_t)(y0 + ly) * im.opitch + x0 + lx0;
            store_quad(im.mx + o, mxb, sel, al_out);
            store_quad(im.mn + o, mnb, sel, al_out);
        }
        write_perimeter_records<T, NT>(sm, rec, tw, th, seam_bits, true);
    } else {
        quad_label<T, NT>(sm, st, tw, th, seam_bits);
        // Number the seam zones (their roots) and fill their fragment slots.
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int q = threadIdx.x + k * NT, p0 = (q / D) * T + 4 * (q % D);
            uint32_t hd = q_heads(st[k]);
            while (hd) {
                const int i = __builtin_ctz(hd);
                hd &= hd - 1u;
                const uint32_t p = p0 + i;
                const uint32_t w = lds_load(&sm.lab[p]);
                if ((w & LMASK) != p || !((w >> FLAG_SHIFT) & F_SEAM)) continue;
                const uint32_t j = atomicAdd(&sm.count, 1u);
                if (j >= (uint32_t)S) {
                    sm.err = 1u;
                    continue;
                }
                sm.lab[p] = w | ((j + 1) << FRAG_SHIFT);
                ws.sbits[gtile * S + j] = (uint8_t)((w >> FLAG_SHIFT) & 3u);
                // a fragment that already has both kinds of evidence is final
                if (((w >> FLAG_SHIFT) & 3u) != 3u) sm.open = 1u;
            }
        }
        __syncthreads();
        // Masks (final for tile-contained zones, provisional for seam zones)
        // and, when the tile holds an open fragment, its per-pixel record
        // plane for the write-back.
        const bool open = sm.open != 0;
        uint8_t* plane = ws.rec8 + gtile * (int64_t)(T * T);
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const uint32_t sel = (st[k] >> Q_VAL) & 0xFu;
            if (!sel) continue;
            const int q = threadIdx.x + k * NT, ly = q / D, lx0 = 4 * (q % D);
            // singletons straight from the evidence bits, the others from their root
            const uint32_t sgl = bytes4((st[k] >> Q_SGL) & 0xFu);
            uint32_t mxw = ~bytes4((st[k] >> Q_G) & 0xFu) & sgl, mnw = ~bytes4((st[k] >> Q_S) & 0xFu) & sgl;
            uint32_t rcw = 0xFFFFFFFFu;
            uint32_t hd = q_heads(st[k]);
            while (hd) {
                const int i = __builtin_ctz(hd);
                hd &= hd - 1u;
                const uint32_t rb = bytes4(q_run(st[k], i));
                const uint32_t rw = sm.lab[sm.lab[ly * T + lx0 + i] & LMASK];
                const uint32_t zf = rw >> FLAG_SHIFT;
                const uint32_t jr = (rw >> FRAG_SHIFT) & 0xFFu;
                if (jr && (zf & 3u) != 3u) rcw ^= (0xFFu ^ (jr - 1u)) * 0x01010101u & rb;
                if (!(zf & F_G)) mxw |= rb;
                if (!(zf & F_S)) mnw |= rb;
            }
            const int64_t o = (int64_t)(y0 + ly) * im.opitch + x0 + lx0;
            store_quad(im.mx + o, mxw, sel, al_out);
            store_quad(im.mn + o, mnw, sel, al_out);
            if (open) *reinterpret_cast<uint32_t*>(plane + ly * T + lx0) = rcw;
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

// Mark and write-back of one tile by one wave (ss_device.hpp, step 5).
// cw is the tile's counts word; only a tile with an open fragment, or a
// uniform tile (whose one fragment may be open), can change.  Lanes resolve
// the tile's fragments against the seam union-find; a fragment whose resolved
// evidence adds a bit to its local evidence changed, and its final mask bits
// go into the wave's LDS table.  If any changed, the wave rewrites only the
// pixels of changed fragments: a uniform tile is one fragment, any other tile
// streams its record plane -- no image read, no relabelling.  Returns whether
// the tile changed.
template <int T>
__device__ __forceinline__ bool tile_writeback(uint8_t* lut, const ImgDev& im, int64_t gtile, uint32_t cw,
                                               const Ws& ws) {
    const int lane = threadIdx.x & 63;
    const uint32_t cnt = cw & COUNT_MASK;
    const uint32_t b = ws.base[gtile];
    bool changed = false;
    for (uint32_t j = lane; j < cnt; j += 64) {
        const uint32_t own = ws.dbits[b + j];
        uint32_t v = 0;
        if ((own & 3u) != 3u) {
            const uint32_t res = ((ws.dbits[gfind_ro(ws.dparent, b + j)] >> 2) | own) & 3u;
            if (res & ~own) {
                v = 0x80u | ((res & F_G) ? 0u : 1u) | ((res & F_S) ? 0u : 2u);
                changed = true;
            }
        }
        lut[j] = (uint8_t)v;
    }
    if (!__ballot(changed)) return false;
    // the table was written by other lanes of this wave: LDS operations of one
    // wave complete in order; the fence keeps the compiler from reordering
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int64_t lt = gtile - im.tile_begin;
    const int tx = (int)(lt % im.tiles_x), ty = (int)(lt / im.tiles_x);
    const int x0 = tx * T, y0 = ty * T;
    const int tw = min(T, im.w - x0), th = min(T, im.h - y0);
    const bool al_out = (((uintptr_t)im.mx | (uintptr_t)im.mn | (uintptr_t)im.opitch) & 3u) == 0;
    if (cw & CNT_UNIFORM) {
        const uint32_t v = lut[0];
        const uint32_t mxb = (v & 1u) ? 0xFFFFFFFFu : 0u, mnb = (v & 2u) ? 0xFFFFFFFFu : 0u;
        for (int q = lane; q < T * T / 4; q += 64) {
            const int ly = q / (T / 4), lx0 = 4 * (q % (T / 4));
            if (ly >= th || lx0 >= tw) continue;
            const uint32_t sel = tw - lx0 >= 4 ? 0xFu : ((1u << (tw - lx0)) - 1u);
            const int64_t o = (int64_t)(y0 + ly) * im.opitch + x0 + lx0;
            store_quad(im.mx + o, mxb, sel, al_out);
            store_quad(im.mn + o, mnb, sel, al_out);
        }
        return true;
    }
    // record plane: 16 bytes per lane and step, 4 quads each
    const uint8_t* plane = ws.rec8 + gtile * (int64_t)(T * T);
    constexpr int VR = T / 16;   // 16-byte vectors per row
    for (int u = lane; u < T * T / 16; u += 64) {
        const int ly = u / VR, c0 = 16 * (u % VR);
        if (ly >= th || c0 >= tw) continue;
        const uint4 rv = *reinterpret_cast<const uint4*>(plane + ly * T + c0);
        const uint32_t rw[4] = {rv.x, rv.y, rv.z, rv.w};
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            const int lx0 = c0 + 4 * d;
            if (lx0 >= tw) break;
            uint32_t mxw = 0, mnw = 0, sel = 0;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t r = (rw[d] >> (8 * i)) & 0xFFu;
                if (lx0 + i >= tw || r == REC_NONE) continue;
                const uint32_t v = lut[r];
                if (!(v & 0x80u)) continue;
                sel |= 1u << i;
                mxw |= ((v & 1u) ? 0xFFu : 0u) << (8 * i);
                mnw |= ((v & 2u) ? 0xFFu : 0u) << (8 * i);
            }
            if (!sel) continue;
            const int64_t o = (int64_t)(y0 + ly) * im.opitch + x0 + lx0;
            store_quad(im.mx + o, mxw, sel, al_out);
            store_quad(im.mn + o, mnw, sel, al_out);
        }
    }
    return true;
}

}  // namespace ssx
