// ss_tile.hpp -- tile-local stage of the dual regional-extrema pipeline:
// LDS staging, Komura flat-zone labelling, evidence, fragments, masks.
//
// Work mapping: a tile of T x T pixels is cut into quads of 4 horizontally
// adjacent pixels; each thread owns the same quads through every phase and
// keeps their per-pixel evidence and edge bits in one register word per quad,
// so LDS holds only the image (for neighbour reads) and the shared labels.
// LDS image rows are dword aligned (pixel (lx, ly) at (ly + 1) * P + 4 + lx),
// so a quad reads its 3 x 6 neighbourhood with nine 32-bit LDS loads.
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

// Per-quad register word, pixel i = 0..3 of the quad:
constexpr int Q_FLAGS = 0;    // bits 3i..3i+2: F_G | F_S | F_SEAM
constexpr int Q_EQL = 12;     // bit 12+2i: equal left neighbour in the tile
constexpr int Q_EQUR = 13;    // bit 13+2i: equal up-right neighbour in the tile
constexpr int Q_SGL = 20;     // bit 20+i: singleton (no equal 8-neighbour)
constexpr int Q_VAL = 24;     // bit 24+i: pixel lies inside the image part of the tile

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
    static constexpr int D = T / 4;                 // quads per row
    static constexpr int Q = T * T / 4;             // quads per tile
    static constexpr int PER = (Q + NT - 1) / NT;   // quads per thread
};

__device__ __forceinline__ uint32_t lds32(const uint8_t* p) { return *reinterpret_cast<const uint32_t*>(p); }

__device__ __forceinline__ uint32_t lfindm(const uint32_t* lab, uint32_t x) {
    uint32_t r = lds_load(lab + x) & LMASK;
    for (;;) {
        const uint32_t n = lds_load(lab + r) & LMASK;
        if (n == r) return r;
        r = n;
    }
}

// Loads the tile (rows 0..th-1) and, with HALO, the halo rows -1 and th and the
// halo columns -1 and tw.  Halo pixels outside the image replicate the nearest
// image pixel: such a value belongs to the pixel itself or to one of its
// genuine 8-neighbours, so the 3x3 greater / smaller tests equal the "absent
// neighbour" rule exactly (equality across the border is masked separately).
template <int T, int NT, bool HALO>
__device__ __forceinline__ void tile_load(TileSmem<T>& sm, const ImgDev& im, int x0, int y0, int tw,
                                          int th) {
    constexpr int P = TileSmem<T>::P, D = T / 4;
    const bool al4 = ((reinterpret_cast<uintptr_t>(im.img) | (uintptr_t)im.pitch) & 3u) == 0;
    const int r0 = HALO ? -1 : 0;
    const int nrows = HALO ? th + 2 : th;
    for (int it = threadIdx.x; it < nrows * D; it += NT) {
        const int r = r0 + it / D, c = 4 * (it % D);
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
    if (HALO) {
        for (int it = threadIdx.x; it < 2 * (th + 2); it += NT) {
            const int r = it / 2 - 1, col = (it & 1) ? tw : -1;
            int gy = y0 + r, gx = x0 + col;
            gy = gy < 0 ? 0 : (gy >= im.h ? im.h - 1 : gy);
            gx = gx < 0 ? 0 : (gx >= im.w ? im.w - 1 : gx);
            sm.img[(r + 1) * P + 4 + col] = im.img[(int64_t)gy * im.pitch + gx];
        }
    }
}

// First pass over the thread's quads: Komura init (link to the first equal
// predecessor in the order up, up-left, up-right, left), the edge bits for the
// union pass, singleton detection and, with CLASSIFY, the pixel evidence and
// the cross-seam equality bits.  Returns whether every pixel equals v0.
template <int T, int NT, bool CLASSIFY>
__device__ __forceinline__ bool quad_init(TileSmem<T>& sm, const ImgDev& im, int x0, int y0, int tw,
                                          int th, uint32_t (&st)[Quads<T, NT>::PER], uint32_t v0) {
    constexpr int P = TileSmem<T>::P, D = Quads<T, NT>::D, Q = Quads<T, NT>::Q;
    const bool hasL = x0 > 0, hasR = x0 + tw < im.w, hasU = y0 > 0, hasD = y0 + th < im.h;
    bool all_eq = true;
#pragma unroll
    for (int k = 0; k < Quads<T, NT>::PER; ++k) {
        st[k] = 0;
        const int q = threadIdx.x + k * NT;
        if (Q % NT != 0 && q >= Q) continue;
        const int ly = q / D, lx0 = 4 * (q % D);
        if (ly >= th || lx0 >= tw) continue;
        const int base = (ly + 1) * P + 4 + lx0;
        uint32_t U[6], C[6], L[6];
        {
            const uint32_t a0 = lds32(&sm.img[base - P - 4]), a1 = lds32(&sm.img[base - P]),
                           a2 = lds32(&sm.img[base - P + 4]);
            const uint32_t b0 = lds32(&sm.img[base - 4]), b1 = lds32(&sm.img[base]),
                           b2 = lds32(&sm.img[base + 4]);
            const uint32_t c0 = lds32(&sm.img[base + P - 4]), c1 = lds32(&sm.img[base + P]),
                           c2 = lds32(&sm.img[base + P + 4]);
            U[0] = a0 >> 24; C[0] = b0 >> 24; L[0] = c0 >> 24;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                U[b + 1] = (a1 >> (8 * b)) & 0xFFu;
                C[b + 1] = (b1 >> (8 * b)) & 0xFFu;
                L[b + 1] = (c1 >> (8 * b)) & 0xFFu;
            }
            U[5] = a2 & 0xFFu; C[5] = b2 & 0xFFu; L[5] = c2 & 0xFFu;
        }
        uint32_t s = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int lx = lx0 + i;
            if (lx >= tw) continue;
            const int p = ly * T + lx;
            const uint32_t v = C[i + 1];
            const uint32_t nu = U[i + 1], nul = U[i], nur = U[i + 2], nl = C[i];
            const uint32_t nr = C[i + 2], ndl = L[i], nd = L[i + 1], ndr = L[i + 2];
            all_eq = all_eq && v == v0;
            const bool eu = ly > 0 && nu == v;
            const bool eul = ly > 0 && lx > 0 && nul == v;
            const bool eur = ly > 0 && lx + 1 < tw && nur == v;
            const bool el = lx > 0 && nl == v;
            const bool er = lx + 1 < tw && nr == v;
            const bool edl = ly + 1 < th && lx > 0 && ndl == v;
            const bool ed = ly + 1 < th && nd == v;
            const bool edr = ly + 1 < th && lx + 1 < tw && ndr == v;
            s |= 1u << (Q_VAL + i);
            uint32_t f = 0;
            bool seam = false;
            if (CLASSIFY) {
                const uint32_t hi = max(max(max(nu, nul), max(nur, nl)), max(max(nr, ndl), max(nd, ndr)));
                const uint32_t lo = min(min(min(nu, nul), min(nur, nl)), min(min(nr, ndl), min(nd, ndr)));
                f = (hi > v ? F_G : 0u) | (lo < v ? F_S : 0u);
                if (lx == 0 || ly == 0 || lx == tw - 1 || ly == th - 1) {
                    // equal neighbour outside the tile but inside the image
                    auto sq = [&](int dx, int dy, uint32_t n) -> uint32_t {
                        const int nx = lx + dx, ny = ly + dy;
                        const bool out = nx < 0 || nx >= tw || ny < 0 || ny >= th;
                        const bool in_img = (nx >= 0 || hasL) && (nx < tw || hasR) && (ny >= 0 || hasU) &&
                                            (ny < th || hasD);
                        return (out && in_img && n == v) ? 1u : 0u;
                    };
                    const uint32_t sul = sq(-1, -1, nul), su = sq(0, -1, nu), sur = sq(1, -1, nur);
                    const uint32_t sl = sq(-1, 0, nl), sr = sq(1, 0, nr);
                    const uint32_t sdl = sq(-1, 1, ndl), sd = sq(0, 1, nd), sdr = sq(1, 1, ndr);
                    seam = (sul | su | sur | sl | sr | sdl | sd | sdr) != 0;
                    if (lx == tw - 1) sm.eqr[ly] = (uint8_t)(sur | (sr << 1) | (sdr << 2));
                    if (ly == th - 1) sm.eqb[lx] = (uint8_t)(sdl | (sd << 1) | (sdr << 2));
                    if (seam) f |= F_SEAM;
                }
            }
            const bool single = !(eu | eul | eur | el | er | edl | ed | edr) && !seam;
            s |= f << (Q_FLAGS + 3 * i);
            if (el) s |= 1u << (Q_EQL + 2 * i);
            if (eur) s |= 1u << (Q_EQUR + 2 * i);
            if (single) {
                s |= 1u << (Q_SGL + i);
            } else {
                uint32_t parent = p;
                if (eu) parent = p - T;
                else if (eul) parent = p - T - 1;
                else if (eur) parent = p - T + 1;
                else if (el) parent = p - 1;
                sm.lab[p] = parent;
            }
        }
        st[k] = s;
    }
    return all_eq;
}

// Labels of the non-singleton pixels: compress, unite the left and up-right
// edges (min-root CAS), compress.  With OR_FLAGS the final compression also
// ORs each pixel's evidence into its root.  Assumes a barrier after quad_init;
// ends with a barrier.
template <int T, int NT, bool OR_FLAGS>
__device__ __forceinline__ void quad_label(TileSmem<T>& sm, const uint32_t (&st)[Quads<T, NT>::PER]) {
    constexpr int D = Quads<T, NT>::D;
#pragma unroll
    for (int k = 0; k < Quads<T, NT>::PER; ++k) {
        const int q = threadIdx.x + k * NT, p0 = (q / D) * T + 4 * (q % D);
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if ((st[k] >> (Q_VAL + i)) & ~(st[k] >> (Q_SGL + i)) & 1u) sm.lab[p0 + i] = lfindm(sm.lab, p0 + i);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < Quads<T, NT>::PER; ++k) {
        const int q = threadIdx.x + k * NT, p0 = (q / D) * T + 4 * (q % D);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if ((st[k] >> (Q_EQL + 2 * i)) & 1u) lunion(sm.lab, p0 + i, p0 + i - 1);
            if ((st[k] >> (Q_EQUR + 2 * i)) & 1u) lunion(sm.lab, p0 + i, p0 + i - T + 1);
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < Quads<T, NT>::PER; ++k) {
        const int q = threadIdx.x + k * NT, p0 = (q / D) * T + 4 * (q % D);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (!((st[k] >> (Q_VAL + i)) & ~(st[k] >> (Q_SGL + i)) & 1u)) continue;
            const uint32_t p = p0 + i;
            const uint32_t r = lfindm(sm.lab, p);
            if (r != p) sm.lab[p] = r;
            if (OR_FLAGS) {
                const uint32_t f = (st[k] >> (Q_FLAGS + 3 * i)) & 7u;
                if (f) __hip_atomic_fetch_or(&sm.lab[r], f << FLAG_SHIFT, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
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
    tile_load<T, NT, true>(sm, im, x0, y0, tw, th);
    __syncthreads();
    const uint32_t v0 = sm.img[TileSmem<T>::P + 4];
    uint32_t st[PER];
    const bool mine_eq = quad_init<T, NT, true>(sm, im, x0, y0, tw, th, st, v0);
    const bool uniform = __syncthreads_and(mine_eq) != 0;
    const bool al_out = (((uintptr_t)im.mx | (uintptr_t)im.mn | (uintptr_t)im.opitch) & 3u) == 0;
    uint8_t* rec = ws.rec + gtile * 4 * T;

    if (uniform) {
        // One zone, root 0: OR the evidence, then everything is known.
        uint32_t f = 0;
#pragma unroll
        for (int k = 0; k < PER; ++k)
#pragma unroll
            for (int i = 0; i < 4; ++i) f |= (st[k] >> (Q_FLAGS + 3 * i)) & 7u;
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
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (!((st[k] >> (Q_FLAGS + 3 * i)) & F_SEAM)) continue;
                const int lx = lx0 + i;
                if (ly == 0) rec[lx] = 0;
                if (ly == th - 1) rec[T + lx] = 0;
                if (lx == 0) rec[2 * T + ly] = 0;
                if (lx == tw - 1) rec[3 * T + ly] = 0;
            }
        }
    } else {
        quad_label<T, NT, true>(sm, st);
        // Number the seam zones (their roots) and fill their fragment slots.
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int q = threadIdx.x + k * NT, p0 = (q / D) * T + 4 * (q % D);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (!((st[k] >> (Q_VAL + i)) & ~(st[k] >> (Q_SGL + i)) & 1u)) continue;
                const uint32_t p = p0 + i;
                const uint32_t w = sm.lab[p];
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
        // Masks (final for tile-contained zones, provisional for seam zones),
        // the perimeter records of seam pixels and, when the tile holds an open
        // fragment, its per-pixel record plane for the write-back.
        const bool open = sm.open != 0;
        uint8_t* plane = ws.rec8 + gtile * (int64_t)(T * T);
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const uint32_t sel = (st[k] >> Q_VAL) & 0xFu;
            if (!sel) continue;
            const int q = threadIdx.x + k * NT, ly = q / D, lx0 = 4 * (q % D);
            uint32_t mxw = 0, mnw = 0, rcw = 0xFFFFFFFFu;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (!((sel >> i) & 1u)) continue;
                const int lx = lx0 + i, p = ly * T + lx;
                uint32_t zf, rw = 0;
                if ((st[k] >> (Q_SGL + i)) & 1u) {
                    zf = (st[k] >> (Q_FLAGS + 3 * i)) & 7u;
                } else {
                    rw = sm.lab[sm.lab[p] & LMASK];
                    zf = rw >> FLAG_SHIFT;
                    const uint32_t jr = (rw >> FRAG_SHIFT) & 0xFFu;
                    if (jr && (zf & 3u) != 3u) rcw ^= (0xFFu ^ (jr - 1u)) << (8 * i);
                }
                mxw |= ((zf & F_G) ? 0u : 0xFFu) << (8 * i);
                mnw |= ((zf & F_S) ? 0u : 0xFFu) << (8 * i);
                if ((st[k] >> (Q_FLAGS + 3 * i)) & F_SEAM) {
                    const uint32_t jr = (rw >> FRAG_SHIFT) & 0xFFu;
                    if (jr == 0u) {
                        sm.err = 2u;
                        continue;
                    }
                    const uint8_t j = (uint8_t)(jr - 1u);
                    if (ly == 0) rec[lx] = j;
                    if (ly == th - 1) rec[T + lx] = j;
                    if (lx == 0) rec[2 * T + ly] = j;
                    if (lx == tw - 1) rec[3 * T + ly] = j;
                }
            }
            const int64_t o = (int64_t)(y0 + ly) * im.opitch + x0 + lx0;
            store_quad(im.mx + o, mxw, sel, al_out);
            store_quad(im.mn + o, mnw, sel, al_out);
            if (open) *reinterpret_cast<uint32_t*>(plane + ly * T + lx0) = rcw;
        }
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

// Write-back of one dirty tile (ss_device.hpp, step 6).  The tile's fragments
// become a small LDS table (changed -> new mask bits); a uniform tile is one
// fragment, any other dirty tile has a record plane, so the pass streams the
// plane and rewrites only the pixels of changed fragments -- no image read,
// no relabelling.
template <int T, int NT>
__device__ __forceinline__ void tile_writeback(TileSmem<T>& sm, const ImgDev& im, int64_t gtile, const Ws& ws) {
    constexpr int D = Quads<T, NT>::D, PER = Quads<T, NT>::PER;
    const int64_t lt = gtile - im.tile_begin;
    const int tx = (int)(lt % im.tiles_x), ty = (int)(lt / im.tiles_x);
    const int x0 = tx * T, y0 = ty * T;
    const int tw = min(T, im.w - x0), th = min(T, im.h - y0);
    const uint32_t cw = ws.counts[gtile];
    const uint32_t cnt = cw & COUNT_MASK;
    const uint32_t b = ws.base[gtile];
    for (uint32_t j = threadIdx.x; j < cnt; j += NT) {
        const uint32_t own = ws.dbits[b + j];
        uint32_t v = 0;
        if (own & DB_CHANGED) {
            const uint32_t res = (ws.dbits[gfind_ro(ws.dparent, b + j)] >> 2) | own;
            v = 0x80u | ((res & F_G) ? 0u : 1u) | ((res & F_S) ? 0u : 2u);
        }
        sm.lut[j] = (uint8_t)v;
    }
    __syncthreads();
    const bool al_out = (((uintptr_t)im.mx | (uintptr_t)im.mn | (uintptr_t)im.opitch) & 3u) == 0;
    const bool uniform = (cw & CNT_UNIFORM) != 0;
    const uint8_t* plane = ws.rec8 + gtile * (int64_t)(T * T);
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int q = threadIdx.x + k * NT;
        if (Quads<T, NT>::Q % NT != 0 && q >= Quads<T, NT>::Q) continue;
        const int ly = q / D, lx0 = 4 * (q % D);
        if (ly >= th || lx0 >= tw) continue;
        const uint32_t rcw = uniform ? 0u : *reinterpret_cast<const uint32_t*>(plane + ly * T + lx0);
        uint32_t mxw = 0, mnw = 0, sel = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t r = (rcw >> (8 * i)) & 0xFFu;
            if (lx0 + i >= tw || r == REC_NONE) continue;
            const uint32_t v = sm.lut[r];
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
    __syncthreads();
}

}  // namespace ssx
