// ss_tile64.hpp -- tile stage for T = 64: one workgroup per region of two
// horizontally adjacent 64 x 64 tiles, one lane per pixel column.
//
// The region's rows are 128 bytes wide, so image loads and mask stores cover
// whole cache lines from one workgroup.  Inside a tile every lane owns one
// pixel column and a wave walks a band of rows (WT waves per tile, 64 / WT
// rows each).  One vector compare per row gives a 64-bit mask of the whole
// row (a strictly greater 8-neighbour exists, the right / upper / ... pixel
// is equal), so the flat-zone structure of a row is handled as bit masks on
// the scalar unit while the vector unit classifies the next row:
//
//   runs      maximal sets of horizontally adjacent equal pixels of a row;
//             each run is one union-find node, numbered row * 64 + its
//             ordinal in the row (an mbcnt of the run-start mask).
//   contacts  the equal pixels of the row above that a run touches form
//             contiguous segments, one per run above; the first segment of a
//             run becomes its parent link, every further one a union.  Which
//             lanes open a new segment follows from the contact masks alone:
//             lane x does unless lane x - 1 of the same run touches the same
//             run above (its up-right or x's up-left contact is shared), and
//             a lane with up-left and up-right but no up contact opens a
//             second one.  The first segment of each run is found with a
//             carry along the run (add of the run-continuation mask).
//   evidence  a run has greater (smaller) evidence if any of its pixels has;
//             a carry carries it to the run's last pixel, whose lane stores
//             the run's flags.
//
// Union-find: 16-bit parents in LDS, min-root links (a node's parent is always
// a smaller id, so roots are minimum ids and labels are deterministic), path
// halving, 32-bit compare-and-swap on the word holding a parent.  Each wave
// builds the forest of its own band as it walks the rows (the band's first
// row starts unlinked, so every walk stays inside the band); after a barrier
// the first rows of the bands are joined to the bands above.  Then roots are
// compressed and collect their zone's evidence, seam roots are numbered as
// the tile's fragments, and every pixel's masks (and, for a tile holding a
// fragment that may still change, its record byte) are written.
//
// Outputs and their layout are those of tile_stage (ss_tile.hpp): both
// masks, counts word, per-fragment local evidence, perimeter records,
// cross-seam equality bits and the record plane.
#pragma once
#include <type_traits>

#include "ss_device.hpp"
#include "ss_tile.hpp"

namespace ssx {
namespace t64 {

constexpr int T = 64;
constexpr int RW = 2 * T;        // region width
constexpr int P = RW + 32;       // staged row pitch in bytes
constexpr int C0 = 16;           // byte of column 0 in a staged row (column -1 at 15)
constexpr int SROWS = T + 2;     // staged rows -1 .. 64
constexpr int S = 4 * T - 4;     // fragment slots per tile
constexpr uint32_t F_NUM = 8u;   // node flag: a numbered seam root (its id is in Smem::fid)

struct Smem {
    // The staged rows are read only until the row pass ends; the fragment ids
    // of the numbered roots are written after it.
    union {
        alignas(16) uint8_t img[SROWS * P];   // row r (-1 .. 64) at (r + 1) * P, column c at C0 + c
        alignas(16) uint8_t fid[2][T * T];    // fragment id of each numbered root
    };
    alignas(16) uint16_t par[2][T * T + 2];   // union-find parents of the run nodes (+ a sink slot)
    alignas(16) uint8_t flg[2][T * T + 4];    // a node's run evidence; at roots the zone's (+ sink)
    alignas(4) uint8_t eqr[2][T];         // cross-seam equality bits, right column
    alignas(4) uint8_t eqb[2][T];         // cross-seam equality bits, bottom row
    uint64_t rsm[2][T];                   // per row: run-start lanes
    uint32_t cnt[2], err[2], open[2], zf[2];
};

// The calling lane's bit of a wave-uniform mask.
__device__ __forceinline__ bool lane_in(uint64_t m) { return __builtin_amdgcn_inverse_ballot_w64(m); }
// b on the lanes of m, a elsewhere.
__device__ __forceinline__ uint32_t lsel(uint64_t m, uint32_t a, uint32_t b) { return lane_in(m) ? b : a; }
// Set bits of m on the lanes below the calling one.
__device__ __forceinline__ uint32_t mbcnt64(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
// Bit x of the result: some seed s < x has p[s .. x-1] all set (the seeds
// carried rightward through p).
__device__ __forceinline__ uint64_t carry_in(uint64_t p, uint64_t seeds) {
    const uint64_t a = seeds & p;
    return (p + a) ^ p ^ a;
}
// Lane x gets lane x - 1's value (lane 0 gets 0) / lane x + 1's (lane 63 gets 0).
__device__ __forceinline__ uint32_t from_left(uint32_t v) { return __builtin_amdgcn_update_dpp(0u, v, 0x138, 0xF, 0xF, false); }
__device__ __forceinline__ uint32_t from_right(uint32_t v) { return __builtin_amdgcn_update_dpp(0u, v, 0x130, 0xF, 0xF, false); }
__device__ __forceinline__ uint32_t max3u(uint32_t a, uint32_t b, uint32_t c) { return max(max(a, b), c); }
__device__ __forceinline__ uint32_t min3u(uint32_t a, uint32_t b, uint32_t c) { return min(min(a, b), c); }

__device__ __forceinline__ uint32_t ld16(const uint16_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void st16(uint16_t* p, uint32_t v) {
    __hip_atomic_store(p, (uint16_t)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t ld8(const uint8_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Root of x, halving the path (a halving store only replaces a non-root's
// parent by one of its ancestors).
__device__ __forceinline__ uint32_t find16(uint16_t* par, uint32_t x) {
    for (;;) {
        const uint32_t p = ld16(par + x);
        const uint32_t gp = ld16(par + p);   // at a root p == x and gp == p
        if (gp == p) return p;
        st16(par + x, gp);
        x = gp;
    }
}

// Root of x without stores (for walks that run beside compression stores).
__device__ __forceinline__ uint32_t find16_ro(const uint16_t* par, uint32_t x) {
    for (;;) {
        const uint32_t p = ld16(par + x);
        if (p == x) return x;
        x = p;
    }
}

// Links root hi under lo if hi is still a root (compare-and-swap of the
// 32-bit word that holds hi's parent); otherwise returns false and hi's
// current parent, a smaller id, in `now`.
__device__ __forceinline__ bool link16(uint16_t* par, uint32_t hi, uint32_t lo, uint32_t& now) {
    uint32_t* w = reinterpret_cast<uint32_t*>(par + (hi & ~1u));
    const uint32_t sh = (hi & 1u) * 16u;
    uint32_t old = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    for (;;) {
        const uint32_t cur = (old >> sh) & 0xFFFFu;
        if (cur != hi) {
            now = cur;
            return false;
        }
        const uint32_t nw = (old & ~(0xFFFFu << sh)) | (lo << sh);
        if (__hip_atomic_compare_exchange_strong(w, &old, nw, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_WORKGROUP))
            return true;
    }
}

// Min-root union; every retry lowers max(a, b), so it terminates.
__device__ __forceinline__ void union16(uint16_t* par, uint32_t a, uint32_t b) {
    for (;;) {
        a = find16(par, a);
        b = find16(par, b);
        if (a == b) return;
        const uint32_t hi = a > b ? a : b, lo = a > b ? b : a;
        uint32_t now;
        if (link16(par, hi, lo, now)) return;
        a = now;
        b = lo;
    }
}

__device__ __forceinline__ void or8(uint8_t* base, uint32_t i, uint32_t f) {
    __hip_atomic_fetch_or(reinterpret_cast<uint32_t*>(base + (i & ~3u)), f << (8u * (i & 3u)), __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Stages rows -1 .. th and columns -1 .. rw of the region (x0, y0): 16-byte
// loads, all issued before the first LDS store.  Rows and columns outside the
// image replicate the nearest image pixel, which is the pixel itself or one of
// its genuine 8-neighbours, so the 3 x 3 greater / smaller tests stay exact;
// equality across the image border is excluded by the seam tests.
template <int NT>
__device__ __forceinline__ void region_load(Smem& sm, const ImgDev& im, int x0, int y0, int rw, int th) {
    constexpr int CPR = RW / 16;            // 16-byte chunks per row
    constexpr int CH = SROWS * CPR;
    constexpr int PER = (CH + NT - 1) / NT;
    const int nrows = th + 2;
    const bool interior = x0 > 0 && y0 > 0 && x0 + RW < im.w && y0 + T < im.h;
    auto row_of = [&](int r) -> const uint8_t* {
        int gy = y0 + r;
        gy = gy < 0 ? 0 : (gy >= im.h ? im.h - 1 : gy);
        return im.img + (int64_t)gy * im.pitch + x0;
    };
    uint4 v[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int c = threadIdx.x + i * NT;
        const int r = c / CPR - 1, cx = 16 * (c % CPR);
        v[i] = make_uint4(0, 0, 0, 0);
        if (c >= CH || r + 1 >= nrows || cx >= rw) continue;
        const uint8_t* src = row_of(r) + cx;
        if (interior || cx + 16 <= rw) {
            __builtin_memcpy(&v[i], src, 16);
        } else {
            uint32_t w[4] = {0, 0, 0, 0};
#pragma unroll
            for (int k = 0; k < 16; ++k)
                if (cx + k < rw) w[k >> 2] |= (uint32_t)src[k] << (8 * (k & 3));
            v[i] = make_uint4(w[0], w[1], w[2], w[3]);
        }
    }
    // halo columns -1 and rw, one byte per staged row
    constexpr int HPER = (2 * SROWS + NT - 1) / NT;
    uint8_t hv[HPER];
#pragma unroll
    for (int i = 0; i < HPER; ++i) {
        const int k = threadIdx.x + i * NT;
        hv[i] = 0;
        if (k >= 2 * nrows) continue;
        int gx = (k & 1) ? x0 + rw : x0 - 1;
        gx = gx < 0 ? 0 : (gx >= im.w ? im.w - 1 : gx);
        hv[i] = row_of((k >> 1) - 1)[gx - x0];
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int c = threadIdx.x + i * NT;
        const int r = c / CPR - 1, cx = 16 * (c % CPR);
        if (c >= CH || r + 1 >= nrows || cx >= rw) continue;
        uint8_t* dst = &sm.img[(r + 1) * P + C0 + cx];
        if (interior || cx + 16 <= rw) {
            *reinterpret_cast<uint4*>(dst) = v[i];
        } else {
            // bytes from rw on stay untouched: column rw is the halo column
            const uint32_t w[4] = {v[i].x, v[i].y, v[i].z, v[i].w};
#pragma unroll
            for (int k = 0; k < 16; ++k)
                if (cx + k < rw) dst[k] = (uint8_t)(w[k >> 2] >> (8 * (k & 3)));
        }
    }
#pragma unroll
    for (int i = 0; i < HPER; ++i) {
        const int k = threadIdx.x + i * NT;
        if (k < 2 * nrows) sm.img[(k >> 1) * P + C0 + ((k & 1) ? rw : -1)] = hv[i];
    }
}

// Node and flag slot that masked-off lanes write to instead of branching.
constexpr uint32_t SINK = T * T;

// Tile stage of one region; WT waves per tile (64 / WT rows each).
template <int WT, bool MULTI>
__device__ __forceinline__ void region_stage(Smem& sm, const Images& ims, const Ws& ws, int64_t reg) {
    constexpr int NT = 2 * 64 * WT, ROWS = T / WT, TPT = 64 * WT;
    constexpr int PPER = (4 * T + TPT - 1) / TPT;   // perimeter slots per lane
    const ImgDev im = MULTI ? ims.table[ws.region_img[reg]] : ims.one;
    const uint32_t lr = (uint32_t)(reg - im.region_begin);
    const int rx = (int)(lr % (uint32_t)im.regions_x), ry = (int)(lr / (uint32_t)im.regions_x);
    const int x0 = rx * RW, y0 = ry * T;
    const int rw = min(RW, im.w - x0), th = min(T, im.h - y0);
    // wave-uniform: the wave's tile and band
    const int wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    const int q = threadIdx.x & (TPT - 1);        // thread within the tile
    const int t = wid / WT;                       // tile of the region
    const int ya = (wid % WT) * ROWS;             // first row of the wave's band
    const int ye = min(ya + ROWS, th);            // end of the band's rows inside the tile
    const int tx = 2 * rx + t;
    const bool tile_ok = tx < im.tiles_x;
    const int x0t = x0 + T * t;
    const int tw = tile_ok ? min(T, im.w - x0t) : 0;
    const int64_t gtile = im.tile_begin + (int64_t)ry * im.tiles_x + tx;
    if (q == 0) {
        sm.cnt[t] = 0;
        sm.err[t] = 0;
        sm.open[t] = 0;
        sm.zf[t] = 0;
    }
    region_load<NT>(sm, im, x0, y0, rw, th);
    __syncthreads();

    uint16_t* par = sm.par[t];
    uint8_t* flg = sm.flg[t];
    uint64_t* rsm = sm.rsm[t];
    const uint64_t VM = tw >= 64 ? ~0ull : ((1ull << tw) - 1ull);   // lanes inside the tile
    const uint64_t VR = VM & (VM >> 1);                              // ... whose right neighbour is too
    const bool active = tile_ok && ya < th;

    // one zone?  (every wave of the tile checks the whole tile)
    bool uniform = false;
    if (tile_ok) {
        const uint32_t v4 = sm.img[P + C0 + T * t] * 0x01010101u;
        bool ok = true;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int k = lane + 64 * i, row = k >> 2, c = (k & 3) * 16;
            if (row >= th || c >= tw) continue;
            const uint4 qv = *reinterpret_cast<const uint4*>(&sm.img[(row + 1) * P + C0 + T * t + c]);
            const uint32_t d[4] = {qv.x ^ v4, qv.y ^ v4, qv.z ^ v4, qv.w ^ v4};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int nb = tw - c - 4 * j;   // valid bytes of this dword
                const uint32_t m = nb >= 4 ? 0xFFFFFFFFu : (nb <= 0 ? 0u : (0xFFFFFFFFu >> (32 - 8 * nb)));
                if (d[j] & m) ok = false;
            }
        }
        uniform = __ballot(!ok) == 0;
    }

    // ---- perimeter: seam pixels, cross-seam equality bits ------------------
    // A perimeter pixel is a seam pixel if an equal 8-neighbour lies outside
    // the tile but inside the image.  The nine staged bytes around it are read
    // unconditionally and the neighbour rules applied as masks.  For a uniform
    // tile only perimeter pixels can have a greater or smaller neighbour, and
    // the staged halo makes that test exact, so they give the zone's evidence.
    uint32_t pbits = 0;   // bit k: slot q + k * TPT is a seam pixel
    if (tile_ok) {
        const bool hasL = x0t > 0, hasR = x0t + tw < im.w, hasU = y0 > 0, hasD = y0 + th < im.h;
        uint32_t zl = 0;
#pragma unroll
        for (int k = 0; k < PPER; ++k) {
            const int e = q + k * TPT;
            int lx, ly;
            if (e >= 4 * T || !perim_pixel<T>(e, tw, th, lx, ly)) continue;
            const uint8_t* c = &sm.img[(ly + 1) * P + C0 + T * t + lx];
            const uint32_t v = c[0];
            const uint32_t nul = c[-P - 1], nu = c[-P], nur = c[-P + 1], nl = c[-1], nrt = c[1];
            const uint32_t ndl = c[P - 1], nd = c[P], ndr = c[P + 1];
            const bool oL = lx == 0, oR = lx == tw - 1, oU = ly == 0, oD = ly == th - 1;
            const bool iL = !oL || hasL, iR = !oR || hasR, iU = !oU || hasU, iD = !oD || hasD;
            const uint32_t sul = (nul == v) & ((oL || oU) && iL && iU), su = (nu == v) & (oU && iU);
            const uint32_t sur = (nur == v) & ((oR || oU) && iR && iU);
            const uint32_t sl = (nl == v) & (oL && iL), sr = (nrt == v) & (oR && iR);
            const uint32_t sdl = (ndl == v) & ((oL || oD) && iL && iD), sd = (nd == v) & (oD && iD);
            const uint32_t sdr = (ndr == v) & ((oR || oD) && iR && iD);
            const bool seam = (sul | su | sur | sl | sr | sdl | sd | sdr) != 0;
            if (seam) pbits |= 1u << k;
            if (oR) sm.eqr[t][ly] = (uint8_t)(sur | (sr << 1) | (sdr << 2));
            if (oD) sm.eqb[t][lx] = (uint8_t)(sdl | (sd << 1) | (sdr << 2));
            if (uniform) {
                const uint32_t hi = max(max3u(nul, nu, nur), max(max(nl, nrt), max3u(ndl, nd, ndr)));
                const uint32_t lo = min(min3u(nul, nu, nur), min(min(nl, nrt), min3u(ndl, nd, ndr)));
                zl |= (hi > v ? F_G : 0u) | (lo < v ? F_S : 0u) | (seam ? F_SEAM : 0u);
            }
        }
        if (uniform) {
            const uint32_t f = (__ballot(zl & F_G) ? F_G : 0u) | (__ballot(zl & F_S) ? F_S : 0u) |
                               (__ballot(zl & F_SEAM) ? F_SEAM : 0u);
            if (lane == 0 && f) atomicOr(&sm.zf[t], f);
        }
    }

    // ---- rows of the band: classification, runs, links, unions ----------
    const uint8_t* col = &sm.img[C0 + T * t + lane];
    uint64_t DN = 0, DX = 0;
    uint32_t du = 0, dur = 0, dnid = 0;
    if (active && !uniform) {
        uint32_t al = col[ya * P - 1], ac = col[ya * P], ar = col[ya * P + 1];
        uint32_t bl = col[(ya + 1) * P - 1], bc = col[(ya + 1) * P], br = col[(ya + 1) * P + 1];
        uint32_t hmax_a = max3u(al, ac, ar), hmin_a = min3u(al, ac, ar);
        uint32_t hmax_b = max3u(bl, bc, br), hmin_b = min3u(bl, bc, br);
        // ordinal of each lane's run in the row above the band
        uint32_t ord_prev = 0;
        if (ya > 0) {
            const uint64_t ERp = __ballot(ac == ar) & VR;
            ord_prev = mbcnt64((VM & ~(ERp << 1)) >> 1);
        }
        // One row; FULL: the tile is 64 wide, TOP: row 0 (nothing above).
        auto row = [&](const int y, auto full_c, auto top_c) {
            constexpr bool FULL = decltype(full_c)::value, TOP = decltype(top_c)::value;
            const uint8_t* c = col + (y + 2) * P;
            const uint32_t cl = c[-1], cc = c[0], cr = c[1];
            const uint32_t hmax_c = max3u(cl, cc, cr), hmin_c = min3u(cl, cc, cr);
            const uint32_t vmax = max3u(hmax_a, hmax_b, hmax_c), vmin = min3u(hmin_a, hmin_b, hmin_c);
            uint64_t G = __ballot(vmax > bc), S_ = __ballot(vmin < bc), ER = __ballot(bc == br);
            uint64_t EU = 0, EUL = 0, EUR = 0;
            if constexpr (!TOP) {
                EU = __ballot(bc == ac);
                EUL = __ballot(bc == al);
                EUR = __ballot(bc == ar);
            }
            if constexpr (FULL) {
                ER &= ~(1ull << 63);
                EUL &= ~1ull;
                EUR &= ~(1ull << 63);
            } else {
                G &= VM;
                S_ &= VM;
                ER &= VR;
                EU &= VM;
                EUL &= VM & (VM << 1);
                EUR &= VR;
            }
            const uint64_t VMx = FULL ? ~0ull : VM;
            const uint64_t RS = VMx & ~(ER << 1), RE = VMx & ~ER;   // run starts / ends
            const uint32_t nid = (uint32_t)(y * T) + mbcnt64(RS >> 1);
            if (lane == 0) rsm[y] = RS;
            // a run's evidence reaches its last pixel as the carry of a sum
            // along the run (at a run end the sum bit is the carry)
            const uint64_t gR = RE & (G | (ER + (G & ER))), sR = RE & (S_ | (ER + (S_ & ER)));
            st16(par + lsel(RS, SINK, nid), nid);   // heads: their own roots for now
            flg[lsel(RE, SINK, nid)] = (uint8_t)(lsel(gR, 0u, F_G) | lsel(sR, 0u, F_S));
            if constexpr (!TOP) {
                // segments of equal pixels above: the first one links, the others unite
                const uint32_t opl = from_left(ord_prev), opr = from_right(ord_prev);
                const uint32_t ub = (uint32_t)((y - 1) * T);
                const uint32_t unode = ub + lsel(EUL, lsel(EU, opr, ord_prev), opl);
                const uint32_t urn = ub + opr;
                const uint64_t CT = EU | EUL | EUR;
                const uint64_t NEW = CT & (RS | ~((CT << 1) & (EUL | (EUR << 1))));
                const uint64_t EXTRA = EUL & EUR & ~EU;
                if (y > ya) {
                    const uint64_t a = NEW & ER;
                    const uint64_t FIRST = NEW & ~((ER + a) ^ ER ^ a);
                    // A run links to the parent of the run above (the root
                    // that run was given, or its own id); a further contact
                    // needs a union only if its label differs from the run's
                    // (mostly the two sides were joined rows above already).
                    const uint32_t lu = ld16(par + unode), lx = ld16(par + urn);
                    st16(par + lsel(FIRST, SINK, nid), lu);
                    const uint32_t lr = ld16(par + nid);
                    const uint64_t UN = NEW & ~FIRST;
                    const uint64_t M1 = __ballot(lane_in(UN | EXTRA) && lsel(UN, lx, lu) != lr);
                    const uint64_t M2 = __ballot(lane_in(UN & EXTRA) && lx != lr);
                    if (M1 | M2) {
                        if (lane_in(M1)) union16(par, lr, lsel(UN, lx, lu));
                        if (lane_in(M2)) union16(par, lr, lx);
                    }
                } else {
                    // the band's first row joins the band above after a barrier
                    DN = NEW;
                    DX = EXTRA;
                    du = unode;
                    dur = urn;
                    dnid = nid;
                }
            }
            ord_prev = nid - (uint32_t)(y * T);
            al = bl; ac = bc; ar = br;
            bl = cl; bc = cc; br = cr;
            hmax_a = hmax_b; hmin_a = hmin_b;
            hmax_b = hmax_c; hmin_b = hmin_c;
        };
        using TrueT = std::integral_constant<bool, true>;
        using FalseT = std::integral_constant<bool, false>;
        if (tw == T) {
            int y = ya;
            if (y == 0) row(y++, TrueT{}, TrueT{});
            for (; y < ye; ++y) row(y, TrueT{}, FalseT{});
        } else {
            int y = ya;
            if (y == 0) row(y++, FalseT{}, TrueT{});
            for (; y < ye; ++y) row(y, FalseT{}, FalseT{});
        }
    }
    __syncthreads();
    if (active && !uniform && ya > 0) {
        if (lane_in(DN)) union16(par, dnid, du);
        if (lane_in(DX)) union16(par, dnid, dur);
    }
    __syncthreads();
    // Rows are revisited in chunks of up to 16 whose run-start masks are read
    // back from LDS at once (wave-uniform; readfirstlane yields a signed int,
    // so each half is zero-extended before the halves are joined), and the
    // chunk's body is unrolled so that the LDS loads of all its rows are in
    // flight together.
    constexpr int CH = ROWS < 16 ? ROWS : 16;
    auto load_masks = [&](int y0c, uint64_t (&rs)[CH]) {
#pragma unroll
        for (int r = 0; r < CH; ++r) {
            const uint64_t v = y0c + r < ye ? rsm[y0c + r] : 0ull;
            const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
            const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
            rs[r] = (uint64_t)hi << 32 | lo;
        }
    };
    // Node of perimeter pixel (lx, ly): its row's run ordinal.
    auto node_at = [&](int lx, int ly) -> uint32_t {
        const uint64_t le = ~0ull >> (63 - lx);   // lanes 0 .. lx
        return (uint32_t)(ly * T) + (uint32_t)__builtin_popcountll(rsm[ly] & le) - 1u;
    };
    // ---- roots: compression, the zone's evidence; seam flags and numbering --
    uint32_t own_root[PPER], own_j[PPER];   // fragment this lane numbered, per perimeter slot
#pragma unroll
    for (int k = 0; k < PPER; ++k) {
        own_root[k] = 0;
        own_j[k] = 0xFFFFFFFFu;
    }
    if (tile_ok && !uniform) {
        for (int c = ya; c < ye; c += CH) {
            uint64_t rs[CH];
            load_masks(c, rs);
            // one lane per run node; read-only walks, so the only stores of
            // this pass are true roots and every entry ends as its root.  The
            // first two steps of every row's walk are loaded together.
            uint32_t n[CH], p[CH], g[CH], root[CH];
#pragma unroll
            for (int r = 0; r < CH; ++r) {
                n[r] = (uint32_t)((c + r) * T + lane);
                p[r] = lane < __builtin_popcountll(rs[r]) ? ld16(par + n[r]) : n[r];
            }
#pragma unroll
            for (int r = 0; r < CH; ++r) g[r] = p[r] != n[r] ? ld16(par + p[r]) : p[r];
#pragma unroll
            for (int r = 0; r < CH; ++r) {
                root[r] = p[r];
                if (g[r] != p[r]) root[r] = find16_ro(par, g[r]);
            }
            uint32_t fn[CH], fr[CH];
#pragma unroll
            for (int r = 0; r < CH; ++r) {
                fn[r] = fr[r] = 0;
                if (root[r] != n[r]) {
                    st16(par + n[r], root[r]);
                    fn[r] = ld8(flg + n[r]);
                    fr[r] = ld8(flg + root[r]);
                }
            }
#pragma unroll
            for (int r = 0; r < CH; ++r)
                if (fn[r] & ~fr[r]) or8(flg, root[r], fn[r]);
        }
        // Seam pixels mark their zone; the lane that sets a root's seam flag
        // first numbers that root as one of the tile's fragments.
#pragma unroll
        for (int k = 0; k < PPER; ++k) {
            int lx, ly;
            if (!((pbits >> k) & 1u)) continue;
            perim_pixel<T>(q + k * TPT, tw, th, lx, ly);
            const uint32_t root = find16_ro(par, node_at(lx, ly));
            if (ld8(flg + root) & F_SEAM) continue;
            const uint32_t old = __hip_atomic_fetch_or(reinterpret_cast<uint32_t*>(flg + (root & ~3u)),
                                                       F_SEAM << (8u * (root & 3u)), __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_WORKGROUP);
            if ((old >> (8u * (root & 3u))) & F_SEAM) continue;
            const uint32_t j = atomicAdd(&sm.cnt[t], 1u);
            if (j >= (uint32_t)S) {
                sm.err[t] = 1u;
                continue;
            }
            sm.fid[t][root] = (uint8_t)j;
            or8(flg, root, F_NUM);
            own_root[k] = root;
            own_j[k] = j;
        }
    }
    __syncthreads();
    // ---- per fragment: local evidence, and whether the tile may change ------
#pragma unroll
    for (int k = 0; k < PPER; ++k) {
        if (own_j[k] == 0xFFFFFFFFu) continue;
        const uint32_t f = ld8(flg + own_root[k]);
        ws.sbits[gtile * S + own_j[k]] = (uint8_t)(f & 3u);
        // a fragment with both kinds of evidence is final
        if ((f & 3u) != 3u) sm.open[t] = 1u;
    }
    __syncthreads();
    // ---- masks, record plane, perimeter records, per-tile words -----------
    if (tile_ok) {
        const bool open = sm.open[t] != 0;
        const uint32_t zf = sm.zf[t];
        const uint32_t cnt = uniform ? ((zf & F_SEAM) ? 1u : 0u) : min(sm.cnt[t], (uint32_t)S);
        if (active) {
            uint8_t* plane = ws.rec8 + gtile * (int64_t)(T * T);
            for (int c = ya; c < ye; c += CH) {
                uint64_t rs[CH];
                if (!uniform) load_masks(c, rs);
                uint32_t f[CH], fid[CH];
#pragma unroll
                for (int r = 0; r < CH; ++r) {
                    f[r] = zf;
                    fid[r] = 0;
                    if (!uniform && c + r < ye) {
                        // every node's entry holds its root
                        const uint32_t root = ld16(par + (uint32_t)((c + r) * T) + mbcnt64(rs[r] >> 1));
                        f[r] = ld8(flg + root);
                        if (f[r] & F_NUM) fid[r] = sm.fid[t][root];
                    }
                }
                if (lane < tw) {
#pragma unroll
                    for (int r = 0; r < CH; ++r) {
                        const int y = c + r;
                        if (y >= ye) break;
                        const uint32_t mx = (f[r] & F_G) ? 0u : 0xFFu, mn = (f[r] & F_S) ? 0u : 0xFFu;
                        const int64_t o = (int64_t)(y0 + y) * im.opitch + x0t + lane;
                        im.mx[o] = (uint8_t)mx;
                        im.mn[o] = (uint8_t)mn;
                        if (open) {
                            const bool opn = (f[r] & F_NUM) && (f[r] & 3u) != 3u;
                            plane[y * T + lane] = (uint8_t)(opn ? fid[r] : (REC_CLOSED | (mx & 1u) | (mn & 2u)));
                        }
                    }
                }
            }
        }
        // perimeter records: the fragment of every seam pixel, in each
        // perimeter array the pixel belongs to
        uint8_t* rec = ws.rec + gtile * 4 * T;
#pragma unroll
        for (int k = 0; k < PPER; ++k) {
            int lx, ly;
            if (!((pbits >> k) & 1u)) continue;
            perim_pixel<T>(q + k * TPT, tw, th, lx, ly);
            uint32_t fid = 0;
            if (!uniform) {
                const uint32_t root = ld16(par + node_at(lx, ly));
                fid = (ld8(flg + root) & F_NUM) ? (uint32_t)sm.fid[t][root] : 0u;
            }
            const uint8_t j = (uint8_t)fid;
            if (ly == 0) rec[lx] = j;
            if (ly == th - 1) rec[T + lx] = j;
            if (lx == 0) rec[2 * T + ly] = j;
            if (lx == tw - 1) rec[3 * T + ly] = j;
        }
        if (ya == 0) {
            if (cnt) {
                uint8_t* eq = ws.eq + gtile * 2 * T;
                eq[lane] = (lane < th && tw == T) ? sm.eqr[t][lane] : 0;
                eq[T + lane] = (lane < tw && th == T) ? sm.eqb[t][lane] : 0;
                if (uniform && lane == 0) ws.sbits[gtile * S] = (uint8_t)(zf & 3u);
            }
            if (lane == 0) {
                const uint32_t flags = (uniform ? CNT_UNIFORM : 0u) | (open ? CNT_OPEN : 0u);
                ws.counts[gtile] = cnt | flags | (sm.err[t] << COUNT_ERR_SHIFT);
            }
        }
    }
}

}  // namespace t64
}  // namespace ssx
