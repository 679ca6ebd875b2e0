// ss_rows.hpp -- 32-bit row-mask logic of the row tile stage (ss_rowtile.hpp).
// Host and device: the kernel and the host checks run the same code.
//
// A wave of the row tile stage holds one half (32 columns) of a 64 x 64 tile,
// one tile row per lane; bit b of a lane's masks is column b of its half-row.
// Flat zones are built from horizontal runs: a run is a maximal segment of
// non-singleton pixels joined by equal right neighbours, and only its first
// pixel (the run start) is a union-find node.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ssx {

__host__ __device__ __forceinline__ uint32_t rbit32(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_bitreverse32(x);
#else
    x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
    x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
    x = ((x >> 4) & 0x0F0F0F0Fu) | ((x & 0x0F0F0F0Fu) << 4);
    x = ((x >> 8) & 0x00FF00FFu) | ((x & 0x00FF00FFu) << 8);
    return (x >> 16) | (x << 16);
#endif
}

// Forward fill along links: c bit b joins columns b and b + 1 (c bit 31 must
// be clear).  Every seed s spreads to the end of its linked segment: the add
// carries from the seed through the segment's ones and stops one past them.
__host__ __device__ __forceinline__ uint32_t fill_fwd(uint32_t s, uint32_t c) {
    return s | (((s & c) + c) ^ c);
}

// Backward fill (to the start of the linked segment), by bit reversal.
__host__ __device__ __forceinline__ uint32_t fill_bwd(uint32_t s, uint32_t c) {
    return rbit32(fill_fwd(rbit32(s), rbit32(c) >> 1));
}

// Highest set bit of m at or below b (m has one there).
__host__ __device__ __forceinline__ int high_at_or_below(uint32_t m, int b) {
    return 31 - __builtin_clz(m & (0xFFFFFFFFu >> (31 - b)));
}

// Union edges from a half-row to the half-row above it.  u, ul, ur: equal up,
// up-left, up-right neighbour inside the tile and the wave's half (bit b of ul
// is the pair (b, b - 1 above), of ur (b, b + 1 above)); h: equal right
// neighbour (bit b joins b and b + 1 in this row).
//
// Each contacting pixel has a primary contact (up, else up-left, else
// up-right), at offset dk = 0, -1, +1.  Two consecutive pixels of one run
// reach the same run above exactly when dk does not increase from the left
// pixel to the right one: their contacts are then adjacent or equal, both of
// the run's value; when it increases, the pixel between the contacts is
// unequal (it would have been the up contact).  So `first` marks the pixels
// whose primary contact starts a new run above (the run's first contacting
// pixel, or a change of run above).  A pixel's other contacts reach the same
// run above as its primary one, except up-left and up-right without up
// (`second`: the up-right contact is a second run above).
struct RowEdges {
    uint32_t first;    // union (run of b) with (run above containing b + dk(b))
    uint32_t second;   // union (run of b) with (run above containing b + 1)
    uint32_t kl, kr;   // dk = -1 / +1 at b (else 0)
};

__host__ __device__ __forceinline__ RowEdges row_edges(uint32_t u, uint32_t ul, uint32_t ur, uint32_t h) {
    RowEdges e;
    const uint32_t kl = ul & ~u, kr = ur & ~u & ~ul, ct = u | ul | ur;
    const uint32_t gt = ((kl << 1) & (u | kr)) | ((u << 1) & kr);
    e.first = ct & (~(h << 1) | ~(ct << 1) | gt);
    e.second = ul & ur & ~u;
    e.kl = kl;
    e.kr = kr;
    return e;
}

// The relations across columns 31 | 32 of a row y: XH (y, 31)-(y, 32), XDR
// (y, 31)-(y + 1, 32), XUR (y, 31)-(y - 1, 32).  They are equalities of pixel
// values, so two of them through a common pixel imply the third; one that
// follows from kept relations and from the vertical relations inside each
// half is dropped:
//   XH(y)  when XH(y - 1) and (y - 1, 31)-(y, 31): then (y - 1, 32)-(y, 32);
//   XDR(y) when (y, 31)-(y + 1, 31): then XH(y + 1); or when XH(y): then
//          (y, 32)-(y + 1, 32);
//   XUR(y) when (y - 1, 31)-(y, 31): then XH(y - 1); or when XH(y): then
//          (y - 1, 32)-(y, 32).
// Only XH relations justify a dropped one, and an XH only through the row
// above, so the kept set implies every relation.  Inputs are 0 / 1: xh, xdr,
// xur of row y; xh_up: XH of row y - 1; d31, d31_up: column 31 equal between
// rows y and y + 1, and between y - 1 and y.  Returns bit 0 = keep XH, bit 1
// = keep XDR, bit 2 = keep XUR.
__host__ __device__ __forceinline__ uint32_t cross_keep(uint32_t xh, uint32_t xdr, uint32_t xur, uint32_t xh_up,
                                                        uint32_t d31, uint32_t d31_up) {
    const uint32_t kh = xh & ~(xh_up & d31_up);
    const uint32_t kdr = xdr & ~d31 & ~xh;
    const uint32_t kur = xur & ~d31_up & ~xh;
    return (kh & 1u) | (kdr & 1u) << 1 | (kur & 1u) << 2;
}

}  // namespace ssx
