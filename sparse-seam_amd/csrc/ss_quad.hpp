// ss_quad.hpp -- the per-quad bit words of the tile stage and the run / link
// logic on them.  Host and device: the same code runs in the kernels and in
// the exhaustive host check of quad_links against the per-run definition.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ssx {

// Per-quad register word: eight 4-bit fields, bit i = pixel i of the quad.
constexpr int Q_G = 0;        // a strictly greater 8-neighbour exists
constexpr int Q_S = 4;        // a strictly smaller 8-neighbour exists
constexpr int Q_EL = 8;       // equal left neighbour in the tile
constexpr int Q_EU = 12;      // equal up neighbour in the tile
constexpr int Q_EUL = 16;     // equal up-left neighbour in the tile
constexpr int Q_EUR = 20;     // equal up-right neighbour in the tile
constexpr int Q_HD = 24;      // run head (see below)
constexpr int Q_UN = 28;      // run head with contacts its Komura link does not cover

// Runs: the non-singleton pixels of a quad split into maximal runs of equal
// horizontal neighbours.  Only a run's first pixel (its head) is a union-find
// node; the other pixels of the run point at the head and never change.  A
// plateau quad is one node instead of four.
__host__ __device__ __forceinline__ uint32_t q_heads(uint32_t st) { return (st >> Q_HD) & 0xFu; }
// Heads the union pass must visit.
__host__ __device__ __forceinline__ uint32_t q_unions(uint32_t st) { return (st >> Q_UN) & 0xFu; }
// Pixels of the run that starts at head h.
__host__ __device__ __forceinline__ uint32_t q_run(uint32_t st, int h) {
    const uint32_t c = (st >> Q_EL) & 0xEu;
    uint32_t r = 1u << h;
    r |= (r << 1) & c;
    r |= (r << 1) & c;
    r |= (r << 1) & c;
    return r;
}
// Row-above contacts of run r: bit j set iff the pixel above at quad column
// j - 1 (j = 0..5) equals the run's value.
__host__ __device__ __forceinline__ uint32_t q_above(uint32_t st, uint32_t r) {
    return (((st >> Q_EUL) & r) | (((st >> Q_EU) & r) << 1) | (((st >> Q_EUR) & r) << 2)) & 0x3Fu;
}

// Nibble -> bytes: bit m to bit m of byte m (diag4), to bit 0 of byte m (spread4).
__host__ __device__ __forceinline__ uint32_t diag4(uint32_t n) { return (n * 0x01010101u) & 0x08040201u; }
__host__ __device__ __forceinline__ uint32_t spread4(uint32_t n) { return (n * 0x00204081u) & 0x01010101u; }

// SWAR helpers on four bytes (one per pixel of a quad).
// Top bit of every byte -> 4-bit mask.
__host__ __device__ __forceinline__ uint32_t movemask4(uint32_t w) {
    return (((w >> 7) & 0x01010101u) * 0x01020408u) >> 24;
}

// Komura labels of a quad (p0 = raster index of its pixel 0 in a T-wide
// tile) and its pass masks, from the quad word s (G, S, EL, EU, EUL, EUR and
// HD fields set).  A head links to its first contact above (the up-left,
// up or up-right neighbour, lowest column first) if its run has one, else to
// the pixel left of the quad if the run continues from there, else to
// itself; every other run pixel links to its head (singletons get a dummy
// entry that nothing reads).  un: heads with a contact segment beyond the
// first, or continuing from the left with contacts that do not include the
// left pixel's own (see quad_label); rt: heads linked to themselves.
//
// All four pixels at once, one byte each.  Byte m of P: the contacts of
// pixel m (bits m, m + 1, m + 2 = up-left, up, up-right, in the 6-bit space
// of q_above); byte m of R: their OR from pixel m to the end of its run, i.e.
// at a head the run's contacts (q_above of q_run).  Bytes stay below 64, so
// the per-byte arithmetic never carries across bytes.
__host__ __device__ __forceinline__ void quad_links(uint32_t s, uint32_t p0, uint32_t T, uint32_t (&lab4)[4],
                                                    uint32_t& un, uint32_t& rt) {
    const uint32_t hd = q_heads(s);
    const uint32_t el4 = (s >> Q_EL) & 0xFu;
    const uint32_t P = diag4((s >> Q_EUL) & 0xFu) | diag4((s >> Q_EU) & 0xFu) << 1 |
                       diag4((s >> Q_EUR) & 0xFu) << 2;
    const uint32_t m1 = spread4(el4 >> 1) * 0xFFu;   // byte m: pixel m + 1 continues the run
    uint32_t R = P | ((P >> 8) & m1);
    R |= (R >> 16) & (m1 & (m1 >> 8));
    // 0x80 per byte: the run has no contact / two contact segments or more
    const uint32_t z80 = ~(((R & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | R) & 0x80808080u;
    const uint32_t sg = R & ~(R << 1);
    const uint32_t s2 = sg & ((sg | 0x80808080u) - 0x01010101u);
    const uint32_t m80 = (((s2 & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | s2) & 0x80808080u;
    const uint32_t left = el4 & 1u;   // pixel 0 continues a run of the quad to the left
    const uint32_t r0 = R & 0xFFu;
    un = (movemask4(m80) | ((left && r0 && !(r0 & 3u)) ? 1u : 0u)) & hd;
    rt = movemask4(z80) & hd & ~left;
    uint32_t head = p0;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const uint32_t rm = (R >> (8 * m)) & 0xFFu;
        const uint32_t self = (m == 0 && left) ? p0 - 1u : p0 + m;
        const uint32_t link = rm ? p0 - T - 1u + __builtin_ctz(rm | 0x40u) : self;
        const bool is_head = (hd >> m) & 1u;
        head = is_head ? p0 + m : head;
        lab4[m] = is_head ? link : head;
    }
}

// quad_links for a quad that is one run (pixels 1..3 continue the run of
// pixel 0, so pixel 0 is the only head): the run's contacts are the OR of the
// three contact fields.  Same results as quad_links for such a quad.
__host__ __device__ __forceinline__ bool q_one_run(uint32_t s) { return ((s >> Q_EL) & 0xEu) == 0xEu; }
__host__ __device__ __forceinline__ void quad_links_one_run(uint32_t s, uint32_t p0, uint32_t T, uint32_t (&lab4)[4],
                                                            uint32_t& un, uint32_t& rt) {
    const uint32_t r0 = (((s >> Q_EUL) & 0xFu) | ((s >> Q_EU) & 0xFu) << 1 | ((s >> Q_EUR) & 0xFu) << 2);
    const uint32_t left = (s >> Q_EL) & 1u;
    const uint32_t sg = r0 & ~(r0 << 1);   // first column of each contact segment
    lab4[0] = r0 ? p0 - T - 1u + __builtin_ctz(r0) : (left ? p0 - 1u : p0);
    lab4[1] = lab4[2] = lab4[3] = p0;
    un = ((sg & (sg - 1u)) != 0u || (left && r0 && !(r0 & 3u))) ? 1u : 0u;
    rt = (!r0 && !left) ? 1u : 0u;
}

}  // namespace ssx
