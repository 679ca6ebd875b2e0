// ss_seam.hpp -- the order of the seam merge's units (tile edges on merge-block
// boundaries).  Host and device: the kernel maps a unit to its edge with this,
// and a host check enumerates it for every grid shape.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ssx {

constexpr int MERGE_BT = 8;   // tiles per side of a block-local seam merge

// Unit u of an image with tiles_x x tiles_y tiles, seam_cols = (tiles_x - 1) /
// MERGE_BT block columns with a block right of them and seam_rows likewise.
// Units [0, seam_cols * tiles_y) are right edges: the right edge of tile
// (MERGE_BT c + MERGE_BT - 1, ty), pixel k against the three pixels right of
// it.  The rest are bottom edges: the bottom edge of tile (tx, MERGE_BT c +
// MERGE_BT - 1).  Within each kind the edges of one seam and one block row
// (column) are consecutive: full block rows first, seam after seam, MERGE_BT
// units each; then the last, partial block row with its hh edges per seam.
// So a workgroup covering MERGE_BT consecutive units holds one block seam.
__host__ __device__ inline bool seam_unit_edge(uint32_t u, uint32_t seam_cols, uint32_t seam_rows,
                                               uint32_t tiles_x, uint32_t tiles_y, int& tx, int& ty) {
    const uint32_t nv = seam_cols * tiles_y;
    const bool right = u < nv;
    const uint32_t v = right ? u : u - nv;
    const uint32_t ns = right ? seam_cols : seam_rows;
    const uint32_t len = right ? tiles_y : tiles_x;
    const uint32_t nfull = len / MERGE_BT, full = nfull * MERGE_BT * ns;
    uint32_t c, e;
    if (v < full) {
        c = (v / MERGE_BT) % ns;
        e = (v / (MERGE_BT * ns)) * MERGE_BT + v % MERGE_BT;
    } else {
        const uint32_t hh = len - nfull * MERGE_BT, r = v - full;
        const uint32_t hd = hh ? hh : 1u;   // no unit reaches here when hh == 0
        c = r / hd;
        e = nfull * MERGE_BT + r % hd;
    }
    const int across = MERGE_BT * (int)c + MERGE_BT - 1;
    tx = right ? across : (int)e;
    ty = right ? (int)e : across;
    return right;
}

}  // namespace ssx
