// Host check of ssx::seam_unit_edge (the seam merge's unit order) for every
// grid of 1..64 x 1..64 tiles: the units map one-to-one onto the tile edges
// on merge-block boundaries (the right edges of the last tile column of each
// block that has a block right of it, the bottom edges of the last tile row of
// each block that has a block below it), and in full block rows (columns)
// every aligned run of MERGE_BT units is the MERGE_BT edges of one block seam.
// Prints the number of grids and failures; exit status 1 if any failed.
#include <cstdio>
#include <set>
#include <tuple>

#include "ss_seam.hpp"

using namespace ssx;

int main() {
    int grids = 0, bad = 0;
    for (uint32_t txn = 1; txn <= 64; ++txn) {
        for (uint32_t tyn = 1; tyn <= 64; ++tyn) {
            ++grids;
            const uint32_t sc = (txn - 1) / MERGE_BT, sr = (tyn - 1) / MERGE_BT;
            const uint32_t n = sc * tyn + sr * txn;
            std::set<std::tuple<bool, int, int>> want, got;
            for (uint32_t ty = 0; ty < tyn; ++ty)
                for (uint32_t c = 0; c < sc; ++c) want.insert({true, (int)(MERGE_BT * c + MERGE_BT - 1), (int)ty});
            for (uint32_t tx = 0; tx < txn; ++tx)
                for (uint32_t c = 0; c < sr; ++c) want.insert({false, (int)tx, (int)(MERGE_BT * c + MERGE_BT - 1)});
            bool ok = true;
            for (uint32_t u = 0; u < n; ++u) {
                int tx, ty;
                const bool right = seam_unit_edge(u, sc, sr, txn, tyn, tx, ty);
                if (!got.insert({right, tx, ty}).second) ok = false;
                if (tx < 0 || ty < 0 || (uint32_t)tx >= txn || (uint32_t)ty >= tyn) ok = false;
                if (right ? (uint32_t)tx + 1 >= txn : (uint32_t)ty + 1 >= tyn) ok = false;
            }
            if (got != want) ok = false;
            // one seam per aligned group of MERGE_BT units in the full part
            const uint32_t parts[2][3] = {{0, sc, tyn}, {sc * tyn, sr, txn}};
            for (const auto& p : parts) {
                const uint32_t full = (p[2] / MERGE_BT) * MERGE_BT * p[1];
                for (uint32_t g = 0; g < full; g += MERGE_BT) {
                    int tx0, ty0;
                    const bool r0 = seam_unit_edge(p[0] + g, sc, sr, txn, tyn, tx0, ty0);
                    for (uint32_t i = 1; i < (uint32_t)MERGE_BT; ++i) {
                        int tx, ty;
                        seam_unit_edge(p[0] + g + i, sc, sr, txn, tyn, tx, ty);
                        const bool same = r0 ? (tx == tx0 && ty == ty0 + (int)i) : (ty == ty0 && tx == tx0 + (int)i);
                        if (!same) ok = false;
                    }
                    if ((r0 ? ty0 : tx0) % MERGE_BT) ok = false;
                }
            }
            if (!ok) {
                if (bad < 5) std::printf("grid %ux%u failed\n", txn, tyn);
                ++bad;
            }
        }
    }
    std::printf("%d grids, %d failed\n", grids, bad);
    return bad ? 1 : 0;
}
