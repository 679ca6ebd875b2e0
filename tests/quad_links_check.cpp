// Exhaustive host check of ssx::quad_links (the SWAR form used by the tile
// stage) against the per-run definition: for every value of the EL, EU, EUL,
// EUR and non-singleton nibbles of a quad, a head's run extends over the
// following pixels with an equal left neighbour; its contacts are the
// up-left / up / up-right equalities of those pixels; it links to the lowest
// contact, else (pixel 0 continuing from the left quad) to the left pixel,
// else to itself; other pixels get the last head at or before them.  The
// one-run form (quad_links_one_run) is checked the same way on every quad it
// applies to (pixels 1..3 continue the run of pixel 0, pixel 0 a head).
// Prints the number of mismatches; exit status 1 if any.
#include <cstdio>

#include "ss_quad.hpp"

using namespace ssx;

static void reference(uint32_t s, uint32_t p0, uint32_t T, uint32_t lab[4], uint32_t& un, uint32_t& rt) {
    const uint32_t hd = q_heads(s);
    const uint32_t el = (s >> Q_EL) & 0xFu, eu = (s >> Q_EU) & 0xFu;
    const uint32_t eul = (s >> Q_EUL) & 0xFu, eur = (s >> Q_EUR) & 0xFu;
    un = rt = 0;
    uint32_t last = p0;
    for (int i = 0; i < 4; ++i) {
        if (!((hd >> i) & 1u)) {
            lab[i] = last;
            continue;
        }
        last = p0 + i;
        uint32_t contacts = 0;   // bit j: the pixel above at quad column j - 1
        for (int m = i; m < 4 && (m == i || ((el >> m) & 1u)); ++m) {
            if ((eul >> m) & 1u) contacts |= 1u << m;
            if ((eu >> m) & 1u) contacts |= 1u << (m + 1);
            if ((eur >> m) & 1u) contacts |= 1u << (m + 2);
        }
        const bool left = i == 0 && (el & 1u);
        if (contacts) {
            int j = 0;
            while (!((contacts >> j) & 1u)) ++j;
            lab[i] = p0 - T + j - 1;
        } else {
            lab[i] = left ? p0 - 1 : p0 + i;
        }
        int segments = 0;
        for (int j = 0; j < 6; ++j)
            if (((contacts >> j) & 1u) && (j == 0 || !((contacts >> (j - 1)) & 1u))) ++segments;
        const bool left_union = left && contacts && !(contacts & 3u);
        if (segments >= 2 || left_union) un |= 1u << i;
        if (!contacts && !left) rt |= 1u << i;
    }
}

int main() {
    const uint32_t T = 64, p0 = 5 * 64 + 8;
    long cases = 0, one_run = 0, bad = 0;
    for (uint32_t v = 0; v < (1u << 20); ++v) {
        const uint32_t el = v & 0xFu, eu = (v >> 4) & 0xFu, eul = (v >> 8) & 0xFu, eur = (v >> 12) & 0xFu;
        const uint32_t nsg = (v >> 16) & 0xFu;
        const uint32_t hd = nsg & ~(el & 0xEu);
        const uint32_t s = el << Q_EL | eu << Q_EU | eul << Q_EUL | eur << Q_EUR | hd << Q_HD;
        uint32_t a[4], b[4], ua, ra, ub, rb;
        reference(s, p0, T, a, ua, ra);
        quad_links(s, p0, T, b, ub, rb);
        ++cases;
        bool ok = ua == ub && ra == rb;
        for (int i = 0; i < 4; ++i) ok = ok && a[i] == b[i];
        if (!ok && bad++ < 5)
            std::printf("mismatch el=%x eu=%x eul=%x eur=%x nsg=%x\n", el, eu, eul, eur, nsg);
        if (q_one_run(s) && hd == 1u) {
            uint32_t c[4], uc, rc;
            quad_links_one_run(s, p0, T, c, uc, rc);
            ++one_run;
            bool ok1 = ua == uc && ra == rc;
            for (int i = 0; i < 4; ++i) ok1 = ok1 && a[i] == c[i];
            if (!ok1 && bad++ < 5)
                std::printf("one-run mismatch el=%x eu=%x eul=%x eur=%x\n", el, eu, eul, eur);
        }
    }
    std::printf("%ld cases (%ld one-run), %ld mismatches\n", cases, one_run, bad);
    return bad ? 1 : 0;
}
