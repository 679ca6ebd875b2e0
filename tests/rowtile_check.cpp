// Host check of the row tile stage's union edges (ss_rows.hpp).  On random
// 64 x 64 tiles (values drawn from few levels so that zones of every shape
// occur, random valid width and height) it builds the run union-find the
// kernel builds -- runs of each 32-column half-row, the row_edges unions to
// the half-row above, the cross-half unions at columns 31 | 32 that
// cross_keep keeps -- and
// compares its partition with a breadth-first labelling of the 8-connected
// equal-value zones of the tile.  Also checks fill_fwd / fill_bwd against a
// per-bit walk.  Prints the counts; exit status 1 on any mismatch.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "ss_rows.hpp"

using namespace ssx;

static uint64_t rng = 0x9E3779B97F4A7C15ull;
static uint32_t next32() {
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return (uint32_t)(rng >> 16);
}

struct UF {
    std::vector<int> p;
    explicit UF(int n) : p(n) {
        for (int i = 0; i < n; ++i) p[i] = i;
    }
    int find(int x) {
        while (p[x] != x) x = p[x] = p[p[x]];
        return x;
    }
    void unite(int a, int b) {
        a = find(a);
        b = find(b);
        if (a != b) p[a < b ? b : a] = a < b ? a : b;
    }
};

static int check_fills() {
    int bad = 0;
    for (int it = 0; it < 200000; ++it) {
        const uint32_t s = next32() & next32(), c = (next32() | next32()) & 0x7FFFFFFFu;
        uint32_t f = 0, g = 0;
        for (int b = 0; b < 32; ++b) {
            // forward: b is reached from a seed at or before it through links
            for (int q = b; q >= 0; --q) {
                if ((s >> q) & 1u) { f |= 1u << b; break; }
                if (q == 0 || !((c >> (q - 1)) & 1u)) break;
            }
            // backward: from a seed at or after it
            for (int q = b; q < 32; ++q) {
                if ((s >> q) & 1u) { g |= 1u << b; break; }
                if (!((c >> q) & 1u)) break;
            }
        }
        if (fill_fwd(s, c) != f || fill_bwd(s, c) != g) ++bad;
    }
    return bad;
}

int main() {
    const int fill_bad = check_fills();
    int tiles = 0, bad = 0;
    long cross_dropped = 0;
    const int levels_set[] = {2, 3, 4, 8, 256};
    for (int it = 0; it < 3000; ++it) {
        const int levels = levels_set[it % 5];
        const int tw = (it % 7 == 0) ? 1 + (int)(next32() % 64) : 64;
        const int th = (it % 11 == 0) ? 1 + (int)(next32() % 64) : 64;
        // blocky content: a random coarse grid upsampled, plus noise
        static uint8_t img[66][66];   // [y + 1][x + 1], halo included
        const int cell = 1 + (int)(next32() % 9);
        for (int y = -1; y <= 64; ++y)
            for (int x = -1; x <= 64; ++x) {
                uint32_t v = ((uint32_t)((y + 64) / cell) * 2654435761u) ^ ((uint32_t)((x + 64) / cell) * 40503u) ^
                             (uint32_t)it * 97u;
                v = (v >> 7) % (uint32_t)levels;
                if (next32() % 13 == 0) v = next32() % (uint32_t)levels;
                img[y + 1][x + 1] = (uint8_t)v;
            }
        auto px = [&](int y, int x) { return img[y + 1][x + 1]; };
        auto valid = [&](int y, int x) { return y >= 0 && y < th && x >= 0 && x < tw; };
        // masks per half and row
        uint32_t H[2][64], D[2][64], DL[2][64], DR[2][64], NS[2][64], RS[2][64];
        uint32_t XH[64], XDR[64], XUR[64];
        for (int h = 0; h < 2; ++h)
            for (int y = 0; y < 64; ++y) {
                uint32_t hh = 0, d = 0, dl = 0, dr = 0;
                for (int b = 0; b < 32; ++b) {
                    const int x = 32 * h + b;
                    if (!valid(y, x)) continue;
                    const uint8_t v = px(y, x);
                    if (b < 31 && valid(y, x + 1) && px(y, x + 1) == v) hh |= 1u << b;
                    if (valid(y + 1, x) && px(y + 1, x) == v) d |= 1u << b;
                    if (b > 0 && valid(y + 1, x - 1) && px(y + 1, x - 1) == v) dl |= 1u << b;
                    if (b < 31 && valid(y + 1, x + 1) && px(y + 1, x + 1) == v) dr |= 1u << b;
                }
                H[h][y] = hh;
                D[h][y] = d;
                DL[h][y] = dl;
                DR[h][y] = dr;
            }
        for (int y = 0; y < 64; ++y) {
            const bool v31 = valid(y, 31) && valid(y, 32);
            XH[y] = v31 && px(y, 31) == px(y, 32);
            XDR[y] = v31 && valid(y + 1, 32) && px(y, 31) == px(y + 1, 32);
            XUR[y] = v31 && valid(y - 1, 32) && px(y, 31) == px(y - 1, 32);
        }
        for (int h = 0; h < 2; ++h)
            for (int y = 0; y < 64; ++y) {
                const uint32_t u = y ? D[h][y - 1] : 0, ul = y ? DR[h][y - 1] << 1 : 0, ur = y ? DL[h][y - 1] >> 1 : 0;
                uint32_t cross = 0;
                if (h == 0 && (XH[y] || XDR[y] || XUR[y])) cross = 1u << 31;
                if (h == 1 && valid(y, 32)) {
                    const uint8_t v = px(y, 32);
                    if (px(y, 31) == v || (valid(y + 1, 31) && px(y + 1, 31) == v) ||
                        (valid(y - 1, 31) && px(y - 1, 31) == v))
                        cross = 1u;
                }
                NS[h][y] = H[h][y] | (H[h][y] << 1) | u | ul | ur | D[h][y] | DL[h][y] | DR[h][y] | cross;
                RS[h][y] = NS[h][y] & ~(H[h][y] << 1);
            }
        UF uf(64 * 64);
        for (int h = 0; h < 2; ++h)
            for (int y = 1; y < 64; ++y) {
                const uint32_t u = D[h][y - 1], ul = DR[h][y - 1] << 1, ur = DL[h][y - 1] >> 1;
                const RowEdges e = row_edges(u, ul, ur, H[h][y]);
                for (int b = 0; b < 32; ++b) {
                    const int a = 64 * y + 32 * h + high_at_or_below(RS[h][y], b);
                    if ((e.first >> b) & 1u) {
                        const int k = b - (int)((e.kl >> b) & 1u) + (int)((e.kr >> b) & 1u);
                        uf.unite(a, 64 * (y - 1) + 32 * h + high_at_or_below(RS[h][y - 1], k));
                    }
                    if ((e.second >> b) & 1u)
                        uf.unite(a, 64 * (y - 1) + 32 * h + high_at_or_below(RS[h][y - 1], b + 1));
                }
            }
        // cross-half relations, the redundant ones dropped (cross_keep)
        auto eqv = [&](int y0, int x0, int y1, int x1) {
            return valid(y0, x0) && valid(y1, x1) && px(y0, x0) == px(y1, x1) ? 1u : 0u;
        };
        int dropped = 0;
        for (int y = 0; y < 64; ++y) {
            if (!(XH[y] || XDR[y] || XUR[y])) continue;
            const uint32_t keep = cross_keep(XH[y], XDR[y], XUR[y], y ? XH[y - 1] : 0u, eqv(y, 31, y + 1, 31),
                                             eqv(y - 1, 31, y, 31));
            dropped += (int)(XH[y] + XDR[y] + XUR[y]) - __builtin_popcount(keep);
            const int a = 64 * y + high_at_or_below(RS[0][y], 31);
            if (keep & 1u) uf.unite(a, 64 * y + 32);
            if (keep & 2u) uf.unite(a, 64 * (y + 1) + 32);
            if (keep & 4u) uf.unite(a, 64 * (y - 1) + 32);
        }
        cross_dropped += dropped;
        // breadth-first zones
        std::vector<int> zone(64 * 64, -1);
        std::vector<int> q;
        for (int s = 0; s < 64 * 64; ++s) {
            if (!valid(s / 64, s % 64) || zone[s] >= 0) continue;
            zone[s] = s;
            q.assign(1, s);
            for (size_t i = 0; i < q.size(); ++i) {
                const int y = q[i] / 64, x = q[i] % 64;
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int ny = y + dy, nx = x + dx;
                        if ((dy || dx) && valid(ny, nx) && zone[ny * 64 + nx] < 0 && px(ny, nx) == px(y, x)) {
                            zone[ny * 64 + nx] = s;
                            q.push_back(ny * 64 + nx);
                        }
                    }
            }
        }
        std::vector<int> size(64 * 64, 0), root_of_zone(64 * 64, -1);
        for (int s = 0; s < 64 * 64; ++s)
            if (zone[s] >= 0) ++size[zone[s]];
        bool ok = true;
        for (int s = 0; s < 64 * 64 && ok; ++s) {
            const int y = s / 64, x = s % 64;
            if (!valid(y, x)) continue;
            const int h = x / 32, b = x % 32;
            const bool ns = (NS[h][y] >> b) & 1u;
            if (ns != (size[zone[s]] > 1)) { ok = false; break; }
            if (!ns) continue;
            const int r = uf.find(64 * y + 32 * h + high_at_or_below(RS[h][y], b));
            if (root_of_zone[zone[s]] < 0) root_of_zone[zone[s]] = r;
            if (root_of_zone[zone[s]] != r) ok = false;
        }
        // distinct zones must have distinct roots
        if (ok) {
            std::vector<int> zone_of_root(64 * 64, -1);
            for (int z = 0; z < 64 * 64 && ok; ++z) {
                const int r = root_of_zone[z];
                if (r < 0) continue;
                if (zone_of_root[r] >= 0 && zone_of_root[r] != z) ok = false;
                zone_of_root[r] = z;
            }
        }
        ++tiles;
        if (!ok) {
            ++bad;
            if (bad < 5) std::printf("mismatch: tile %d levels %d tw %d th %d cell %d\n", it, levels, tw, th, cell);
        }
    }
    std::printf("fills: %d mismatches; %d tiles, %d mismatches (%ld cross-half relations dropped)\n", fill_bad, tiles, bad,
                cross_dropped);
    return (fill_bad || bad) ? 1 : 0;
}
