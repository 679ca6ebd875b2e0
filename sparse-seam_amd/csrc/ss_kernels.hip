// ss_kernels.hip -- gfx950 kernels of the exact dual regional-extrema path and
// the host-side sequencer that enqueues them.
//
// Reference behaviour followed (harness/primary/snapshot_sparse_seam_cuda_v1_0.cu):
//   tile-local Komura labels        :194-289  -> tile_label()
//   classify / number / emit        :291-475  -> tile_stage()
//   seam merge (3 pairs per seam px):508-581  -> k_block_merge, k_group_merge, k_seam_merge
//   compress + reduce               :583-619  -> k_resolve, k_final
//   materialise seam records        :477-506  -> k_writeback
// The reference writes a 1-byte record for every pixel and re-reads it for
// every pixel; here masks are written once (provisionally for seam zones), the
// record plane is written only by tiles holding a fragment that can still
// change, and only the tiles whose fragments did change are streamed again.
#include "ss_device.hpp"
#include "ss_internal.hpp"
#include "ss_rowtile.hpp"
#include "ss_tile.hpp"

namespace ssx {

// Set by the tile stage (one thread per tile), the first kernel of a run that
// uses the workspace header: the header's per-run words (tile 0), and the
// look-back state of each scan block (cleared by its first tile).
__device__ __forceinline__ void run_prologue(int64_t gtile, const Ws& ws) {
    if (gtile == 0) {
        WsHeader* h = ws.hdr;
        h->magic = WS_MAGIC;
        h->total = 0;
        h->error = 0;
        h->dirty_count = 0;
        h->retries = 0;
        h->tiles = (uint64_t)ws.ntiles;
        h->tile = ws.tile;
        h->threads = ws.threads;
    }
    if (gtile % SCAN_CHUNK == 0) ws.bstate[gtile / SCAN_CHUNK] = 0;
}

// Plain form: one workgroup per tile, XCD-aware tile order.
template <int T, int NT, bool MULTI>
__global__ __launch_bounds__(NT) void k_tile_stage(Images ims, Ws ws) {
    __shared__ TileSmem<T> sm;
    const int64_t gtile = xcd_block(blockIdx.x, gridDim.x);
    if (gtile >= ws.ntiles) return;
    const ImgDev im = locate<MULTI>(ims, ws.tile_img, gtile);
    if (threadIdx.x == 0) run_prologue(gtile, ws);
    tile_stage<T, NT>(sm, im, gtile, ws);
}

// Row form (T = 64, 128 threads per tile): a workgroup holds RT_NTW
// consecutive tiles (horizontal neighbours within an image row), two waves
// each, so that its loads and stores cover whole 128-byte lines.
template <bool MULTI>
__global__ __launch_bounds__(RT_THREADS * RT_NTW) void k_row_tile(Images ims, Ws ws) {
    __shared__ RowTileSmem sm[RT_NTW];
    const int sub = RT_NTW == 1 ? 0 : (int)(threadIdx.x / RT_THREADS);   // 0: the tile index stays scalar
    const int64_t gtile = xcd_block(blockIdx.x, gridDim.x) * RT_NTW + sub;
    const bool live = gtile < ws.ntiles;
    const ImgDev im = live ? locate<MULTI>(ims, ws.tile_img, gtile) : ims.one;
    if (live && threadIdx.x % RT_THREADS == 0) run_prologue(gtile, ws);
    row_tile_stage<MULTI>(sm[sub], im, gtile, ws, live, RT_NTW == 1 ? (int)threadIdx.x : (int)(threadIdx.x % RT_THREADS));
}

// Per-tile, per-merge-block and per-seam-unit image index of a batch: one
// workgroup per image.
__global__ __launch_bounds__(256) void k_tile_map(Images ims, Ws ws) {
    const ImgDev im = ims.table[blockIdx.x];
    const int64_t n = (int64_t)im.tiles_x * im.tiles_y;
    for (int64_t t = threadIdx.x; t < n; t += blockDim.x) ws.tile_img[im.tile_begin + t] = blockIdx.x;
    const int64_t nb = (int64_t)im.blocks_x * im.blocks_y;
    for (int64_t b = threadIdx.x; b < nb; b += blockDim.x) ws.block_img[im.block_begin + b] = blockIdx.x;
    const int64_t ns = (int64_t)im.seam_cols * im.tiles_y + (int64_t)im.seam_rows * im.tiles_x;
    for (int64_t u = threadIdx.x; u < ns; u += blockDim.x) ws.seam_img[im.seam_begin + u] = blockIdx.x;
    const int64_t ng = (int64_t)((im.blocks_x + 1) / 2) * ((im.blocks_y + 1) / 2);
    for (int64_t g = threadIdx.x; g < ng; g += blockDim.x) ws.group_img[im.group_begin + g] = blockIdx.x;
}

// ---- compaction: exclusive scan of the per-tile counts --------------------

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_wave, uint32_t& total) {
    constexpr int W = SCAN_BLOCK / 64;
    const uint32_t inc = wave_incl_scan(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 63) s_wave[wid] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int i = 0; i < W; ++i) {
            const uint32_t t = s_wave[i];
            s_wave[i] = acc;
            acc += t;
        }
        s_wave[W] = acc;
    }
    __syncthreads();
    total = s_wave[W];
    const uint32_t r = s_wave[wid] + inc - v;
    __syncthreads();
    return r;
}

// Exclusive prefix of the scan blocks' sums before block b (in launch order;
// `own` is block b's sum), by a decoupled look-back, one wave: block b
// publishes its own sum, then reads the states of the blocks before it, 64 at
// a time, back to the nearest one that has published its inclusive prefix, and
// publishes its own inclusive prefix.  A block waits only after publishing,
// and blocks are dispatched in order, so a block it waits for runs or has run.
// A wait that has not ended after LOOKBACK_SPINS reads sets invariant bit
// 0x100 and returns (the run is then reported as not to be trusted) instead of
// waiting on.
constexpr uint32_t LOOKBACK_SPINS = 1u << 20;
__device__ __forceinline__ uint32_t scan_lookback(const Ws& ws, uint32_t b, uint32_t own) {
    constexpr uint64_t F_OWN = 1ull << 32, F_INC = 2ull << 32;
    const int lane = threadIdx.x & 63;
    if (lane == 0)
        __hip_atomic_store(&ws.bstate[b], (b == 0 ? F_INC : F_OWN) | own, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t pre = 0, spins = 0;
    int64_t j0 = (int64_t)b - 1;   // the window holds blocks j0 - lane
    while (j0 >= 0) {
        const int64_t j = j0 - lane;
        const uint64_t st =
            j >= 0 ? __hip_atomic_load(&ws.bstate[j], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) : F_INC;
        const uint32_t fl = (uint32_t)(st >> 32);
        const uint64_t inc = __ballot(fl == 2u), none = __ballot(fl == 0u);
        // lanes up to the nearest inclusive state (all, if the window has none)
        const uint64_t upto = inc ? (inc & (~inc + 1ull)) * 2ull - 1ull : ~0ull;
        if (none & upto) {
            if (++spins > LOOKBACK_SPINS) {
                if (lane == 0) atomicOr(&ws.hdr->error, 0x100u);
                return pre;
            }
            __builtin_amdgcn_s_sleep(2);
            continue;
        }
        pre += wave_sum(((upto >> lane) & 1ull) ? (uint32_t)st : 0u);
        if (inc) break;
        j0 -= 64;
    }
    if (lane == 0)
        __hip_atomic_store(&ws.bstate[b], F_INC | (pre + own), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    return pre;
}

// One pass: each workgroup scans SCAN_CHUNK tile counts, takes its offset by
// the look-back, and writes its tiles' bases; the tile-stage invariant codes
// are folded into the run's word and the last block writes the total.
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan(Ws ws) {
    __shared__ uint32_t s_wave[SCAN_BLOCK / 64 + 1];
    __shared__ uint32_t s_pre;
    const int64_t t0 = (int64_t)blockIdx.x * SCAN_CHUNK + threadIdx.x * SCAN_ITEMS;
    uint32_t c[SCAN_ITEMS];
    uint32_t sum = 0, err = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        const uint32_t cw = (t0 + i < ws.ntiles) ? ws.counts[t0 + i] : 0u;
        c[i] = cw & COUNT_MASK;
        err |= cw >> COUNT_ERR_SHIFT;
        sum += c[i];
    }
    if (err) atomicOr(&ws.hdr->error, err);
    uint32_t total;
    uint32_t run = block_excl_scan(sum, s_wave, total);
    if (threadIdx.x < 64) {
        const uint32_t pre = scan_lookback(ws, blockIdx.x, total);
        if (threadIdx.x == 0) {
            s_pre = pre;
            if (blockIdx.x == gridDim.x - 1) ws.hdr->total = pre + total;
        }
    }
    __syncthreads();
    run += s_pre;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        if (t0 + i < ws.ntiles) ws.base[t0 + i] = run;
        run += c[i];
    }
}

// ---- block-local seam merge ----------------------------------------------

// Fragments a merge block keeps in LDS (every block fits at T <= 32).
template <int T>
struct MergeCap {
    static constexpr int all = MERGE_BT * MERGE_BT * (4 * T - 4);
    static constexpr int value = all < 12288 ? all : 12288;
};

// Block tile holding local fragment l: lb[i] is tile i's first local id and
// lb[NB] the block's total, so lb[lo] <= l < lb[lo + 1] at the end.
__device__ __forceinline__ int block_tile_of(const uint32_t* lb, uint32_t l) {
    int lo = 0, hi = MERGE_BT * MERGE_BT;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (lb[mid] <= l) lo = mid;
        else hi = mid;
    }
    return lo;
}

// One workgroup per block of MERGE_BT x MERGE_BT tiles of one image.  The
// block's fragments become a union-find in LDS, the seams inside the block
// are merged there, and each fragment's dense parent is written as its
// block-local root, so the global merge (which sees only the seams between
// blocks) starts from trees of depth one.  Local ids follow the dense ids'
// order, so a local root is also the smallest dense id of its component.
// Also initialises the dense evidence bits: a fragment's own local evidence,
// and at a block root the OR over its block component.  A block with more
// fragments than LDS holds merges its inner seams on the dense arrays instead.
template <int T, bool MULTI>
__global__ __launch_bounds__(MERGE_THREADS) void k_block_merge(Images ims, Ws ws) {
    constexpr int S = 4 * T - 4, NB = MERGE_BT * MERGE_BT, CAP = MergeCap<T>::value;
    static_assert(NB == 64, "block tile scan is one wave");
    __shared__ uint32_t par[CAP];
    __shared__ uint32_t lb[NB + 1];
    __shared__ uint32_t gb[NB];
    __shared__ alignas(16) uint8_t srec[NB][4 * T];
    __shared__ alignas(16) uint8_t seq[NB][2 * T];
    __shared__ uint32_t own2[CAP / 16];   // 2-bit local evidence per fragment
    __shared__ uint32_t acc2[CAP / 16];   // 2-bit OR over the component, at its root
    constexpr int NRS = (MERGE_BT - 1) * MERGE_BT;   // right-edge inner seams (bottom likewise)
    __shared__ uint8_t slist[2 * NRS];    // inner seams to visit
    __shared__ uint32_t s_nseam;
    const ImgDev im = MULTI ? ims.table[ws.block_img[blockIdx.x]] : ims.one;
    const uint32_t bl = (uint32_t)((int64_t)blockIdx.x - im.block_begin);
    const int tx0 = (int)(bl % (uint32_t)im.blocks_x) * MERGE_BT, ty0 = (int)(bl / (uint32_t)im.blocks_x) * MERGE_BT;
    const int ntx = min(MERGE_BT, im.tiles_x - tx0), nty = min(MERGE_BT, im.tiles_y - ty0);
    auto gtile = [&](int i) -> int64_t {
        return im.tile_begin + (int64_t)(ty0 + i / MERGE_BT) * im.tiles_x + tx0 + i % MERGE_BT;
    };
    if (threadIdx.x < NB) {
        const int i = threadIdx.x;
        uint32_t c = 0, g = 0;
        if (i % MERGE_BT < ntx && i / MERGE_BT < nty) {
            const int64_t t = gtile(i);
            c = ws.counts[t] & COUNT_MASK;
            g = ws.base[t];
        }
        const uint32_t inc = wave_incl_scan(c);
        lb[i] = inc - c;
        gb[i] = g;
        if (i == NB - 1) lb[NB] = inc;
    }
    __syncthreads();
    const uint32_t L = lb[NB];
    if (L == 0) return;
    const bool in_lds = L <= (uint32_t)CAP;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // Block tile of each local fragment (blocks held in LDS), one byte each,
    // in the perimeter-record array while that is not loaded: filled here for
    // the fragment init and again for the dense write-out.
    static_assert(CAP <= NB * 4 * T, "the tile map must fit in the record array");
    uint8_t* const tof = &srec[0][0];
    auto fill_tof = [&]() {
        for (int i = wave; i < NB; i += MERGE_THREADS / 64)
            for (uint32_t l = lb[i] + lane; l < lb[i + 1]; l += 64) tof[l] = (uint8_t)i;
    };
    // Perimeter records (4T bytes) and cross-seam bits (2T bytes) of the
    // block's tiles, 16 bytes per load, all loads of a thread issued together
    // (before the fragment init, so that they are in flight meanwhile).
    constexpr int RV = T / 4, EV = T / 8;   // 16-byte vectors per tile
    constexpr int NV = NB * (RV + EV);
    constexpr int PER = (NV + MERGE_THREADS - 1) / MERGE_THREADS;
    uint4 v[PER];
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        const int w = threadIdx.x + m * MERGE_THREADS;
        v[m] = make_uint4(0, 0, 0, 0);
        if (w >= NV) continue;
        const int i = w / (RV + EV), d = w % (RV + EV);
        if (lb[i + 1] == lb[i]) continue;
        v[m] = d < RV ? reinterpret_cast<const uint4*>(ws.rec + gtile(i) * 4 * T)[d]
                      : reinterpret_cast<const uint4*>(ws.eq + gtile(i) * 2 * T)[d - RV];
    }
    for (uint32_t w = threadIdx.x; w < CAP / 16; w += MERGE_THREADS) {
        own2[w] = 0;
        acc2[w] = 0;
    }
    if (in_lds) fill_tof();
    __syncthreads();
    // Fragment init, four fragments per thread and pass so that their
    // evidence loads are in flight together.
    for (uint32_t l0 = 0; l0 < L; l0 += 4 * MERGE_THREADS) {
        uint32_t sb[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t l = l0 + u * MERGE_THREADS + threadIdx.x;
            sb[u] = 0;
            if (l < L) {
                const int t = in_lds ? (int)tof[l] : block_tile_of(lb, l);
                sb[u] = ws.sbits[gtile(t) * S + (l - lb[t])] & 3u;
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t l = l0 + u * MERGE_THREADS + threadIdx.x;
            if (l >= L) continue;
            if (in_lds) {
                par[l] = l;
                if (sb[u]) atomicOr(&own2[l >> 4], sb[u] << (2 * (l & 15u)));
            } else {
                // every fragment is its own block root; the inner seams are
                // merged on the dense arrays below
                const int t = block_tile_of(lb, l);
                const uint32_t g = gb[t] + (l - lb[t]);
                ws.dparent[g] = g;
                ws.dbits[g] = sb[u] | (sb[u] << 2) | DB_BROOT;
            }
        }
    }
    __syncthreads();   // the tile map is read no more
    if (wave == 0) {
        // list of the inner seams to visit: slots [0, NRS) right edges, [NRS,
        // 2 NRS) bottom edges (as in the loop below)
        uint32_t nsl = 0;
        for (int s0 = 0; s0 < 2 * NRS; s0 += 64) {
            const int sl = s0 + lane;
            bool ok = false;
            if (sl < 2 * NRS) {
                const bool r = sl < NRS;
                const int q = r ? sl : sl - NRS;
                const int ix = r ? q % (MERGE_BT - 1) : q % MERGE_BT, iy = r ? q / (MERGE_BT - 1) : q / MERGE_BT;
                const int i = iy * MERGE_BT + ix;
                ok = (r ? (ix + 1 < ntx && iy < nty) : (ix < ntx && iy + 1 < nty)) && lb[i + 1] != lb[i];
            }
            const uint64_t m = __ballot(ok);
            if (ok) slist[nsl + __builtin_popcountll(m & ((1ull << lane) - 1ull))] = (uint8_t)sl;
            nsl += (uint32_t)__builtin_popcountll(m);
        }
        if (lane == 0) s_nseam = nsl;
    }
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        const int w = threadIdx.x + m * MERGE_THREADS;
        if (w >= NV) continue;
        const int i = w / (RV + EV), d = w % (RV + EV);
        if (d < RV) reinterpret_cast<uint4*>(srec[i])[d] = v[m];
        else reinterpret_cast<uint4*>(seq[i])[d - RV] = v[m];
    }
    __syncthreads();
    // Inner seams: slots s in [0, NRS) are right columns (between block tile
    // (s % (BT-1), s / (BT-1)) and the tile right of it), s in [NRS, 2 NRS)
    // bottom rows (between tile ((s - NRS) % BT, (s - NRS) / BT) and the tile
    // below); T positions per seam.
    // Position k of a seam pairs pixel A(k) of tile i's edge with the pixels
    // B(k - 1), B(k), B(k + 1) of the facing edge (equality bits 0, 1, 2).
    // Most pairs are implied by another one exactly, with the same two
    // fragments, because equal pixels adjacent along one tile edge share a
    // fragment:
    //   (A(k), B(k)) with bit 0 at k and bit 1 at k - 1: A(k-1) = B(k-1) =
    //     A(k) = B(k), the pair of position k - 1;
    //   (A(k), B(k-1)), k > 0, with bit 1 at k: B(k-1) = A(k) = B(k), the
    //     pair (A(k), B(k)); with bit 1 at k - 1: A(k-1) = B(k-1) = A(k), the
    //     pair of position k - 1;
    //   (A(k), B(k+1)), k < T - 1, likewise with bit 1 at k or k + 1.
    // So a position unites its straight pair unless the one before holds it,
    // and a diagonal pair only at a tile corner or where neither straight
    // pair exists.
    bool bad = false;
    uint32_t retries = 0;
    auto unite = [&](uint32_t a, int ta, uint32_t b, int tb) {
        if (in_lds) lunion_pair(par, a, b);
        else gunion(ws.dparent, gb[ta] + (a - lb[ta]), gb[tb] + (b - lb[tb]), retries);
    };
    // Only the inner seams that exist (a block at an image's right or bottom
    // edge may be partial) and whose first tile holds a fragment are visited:
    // their slots, in order, were listed by wave 0.
    const int NP = (int)s_nseam * T;
    // The body is kept free of divergent branches up to the unions (LDS reads
    // at clamped addresses, conditions folded into the pair mask): every
    // divergent branch costs the wave scalar instructions for its exec mask.
#pragma unroll 1
    for (int P0 = 0; P0 < NP; P0 += MERGE_THREADS) {
        const int P = P0 + threadIdx.x;
        const bool inp = P < NP;
        const int Pc = inp ? P : 0;
        const int k = Pc % T;
        const int sl = (int)slist[Pc / T];
        const bool right = sl < NRS;
        const int sidx = right ? sl : sl - NRS;
        const int ix = right ? sidx % (MERGE_BT - 1) : sidx % MERGE_BT;
        const int iy = right ? sidx / (MERGE_BT - 1) : sidx / MERGE_BT;
        const int i = iy * MERGE_BT + ix;
        const uint8_t* eqs = seq[i] + (right ? 0 : T);
        const uint32_t e = inp ? (uint32_t)eqs[k] : 0u;
        const uint32_t ep = k > 0 ? (uint32_t)eqs[k > 0 ? k - 1 : 0] : 0u;
        const uint32_t en = k < T - 1 ? (uint32_t)eqs[k < T - 1 ? k + 1 : k] : 0u;
        const bool s1 = (e & 2u) != 0;
        const bool u1 = s1 && !((e & 1u) && (ep & 2u));
        const bool u0 = (e & 1u) && (k == 0 || (!s1 && !(ep & 2u)));
        const bool u2 = (e & 4u) && (k == T - 1 || (!s1 && !(en & 2u)));
        const uint32_t ci = lb[i + 1] - lb[i];
        const uint32_t ja = srec[i][right ? 3 * T + k : T + k];
        bad = bad || (e != 0u && ja >= ci);
        const uint32_t a = lb[i] + ja;
        // one union site for the (up to three) pairs of the position
        uint32_t todo = (e != 0u && ja < ci) ? ((u0 ? 1u : 0u) | (u1 ? 2u : 0u) | (u2 ? 4u : 0u)) : 0u;
        while (todo) {
            const int d = __builtin_ctz(todo);
            todo &= todo - 1u;
            const int nk = k + d - 1, dn = nk < 0 ? -1 : (nk >= T ? 1 : 0);
            const int niy = right ? iy + dn : iy + 1, nix = right ? ix + 1 : ix + dn;
            const bool inb = niy >= 0 && niy < nty && nix >= 0 && nix < ntx;   // else another block's pair
            const int nt = inb ? niy * MERGE_BT + nix : i;
            const uint32_t jb = srec[nt][(right ? 2 * T : 0) + nk - dn * T];
            const uint32_t cn = lb[nt + 1] - lb[nt];
            bad = bad || (inb && jb >= cn);
            if (inb && jb < cn) unite(a, i, lb[nt] + jb, nt);
        }
    }
    __syncthreads();
    if (in_lds) {
        fill_tof();   // the records are read no more
        // evidence of each block component at its root; labels compressed
        for (uint32_t l = threadIdx.x; l < L; l += MERGE_THREADS) {
            const uint32_t r = lfind(par, l);
            if (r != l) par[l] = r;
            const uint32_t b = (own2[l >> 4] >> (2 * (l & 15u))) & 3u;
            if (b & ~(lds_load(&acc2[r >> 4]) >> (2 * (r & 15u))))
                atomicOr(&acc2[r >> 4], b << (2 * (r & 15u)));
        }
        __syncthreads();
        // dense arrays: parent = block root; a block root also keeps its
        // component's evidence in bits 2-3
        for (uint32_t l = threadIdx.x; l < L; l += MERGE_THREADS) {
            const int t = tof[l];
            const uint32_t g = gb[t] + (l - lb[t]);
            const uint32_t r = par[l];
            const int tr = tof[r];
            ws.dparent[g] = gb[tr] + (r - lb[tr]);
            uint32_t d = (own2[l >> 4] >> (2 * (l & 15u))) & 3u;
            if (r == l) d |= (((acc2[l >> 4] >> (2 * (l & 15u))) & 3u) << 2) | DB_BROOT;
            ws.dbits[g] = d;
        }
    }
    if (retries) atomicAdd((unsigned long long*)&ws.hdr->retries, (unsigned long long)retries);
    if (bad) atomicOr(&ws.hdr->error, 2u);
}

// ---- pairs across a tile edge (group and seam merges) ---------------------

// Words of position k of a tile edge (right edge of tile gtile if `right`,
// else its bottom edge) and of the facing tile nt1, loaded together (one round
// trip); those of a tile without fragments are stale and masked by its count.
struct EdgeWords {
    uint32_t e = 0, cnt = 0, ja = 0, base = 0, jb1 = 0, cnt1 = 0, base1 = 0;
};
template <int T>
__device__ __forceinline__ EdgeWords load_edge(const Ws& ws, int64_t gtile, int64_t nt1, bool right, int k) {
    EdgeWords w;
    w.cnt = ws.counts[gtile] & COUNT_MASK;
    w.e = ws.eq[gtile * 2 * T + (right ? k : T + k)];
    w.ja = ws.rec[gtile * 4 * T + (right ? 3 * T + k : T + k)];
    w.base = ws.base[gtile];
    w.jb1 = ws.rec[nt1 * 4 * T + (right ? 2 * T + k : k)];
    w.cnt1 = ws.counts[nt1] & COUNT_MASK;
    w.base1 = ws.base[nt1];
    if (!w.cnt) w.e = 0;
    return w;
}

// Which of the three pairs (A(k), B(k - 1 + d)) of an edge position to unite.
// Equal pixels adjacent along one tile edge share a fragment, so most pairs
// join the same two fragments as another one: the straight pair is skipped
// when position k - 1 holds the same one (bit 0 at k and bit 1 at k - 1), a
// diagonal pair whenever this position or the neighbouring one has a straight
// pair, except at a tile corner.  A unit's T positions are consecutive lanes
// (every lane of the wave calls this: it reads the neighbours' bits).
template <int T>
__device__ __forceinline__ void edge_pairs_used(uint32_t e, int k, bool use[3]) {
    uint32_t ep = __shfl_up(e, 1, 64), en = __shfl_down(e, 1, 64);
    if (k == 0) ep = 0;
    if (k == T - 1) en = 0;
    const bool s1 = (e & 2u) != 0;
    use[0] = (e & 1u) && (k == 0 || (!s1 && !(ep & 2u)));
    use[1] = s1 && !((e & 1u) && (ep & 2u));
    use[2] = (e & 4u) && (k == T - 1 || (!s1 && !(en & 2u)));
}

// Block roots of the used pairs: ra for A(k), rb[d] for B(k - 1 + d) (NONE if
// unused).  A perimeter record out of range (a broken tile stage) sets bad.
template <int T>
__device__ __forceinline__ void edge_pair_roots(const Ws& ws, const ImgDev& im, const EdgeWords& w, int tx, int ty,
                                                bool right, int k, const bool use[3], uint32_t& ra, uint32_t rb[3],
                                                bool& bad) {
    constexpr uint32_t NONE = 0xFFFFFFFFu;
    uint32_t fa = NONE, fb[3] = {NONE, NONE, NONE};
    ra = NONE;
    rb[0] = rb[1] = rb[2] = NONE;
    if (!w.e) return;
    if (w.ja < w.cnt) fa = w.base + w.ja; else bad = true;
    if (use[1] && fa != NONE) {
        if (w.jb1 < w.cnt1) fb[1] = w.base1 + w.jb1; else bad = true;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (d == 1 || !use[d] || fa == NONE) continue;   // the diagonals
        int64_t nt;
        uint32_t slot;
        if (right) {
            const int ny = ty * T + k + d - 1, nty = ny / T;
            nt = im.tile_begin + (int64_t)nty * im.tiles_x + tx + 1;
            slot = 2 * T + (ny - nty * T);   // left column of the tile to the right
        } else {
            const int nx = tx * T + k + d - 1, ntx = nx / T;
            nt = im.tile_begin + (int64_t)(ty + 1) * im.tiles_x + ntx;
            slot = nx - ntx * T;             // top row of the tile below
        }
        const uint32_t jb = ws.rec[nt * 4 * T + slot];
        if (jb >= (ws.counts[nt] & COUNT_MASK)) { bad = true; continue; }
        fb[d] = ws.base[nt] + jb;
    }
    // after the block merge a fragment's parent is its block root (or, once
    // that root is linked, an ancestor of it)
    if (fa != NONE) {
        ra = ws.dparent[fa];
#pragma unroll
        for (int d = 0; d < 3; ++d)
            if (fb[d] != NONE) rb[d] = ws.dparent[fb[d]];
    }
}

// ---- group merge: the block seams inside 2 x 2 groups of merge blocks -----

// Every union of the global seam merge ends in a device-scope CAS, and their
// number, not the walks, sets that kernel's time (with plain stores in place
// of the CAS it ran in 40 % of the time).  Merge blocks are paired up: in each
// group of 2 x 2 blocks (GROUP_T x GROUP_T tiles) the two block seams inside
// it are merged here, in LDS, by one workgroup that alone touches the group's
// block roots, so its links are plain stores; the seam merge then takes only
// the seams between groups (and the diagonal contacts that leave a group at a
// tile corner).  One workgroup per group works on it: its edges are the right edges of the tiles in the last
// tile column of its block column (vertical seam, GROUP_T tiles) and the
// bottom edges of the tiles in the last tile row of its block row (horizontal
// seam), T positions each.  The group's distinct block roots get LDS slots in
// an open-addressed hash table and are united there; then each joins the
// smallest id of its component (a smaller id, as the seam merge's links).  A
// group with more roots than the table holds unites all its pairs with the
// seam merge's CAS instead.
constexpr int GROUP_T = 2 * MERGE_BT;   // tiles per group side
constexpr int GROUP_THREADS = 1024;
constexpr int GSET = 4096;              // LDS slots for the distinct block roots of a group

template <int T, bool MULTI>
__global__ __launch_bounds__(GROUP_THREADS) void k_group_merge(Images ims, Ws ws) {
    constexpr uint32_t EMPTY = 0xFFFFFFFFu;
    constexpr int NPOS = 2 * GROUP_T * T;             // edge positions of a group
    constexpr int PER = (NPOS + GROUP_THREADS - 1) / GROUP_THREADS;   // per thread
    static_assert(GROUP_THREADS % T == 0, "whole edges per thread pass");
    __shared__ uint32_t key[GSET];    // block-root id of a slot
    __shared__ uint32_t gpar[GSET];   // union-find over slots
    __shared__ uint32_t gmin[GSET];   // smallest id of a component, at its root slot
    __shared__ uint32_t used[GSET];   // the slots in use
    __shared__ uint32_t s_over, s_used;
    // one workgroup per group
    const ImgDev im = MULTI ? ims.table[ws.group_img[blockIdx.x]] : ims.one;
    const uint32_t gl = (uint32_t)((int64_t)blockIdx.x - im.group_begin), gx = (uint32_t)(im.blocks_x + 1) / 2;
    const int bx = 2 * (int)(gl % gx), by = 2 * (int)(gl / gx);
    const bool hasV = bx + 1 < im.blocks_x, hasH = by + 1 < im.blocks_y;
    if (!hasV && !hasH) return;
    for (int i = threadIdx.x; i < GSET; i += GROUP_THREADS) {
        key[i] = EMPTY;
        gpar[i] = (uint32_t)i;
        gmin[i] = EMPTY;
    }
    if (threadIdx.x == 0) {
        s_over = 0;
        s_used = 0;
    }
    bool bad = false;
    uint32_t ra[PER], rb[PER][3];
    bool keep[PER][3];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int p = (int)threadIdx.x + q * GROUP_THREADS;
        const int edge = p / T, k = p % T;
        const bool right = edge < GROUP_T;
        const int tx = right ? MERGE_BT * bx + MERGE_BT - 1 : MERGE_BT * bx + (edge - GROUP_T);
        const int ty = right ? MERGE_BT * by + edge : MERGE_BT * by + MERGE_BT - 1;
        const bool valid = p < NPOS && (right ? (hasV && ty < im.tiles_y) : (hasH && tx < im.tiles_x));
        EdgeWords w;
        const int64_t gtile = im.tile_begin + (int64_t)ty * im.tiles_x + tx;
        if (valid) w = load_edge<T>(ws, gtile, right ? gtile + 1 : gtile + im.tiles_x, right, k);
        bool use[3];
        edge_pairs_used<T>(w.e, k, use);
        // a diagonal pair at a tile corner may leave the group: the seam
        // merge takes it
        const int along = right ? ty : tx, lim = right ? im.tiles_y : im.tiles_x;
        const int first = right ? MERGE_BT * by : MERGE_BT * bx;
        if (k == 0 && along - 1 < first) use[0] = false;
        if (k == T - 1 && (along + 1 >= first + GROUP_T || along + 1 >= lim)) use[2] = false;
        edge_pair_roots<T>(ws, im, w, tx, ty, right, k, use, ra[q], rb[q], bad);
#pragma unroll
        for (int d = 0; d < 3; ++d) keep[q][d] = rb[q][d] != 0xFFFFFFFFu;
    }
    __syncthreads();   // table initialised
    auto slot_of = [&](uint32_t id) -> int {
        const uint32_t h = (id * 0x9E3779B1u) >> 20;
        static_assert(GSET == 4096, "the hash above yields 12 bits");
        for (int pr = 0; pr < 64; ++pr) {
            const int sl = (int)((h + (uint32_t)pr) & (GSET - 1));
            const uint32_t old = atomicCAS(&key[sl], EMPTY, id);
            if (old == EMPTY) used[atomicAdd(&s_used, 1u)] = (uint32_t)sl;
            if (old == EMPTY || old == id) return sl;
        }
        return -1;
    };
#pragma unroll
    for (int q = 0; q < PER; ++q) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (!keep[q][d]) continue;
            const int sa = slot_of(ra[q]), sb = slot_of(rb[q][d]);
            if (sa < 0 || sb < 0) {
                s_over = 1u;
                continue;
            }
            lunion_pair(gpar, (uint32_t)sa, (uint32_t)sb);
        }
    }
    __syncthreads();
    const uint32_t nused = s_used;
    for (uint32_t u = threadIdx.x; u < nused; u += GROUP_THREADS) {
        const uint32_t i = used[u];
        atomicMin(&gmin[lfind(gpar, i)], key[i]);
    }
    __syncthreads();
    for (uint32_t u = threadIdx.x; u < nused; u += GROUP_THREADS) {
        const uint32_t i = used[u], id = key[i];
        const uint32_t m = gmin[lfind(gpar, i)];
        if (m != id) ws.dparent[id] = m;   // id was a block root: no one else writes it meanwhile
    }
    uint32_t retries = 0;
    if (s_over) {
        // the table overflowed: every pair of the group through the CAS union
        __threadfence();
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PER; ++q)
#pragma unroll
            for (int d = 0; d < 3; ++d)
                if (keep[q][d]) gunion(ws.dparent, ra[q], rb[q][d], retries);
    }
    if (retries) atomicAdd((unsigned long long*)&ws.hdr->retries, (unsigned long long)retries);
    if (bad) atomicOr(&ws.hdr->error, 2u);
}

// ---- seam merge -----------------------------------------------------------

// One thread per (seam unit, k).  A unit is a tile edge on a merge-block
// boundary: the right edge of a tile in the last tile column of a block
// (pixel k of its right column against the three pixels right of it), or the
// bottom edge of a tile in the last tile row of a block (pixel k of its bottom
// row against the three below it).  Every equal pixel pair that straddles two
// blocks is among these: a pair across a block column is on a right edge, a
// pair across a block row on a bottom edge, and a diagonal pair over a tile
// corner is both the up-right contact of one tile's right edge and the
// down-left contact of the bottom edge of the tile above-right of it (resp.
// the down-right contact of both edges of one tile).  Pairs inside one block
// were united by k_block_merge, pairs inside a 2 x 2 group of blocks by
// k_group_merge; the edges on seams inside a group have nothing left (see
// below).
//
// A plateau crossing a block seam yields the same pair of block components at
// many pixels of the seam, and every union ends in a device-scope CAS (or
// several, when they collide), which is what this kernel's time goes to.  So
// the units of one block seam are consecutive (a workgroup of SEAM_THREADS
// covers the MERGE_BT edges of one seam at T = 64), and each distinct pair of
// block roots is united once per workgroup: an LDS hash set drops the pairs
// another lane of the workgroup already holds (a wave-level leader loop that
// found the distinct pairs first cost more than it saved).
constexpr int SEAM_THREADS = 512;
constexpr int SEAM_SET = 2048;   // LDS hash-set slots (pairs a workgroup may keep: 3 per thread)

template <int T, bool MULTI>
__global__ __launch_bounds__(SEAM_THREADS) void k_seam_merge(Images ims, Ws ws) {
    constexpr unsigned long long EMPTY = ~0ull;
    __shared__ unsigned long long set[SEAM_SET];
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t unit = g / T;
    const int k = (int)(g % T);
    uint32_t retries = 0;
    bool bad = false;
    EdgeWords w;
    ImgDev im{};
    int tx = 0, ty = 0;
    bool right = false, inner = true;
    int64_t gtile = 0;
    if (unit < ws.nseam) {
        im = MULTI ? ims.table[ws.seam_img[unit]] : ims.one;
        const uint32_t u = (uint32_t)(unit - im.seam_begin);
        right = seam_unit_edge(u, (uint32_t)im.seam_cols, (uint32_t)im.seam_rows, (uint32_t)im.tiles_x,
                               (uint32_t)im.tiles_y, tx, ty);
        // a seam inside a group: between block columns (rows) 2 i and 2 i + 1
        inner = (((right ? tx : ty) / MERGE_BT) & 1) == 0;
        // by construction the tile right of (below) the edge exists
        gtile = im.tile_begin + (int64_t)ty * im.tiles_x + tx;
    }
    // An edge on a seam inside a group has nothing left: k_group_merge united
    // the pairs inside the group, and each diagonal contact at one of its tile
    // corners that reaches another group is also the corner contact of an
    // edge on a seam between groups (the right edge of the tile below-left of
    // it, the bottom edge of the tile above-right, or the other edge of the
    // same tile), which this kernel takes.  A workgroup of such edges only
    // (one block seam at T = 64) leaves at once.
    if (__syncthreads_and(inner)) return;
    for (int i = threadIdx.x; i < SEAM_SET; i += SEAM_THREADS) set[i] = EMPTY;
    __syncthreads();
    if (!inner) w = load_edge<T>(ws, gtile, right ? gtile + 1 : gtile + im.tiles_x, right, k);
    bool use[3];
    edge_pairs_used<T>(w.e, k, use);
    uint32_t ra, rb[3];
    edge_pair_roots<T>(ws, im, w, tx, ty, right, k, use, ra, rb, bad);
    bool keep[3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
        keep[d] = rb[d] != 0xFFFFFFFFu && !(d > 0 && rb[d] == rb[d - 1]) && !(d > 1 && rb[d] == rb[d - 2]);
    // Workgroup set: each distinct pair of block roots is united once per
    // workgroup (a pair found there was kept by another lane).  A pair that
    // finds no slot within 8 probes is simply united (a repeated union is
    // harmless).
    static_assert(SEAM_SET == 2048, "the hash below yields 11 bits");
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (!keep[d]) continue;
        const uint32_t lo = min(ra, rb[d]), hi = max(ra, rb[d]);
        const unsigned long long key = (unsigned long long)lo << 32 | hi;
        const uint32_t h = (lo * 0x9E3779B1u ^ hi * 0x85EBCA77u) >> 21;
        for (int p = 0; p < 8; ++p) {
            const unsigned long long old = atomicCAS(&set[(h + p) & (SEAM_SET - 1)], EMPTY, key);
            if (old == EMPTY) break;
            if (old == key) { keep[d] = false; break; }
        }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d)
        if (keep[d]) gunion(ws.dparent, ra, rb[d], retries);
    const uint32_t r = wave_sum(retries);
    if ((threadIdx.x & 63) == 0 && r) atomicAdd((unsigned long long*)&ws.hdr->retries, (unsigned long long)r);
    // a perimeter record out of range would mean a broken tile stage: report
    // it instead of following it
    if (bad) atomicOr(&ws.hdr->error, 2u);
}

// ---- resolve: compress + OR block evidence into the roots ----------------

// Only block roots take part: every other fragment's parent is its block root,
// which carries the evidence of its whole block component.

// A component can hold millions of fragments (a near-percolating plateau), so
// OR-ing every fragment's evidence into the one root word would serialise on
// that address.  Each wave therefore issues one atomic per distinct root and
// only for bits the (possibly stale, never wrong) cached root word lacks.
__global__ __launch_bounds__(256) void k_resolve(Ws ws) {
    const uint32_t total = ws.hdr->total;
    const int lane = threadIdx.x & 63;
    // 64-bit position: a total near 2^32 must not wrap the loop counter
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t wbase = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u); wbase < total; wbase += stride) {
        const uint32_t i = (uint32_t)wbase + lane;
        uint32_t r = 0, lb = 0;
        const uint32_t d = i < total ? ws.dbits[i] : 0u;
        if (d & DB_BROOT) {
            // halving walk: compresses the trees for the other threads; a later
            // halving store may replace this entry with an ancestor, so
            // consumers finish the walk with gfind_ro
            r = gfind(ws.dparent, i);
            if (r != i) ws.dparent[i] = r;
            lb = (d >> 2) & 3u;
            if (lb && ((ws.dbits[r] >> 2) & lb) == lb) lb = 0;
        }
        uint64_t pending = __ballot(lb != 0);
        while (pending) {
            const int leader = __ffsll((unsigned long long)pending) - 1;
            const uint32_t rl = __shfl(r, leader, 64);
            const bool mine = lb != 0 && r == rl;
            const uint64_t m = __ballot(mine);
            const uint32_t bits = (__ballot(mine && (lb & 1u)) ? 1u : 0u) | (__ballot(mine && (lb & 2u)) ? 2u : 0u);
            if (lane == leader)
                __hip_atomic_fetch_or(ws.dbits + rl, bits << 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (mine) lb = 0;
            pending &= ~m;
        }
    }
}

// ---- final per-fragment results -------------------------------------------

// One thread per fragment (dense id), once every root holds its component's
// evidence: the fragment's final masks and which of them differ from what the
// tile stage wrote (it wrote the masks of the fragment's own evidence, so a
// mask changes where the zone has evidence the fragment lacks).  Computing
// these densely keeps the union-find walks off the write-back's per-tile
// critical path: there a tile's bytes are one load.
__global__ __launch_bounds__(256) void k_final(Ws ws) {
    const uint32_t total = ws.hdr->total;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const uint32_t own = ws.dbits[i];
        uint32_t res = own & 3u;
        if (res != 3u) res |= (ws.dbits[gfind_ro(ws.dparent, (uint32_t)i)] >> 2) & 3u;
        const uint32_t gain = res & ~own & 3u;   // bit 0: greater evidence gained, bit 1: smaller
        ws.flut[i] = (uint8_t)(gain << 6 | ((res & F_G) ? 0u : 1u) | ((res & F_S) ? 0u : 2u));
    }
}

// ---- mark + write-back of changed seam zones -----------------------------

// One wave per tile, a fixed grid of waves over contiguous tile ranges (the
// ranges of one XCD's waves are adjacent, so neighbouring tiles share an L2).
// Tiles without an open fragment are skipped on their counts word alone.
template <int T, bool MULTI>
__global__ __launch_bounds__(256) void k_writeback(Images ims, Ws ws) {
    // per wave: the resolved byte of every fragment id (0 .. S - 1) and, at
    // REC_CLOSED + c, the closed code c itself (final masks c, no change)
    __shared__ alignas(256) uint8_t lut[4 * 256];
    __shared__ uint32_t s_dirty;
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (threadIdx.x == 0) s_dirty = 0;
    if (lane < 4) lut[wid * 256 + REC_CLOSED + lane] = (uint8_t)lane;
    __syncthreads();
    const int64_t nw = (int64_t)gridDim.x * 4;
    const int64_t w = xcd_block(blockIdx.x, gridDim.x) * 4 + wid;
    const int64_t t0 = ws.ntiles * w / nw, t1 = ws.ntiles * (w + 1) / nw;
    uint32_t nd = 0;
    // the counts and base words of up to 64 tiles in one load each; the tiles
    // that hold a fragment which may change are then taken one by one
    for (int64_t tb = t0; tb < t1; tb += 64) {
        const int64_t tl = tb + lane;
        const uint32_t cwl = tl < t1 ? ws.counts[tl] : 0u;
        const uint32_t bl = tl < t1 ? ws.base[tl] : 0u;
        uint64_t todo = __ballot((cwl & COUNT_MASK) && (cwl & (CNT_OPEN | CNT_UNIFORM)));
        while (todo) {
            const int i = __ffsll((unsigned long long)todo) - 1;
            todo &= todo - 1;
            const int64_t t = tb + i;
            const uint32_t cw = __shfl(cwl, i, 64), b = __shfl(bl, i, 64);
            const ImgDev im = locate<MULTI>(ims, ws.tile_img, t);
            nd += tile_writeback<T>(lut, (uint32_t)wid, im, t, cw, b, ws) ? 1u : 0u;
        }
    }
    if (lane == 0 && nd) atomicAdd(&s_dirty, nd);
    __syncthreads();
    if (threadIdx.x == 0 && s_dirty) atomicAdd(&ws.hdr->dirty_count, s_dirty);
}

// ---- seam-sparsity profile (diagnostic) ------------------------------------

// One workgroup per tile; per-tile sums in LDS, one 64-bit atomic per counter
// and workgroup into the workspace header.
template <int T>
__global__ __launch_bounds__(256) void k_seam_profile(Images ims, Ws ws) {
    __shared__ TileSmem<T> sm;
    __shared__ uint32_t c[5];
    if (threadIdx.x < 5) c[threadIdx.x] = 0;   // ordered by the barrier after the tile load
    tile_profile<T, 256>(sm, ims.one, (int64_t)blockIdx.x, c);
    __syncthreads();
    if (threadIdx.x < 5 && c[threadIdx.x])
        atomicAdd((unsigned long long*)&ws.hdr->profile[threadIdx.x], (unsigned long long)c[threadIdx.x]);
}

hipError_t launch_profile(const Images& ims, const Ws& ws, int T, hipStream_t s) {
    hipError_t e = hipMemsetAsync(&ws.hdr->profile[0], 0, sizeof(ws.hdr->profile), s);
    if (e != hipSuccess) return e;
    const unsigned grid = (unsigned)ws.ntiles;
    switch (T) {
    case 16: k_seam_profile<16><<<grid, 256, 0, s>>>(ims, ws); break;
    case 32: k_seam_profile<32><<<grid, 256, 0, s>>>(ims, ws); break;
    case 64: k_seam_profile<64><<<grid, 256, 0, s>>>(ims, ws); break;
    case 128: k_seam_profile<128><<<grid, 256, 0, s>>>(ims, ws); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---- mask digest (autotuner admission) ------------------------------------

// 64-bit finaliser (an avalanche mix: every input bit reaches every output bit).
__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x ^= x >> 31;
    x *= 0x7FB5D329728EA185ull;
    x ^= x >> 27;
    x *= 0x81DADEF4BC2DD44Dull;
    x ^= x >> 33;
    return x;
}

// Order-independent digest of a w x h byte image: the sum over 8-byte chunks
// of mix64(chunk ^ mix64(salt + chunk position)), the position taken in the
// packed w x h layout (the pitch does not enter).  Equal images give equal
// digests; one differing byte changes the digest except with probability
// about 2^-64.  One thread per chunk, one 64-bit atomic per wave.
__global__ __launch_bounds__(256) void k_mask_digest(const uint8_t* p, int64_t pitch, int w, int h, uint64_t salt,
                                                     unsigned long long* out) {
    const int64_t cpr = (w + 7) / 8;   // chunks per row
    const int64_t n = cpr * h;
    uint64_t acc = 0;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        const int64_t y = c / cpr;
        const int x0 = (int)(c - y * cpr) * 8;
        const uint8_t* row = p + y * pitch + x0;
        uint64_t v = 0;
        const int nb = min(8, w - x0);
        for (int b = 0; b < nb; ++b) v |= (uint64_t)row[b] << (8 * b);
        acc += mix64(v ^ mix64(salt + (uint64_t)c));
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d, 64);
    if ((threadIdx.x & 63) == 0 && acc) atomicAdd(out, (unsigned long long)acc);
}

hipError_t launch_digest(const uint8_t* p, int64_t pitch, int w, int h, uint64_t salt, uint64_t* d_out,
                         hipStream_t s) {
    const int64_t chunks = (int64_t)((w + 7) / 8) * h;
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>((chunks + 255) / 256, 4096));
    k_mask_digest<<<grid, 256, 0, s>>>(p, pitch, w, h, salt, reinterpret_cast<unsigned long long*>(d_out));
    return hipGetLastError();
}

// ---- host sequencer -------------------------------------------------------

// One workgroup per tile over the combined tile space of all images: for a
// batch the hardware dispatcher is the tile work queue (a persistent grid
// draining per-XCD queues needed 148 VGPRs per lane and ran slower).
template <int T, int NT, bool MULTI>
static void launch_tile_kernels(const Images& ims, const Ws& ws, hipStream_t s) {
    if constexpr (T == RT && NT == RT_THREADS)
        k_row_tile<MULTI><<<(unsigned)((ws.ntiles + RT_NTW - 1) / RT_NTW), RT_THREADS * RT_NTW, 0, s>>>(ims, ws);
    else
        k_tile_stage<T, NT, MULTI><<<(unsigned)ws.ntiles, NT, 0, s>>>(ims, ws);
}

template <int T, bool MULTI>
static void launch_writeback(const Images& ims, const Ws& ws, int grid_cap, hipStream_t s) {
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((ws.ntiles + 3) / 4, grid_cap));
    k_writeback<T, MULTI><<<(unsigned)grid, 256, 0, s>>>(ims, ws);
}

template <int T, bool MULTI>
static void launch_merge(const Images& ims, const Ws& ws, hipStream_t s) {
    k_block_merge<T, MULTI><<<(unsigned)ws.nblk, MERGE_THREADS, 0, s>>>(ims, ws);
    const int64_t groups = MULTI ? ws.ngroup : (int64_t)((ims.one.blocks_x + 1) / 2) * ((ims.one.blocks_y + 1) / 2);
    k_group_merge<T, MULTI><<<(unsigned)groups, GROUP_THREADS, 0, s>>>(ims, ws);
    const int64_t merge_threads = ws.nseam * T;
    if (merge_threads > 0) {
        const unsigned merge_grid = (unsigned)((merge_threads + SEAM_THREADS - 1) / SEAM_THREADS);
        k_seam_merge<T, MULTI><<<merge_grid, SEAM_THREADS, 0, s>>>(ims, ws);
    }
}

#define SSX_DISPATCH(T_, NT_, CALL)                                        \
    switch ((T_) * 1000 + (NT_)) {                                         \
    case 16064: CALL(16, 64); break;                                       \
    case 16128: CALL(16, 128); break;                                      \
    case 16256: CALL(16, 256); break;                                      \
    case 32064: CALL(32, 64); break;                                       \
    case 32128: CALL(32, 128); break;                                      \
    case 32256: CALL(32, 256); break;                                      \
    case 32512: CALL(32, 512); break;                                      \
    case 64064: CALL(64, 64); break;                                       \
    case 64128: CALL(64, 128); break;                                      \
    case 64256: CALL(64, 256); break;                                      \
    case 64512: CALL(64, 512); break;                                      \
    default: return hipErrorInvalidValue;                                  \
    }

template <bool MULTI>
static hipError_t launch_all(const Images& ims, const Ws& ws, int T, int NT, int cus, hipStream_t s) {
    const int grid_cap = std::max(1, cus) * 8;
    if (MULTI) k_tile_map<<<(unsigned)ims.count, 256, 0, s>>>(ims, ws);
    const ProfileHook& hook = profile_hook();
    if (hook.tile_begin) (void)hipEventRecord(hook.tile_begin, s);
#define CALL_TILE(a, b) launch_tile_kernels<a, b, MULTI>(ims, ws, s)
    SSX_DISPATCH(T, NT, CALL_TILE)
#undef CALL_TILE
    if (hook.tile_end) (void)hipEventRecord(hook.tile_end, s);
    k_scan<<<ws.nblocks, SCAN_BLOCK, 0, s>>>(ws);
    switch (T) {
    case 16: launch_merge<16, MULTI>(ims, ws, s); break;
    case 32: launch_merge<32, MULTI>(ims, ws, s); break;
    case 64: launch_merge<64, MULTI>(ims, ws, s); break;
    default: return hipErrorInvalidValue;
    }
    const int64_t max_frag = ws.ntiles * (int64_t)ws.slots;
    const unsigned resolve_grid =
        (unsigned)std::max<int64_t>(1, std::min<int64_t>((max_frag + 255) / 256, (int64_t)grid_cap * 2));
    k_resolve<<<resolve_grid, 256, 0, s>>>(ws);
    k_final<<<resolve_grid, 256, 0, s>>>(ws);
    switch (T) {
    case 16: launch_writeback<16, MULTI>(ims, ws, grid_cap, s); break;
    case 32: launch_writeback<32, MULTI>(ims, ws, grid_cap, s); break;
    case 64: launch_writeback<64, MULTI>(ims, ws, grid_cap, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_pipeline(const Images& ims, const Ws& ws, int T, int NT, int cus, hipStream_t s) {
    if (ims.count > 1) return launch_all<true>(ims, ws, T, NT, cus, s);
    return launch_all<false>(ims, ws, T, NT, cus, s);
}

}  // namespace ssx
