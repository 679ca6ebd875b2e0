/*
 * ss_oracle.c -- CPU oracle for the exact dual regional-extrema path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by the
 * pytest suite, by __graft_entry__.smoke() and by the cpu_baseline leg of
 * bench.py.  The product path (the HIP library) never links or calls it.
 *
 * Three independent CPU routes produce the same two masks (uint8 0/255,
 * row-major, 8-connectivity, out-of-image neighbours absent):
 *
 *   ssor_bruteforce    closed form: breadth-first search over every flat zone,
 *                      then "no strictly greater / smaller neighbour" per zone.
 *                      Pinned against the 206 MATLAB imregionalmax/min golden
 *                      triples (oracles/oracle_v0_{1,2,3}*).
 *   ssor_global_uf     restatement of the reference's dense Global-UF route
 *                      (harness/primary/snapshot_gpu_benchmark_globaluf_ole_v0_1.cu:
 *                      105-141 flags, 185-190 init, 200-247 min-root CAS union of
 *                      the left / up / up-left / up-right equal edges, 249-259
 *                      full compression, 143-162 root OR, 164-179 materialise;
 *                      sequence 341-385).  OpenMP over row bands with atomic CAS.
 *                      This is also the reported CPU baseline.
 *   ssor_sparse_seam   restatement of the reference's tiled Sparse-Seam route
 *                      (harness/primary/snapshot_sparse_seam_cuda_v1_0.cu:
 *                      194-289 Komura tile labels, 291-475 classify / number /
 *                      emit, 162-192 boundary slots, 508-581 seam merge,
 *                      583-619 compress + reduce, 477-506 materialise).  It also
 *                      returns the per-tile seam-fragment counts, the one piece
 *                      of intermediate state that is deterministic.
 *
 * ssor_quantize restates quantize_grayscale_in_place
 * (harness/quantization_and_scaling/gpu_common_quant_v0_7.cu:479-495).
 *
 * ssor_seam_profile restates the seam-sparsity counters of
 * harness/seam_sparsity_profiler/gpu_seam_sparsity_profiler_v0_1.cu
 * (139-171 tile-local unions, 173-183 compression, 185-249 equal forward
 * edges and seam marks, 251-294 fragment and pixel counts).
 *
 * Every route is pinned by the golden fixtures and by agreement with the
 * brute-force closed form on random images (tests/test_oracle.py).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define SSOR_OK 0
#define SSOR_ERR_ARG -1
#define SSOR_ERR_NOMEM -2
#define SSOR_ERR_INVARIANT -3

/* ------------------------------------------------------------------------ */
/* Closed form: BFS over flat zones.                                         */
/* ------------------------------------------------------------------------ */

int ssor_bruteforce(const uint8_t* img, int w, int h, uint8_t* mx, uint8_t* mn)
{
    if (w <= 0 || h <= 0 || !img || !mx || !mn) return SSOR_ERR_ARG;
    const size_t n = (size_t)w * (size_t)h;
    int32_t* queue = (int32_t*)malloc(n * sizeof(int32_t));
    uint8_t* seen = (uint8_t*)calloc(n, 1);
    if (!queue || !seen) { free(queue); free(seen); return SSOR_ERR_NOMEM; }
    for (size_t start = 0; start < n; ++start) {
        if (seen[start]) continue;
        const uint8_t v = img[start];
        size_t head = 0, tail = 0;
        queue[tail++] = (int32_t)start;
        seen[start] = 1;
        int greater = 0, smaller = 0;
        while (head < tail) {
            const int32_t p = queue[head++];
            const int x = p % w, y = p / w;
            for (int dy = -1; dy <= 1; ++dy) {
                const int ny = y + dy;
                if (ny < 0 || ny >= h) continue;
                for (int dx = -1; dx <= 1; ++dx) {
                    if (!dx && !dy) continue;
                    const int nx = x + dx;
                    if (nx < 0 || nx >= w) continue;
                    const size_t q = (size_t)ny * w + nx;
                    const uint8_t u = img[q];
                    if (u > v) greater = 1;
                    else if (u < v) smaller = 1;
                    else if (!seen[q]) { seen[q] = 1; queue[tail++] = (int32_t)q; }
                }
            }
        }
        const uint8_t omax = greater ? 0 : 255, omin = smaller ? 0 : 255;
        for (size_t i = 0; i < tail; ++i) { mx[queue[i]] = omax; mn[queue[i]] = omin; }
    }
    free(queue);
    free(seen);
    return SSOR_OK;
}

/* ------------------------------------------------------------------------ */
/* Global-UF restatement (dense flat-zone union-find over all pixels).       */
/* ------------------------------------------------------------------------ */

static inline int32_t guf_find(int32_t* parent, int32_t node)
{
    int32_t root = __atomic_load_n(parent + node, __ATOMIC_RELAXED);
    for (;;) {
        const int32_t next = __atomic_load_n(parent + root, __ATOMIC_RELAXED);
        if (next == root) return root;
        root = next;
    }
}

static inline void guf_union(int32_t* parent, int32_t a, int32_t b)
{
    for (;;) {
        const int32_t ra = guf_find(parent, a), rb = guf_find(parent, b);
        if (ra == rb) return;
        const int32_t high = ra > rb ? ra : rb, low = ra > rb ? rb : ra;
        int32_t expected = high;
        if (__atomic_compare_exchange_n(parent + high, &expected, low, 0,
                                        __ATOMIC_ACQ_REL, __ATOMIC_RELAXED))
            return;
    }
}

/* work: optional caller buffer of at least 6*w*h bytes (int32 parent + 2 flag
 * bytes per pixel); NULL allocates.  threads <= 0 means "all". */
int ssor_global_uf(const uint8_t* img, int w, int h, uint8_t* mx, uint8_t* mn,
                   int threads, void* work)
{
    if (w <= 0 || h <= 0 || !img || !mx || !mn) return SSOR_ERR_ARG;
    if ((int64_t)w * (int64_t)h > INT32_MAX) return SSOR_ERR_ARG;
    const int32_t n = w * h;
    void* owned = NULL;
    if (!work) {
        owned = malloc((size_t)n * 6);
        if (!owned) return SSOR_ERR_NOMEM;
        work = owned;
    }
    int32_t* parent = (int32_t*)work;
    uint8_t* flags = (uint8_t*)(parent + n);   /* bit0 has_greater, bit1 has_smaller */
    uint8_t* root_flags = flags + n;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#else
    (void)threads;
#endif

    /* clear root flags + init parent (…:336-339, 185-190) */
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int32_t i = 0; i < n; ++i) { parent[i] = i; root_flags[i] = 0; }

    /* per-pixel flags over the valid 8-neighbourhood (…:105-141) */
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            const uint8_t v = img[(size_t)y * w + x];
            uint8_t f = 0;
            for (int dy = -1; dy <= 1; ++dy) {
                const int ny = y + dy;
                if (ny < 0 || ny >= h) continue;
                for (int dx = -1; dx <= 1; ++dx) {
                    if (!dx && !dy) continue;
                    const int nx = x + dx;
                    if (nx < 0 || nx >= w) continue;
                    const uint8_t u = img[(size_t)ny * w + nx];
                    f |= (u > v) ? 1 : 0;
                    f |= (u < v) ? 2 : 0;
                }
            }
            flags[(size_t)y * w + x] = f;
        }
    }

    /* union of equal left / up / up-left / up-right edges (…:220-247) */
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            const int32_t i = y * w + x;
            const uint8_t v = img[i];
            if (x > 0 && img[i - 1] == v) guf_union(parent, i, i - 1);
            if (y > 0) {
                if (img[i - w] == v) guf_union(parent, i, i - w);
                if (x > 0 && img[i - w - 1] == v) guf_union(parent, i, i - w - 1);
                if (x + 1 < w && img[i - w + 1] == v) guf_union(parent, i, i - w + 1);
            }
        }
    }

    /* full compression (…:249-259) */
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int32_t i = 0; i < n; ++i) parent[i] = guf_find(parent, i);

    /* OR flags into roots (…:143-162) */
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int32_t i = 0; i < n; ++i) {
        const uint8_t f = flags[i];
        if (f) __atomic_fetch_or(root_flags + parent[i], f, __ATOMIC_RELAXED);
    }

    /* materialise (…:164-179) */
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int32_t i = 0; i < n; ++i) {
        const uint8_t f = root_flags[parent[i]];
        mx[i] = (f & 1) ? 0 : 255;
        mn[i] = (f & 2) ? 0 : 255;
    }
    free(owned);
    return SSOR_OK;
}

/* ------------------------------------------------------------------------ */
/* Sparse-Seam restatement (tile-local flat zones + seam fragment DSU).      */
/* ------------------------------------------------------------------------ */

#define SS_HAS_GREATER 1u
#define SS_HAS_SMALLER 2u
#define SS_TOUCHES_SEAM 4u

static int32_t ss_find(const int32_t* parent, int32_t node)
{
    int32_t root = parent[node];
    while (root != parent[root]) root = parent[root];
    return root;
}

static void ss_union(int32_t* parent, int32_t a, int32_t b)
{
    const int32_t ra = ss_find(parent, a), rb = ss_find(parent, b);
    if (ra == rb) return;
    if (ra > rb) parent[ra] = rb; else parent[rb] = ra;
}

/* Perimeter slot of pixel (x, y) inside its tile (…:162-192): top row, then
 * bottom row, then the left and right columns without their first row. */
static int64_t ss_boundary_slot(int x, int y, int w, int h, int tiles_x, int T)
{
    const int tx = x / T, ty = y / T;
    const int sx = tx * T, sy = ty * T;
    const int tw = (w - sx) < T ? (w - sx) : T;
    const int th = (h - sy) < T ? (h - sy) : T;
    const int lx = x - sx, ly = y - sy;
    const int64_t base = ((int64_t)ty * tiles_x + tx) * (4 * T);
    if (ly == 0) return base + lx;
    if (ly == th - 1) return base + T + lx;
    if (lx == 0) return base + 2 * T + (ly - 1);
    if (lx == tw - 1) return base + 3 * T + (ly - 1);
    return -1;
}

static void ss_merge_pair(const uint8_t* img, const int64_t* b2f, int32_t* parent,
                          int x, int y, int nx, int ny, int w, int h, int tiles_x, int T)
{
    if (nx < 0 || nx >= w || ny < 0 || ny >= h) return;
    if (img[(size_t)y * w + x] != img[(size_t)ny * w + nx]) return;
    const int64_t s1 = ss_boundary_slot(x, y, w, h, tiles_x, T);
    const int64_t s2 = ss_boundary_slot(nx, ny, w, h, tiles_x, T);
    if (s1 < 0 || s2 < 0) return;
    const int64_t f1 = b2f[s1], f2 = b2f[s2];
    if (f1 >= 0 && f2 >= 0) ss_union(parent, (int32_t)f1, (int32_t)f2);
}

/* tile in [1, 64]; tile_counts (optional) receives tiles_x*tiles_y entries. */
int ssor_sparse_seam(const uint8_t* img, int w, int h, int T, uint8_t* mx, uint8_t* mn,
                     uint16_t* tile_counts)
{
    if (w <= 0 || h <= 0 || !img || !mx || !mn || T < 1 || T > 64) return SSOR_ERR_ARG;
    const int tiles_x = (w + T - 1) / T, tiles_y = (h + T - 1) / T;
    const int64_t tiles = (int64_t)tiles_x * tiles_y;
    const int64_t slots = tiles * 4 * T;
    if (slots > INT32_MAX) return SSOR_ERR_ARG;
    const size_t n = (size_t)w * h;
    uint8_t* record = (uint8_t*)malloc(n);
    int64_t* b2f = (int64_t*)malloc((size_t)slots * sizeof(int64_t));
    int32_t* fparent = (int32_t*)malloc((size_t)slots * sizeof(int32_t));
    uint32_t* flocal = (uint32_t*)calloc((size_t)slots, sizeof(uint32_t));
    uint32_t* froot = (uint32_t*)calloc((size_t)slots, sizeof(uint32_t));
    uint16_t* counts = (uint16_t*)calloc((size_t)tiles, sizeof(uint16_t));
    int32_t* labels = (int32_t*)malloc(sizeof(int32_t) * T * T);
    uint32_t* rbits = (uint32_t*)malloc(sizeof(uint32_t) * T * T);
    uint8_t* timg = (uint8_t*)malloc((size_t)T * T);
    int rc = SSOR_OK;
    if (!record || !b2f || !fparent || !flocal || !froot || !counts || !labels || !rbits || !timg) {
        rc = SSOR_ERR_NOMEM;
        goto done;
    }
    for (int64_t s = 0; s < slots; ++s) { b2f[s] = -1; fparent[s] = (int32_t)s; }

    for (int ty = 0; ty < tiles_y; ++ty) {
        for (int tx = 0; tx < tiles_x; ++tx) {
            const int sx = tx * T, sy = ty * T;
            const int tw = (w - sx) < T ? (w - sx) : T;
            const int th = (h - sy) < T ? (h - sy) : T;
            const int64_t tile_id = (int64_t)ty * tiles_x + tx;
            const int64_t fbase = tile_id * 4 * T;
            /* load (…:205-217) */
            for (int l = 0; l < T * T; ++l) {
                const int lx = l % T, ly = l / T;
                if (lx < tw && ly < th) { timg[l] = img[(size_t)(sy + ly) * w + sx + lx]; labels[l] = l; }
                else { timg[l] = 0; labels[l] = -1; }
                rbits[l] = 0;
            }
            /* Komura init: first equal predecessor in the order up, up-left,
             * up-right, left (…:225-246) */
            for (int l = 0; l < T * T; ++l) {
                const int lx = l % T, ly = l / T;
                if (lx >= tw || ly >= th) continue;
                const uint8_t v = timg[l];
                int p = l;
                if (ly > 0 && timg[l - T] == v) p = l - T;
                else if (ly > 0 && lx > 0 && timg[l - T - 1] == v) p = l - T - 1;
                else if (ly > 0 && lx + 1 < tw && timg[l - T + 1] == v) p = l - T + 1;
                else if (lx > 0 && timg[l - 1] == v) p = l - 1;
                labels[l] = p;
            }
            for (int l = 0; l < T * T; ++l)
                if (l % T < tw && l / T < th) labels[l] = ss_find(labels, l);
            /* union of the left and up-right edges (…:260-276) */
            for (int l = 0; l < T * T; ++l) {
                const int lx = l % T, ly = l / T;
                if (lx >= tw || ly >= th) continue;
                const uint8_t v = timg[l];
                if (lx > 0 && timg[l - 1] == v) ss_union(labels, l, l - 1);
                if (ly > 0 && lx + 1 < tw && timg[l - T + 1] == v) ss_union(labels, l, l - T + 1);
            }
            for (int l = 0; l < T * T; ++l)
                if (l % T < tw && l / T < th) labels[l] = ss_find(labels, l);
            /* classify against the global 8-neighbourhood (…:341-389) */
            for (int l = 0; l < T * T; ++l) {
                const int lx = l % T, ly = l / T;
                if (lx >= tw || ly >= th) continue;
                const int x = sx + lx, y = sy + ly;
                const uint8_t v = timg[l];
                uint32_t bits = 0;
                for (int dy = -1; dy <= 1; ++dy) {
                    const int ny = y + dy;
                    if (ny < 0 || ny >= h) continue;
                    for (int dx = -1; dx <= 1; ++dx) {
                        if (!dx && !dy) continue;
                        const int nx = x + dx;
                        if (nx < 0 || nx >= w) continue;
                        const uint8_t u = img[(size_t)ny * w + nx];
                        bits |= u > v ? SS_HAS_GREATER : 0u;
                        bits |= u < v ? SS_HAS_SMALLER : 0u;
                        const int crosses = nx < sx || nx >= sx + tw || ny < sy || ny >= sy + th;
                        if (u == v && crosses) bits |= SS_TOUCHES_SEAM;
                    }
                }
                rbits[labels[l]] |= bits;
            }
            /* number seam-touching roots (raster order here; the reference
             * uses shared atomicAdd order, …:393-411) */
            int count = 0;
            for (int l = 0; l < T * T; ++l) {
                const int lx = l % T, ly = l / T;
                if (lx >= tw || ly >= th || labels[l] != l || !(rbits[l] & SS_TOUCHES_SEAM)) continue;
                if (count >= 255) { rc = SSOR_ERR_INVARIANT; goto done; }
                rbits[l] = (rbits[l] & 7u) | ((uint32_t)(count + 1) << 3);
                ++count;
            }
            /* write masks of tile-contained zones, records and fragments (…:414-450) */
            for (int l = 0; l < T * T; ++l) {
                const int lx = l % T, ly = l / T;
                if (lx >= tw || ly >= th) continue;
                const int root = labels[l];
                const uint32_t packed = rbits[root];
                const size_t idx = (size_t)(sy + ly) * w + sx + lx;
                if (!(packed & SS_TOUCHES_SEAM)) {
                    record[idx] = 255;
                    mx[idx] = (packed & SS_HAS_GREATER) ? 0 : 255;
                    mn[idx] = (packed & SS_HAS_SMALLER) ? 0 : 255;
                    continue;
                }
                const int rec = (int)(packed >> 3) - 1;
                if (rec < 0 || rec >= 4 * T) { rc = SSOR_ERR_INVARIANT; goto done; }
                record[idx] = (uint8_t)rec;
                if (l == root) {
                    fparent[fbase + rec] = (int32_t)(fbase + rec);
                    flocal[fbase + rec] = packed & 3u;
                    froot[fbase + rec] = 0;
                }
            }
            counts[tile_id] = (uint16_t)count;
            /* boundary map (…:452-474) */
            for (int l = 0; l < T * T; ++l) {
                const int lx = l % T, ly = l / T;
                if (lx >= tw || ly >= th) continue;
                const uint32_t packed = rbits[labels[l]];
                if (!(packed & SS_TOUCHES_SEAM)) continue;
                const int64_t slot = ss_boundary_slot(sx + lx, sy + ly, w, h, tiles_x, T);
                if (slot >= 0) b2f[slot] = fbase + (int)(packed >> 3) - 1;
            }
        }
    }

    /* merge equal fragments across the vertical, then horizontal seams (…:538-581) */
    for (int seam = 1; seam < tiles_x; ++seam) {
        const int x = seam * T - 1;
        for (int y = 0; y < h; ++y) {
            ss_merge_pair(img, b2f, fparent, x, y, x + 1, y - 1, w, h, tiles_x, T);
            ss_merge_pair(img, b2f, fparent, x, y, x + 1, y, w, h, tiles_x, T);
            ss_merge_pair(img, b2f, fparent, x, y, x + 1, y + 1, w, h, tiles_x, T);
        }
    }
    for (int seam = 1; seam < tiles_y; ++seam) {
        const int y = seam * T - 1;
        for (int x = 0; x < w; ++x) {
            ss_merge_pair(img, b2f, fparent, x, y, x - 1, y + 1, w, h, tiles_x, T);
            ss_merge_pair(img, b2f, fparent, x, y, x, y + 1, w, h, tiles_x, T);
            ss_merge_pair(img, b2f, fparent, x, y, x + 1, y + 1, w, h, tiles_x, T);
        }
    }
    /* compress and reduce over the active slots (…:583-619) */
    for (int64_t s = 0; s < slots; ++s) {
        if (s % (4 * T) < counts[s / (4 * T)]) fparent[s] = ss_find(fparent, (int32_t)s);
    }
    for (int64_t s = 0; s < slots; ++s) {
        if (s % (4 * T) < counts[s / (4 * T)]) froot[fparent[s]] |= flocal[s] & 3u;
    }
    /* materialise pixels whose record is a fragment (…:477-506) */
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            const size_t idx = (size_t)y * w + x;
            const uint8_t rec = record[idx];
            if (rec == 255) continue;
            const int64_t tile_id = (int64_t)(y / T) * tiles_x + (x / T);
            const uint32_t bits = froot[fparent[tile_id * 4 * T + rec]];
            mx[idx] = (bits & SS_HAS_GREATER) ? 0 : 255;
            mn[idx] = (bits & SS_HAS_SMALLER) ? 0 : 255;
        }
    }
    if (tile_counts) memcpy(tile_counts, counts, (size_t)tiles * sizeof(uint16_t));
done:
    free(record); free(b2f); free(fparent); free(flocal); free(froot); free(counts);
    free(labels); free(rbits); free(timg);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Reference quantiser (gpu_common_quant_v0_7.cu:479-495).                  */
/* ------------------------------------------------------------------------ */

void ssor_quantize(uint8_t* img, size_t n, int levels)
{
    if (levels == 256 || levels < 2) return;
    for (size_t i = 0; i < n; ++i) {
        const int bin = ((int)img[i] * levels) >> 8;
        img[i] = (uint8_t)((bin * 255 + (levels - 1) / 2) / (levels - 1));
    }
}

/* ------------------------------------------------------------------------ */
/* Seam-sparsity counters (gpu_seam_sparsity_profiler_v0_1.cu:133-294).     */
/* ------------------------------------------------------------------------ */

/* out[0] tile-local zones ("fragments", singletons included), out[1] zones
 * with an equal 8-neighbour in another tile, out[2] pixels of those zones,
 * out[3] equal 8-neighbour pairs (each once: right, down-left, down,
 * down-right, both pixels in the image), out[4] those pairs whose pixels lie
 * in different tiles. */
int ssor_seam_profile(const uint8_t* img, int w, int h, int T, uint64_t* out)
{
    if (!img || !out || w <= 0 || h <= 0 || T <= 0) return SSOR_ERR_ARG;
    const int64_t n = (int64_t)w * h;
    int32_t* parent = (int32_t*)malloc((size_t)n * sizeof(int32_t));
    uint8_t* seam = (uint8_t*)calloc((size_t)n, 1);
    if (!parent || !seam) {
        free(parent);
        free(seam);
        return SSOR_ERR_NOMEM;
    }
    for (int64_t i = 0; i < n; ++i) parent[i] = (int32_t)i;
    /* 139-171: unions of the equal left / up / up-left / up-right pairs that
     * stay inside one tile */
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            const int64_t i = (int64_t)y * w + x;
            const uint8_t v = img[i];
            const int tx = x / T, ty = y / T;
            if (x > 0 && (x - 1) / T == tx && img[i - 1] == v) ss_union(parent, (int32_t)i, (int32_t)(i - 1));
            if (y > 0 && (y - 1) / T == ty) {
                if (img[i - w] == v) ss_union(parent, (int32_t)i, (int32_t)(i - w));
                if (x > 0 && (x - 1) / T == tx && img[i - w - 1] == v)
                    ss_union(parent, (int32_t)i, (int32_t)(i - w - 1));
                if (x + 1 < w && (x + 1) / T == tx && img[i - w + 1] == v)
                    ss_union(parent, (int32_t)i, (int32_t)(i - w + 1));
            }
        }
    }
    for (int64_t i = 0; i < n; ++i) parent[i] = ss_find(parent, (int32_t)i);
    /* 185-249: equal forward pairs; a pair across tiles marks both zones */
    uint64_t equal = 0, seam_edges = 0;
    static const int dx[4] = {1, -1, 0, 1}, dy[4] = {0, 1, 1, 1};
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            const int64_t i = (int64_t)y * w + x;
            for (int d = 0; d < 4; ++d) {
                const int nx = x + dx[d], ny = y + dy[d];
                if (nx < 0 || nx >= w || ny >= h) continue;
                const int64_t j = (int64_t)ny * w + nx;
                if (img[j] != img[i]) continue;
                ++equal;
                if (nx / T != x / T || ny / T != y / T) {
                    ++seam_edges;
                    seam[parent[i]] = 1;
                    seam[parent[j]] = 1;
                }
            }
        }
    }
    /* 251-294: zones are the roots; seam zones and their pixels */
    uint64_t zones = 0, seam_zones = 0, seam_px = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (parent[i] == (int32_t)i) {
            ++zones;
            seam_zones += seam[i];
        }
        seam_px += seam[parent[i]];
    }
    out[0] = zones;
    out[1] = seam_zones;
    out[2] = seam_px;
    out[3] = equal;
    out[4] = seam_edges;
    free(parent);
    free(seam);
    return SSOR_OK;
}
