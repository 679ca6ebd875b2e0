/*
 * ss_extrema.h -- C ABI of the MI355X (gfx950) exact dual regional-extrema
 * operator: an 8-bit grayscale image in, its exact 8-connected regional-maximum
 * and regional-minimum masks (uint8 0/255) out.
 *
 * Semantics (identical to the reference's Sparse-Seam and Global-UF routes):
 *   - flat zones are 8-connected sets of equal-valued pixels;
 *   - a zone is a regional maximum iff no pixel of it has a strictly greater
 *     8-neighbour, a regional minimum iff none has a strictly smaller one;
 *   - neighbours outside the image are absent (never compared);
 *   - masks are 255 on extremal zones and 0 elsewhere.
 *
 * The reference has no ABI: its operator is a C++ source-level API pulled into
 * harness executables by #include (harness/primary/gpu_common_v0_6.cu:11-55).
 * Each entry point below names the reference function it replaces.
 *
 * Conventions: every device pointer is caller-owned HIP device memory; no entry
 * point allocates device memory or synchronises unless its comment says so;
 * no entry point throws.  Work is enqueued on the caller's stream and is
 * capturable into a hipGraph.  Calls that share a workspace must be ordered
 * on one stream; distinct workspaces may run concurrently on distinct streams.
 */
#ifndef SS_EXTREMA_H
#define SS_EXTREMA_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ss_status {
    SS_OK = 0,
    SS_ERR_ARG = 1,        /* bad dimensions, pitch, config or null pointer  */
    SS_ERR_WORKSPACE = 2,  /* workspace too small or not 256-byte aligned    */
    SS_ERR_HIP = 3,        /* a HIP runtime call failed                      */
    SS_ERR_INVARIANT = 4,  /* a device invariant failed (see ss_stats.invariant) */
    SS_ERR_CAPACITY = 5,   /* batch / serving capacity exceeded or admission refused */
    SS_ERR_NO_DEVICE = 6   /* no gfx950 device / kernels not loadable        */
} ss_status;

/* Tile side T in {16, 32, 64} and threads per tile in {64, 128, 256, 512}
 * (threads <= T*T).  Replaces the reference's compile-time SS_TILE /
 * SS_CLASSIFY_THREADS (snapshot_sparse_seam_cuda_v1_0.cu:34-46,
 * generated_sparse_seam_kernel_v0_7.cu:782-785). */
typedef struct ss_config {
    int tile;
    int threads;
} ss_config;

/* Per-run counters, read back by ss_read_stats (synchronous). */
typedef struct ss_stats {
    uint32_t invariant;      /* 0 ok; 1 = more than 4T-4 seam fragments in a tile,
                                2 = bad fragment record (reference flags 1 and 2,
                                snapshot_sparse_seam_cuda_v1_0.cu:404-406,432-434) */
    uint32_t dirty_tiles;    /* tiles whose seam zones changed after the global merge */
    uint64_t fragments;      /* seam fragments in the dense array (the reference's
                                active_fragments, snapshot_sparse_seam_cuda_v1_0.cu:852-863) */
    uint64_t dsu_retries;    /* failed CAS attempts in the seam union-find */
    uint64_t tiles;          /* tiles processed */
} ss_stats;

/* Human-readable name of a status code (static storage). */
const char* ss_status_str(ss_status status);

/* Bytes of device workspace needed for one w x h image at cfg (NULL = the
 * default / autotuned config).  Replaces allocate_buffers
 * (snapshot_sparse_seam_cuda_v1_0.cu:710-737): the reference allocates image,
 * masks, a per-pixel record and 16 B per tile slot; here the caller owns image
 * and masks and the workspace holds only per-tile and per-fragment state. */
ss_status ss_workspace_size(int width, int height, const ss_config* cfg, size_t* bytes);

/* Exact dual regional extrema of one image.  Replaces run_sparse_seam /
 * run_sparse_resident (snapshot_sparse_seam_cuda_v1_0.cu:753-878,
 * gpu_common_v0_6.cu:137-206) with the image already resident in HBM.
 *   d_img    : width x height uint8 pixels, row r at d_img + r*pitch (pitch >= width;
 *              any pitch and base address, no alignment required)
 *   d_max/min: outputs, row r at d + r*out_pitch (out_pitch >= width); both masks
 *              are fully written (every pixel 0 or 255)
 *   d_ws     : workspace of ws_bytes >= ss_workspace_size(...), 256-byte aligned
 *   cfg      : NULL = default / autotuned
 * Asynchronous on `stream`; invariant failures are reported by ss_read_stats. */
ss_status ss_extrema_u8(const uint8_t* d_img, size_t pitch, int width, int height,
                        uint8_t* d_max, uint8_t* d_min, size_t out_pitch,
                        void* d_ws, size_t ws_bytes, const ss_config* cfg,
                        hipStream_t stream);

/* Allocation-free form of ss_extrema_u8.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under hipGraph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * workspace is a library-owned device array of SS_NOALLOC_SCRATCH_BYTES per GPU
 * (held until the process ends); images whose workspace
 * (ss_workspace_size) exceeds it are refused with SS_ERR_CAPACITY (it holds
 * 4096 x 4096 at every tile size, 6144 x 6144 at T = 64).
 * The call allocates nothing.  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (several streams or threads) race and must be
 * ordered by the caller.  d_invariant (optional, device memory) receives the
 * run's invariant code (0 = ok; see ss_stats.invariant), since the workspace
 * is not the caller's to pass to ss_read_stats.  Asynchronous on `stream`. */
#define SS_NOALLOC_SCRATCH_BYTES (64u << 20)
ss_status ss_extrema_u8_noalloc(const uint8_t* d_img, size_t pitch, int width, int height,
                                uint8_t* d_max, uint8_t* d_min, size_t out_pitch,
                                const ss_config* cfg, uint32_t* d_invariant, hipStream_t stream);

/* Synchronises `stream` and reads the counters of the last run that used d_ws.
 * Returns SS_ERR_INVARIANT when the device invariant word is non-zero (the
 * reference throws "Sparse-seam device invariant failed",
 * snapshot_sparse_seam_cuda_v1_0.cu:840-850). */
ss_status ss_read_stats(const void* d_ws, size_t ws_bytes, ss_stats* out, hipStream_t stream);

/* Co-reference route: the dense Global-UF operator (one union-find node per
 * pixel, equal left / up-left / up / up-right unions with min-root CAS, full
 * compression, greater / smaller flags ORed into the roots).  Replaces
 * run_global_uf (snapshot_gpu_benchmark_globaluf_ole_v0_1.cu:341-385, kernels
 * :105-259), which the reference's harness runs beside Sparse-Seam to
 * cross-check every timed image (gpu_common_v0_6.cu:659-669).  Same masks as
 * ss_extrema_u8; a validator, not tuned for speed.  Images below 2^30 pixels;
 * the workspace (4 bytes per pixel) must be 256-byte aligned.  Asynchronous. */
ss_status ss_global_uf_workspace_size(int width, int height, size_t* bytes);
ss_status ss_global_uf_u8(const uint8_t* d_img, size_t pitch, int width, int height,
                          uint8_t* d_max, uint8_t* d_min, size_t out_pitch,
                          void* d_ws, size_t ws_bytes, hipStream_t stream);

/* Per-tile seam-fragment counts of the last run (tiles_y x tiles_x uint32),
 * copied to host memory `out` (synchronous).  The reference reads the same
 * counts back to sum active_fragments (snapshot_sparse_seam_cuda_v1_0.cu:852-863). */
ss_status ss_read_tile_counts(const void* d_ws, size_t ws_bytes, uint32_t* out, size_t count,
                              hipStream_t stream);

/* Seam-sparsity counters of one image at tile size cfg->tile (16, 32, 64 or
 * 128 -- the reference profiler's default tiles are 32, 64 and 128; threads
 * are fixed at 256; NULL = the process default tile), the quantities the
 * reference's seam profiler reports per image
 * (harness/seam_sparsity_profiler/gpu_seam_sparsity_profiler_v0_1.cu:133-294,
 * CSV columns :391-396; fractions and per-megapixel rates are ratios of these).
 * A diagnostic, not part of the extrema path: it stages and labels every tile
 * but writes no masks.  Needs only a 256-byte aligned workspace of at least 256
 * bytes (its counters; any ss_workspace_size workspace will do); synchronous. */
typedef struct ss_seam_profile {
    uint64_t fragments;       /* tile-local zones, singletons included */
    uint64_t seam_fragments;  /* zones with an equal 8-neighbour in another tile */
    uint64_t seam_pixels;     /* pixels of those zones */
    uint64_t equal_edges;     /* equal 8-neighbour pairs inside the image, each once */
    uint64_t seam_edges;      /* those pairs whose two pixels lie in different tiles */
} ss_seam_profile;

ss_status ss_seam_profile_u8(const uint8_t* d_img, size_t pitch, int width, int height,
                             const ss_config* cfg, void* d_ws, size_t ws_bytes, hipStream_t stream,
                             ss_seam_profile* out);

/* ------------------------------------------------------------------------ */
/* Batched form: many images of uneven sizes in one pipeline.               */
/* No reference counterpart: the reference loops over images, one          */
/* allocate / run / free per image (gpu_common_v0_6.cu:584-745).            */
/* ------------------------------------------------------------------------ */

typedef struct ss_image_desc {
    const uint8_t* img;   /* device pointer */
    uint8_t* max;         /* device pointer */
    uint8_t* min;         /* device pointer */
    uint64_t pitch;       /* input row pitch in bytes  (>= width) */
    uint64_t out_pitch;   /* output row pitch in bytes (>= width) */
    int32_t width;
    int32_t height;
} ss_image_desc;

/* Workspace for a batch: n images described on the HOST (only their sizes are
 * read). */
ss_status ss_batch_workspace_size(const ss_image_desc* images, int n, const ss_config* cfg,
                                  size_t* bytes);

/* Runs the whole batch: one tile-stage launch over the tiles of all images
 * (one workgroup per tile of the combined tile space), then the seam stages
 * over all images at once.  `images` is
 * a HOST array; its descriptors are staged into the workspace with one
 * asynchronous copy.  Not capturable (the copy reads pageable memory); use
 * ss_batch_plan_* for graph replay. */
ss_status ss_extrema_batch_u8(const ss_image_desc* images, int n, void* d_ws, size_t ws_bytes,
                              const ss_config* cfg, hipStream_t stream);

/* Graph-replayable batch plan: descriptors are staged once, the pipeline is
 * captured into a hipGraph, and each ss_batch_plan_run replays it. */
typedef struct ss_batch_plan ss_batch_plan;
ss_status ss_batch_plan_create(const ss_image_desc* images, int n, void* d_ws, size_t ws_bytes,
                               const ss_config* cfg, hipStream_t stream, ss_batch_plan** plan);
ss_status ss_batch_plan_run(ss_batch_plan* plan, hipStream_t stream);
ss_status ss_batch_plan_destroy(ss_batch_plan* plan);

/* ------------------------------------------------------------------------ */
/* Autotuner: picks (tile, threads) by timing the candidates on the device. */
/* Replaces the reference's offline 9-build sweep and its selection metric, */
/* sum(pixels) / sum(per-image median) (run_config_study_v0_8.ps1:85-109).  */
/* ------------------------------------------------------------------------ */

/* Times every candidate (tile, threads) on the given resident images
 * (reps per candidate, median per image), writes the best to *best and makes it
 * the process-wide default used when cfg == NULL.  Allocates a scratch
 * workspace with hipMalloc and synchronises.  `throughput` (optional) receives
 * the winner's pixels/s. */
ss_status ss_autotune(const ss_image_desc* images, int n, int reps, hipStream_t stream,
                      ss_config* best, double* throughput);

/* Same, scoring each candidate by the whole batch: total pixels / median time
 * of ss_extrema_batch_u8 over all n images (the launch a batch or a plan
 * will use; per-image launch overhead does not enter).  No counterpart in the
 * reference, whose study times single images only. */
ss_status ss_autotune_batch(const ss_image_desc* images, int n, int reps, hipStream_t stream,
                            ss_config* best, double* throughput);

/* Sets / reads the process-wide default config (used when cfg == NULL). */
ss_status ss_set_default_config(const ss_config* cfg);
ss_status ss_get_default_config(ss_config* cfg);

/* ------------------------------------------------------------------------ */
/* Serving: a continuous-batching request loop.  No reference counterpart   */
/* (the reference processes one image at a time, gpu_common_v0_6.cu:584).   */
/*                                                                          */
/* Requests carry host images.  A scheduler thread admits pending requests  */
/* in arrival order while their device footprint (image + two masks + their */
/* share of the tile / seam-DSU workspace) fits under the HBM cap, folds    */
/* every admitted request into the next batched launch, and keeps up to     */
/* `streams` launches in flight, so one batch's copies overlap another's    */
/* kernels.  Requests that arrive while batches run join the next launch.   */
/* The device arena is allocated once by ss_server_create.                  */
/* ------------------------------------------------------------------------ */

typedef struct ss_server ss_server;

typedef struct ss_server_config {
    uint64_t hbm_cap_bytes;  /* device arena = admission cap (bytes)            */
    int max_batch;           /* requests per launch, 0 = no limit               */
    int streams;             /* launches in flight, >= 1 (default 2)            */
    int tile;                /* 0 = default config                               */
    int threads;
    uint64_t max_batch_pixels; /* stop admitting into a launch past this many
                                  pixels (0 = no limit); smaller launches let
                                  one launch's copies overlap another's kernels */
} ss_server_config;

typedef struct ss_request {
    const uint8_t* host_img; /* width x height, row-major, pitch = width        */
    uint8_t* host_max;       /* outputs, same layout                            */
    uint8_t* host_min;
    int32_t width;
    int32_t height;
    uint64_t tag;            /* caller's value, returned in the completion      */
} ss_request;

typedef struct ss_completion {
    uint64_t id;
    uint64_t tag;
    int32_t status;          /* ss_status of the request                        */
    int32_t batch_size;      /* requests in the launch that carried it          */
    double submit_s;         /* host steady-clock seconds                       */
    double launch_s;
    double done_s;
} ss_completion;

typedef struct ss_server_stats {
    uint64_t submitted, completed, rejected, launches;
    uint64_t pixels;
    uint64_t peak_inflight_bytes;
} ss_server_stats;

ss_status ss_server_create(const ss_server_config* cfg, ss_server** server);
/* Non-blocking.  SS_ERR_CAPACITY if the request alone exceeds the cap. */
ss_status ss_server_submit(ss_server* s, const ss_request* req, uint64_t* id);
/* Blocks until request `id` completes and returns its record. */
ss_status ss_server_wait(ss_server* s, uint64_t id, ss_completion* out);
/* Blocks until every submitted request has completed. */
ss_status ss_server_drain(ss_server* s);
ss_status ss_server_get_stats(ss_server* s, ss_server_stats* out);
ss_status ss_server_destroy(ss_server* s);

/* Open-loop load generator: submits reqs[i] at Poisson arrival times of rate
 * `qps` (seeded), drains, and writes one completion per request into out[i]
 * (host clock, so latency = done_s - submit_s). */
ss_status ss_server_run_open_loop(ss_server* s, const ss_request* reqs, int n, double qps,
                                  uint64_t seed, ss_completion* out);

/* ------------------------------------------------------------------------ */
/* Measurement aid (no reference counterpart; the reference times the whole */
/* kernel sequence with one cudaEvent pair, gpu_common_v0_6.cu:338-347).    */
/* ------------------------------------------------------------------------ */

/* Records tile_begin / tile_end around the tile-stage kernel of every pipeline
 * enqueued afterwards by the calling thread, on that pipeline's stream, so the
 * dominant kernel can be timed in place.  Pass two nulls to turn it off. */
ss_status ss_profile_events(hipEvent_t tile_begin, hipEvent_t tile_end);

#ifdef __cplusplus
}
#endif

#endif /* SS_EXTREMA_H */
