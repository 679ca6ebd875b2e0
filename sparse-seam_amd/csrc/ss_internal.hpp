// ss_internal.hpp -- host-side declarations shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cstddef>

#include "ss_device.hpp"
#include "ss_extrema.h"

namespace ssx {

// Byte offsets of the workspace regions (all 256-byte aligned).
struct Layout {
    int64_t ntiles = 0;
    int64_t nblk = 0;
    int64_t nseam = 0;
    int64_t ngroup = 0;
    int tile = 0;
    int slots = 0;
    uint32_t nblocks = 0;
    size_t off_counts = 0, off_base = 0, off_bstate = 0, off_sbits = 0,
           off_rec8 = 0, off_orows = 0, off_rec = 0, off_eq = 0, off_dparent = 0, off_dbits = 0,
           off_descs = 0, off_tile_img = 0, off_block_img = 0, off_seam_img = 0, off_group_img = 0;
    size_t bytes = 0;
};

inline size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

inline Layout make_layout(int64_t ntiles, int64_t nblk, int64_t nseam, int T, int ndesc, int64_t ngroup = 0) {
    Layout L;
    L.ntiles = ntiles;
    L.nblk = nblk;
    L.nseam = nseam;
    L.ngroup = ngroup;
    L.tile = T;
    L.slots = 4 * T - 4;
    L.nblocks = (uint32_t)((ntiles + SCAN_CHUNK - 1) / SCAN_CHUNK);
    const size_t nt = (size_t)ntiles, S = (size_t)L.slots;
    size_t o = sizeof(WsHeader);
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = align256(o + bytes);
        return at;
    };
    L.off_counts = take(nt * 4);
    L.off_base = take(nt * 4);
    L.off_bstate = take((size_t)L.nblocks * 8);
    L.off_sbits = take(nt * S);
    L.off_rec8 = take(nt * (size_t)T * (size_t)T);
    L.off_orows = take(nt * 8);
    L.off_rec = take(nt * 4 * (size_t)T);
    L.off_eq = take(nt * 2 * (size_t)T);
    L.off_dparent = take(nt * S * 4);
    L.off_dbits = take(nt * S * 4);
    L.off_descs = take((size_t)(ndesc > 1 ? ndesc : 0) * sizeof(ImgDev));
    L.off_tile_img = take(ndesc > 1 ? nt * 4 : 0);
    L.off_block_img = take(ndesc > 1 ? (size_t)nblk * 4 : 0);
    L.off_seam_img = take(ndesc > 1 ? (size_t)nseam * 4 : 0);
    L.off_group_img = take(ndesc > 1 ? (size_t)ngroup * 4 : 0);
    L.bytes = o;
    return L;
}

inline Ws make_ws(void* base, const Layout& L, int threads) {
    char* b = static_cast<char*>(base);
    Ws w;
    w.hdr = reinterpret_cast<WsHeader*>(b);
    w.counts = reinterpret_cast<uint32_t*>(b + L.off_counts);
    w.base = reinterpret_cast<uint32_t*>(b + L.off_base);
    w.bstate = reinterpret_cast<uint64_t*>(b + L.off_bstate);
    w.sbits = reinterpret_cast<uint8_t*>(b + L.off_sbits);
    w.flut = w.sbits;
    w.rec8 = reinterpret_cast<uint8_t*>(b + L.off_rec8);
    w.orows = reinterpret_cast<uint64_t*>(b + L.off_orows);
    w.rec = reinterpret_cast<uint8_t*>(b + L.off_rec);
    w.eq = reinterpret_cast<uint8_t*>(b + L.off_eq);
    w.dparent = reinterpret_cast<uint32_t*>(b + L.off_dparent);
    w.dbits = reinterpret_cast<uint32_t*>(b + L.off_dbits);
    w.tile_img = reinterpret_cast<uint32_t*>(b + L.off_tile_img);
    w.block_img = reinterpret_cast<uint32_t*>(b + L.off_block_img);
    w.seam_img = reinterpret_cast<uint32_t*>(b + L.off_seam_img);
    w.group_img = reinterpret_cast<uint32_t*>(b + L.off_group_img);
    w.ntiles = L.ntiles;
    w.nblk = L.nblk;
    w.nseam = L.nseam;
    w.ngroup = L.ngroup;
    w.nblocks = L.nblocks;
    w.tile = L.tile;
    w.threads = threads;
    w.slots = L.slots;
    return w;
}

inline bool config_valid(int T, int NT) {
    if (T != 16 && T != 32 && T != 64) return false;
    if (NT != 64 && NT != 128 && NT != 256 && NT != 512) return false;
    return NT <= T * T;
}

inline int64_t tiles_of(int w, int h, int T) {
    return (int64_t)((w + T - 1) / T) * (int64_t)((h + T - 1) / T);
}

// Descriptor of one image: tile grid, merge-block grid and seam-merge units.
inline ImgDev make_imgdev(const uint8_t* img, uint8_t* mx, uint8_t* mn, int64_t pitch, int64_t opitch, int w,
                          int h, int T, int64_t tile_begin, int64_t block_begin, int64_t seam_begin) {
    ImgDev d{};
    d.img = img;
    d.mx = mx;
    d.mn = mn;
    d.pitch = pitch;
    d.opitch = opitch;
    d.w = w;
    d.h = h;
    d.tiles_x = (w + T - 1) / T;
    d.tiles_y = (h + T - 1) / T;
    d.tile_begin = tile_begin;
    d.block_begin = block_begin;
    d.blocks_x = (d.tiles_x + MERGE_BT - 1) / MERGE_BT;
    d.blocks_y = (d.tiles_y + MERGE_BT - 1) / MERGE_BT;
    d.seam_begin = seam_begin;
    // tile columns MERGE_BT k + MERGE_BT - 1 that have a column right of them
    d.seam_cols = (d.tiles_x - 1) / MERGE_BT;
    d.seam_rows = (d.tiles_y - 1) / MERGE_BT;
    return d;
}
inline int64_t blocks_of(const ImgDev& d) { return (int64_t)d.blocks_x * d.blocks_y; }
inline int64_t groups_of(const ImgDev& d) { return (int64_t)((d.blocks_x + 1) / 2) * ((d.blocks_y + 1) / 2); }
inline int64_t seams_of(const ImgDev& d) {
    return (int64_t)d.seam_cols * d.tiles_y + (int64_t)d.seam_rows * d.tiles_x;
}

// Resolved config: cfg, else the process default.
ss_config resolve_config(const ss_config* cfg);

// Devices whose per-device state the library caches.
constexpr int MAX_DEVICES = 64;

// Compute units of the current device (cached per device).
int device_cus();

// Events recorded around the tile stage of every pipeline enqueued by the
// calling thread (both null = off).  Set by ss_profile_events.
struct ProfileHook {
    hipEvent_t tile_begin = nullptr;
    hipEvent_t tile_end = nullptr;
};
ProfileHook& profile_hook();

// The single-image operator on a given workspace (ss_extrema_u8's body).  If
// `need` is not null it receives the workspace bytes the call requires, also
// when the call fails for lack of them.
ss_status extrema_single(const uint8_t* d_img, size_t pitch, int width, int height, uint8_t* d_max,
                         uint8_t* d_min, size_t out_pitch, void* d_ws, size_t ws_bytes, const ss_config* cfg,
                         hipStream_t stream, size_t* need);

// Enqueues the whole pipeline on `s`.
hipError_t launch_pipeline(const Images& ims, const Ws& ws, int T, int NT, int cus, hipStream_t s);

// Adds the digest of a w x h byte image (salted by `salt`) to *d_out (device
// memory, cleared by the caller).  Used to compare candidates' masks.
hipError_t launch_digest(const uint8_t* p, int64_t pitch, int w, int h, uint64_t salt, uint64_t* d_out,
                         hipStream_t s);

// Enqueues the seam-sparsity profile of one image (counters zeroed first, in
// WsHeader::profile).
hipError_t launch_profile(const Images& ims, const Ws& ws, int T, hipStream_t s);

}  // namespace ssx
