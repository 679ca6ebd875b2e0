"""Deterministic synthetic image generators for the parity tests and the benchmark.

All arithmetic is integer (torch int64 on CPU or GPU gives identical bytes):
  * hash      : mix32 (a 32-bit avalanche hash) of (seed, stream, pixel index)
  * iid       : high byte of the hash, uniform 0..255
  * smooth    : iid noise, separable Gaussian with integer weights
                w_k = round(65536 * exp(-k^2 / (2 sigma^2)) / sum), radius ceil(3 sigma),
                mirror padding without edge repeat, min-max stretch to 0..255 with
                round-half-up, then the reference quantiser
                (harness/quantization_and_scaling/gpu_common_quant_v0_7.cu:479-495)
  * rects     : constant rectangles on a block grid whose first boundary is shifted by
                `shift` pixels (T/2 puts every boundary mid-tile); cell sides 8..512
  * serpentine: a 1-px flat path (value 254) over iid noise in 0..253, rows every 4 px,
                joined at alternating ends; a 255 "leak" pixel next to its start makes
                the whole path non-maximal, so the fact has to cross every tile
  * checker   : 2-level checkerboard; same-coloured cells touch at corners, so each
                colour is one 8-connected zone spanning the image
  * edt       : Euclidean distance transform of non-overlapping disks scaled by
                rint(d * 255 / max) (the rule of datasets/bbbc038_distance/
                derive_bbbc038_distance_markers_v0_2.py:173-181)
"""
import math

import numpy as np
import torch

M32 = 0xFFFFFFFF


def mix32(x):
    """32-bit avalanche hash on an int64 tensor holding values in [0, 2^32)."""
    x = x & M32
    x = x ^ (x >> 16)
    x = (x * 0x7FEB352D) & M32
    x = x ^ (x >> 15)
    x = (x * 0x846CA68B) & M32
    x = x ^ (x >> 16)
    return x


def _key(seed, stream):
    return (seed * 0x9E3779B9 + stream * 0x85EBCA6B + 0x27D4EB2F) & M32


def iid(h, w, seed=0, stream=0, device="cpu"):
    idx = torch.arange(h * w, dtype=torch.int64, device=device)
    v = mix32(idx + _key(seed, stream)) >> 24
    return v.to(torch.uint8).reshape(h, w)


def quantize(img, levels):
    """Reference quantiser: bin = (v*L) >> 8; out = (bin*255 + (L-1)/2) / (L-1)."""
    if levels >= 256:
        return img
    v = img.to(torch.int64)
    b = (v * levels) >> 8
    return ((b * 255 + (levels - 1) // 2) // (levels - 1)).to(torch.uint8)


def gaussian_weights(sigma):
    r = int(math.ceil(3.0 * sigma))
    g = [math.exp(-(k * k) / (2.0 * sigma * sigma)) for k in range(-r, r + 1)]
    s = sum(g)
    return [int(math.floor(65536.0 * g[k + r] / s + 0.5)) for k in range(0, r + 1)]


def _mirror_index(n, r, device):
    i = torch.arange(-r, n + r, dtype=torch.int64, device=device)
    if n == 1:
        return torch.zeros_like(i)
    period = 2 * (n - 1)
    i = torch.remainder(i, period)
    return torch.where(i >= n, period - i, i)


def _smooth_axis(x, weights, dim):
    n = x.shape[dim]
    r = len(weights) - 1
    xp = x.index_select(dim, _mirror_index(n, r, x.device))
    acc = xp.narrow(dim, r, n) * weights[0]
    for k in range(1, r + 1):
        acc += (xp.narrow(dim, r - k, n) + xp.narrow(dim, r + k, n)) * weights[k]
    return acc


def smooth(h, w, sigma, levels, seed=0, device="cpu"):
    x = iid(h, w, seed, 1, device).to(torch.int64)
    wt = gaussian_weights(sigma)
    x = _smooth_axis(x, wt, 1)
    x = _smooth_axis(x, wt, 0)
    lo, hi = int(x.min()), int(x.max())
    if hi == lo:
        s = torch.zeros((h, w), dtype=torch.uint8, device=device)
    else:
        s = (((x - lo) * 255 + (hi - lo) // 2) // (hi - lo)).to(torch.uint8)
    return quantize(s, levels)


def rects(h, w, seed=0, shift=16, device="cpu"):
    rng = np.random.default_rng([seed, 3])

    def ids(n):
        out = np.empty(n, np.int64)
        pos, cell = 0, 0
        first = shift if shift > 0 else int(rng.integers(8, 513))
        while pos < n:
            size = first if cell == 0 else int(rng.integers(8, 513))
            out[pos:pos + size] = cell
            pos += size
            cell += 1
        return out

    cx = torch.as_tensor(ids(w), device=device)
    cy = torch.as_tensor(ids(h), device=device)
    key = (cy[:, None] * 0x9E3779B1 + cx[None, :] * 0x85EBCA77 + _key(seed, 4)) & M32
    return (mix32(key) >> 24).to(torch.uint8)


def serpentine(h, w, seed=0, vertical=False, device="cpu"):
    if vertical:
        return serpentine(w, h, seed, False, device).t().contiguous()
    img = (iid(h, w, seed, 5, device).to(torch.int64) % 254).to(torch.uint8)
    y = torch.arange(h, device=device)[:, None]
    x = torch.arange(w, device=device)[None, :]
    k = y // 4
    has_next = (k + 1) * 4 < h
    conn_x = torch.where(k % 2 == 0, torch.full_like(k, w - 1), torch.zeros_like(k))
    path = (y % 4 == 0) | (has_next & (x == conn_x))
    img = torch.where(path, torch.full_like(img, 254), img)
    if h > 1 and w > 1:
        img[1, 0] = 255
    return img


def checker(h, w, cell=7, lo=64, hi=192, device="cpu"):
    y = torch.arange(h, device=device)[:, None] // cell
    x = torch.arange(w, device=device)[None, :] // cell
    return torch.where((x + y) % 2 == 0, torch.tensor(lo, device=device),
                       torch.tensor(hi, device=device)).to(torch.uint8)


def constant(h, w, value=77, device="cpu"):
    return torch.full((h, w), value, dtype=torch.uint8, device=device)


def seam_comb(h, w, tile, device="cpu"):
    """Every perimeter pixel of every tile its own seam fragment (up to 4T-4
    per tile).  Rows next to a horizontal seam cycle through 10, 20, 30 along
    x, so a pixel equals only the pixel straight across the seam; columns next
    to a vertical seam cycle through 40, 50, 60 along y; the rows and columns
    one step inside (200, 210) keep each perimeter pixel apart from the
    tile's interior (128)."""
    y = torch.arange(h, device=device)[:, None]
    x = torch.arange(w, device=device)[None, :]
    ry, rx = y % tile, x % tile
    v = torch.full((h, w), 128, dtype=torch.int64, device=device)
    v = torch.where((rx == 1) | (rx == tile - 2), torch.full_like(v, 210), v)
    v = torch.where((rx == 0) | (rx == tile - 1), (40 + 10 * (y % 3)).expand(h, w), v)
    v = torch.where((ry == 1) | (ry == tile - 2), torch.full_like(v, 200), v)
    v = torch.where((ry == 0) | (ry == tile - 1), (10 + 10 * (x % 3)).expand(h, w), v)
    return v.to(torch.uint8)


_TEMPLATES = {}


def _disk_d2(r):
    """Exact squared EDT inside a lone digital disk of radius r (pixels with
    dx^2 + dy^2 <= r^2), 0 outside."""
    if r not in _TEMPLATES:
        from scipy import ndimage
        yy, xx = np.mgrid[-r - 1:r + 2, -r - 1:r + 2]
        mask = yy * yy + xx * xx <= r * r
        d = ndimage.distance_transform_edt(mask)
        _TEMPLATES[r] = (np.rint(d * d).astype(np.int64), mask)
    return _TEMPLATES[r]


def edt_disks(h, w, seed=0):
    """numpy uint8 distance map of 20..300 non-overlapping disks (radii 5..40).
    Disks are kept 3 px apart, so each disk's nearest background pixel is the
    one of the lone disk and the union's EDT is a paste of per-disk EDTs."""
    rng = np.random.default_rng([seed, 6])
    want = int(rng.integers(20, 301))
    d2 = np.zeros((h, w), np.int64)
    # centres and radii of the placed disks (first n entries)
    pl = np.zeros((3, want), np.int64)
    n = 0
    for _ in range(want * 20):
        if n >= want:
            break
        r = int(rng.integers(5, 41))
        if 2 * r + 3 > min(h, w):
            continue
        cy = int(rng.integers(r + 1, h - r - 1))
        cx = int(rng.integers(r + 1, w - r - 1))
        py, px, pr = pl[0, :n], pl[1, :n], pl[2, :n]
        if n and ((cy - py) ** 2 + (cx - px) ** 2 < (r + pr + 3) ** 2).any():
            continue
        pl[:, n] = (cy, cx, r)
        n += 1
        t, m = _disk_d2(r)
        sl = (slice(cy - r - 1, cy + r + 2), slice(cx - r - 1, cx + r + 2))
        np.copyto(d2[sl], t, where=m)
    m2 = int(d2.max())
    if m2 == 0:
        return np.zeros((h, w), np.uint8)
    return np.rint(np.sqrt(d2.astype(np.float64)) * 255.0 / math.sqrt(m2)).astype(np.uint8)


def config1(device="cpu"):
    """512 x 512, seed 0: iid noise, Gaussian sigma 2, 32 levels."""
    return smooth(512, 512, 2.0, 32, seed=0, device=device)


def sweep_images(side, seed=0, device="cpu", sigmas=(1, 4, 16), levels=(2, 4, 8, 16, 64, 256)):
    """(name, image) of the resolution / gray-level sweep at one side and seed."""
    yield "iid", iid(side, side, seed, 0, device)
    for s in sigmas:
        base = smooth(side, side, s, 256, seed, device)
        for lv in levels:
            yield "smooth_s%d_q%d" % (s, lv), quantize(base, lv)


def adversarial_images(side=4096, tile=32, seed=0, device="cpu"):
    yield "rects_shift%d" % (tile // 2), rects(side, side, seed, tile // 2, device)
    yield "serpentine_h", serpentine(side, side, seed, False, device)
    yield "serpentine_v", serpentine(side, side, seed, True, device)
    yield "checker2", checker(side, side, device=device)
    yield "seam_comb", seam_comb(side, side, tile, device=device)
    yield "constant", constant(side, side, device=device)


def mixed_sizes(n, seed=0, lo=256, hi=1024):
    rng = np.random.default_rng([seed, 7])
    return [(int(rng.integers(lo, hi + 1)), int(rng.integers(lo, hi + 1))) for _ in range(n)]


def mixed_image(k, h, w, seed=0, device="cpu"):
    """Content k % 3: iid, smoothed (sigma 3, 16 levels), EDT of disks."""
    kind = k % 3
    s = seed * 1000003 + k
    if kind == 0:
        return iid(h, w, s, 0, device)
    if kind == 1:
        return smooth(h, w, 3.0, 16, s, device)
    return torch.as_tensor(edt_disks(h, w, s)).to(device)


def serving_sizes(n, seed=0, lo=64, hi=8192):
    """(h, w) of n serving requests: side log-uniform in lo..hi, aspect ratio
    log-uniform in 0.5..2 (w / h), the area kept at side^2."""
    rng = np.random.default_rng([seed, 9])
    out = []
    for _ in range(n):
        s = float(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        a = float(np.exp(rng.uniform(np.log(0.5), np.log(2.0))))
        out.append((max(1, int(round(s / a ** 0.5))), max(1, int(round(s * a ** 0.5)))))
    return out
