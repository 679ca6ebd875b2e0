"""Shared helpers for the test suite: PGM fixtures and the CPU oracle binding.

The oracle (oracle/libss_oracle.so) is the checker; nothing here is used by the
product path.
"""
import ctypes
import csv
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")
SUITES = ("oracle_v0_1", "oracle_v0_2_fuzz", "oracle_v0_3_sparse_seam_adversarial")
ORACLE_DIR = os.path.join(REPO, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "libss_oracle.so")


def read_pgm(path):
    """Reads a binary P5 PGM (maxval 255).  The header is parsed token by token:
    a pixel byte right after the header may itself be a whitespace byte."""
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    tokens = []
    while len(tokens) < 4:
        while data[pos:pos + 1].isspace():
            pos += 1
        if data[pos:pos + 1] == b"#":
            while data[pos:pos + 1] not in (b"\n", b""):
                pos += 1
            continue
        start = pos
        while not data[pos:pos + 1].isspace():
            pos += 1
        tokens.append(data[start:pos])
    pos += 1  # exactly one whitespace byte ends the header
    if tokens[0] != b"P5" or int(tokens[3]) != 255:
        raise ValueError("unsupported PGM %s" % path)
    w, h = int(tokens[1]), int(tokens[2])
    arr = np.frombuffer(data, dtype=np.uint8, count=w * h, offset=pos)
    return arr.reshape(h, w).copy()


def golden_cases(suite=None):
    """Yields (suite, case_id, image, expected_max, expected_min, manifest_row)."""
    for s in SUITES if suite is None else (suite,):
        root = os.path.join(GOLDEN, s)
        with open(os.path.join(root, "oracle_manifest.csv"), newline="") as f:
            rows = {r["case_id"]: r for r in csv.DictReader(f)}
        for name in sorted(os.listdir(os.path.join(root, "inputs"))):
            case = name[:-4]
            img = read_pgm(os.path.join(root, "inputs", name))
            emax = read_pgm(os.path.join(root, "expected_max", name))
            emin = read_pgm(os.path.join(root, "expected_min", name))
            yield s, case, img, emax, emin, rows.get(case)


_oracle = None


def oracle():
    """Loads (building on first use if needed) the CPU oracle library."""
    global _oracle
    if _oracle is not None:
        return _oracle
    if not os.path.exists(ORACLE_SO):
        subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])
    lib = ctypes.CDLL(ORACLE_SO)
    u8p = ctypes.POINTER(ctypes.c_uint8)
    lib.ssor_bruteforce.argtypes = [u8p, ctypes.c_int, ctypes.c_int, u8p, u8p]
    lib.ssor_global_uf.argtypes = [u8p, ctypes.c_int, ctypes.c_int, u8p, u8p, ctypes.c_int,
                                   ctypes.c_void_p]
    lib.ssor_sparse_seam.argtypes = [u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, u8p, u8p,
                                     ctypes.POINTER(ctypes.c_uint16)]
    lib.ssor_quantize.argtypes = [u8p, ctypes.c_size_t, ctypes.c_int]
    lib.ssor_quantize.restype = None
    lib.ssor_seam_profile.argtypes = [u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_uint64)]
    _oracle = lib
    return lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def oracle_bruteforce(img):
    img = np.ascontiguousarray(img, dtype=np.uint8)
    h, w = img.shape
    mx = np.empty_like(img)
    mn = np.empty_like(img)
    rc = oracle().ssor_bruteforce(_p(img), w, h, _p(mx), _p(mn))
    assert rc == 0, rc
    return mx, mn


def oracle_global_uf(img, threads=0):
    img = np.ascontiguousarray(img, dtype=np.uint8)
    h, w = img.shape
    mx = np.empty_like(img)
    mn = np.empty_like(img)
    rc = oracle().ssor_global_uf(_p(img), w, h, _p(mx), _p(mn), threads, None)
    assert rc == 0, rc
    return mx, mn


def oracle_sparse_seam(img, tile):
    """Returns (max, min, per-tile seam-fragment counts [tiles_y, tiles_x])."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    h, w = img.shape
    mx = np.empty_like(img)
    mn = np.empty_like(img)
    ty, tx = -(-h // tile), -(-w // tile)
    counts = np.zeros((ty, tx), dtype=np.uint16)
    rc = oracle().ssor_sparse_seam(_p(img), w, h, tile, _p(mx), _p(mn),
                                   counts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)))
    assert rc == 0, rc
    return mx, mn, counts


SEAM_PROFILE_KEYS = ("fragments", "seam_fragments", "seam_pixels", "equal_edges", "seam_edges")


def oracle_seam_profile(img, tile):
    """Seam-sparsity counters of the reference profiler (dict of SEAM_PROFILE_KEYS)."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    h, w = img.shape
    out = (ctypes.c_uint64 * 5)()
    rc = oracle().ssor_seam_profile(_p(img), w, h, tile, out)
    assert rc == 0, rc
    return dict(zip(SEAM_PROFILE_KEYS, (int(v) for v in out)))


def oracle_quantize(img, levels):
    out = np.ascontiguousarray(img, dtype=np.uint8).copy()
    oracle().ssor_quantize(_p(out), out.size, levels)
    return out


def mismatches(a, b):
    """Reference gate count: pixels where (a != 0) != (b != 0)."""
    return int(np.count_nonzero((np.asarray(a) != 0) != (np.asarray(b) != 0)))
