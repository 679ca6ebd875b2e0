"""CPU-side checks of the C ABI: the library loads, exports every declared
symbol, and rejects bad arguments before touching a device."""
import ctypes
import re

import pytest

from ssx import sparse_seam as ss


def declared_symbols():
    text = open(ss.HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(ss_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = ss.lib()
    names = declared_symbols()
    assert len(names) >= 13
    for n in names:
        assert hasattr(lib, n), n


def test_status_strings():
    lib = ss.lib()
    for code in range(7):
        assert lib.ss_status_str(code)
    assert lib.ss_status_str(99) == b"unknown status"


def test_workspace_size_validation():
    for t, n in ss.CONFIGS:
        assert ss.workspace_size(512, 512, t, n) > 0
    with pytest.raises(ss.SSError) as e:
        ss.workspace_size(0, 5)
    assert e.value.status == ss.SS_ERR_ARG
    for bad in ((24, 256), (32, 96), (16, 512), (128, 256)):
        with pytest.raises(ss.SSError):
            ss.workspace_size(100, 100, *bad)


def test_workspace_grows_with_image():
    a = ss.workspace_size(512, 512, 32, 256)
    b = ss.workspace_size(1024, 1024, 32, 256)
    assert b > 3 * a
    # record plane (1 B/px) + per-tile and per-fragment-slot state: below the
    # reference's 3.002 B/px of auxiliary state at T=32, 2.0005 at T=64
    assert ss.workspace_size(16384, 16384, 32, 256) / (16384 * 16384) < 3.0
    assert ss.workspace_size(16384, 16384, 64, 256) / (16384 * 16384) < 2.0


def test_extrema_rejects_bad_arguments_without_device():
    lib = ss.lib()
    cfg = ss.Config(32, 256)
    # null image
    assert lib.ss_extrema_u8(None, 64, 64, 64, 1, 1, 64, 256, 1 << 20, ctypes.byref(cfg), None) == ss.SS_ERR_ARG
    # pitch < width
    assert lib.ss_extrema_u8(256, 10, 64, 64, 256, 256, 64, 256, 1 << 20, ctypes.byref(cfg), None) == ss.SS_ERR_ARG
    # bad config
    bad = ss.Config(48, 256)
    assert lib.ss_extrema_u8(256, 64, 64, 64, 256, 256, 64, 256, 1 << 20, ctypes.byref(bad), None) == ss.SS_ERR_ARG
    # undersized / misaligned workspace
    assert lib.ss_extrema_u8(256, 64, 64, 64, 256, 256, 64, 256, 16, ctypes.byref(cfg), None) == ss.SS_ERR_WORKSPACE
    assert lib.ss_extrema_u8(256, 64, 64, 64, 256, 256, 64, 260, 1 << 20, ctypes.byref(cfg), None) == ss.SS_ERR_WORKSPACE


def test_batch_validation_without_device():
    arr = (ss.ImageDesc * 2)()
    for d in arr:
        d.img, d.max, d.min, d.pitch, d.out_pitch, d.width, d.height = 256, 512, 768, 64, 64, 64, 64
    assert ss.batch_workspace_size(arr) > ss.workspace_size(64, 64)
    arr[1].width = 0
    with pytest.raises(ss.SSError):
        ss.batch_workspace_size(arr)


def test_default_config_roundtrip():
    old = ss.default_config()
    ss.set_default_config(64, 128)
    assert ss.default_config() == (64, 128)
    with pytest.raises(ss.SSError):
        ss.set_default_config(64, 1024)
    ss.set_default_config(*old)
