"""ctypes binding of libss_extrema.so (the C ABI in include/ss_extrema.h).

PyTorch is used only for device memory and streams.  Every call goes through
the HIP library; there is no CPU fallback: if the library or a gfx950 device is
missing the calls raise.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libss_extrema.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "ss_extrema.h")

SS_OK, SS_ERR_ARG, SS_ERR_WORKSPACE, SS_ERR_HIP, SS_ERR_INVARIANT, SS_ERR_CAPACITY, SS_ERR_NO_DEVICE = range(7)
TILES = (16, 32, 64)
THREADS = (64, 128, 256, 512)
CONFIGS = tuple((t, n) for t in TILES for n in THREADS if n <= t * t)


class SSError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        super().__init__("%s: %s (status %d)" % (what, _lib().ss_status_str(status).decode(), status))


class Config(ctypes.Structure):
    _fields_ = [("tile", ctypes.c_int), ("threads", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("invariant", ctypes.c_uint32), ("dirty_tiles", ctypes.c_uint32),
                ("fragments", ctypes.c_uint64), ("dsu_retries", ctypes.c_uint64),
                ("tiles", ctypes.c_uint64)]


class SeamProfile(ctypes.Structure):
    _fields_ = [("fragments", ctypes.c_uint64), ("seam_fragments", ctypes.c_uint64),
                ("seam_pixels", ctypes.c_uint64), ("equal_edges", ctypes.c_uint64),
                ("seam_edges", ctypes.c_uint64)]


class ImageDesc(ctypes.Structure):
    _fields_ = [("img", ctypes.c_void_p), ("max", ctypes.c_void_p), ("min", ctypes.c_void_p),
                ("pitch", ctypes.c_uint64), ("out_pitch", ctypes.c_uint64),
                ("width", ctypes.c_int32), ("height", ctypes.c_int32)]


_LIB = None


def _lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libss_extrema.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    vp, sz, i = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
    cfgp = ctypes.POINTER(Config)
    lib.ss_status_str.argtypes = [i]
    lib.ss_status_str.restype = ctypes.c_char_p
    lib.ss_workspace_size.argtypes = [i, i, cfgp, ctypes.POINTER(sz)]
    lib.ss_extrema_u8.argtypes = [vp, sz, i, i, vp, vp, sz, vp, sz, cfgp, vp]
    lib.ss_read_stats.argtypes = [vp, sz, ctypes.POINTER(Stats), vp]
    lib.ss_read_tile_counts.argtypes = [vp, sz, ctypes.POINTER(ctypes.c_uint32), sz, vp]
    lib.ss_seam_profile_u8.argtypes = [vp, sz, i, i, cfgp, vp, sz, vp, ctypes.POINTER(SeamProfile)]
    descp = ctypes.POINTER(ImageDesc)
    lib.ss_batch_workspace_size.argtypes = [descp, i, cfgp, ctypes.POINTER(sz)]
    lib.ss_extrema_batch_u8.argtypes = [descp, i, vp, sz, cfgp, vp]
    lib.ss_batch_plan_create.argtypes = [descp, i, vp, sz, cfgp, vp, ctypes.POINTER(vp)]
    lib.ss_batch_plan_run.argtypes = [vp, vp]
    lib.ss_batch_plan_destroy.argtypes = [vp]
    lib.ss_autotune.argtypes = [descp, i, i, vp, cfgp, ctypes.POINTER(ctypes.c_double)]
    lib.ss_autotune_batch.argtypes = [descp, i, i, vp, cfgp, ctypes.POINTER(ctypes.c_double)]
    lib.ss_profile_events.argtypes = [vp, vp]
    lib.ss_set_default_config.argtypes = [cfgp]
    lib.ss_get_default_config.argtypes = [cfgp]
    lib.ss_global_uf_workspace_size.argtypes = [i, i, ctypes.POINTER(sz)]
    lib.ss_global_uf_u8.argtypes = [vp, sz, i, i, vp, vp, sz, vp, sz, vp]
    for name in ("ss_global_uf_workspace_size", "ss_global_uf_u8", "ss_workspace_size", "ss_extrema_u8", "ss_read_stats", "ss_read_tile_counts",
                 "ss_batch_workspace_size", "ss_extrema_batch_u8", "ss_batch_plan_create",
                 "ss_batch_plan_run", "ss_batch_plan_destroy", "ss_autotune", "ss_autotune_batch",
                 "ss_set_default_config", "ss_get_default_config", "ss_profile_events"):
        getattr(lib, name).restype = i
    _LIB = lib
    return lib


def lib():
    return _lib()


def _check(status, what):
    if status != SS_OK:
        raise SSError(status, what)


def _cfg(tile, threads):
    if tile is None and threads is None:
        return None
    c = Config()
    d = default_config()
    c.tile = d[0] if tile is None else tile
    c.threads = d[1] if threads is None else threads
    return ctypes.byref(c)


def default_config():
    c = Config()
    _check(_lib().ss_get_default_config(ctypes.byref(c)), "ss_get_default_config")
    return c.tile, c.threads


def set_default_config(tile, threads):
    c = Config(tile, threads)
    _check(_lib().ss_set_default_config(ctypes.byref(c)), "ss_set_default_config")


def workspace_size(width, height, tile=None, threads=None):
    out = ctypes.c_size_t()
    _check(_lib().ss_workspace_size(width, height, _cfg(tile, threads), ctypes.byref(out)),
           "ss_workspace_size")
    return out.value


def _stream_handle(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def new_workspace(nbytes, device="cuda"):
    """Device workspace as a uint8 tensor (256-byte aligned by the allocator)."""
    import torch
    ws = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)
    assert ws.data_ptr() % 256 == 0
    return ws


def extrema(img, tile=None, threads=None, ws=None, out=None, stream=None):
    """Dual regional extrema of a 2-D uint8 CUDA tensor (rows may be strided).
    Returns (max_mask, min_mask, ws); masks are uint8 0/255 tensors."""
    import torch
    if img.dim() != 2 or img.dtype != torch.uint8 or not img.is_cuda or img.stride(1) != 1:
        raise ValueError("img must be a 2-D row-major uint8 CUDA tensor")
    h, w = img.shape
    need = workspace_size(w, h, tile, threads)
    if ws is None or ws.numel() < need:
        ws = new_workspace(need, img.device)
    if out is None:
        mx = torch.empty((h, w), dtype=torch.uint8, device=img.device)
        mn = torch.empty((h, w), dtype=torch.uint8, device=img.device)
    else:
        mx, mn = out
    assert mx.stride() == mn.stride() and mx.stride(1) == 1
    _check(_lib().ss_extrema_u8(img.data_ptr(), img.stride(0), w, h, mx.data_ptr(), mn.data_ptr(),
                                mx.stride(0), ws.data_ptr(), ws.numel(), _cfg(tile, threads),
                                _stream_handle(stream)), "ss_extrema_u8")
    return mx, mn, ws


def global_uf(img, ws=None, out=None, stream=None):
    """The dense Global-UF route (ss_global_uf_u8) on a 2-D uint8 CUDA tensor:
    the co-reference route the harness cross-checks the operator against.
    Returns (max_mask, min_mask, ws)."""
    import torch
    if img.dim() != 2 or img.dtype != torch.uint8 or not img.is_cuda or img.stride(1) != 1:
        raise ValueError("img must be a 2-D row-major uint8 CUDA tensor")
    h, w = img.shape
    need = ctypes.c_size_t()
    _check(_lib().ss_global_uf_workspace_size(w, h, ctypes.byref(need)), "ss_global_uf_workspace_size")
    if ws is None or ws.numel() < need.value:
        ws = new_workspace(need.value, img.device)
    if out is None:
        mx = torch.empty((h, w), dtype=torch.uint8, device=img.device)
        mn = torch.empty((h, w), dtype=torch.uint8, device=img.device)
    else:
        mx, mn = out
    assert mx.stride() == mn.stride() and mx.stride(1) == 1
    _check(_lib().ss_global_uf_u8(img.data_ptr(), img.stride(0), w, h, mx.data_ptr(), mn.data_ptr(),
                                  mx.stride(0), ws.data_ptr(), ws.numel(), _stream_handle(stream)),
           "ss_global_uf_u8")
    return mx, mn, ws


def stats(ws, stream=None, check=True):
    st = Stats()
    status = _lib().ss_read_stats(ws.data_ptr(), ws.numel(), ctypes.byref(st), _stream_handle(stream))
    if check:
        _check(status, "ss_read_stats")
    return {"status": status, "invariant": st.invariant, "dirty_tiles": st.dirty_tiles,
            "fragments": st.fragments, "dsu_retries": st.dsu_retries, "tiles": st.tiles}


SEAM_PROFILE_KEYS = ("fragments", "seam_fragments", "seam_pixels", "equal_edges", "seam_edges")


def seam_profile(img, tile=None, ws=None, stream=None):
    """Seam-sparsity counters of a 2-D uint8 CUDA tensor at tile size `tile`
    (dict of SEAM_PROFILE_KEYS; the reference profiler's per-image counts)."""
    import torch
    if img.dim() != 2 or img.dtype != torch.uint8 or not img.is_cuda or img.stride(1) != 1:
        raise ValueError("img must be a 2-D row-major uint8 CUDA tensor")
    h, w = img.shape
    if ws is None:
        ws = new_workspace(256, img.device)   # the counters are all it needs
    out = SeamProfile()
    _check(_lib().ss_seam_profile_u8(img.data_ptr(), img.stride(0), w, h, _cfg(tile, 256), ws.data_ptr(),
                                     ws.numel(), _stream_handle(stream), ctypes.byref(out)),
           "ss_seam_profile_u8")
    return {k: int(getattr(out, k)) for k in SEAM_PROFILE_KEYS}


def seam_profile_row(counts, width, height):
    """The reference profiler's derived columns from the counts
    (gpu_seam_sparsity_profiler_v0_1.cu:482-517)."""
    mp = width * height / 1e6
    f, sf = counts["fragments"], counts["seam_fragments"]
    return dict(counts, interior_fragments=f - sf,
                seam_fragment_fraction=sf / f if f else 0.0,
                seam_pixel_fraction=counts["seam_pixels"] / (width * height),
                seam_edge_fraction=counts["seam_edges"] / counts["equal_edges"] if counts["equal_edges"] else 0.0,
                seam_fragments_per_mpix=sf / mp, equal_seam_edges_per_mpix=counts["seam_edges"] / mp)


def tile_counts(ws, tiles_y, tiles_x, stream=None):
    import numpy as np
    out = np.zeros(tiles_y * tiles_x, dtype=np.uint32)
    _check(_lib().ss_read_tile_counts(ws.data_ptr(), ws.numel(),
                                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), out.size,
                                      _stream_handle(stream)), "ss_read_tile_counts")
    return out.reshape(tiles_y, tiles_x)


def descs(images, outs):
    """ImageDesc array for lists of input tensors and (max, min) output pairs."""
    arr = (ImageDesc * len(images))()
    for k, (img, (mx, mn)) in enumerate(zip(images, outs)):
        h, w = img.shape
        arr[k].img = img.data_ptr()
        arr[k].max = mx.data_ptr()
        arr[k].min = mn.data_ptr()
        arr[k].pitch = img.stride(0)
        arr[k].out_pitch = mx.stride(0)
        arr[k].width = w
        arr[k].height = h
    return arr


def batch_workspace_size(desc_array, tile=None, threads=None):
    out = ctypes.c_size_t()
    _check(_lib().ss_batch_workspace_size(desc_array, len(desc_array), _cfg(tile, threads),
                                          ctypes.byref(out)), "ss_batch_workspace_size")
    return out.value


def extrema_batch(desc_array, ws, tile=None, threads=None, stream=None):
    _check(_lib().ss_extrema_batch_u8(desc_array, len(desc_array), ws.data_ptr(), ws.numel(),
                                      _cfg(tile, threads), _stream_handle(stream)),
           "ss_extrema_batch_u8")


class BatchPlan:
    """hipGraph-captured batch pipeline (ss_batch_plan_*)."""

    def __init__(self, desc_array, ws, tile=None, threads=None, stream=None):
        self._h = ctypes.c_void_p()
        self.ws = ws
        self.descs = desc_array
        _check(_lib().ss_batch_plan_create(desc_array, len(desc_array), ws.data_ptr(), ws.numel(),
                                           _cfg(tile, threads), _stream_handle(stream),
                                           ctypes.byref(self._h)), "ss_batch_plan_create")

    def run(self, stream=None):
        _check(_lib().ss_batch_plan_run(self._h, _stream_handle(stream)), "ss_batch_plan_run")

    def close(self):
        if self._h:
            _lib().ss_batch_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_cap_bytes", ctypes.c_uint64), ("max_batch", ctypes.c_int),
                ("streams", ctypes.c_int), ("tile", ctypes.c_int), ("threads", ctypes.c_int),
                ("max_batch_pixels", ctypes.c_uint64)]


class Request(ctypes.Structure):
    _fields_ = [("host_img", ctypes.c_void_p), ("host_max", ctypes.c_void_p),
                ("host_min", ctypes.c_void_p), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("tag", ctypes.c_uint64)]


class Completion(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("tag", ctypes.c_uint64), ("status", ctypes.c_int32),
                ("batch_size", ctypes.c_int32), ("submit_s", ctypes.c_double),
                ("launch_s", ctypes.c_double), ("done_s", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("rejected", ctypes.c_uint64), ("launches", ctypes.c_uint64),
                ("pixels", ctypes.c_uint64), ("peak_inflight_bytes", ctypes.c_uint64)]


def _server_protos(lib):
    vp = ctypes.c_void_p
    lib.ss_server_create.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]
    lib.ss_server_submit.argtypes = [vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]
    lib.ss_server_wait.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(Completion)]
    lib.ss_server_drain.argtypes = [vp]
    lib.ss_server_get_stats.argtypes = [vp, ctypes.POINTER(ServerStats)]
    lib.ss_server_destroy.argtypes = [vp]
    lib.ss_server_run_open_loop.argtypes = [vp, ctypes.POINTER(Request), ctypes.c_int, ctypes.c_double,
                                            ctypes.c_uint64, ctypes.POINTER(Completion)]
    for n in ("ss_server_create", "ss_server_submit", "ss_server_wait", "ss_server_drain",
              "ss_server_get_stats", "ss_server_destroy", "ss_server_run_open_loop"):
        getattr(lib, n).restype = ctypes.c_int


class Server:
    """Continuous-batching request loop (ss_server_*).  Requests are host
    numpy arrays; outputs are written into host arrays owned by the caller."""

    def __init__(self, hbm_cap_bytes, max_batch=0, streams=2, tile=0, threads=0, max_batch_pixels=0):
        lib = _lib()
        if not hasattr(lib, "_server_ready"):
            _server_protos(lib)
            lib._server_ready = True
        self._h = ctypes.c_void_p()
        cfg = ServerConfig(hbm_cap_bytes, max_batch, streams, tile, threads, max_batch_pixels)
        _check(lib.ss_server_create(ctypes.byref(cfg), ctypes.byref(self._h)), "ss_server_create")
        self._keep = {}

    @staticmethod
    def request(img, mx, mn, tag=0):
        h, w = img.shape
        for a in (img, mx, mn):
            assert a.flags["C_CONTIGUOUS"] and a.shape == (h, w)
        return Request(img.ctypes.data, mx.ctypes.data, mn.ctypes.data, w, h, tag)

    def submit(self, img, mx, mn, tag=0):
        rid = ctypes.c_uint64()
        req = self.request(img, mx, mn, tag)
        _check(_lib().ss_server_submit(self._h, ctypes.byref(req), ctypes.byref(rid)), "ss_server_submit")
        self._keep[rid.value] = (img, mx, mn)
        return rid.value

    def wait(self, rid):
        c = Completion()
        st = _lib().ss_server_wait(self._h, rid, ctypes.byref(c))
        self._keep.pop(rid, None)
        _check(st, "ss_server_wait")
        return c

    def drain(self):
        _check(_lib().ss_server_drain(self._h), "ss_server_drain")

    def stats(self):
        s = ServerStats()
        _check(_lib().ss_server_get_stats(self._h, ctypes.byref(s)), "ss_server_get_stats")
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}

    def run_open_loop(self, reqs, qps, seed=0):
        """reqs: ctypes array of Request.  Returns a ctypes array of Completion."""
        out = (Completion * len(reqs))()
        _check(_lib().ss_server_run_open_loop(self._h, reqs, len(reqs), qps, seed, out),
               "ss_server_run_open_loop")
        return out

    def close(self):
        if self._h:
            _lib().ss_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_LOOP = None


def per_image_loop(desc_array, ws, tile=None, threads=None, reps=3, stream=None):
    """Times the plain path image by image from C++ (libss_loop.so: one
    ss_extrema_u8 call per image, back to back on one stream, the reference's
    per-image harness loop): one untimed pass, then `reps` timed passes.
    Returns the per-pass device times in ms."""
    global _LOOP
    if _LOOP is None:
        path = os.path.join(os.path.dirname(LIB_PATH), "libss_loop.so")
        if not os.path.exists(path):
            raise RuntimeError("libss_loop.so is not built (run __graft_entry__.build())")
        _lib()   # the loop library links the operator library
        lp = ctypes.CDLL(path)
        lp.ss_loop_per_image.argtypes = [ctypes.POINTER(ImageDesc), ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                         ctypes.POINTER(Config), ctypes.c_void_p, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_float)]
        lp.ss_loop_per_image.restype = ctypes.c_int
        _LOOP = lp
    out = (ctypes.c_float * reps)()
    _check(_LOOP.ss_loop_per_image(desc_array, len(desc_array), ws.data_ptr(), ws.numel(), _cfg(tile, threads),
                                   _stream_handle(stream), reps, out), "ss_loop_per_image")
    return list(out)


def profile_events(begin=None, end=None):
    """Route the tile-stage kernel through a pair of torch.cuda.Event objects
    (recorded on the pipeline's own stream); None, None turns it off."""
    b = ctypes.c_void_p(begin.cuda_event) if begin is not None else None
    e = ctypes.c_void_p(end.cuda_event) if end is not None else None
    _check(_lib().ss_profile_events(b, e), "ss_profile_events")


def autotune(desc_array, reps=5, stream=None, batch=False):
    """Times every (tile, threads) candidate and installs the best as the
    process default.  batch=False: per-image launches scored by
    sum(px) / sum(per-image median) (the reference's rule); batch=True: one
    batched launch over all the images, scored by sum(px) / median."""
    best = Config()
    rate = ctypes.c_double()
    fn = "ss_autotune_batch" if batch else "ss_autotune"
    _check(getattr(_lib(), fn)(desc_array, len(desc_array), reps, _stream_handle(stream),
                               ctypes.byref(best), ctypes.byref(rate)), fn)
    return (best.tile, best.threads), rate.value
