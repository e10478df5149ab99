"""neo_fft3d for AMD Instinct MI355X: Python host side of the C ABI (include/nf3d.h).

The product is ``libnf3d.so`` (gfx950 kernels + C ABI).  This module binds it
with ctypes and mirrors the reference's FFT3D filter interface
(src/fft3d.hpp:12-364): ``FFT3D(args, clip_format, ...)`` parses the same
parameters with the same defaults, ``request_frames(n)`` is
RequestReferenceFrames, ``get_frame(n, frames)`` is GetFrame.  ``Engine`` is
one per-plane engine (FFT3DEngine) for device-resident work.

There is no CPU fallback: importing the bindings raises if the library has not
been built (``python -m neo_fft3d_amd.build``).
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
# NF3D_LIBRARY selects another build of the library (e.g. a tuned variant)
LIB_PATH = os.environ.get("NF3D_LIBRARY") or os.path.join(PKG, "libnf3d.so")
SYNTH_PATH = os.path.join(PKG, "libnf3d_synth.so")


class Nf3dError(RuntimeError):
    pass


class Params(ctypes.Structure):
    """nf3d_params: every knob of EngineParams (src/fft3d_common.h:48-77)."""
    _fields_ = [
        ("sigma", ctypes.c_float), ("beta", ctypes.c_float),
        ("bw", ctypes.c_int), ("bh", ctypes.c_int), ("bt", ctypes.c_int),
        ("ow", ctypes.c_int), ("oh", ctypes.c_int),
        ("kratio", ctypes.c_float), ("sharpen", ctypes.c_float), ("scutoff", ctypes.c_float),
        ("svr", ctypes.c_float), ("smin", ctypes.c_float), ("smax", ctypes.c_float),
        ("measure", ctypes.c_int), ("interlaced", ctypes.c_int), ("wintype", ctypes.c_int),
        ("pframe", ctypes.c_int), ("px", ctypes.c_int), ("py", ctypes.c_int), ("pshow", ctypes.c_int),
        ("pcutoff", ctypes.c_float), ("pfactor", ctypes.c_float), ("sigma2", ctypes.c_float),
        ("sigma3", ctypes.c_float), ("sigma4", ctypes.c_float), ("degrid", ctypes.c_float),
        ("dehalo", ctypes.c_float), ("hr", ctypes.c_float), ("ht", ctypes.c_float),
        ("l", ctypes.c_int), ("t", ctypes.c_int), ("r", ctypes.c_int), ("b", ctypes.c_int),
        ("opt", ctypes.c_int),
    ]


class Format(ctypes.Structure):
    """nf3d_format (DSFormat, include/dualsynth/ds_format.hpp:10-26)."""
    _fields_ = [("color_family", ctypes.c_int), ("is_float", ctypes.c_int),
                ("bits_per_sample", ctypes.c_int), ("ssw", ctypes.c_int), ("ssh", ctypes.c_int),
                ("num_planes", ctypes.c_int)]


class FrameRef(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("data", ctypes.c_void_p), ("stride", ctypes.c_int)]


class RunItem(ctypes.Structure):
    """nf3d_run_item: one engine's frame range in a batched call."""
    _fields_ = [("engine", ctypes.c_void_p), ("first", ctypes.c_int), ("count", ctypes.c_int),
                ("src", ctypes.POINTER(ctypes.c_void_p)), ("src_stride", ctypes.c_int),
                ("dst", ctypes.POINTER(ctypes.c_void_p)), ("dst_stride", ctypes.c_int)]


_lib = None


def lib() -> ctypes.CDLL:
    """The product library; raises Nf3dError when it is not built."""
    global _lib
    if _lib is None:
        # One HIP runtime per process: torch ships its own libamdhip64.so.7;
        # loaded first, it also serves this library (same soname).  Loaded
        # after the system runtime, torch finds no device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise Nf3dError(f"{LIB_PATH} is missing: build it with `python -m neo_fft3d_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        vp, ip, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t
        L.nf3d_params_defaults.argtypes = [ctypes.POINTER(Params), ctypes.c_float]
        L.nf3d_engine_create.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(Format), ip, ip, ip, ip,
                                         ctypes.POINTER(vp), ctypes.c_char_p, sz]
        L.nf3d_engine_get_frame.argtypes = [vp, ip, ctypes.POINTER(FrameRef), ip, vp, ip, ctypes.c_char_p, sz]
        L.nf3d_engine_get_frame_device.argtypes = [vp, ip, ctypes.POINTER(FrameRef), ip, vp, ip, vp,
                                                   ctypes.c_char_p, sz]
        L.nf3d_engine_run_device.argtypes = [vp, ip, ip, ctypes.POINTER(vp), ip, ctypes.POINTER(vp), ip, vp,
                                             ctypes.c_char_p, sz]
        L.nf3d_run_device_batch.argtypes = [ctypes.POINTER(RunItem), ip, vp, ctypes.c_char_p, sz]
        L.nf3d_run_host_batch.argtypes = [ctypes.POINTER(RunItem), ip, ctypes.c_char_p, sz]
        L.nf3d_run_device_batch_graph.argtypes = [ctypes.POINTER(RunItem), ip, vp, ctypes.POINTER(vp),
                                                  ctypes.c_char_p, sz]
        L.nf3d_run_device_batch_graph_ring.argtypes = L.nf3d_run_device_batch_graph.argtypes
        L.nf3d_graph_launch.argtypes = [vp, vp]
        L.nf3d_graph_destroy.argtypes = [vp]
        L.nf3d_graph_destroy.restype = None
        L.nf3d_engine_cache_refresh.argtypes = [vp, ip]
        L.nf3d_engine_geometry.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
        L.nf3d_engine_profile.argtypes = [vp, ip]
        L.nf3d_engine_profile_read.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]
        L.nf3d_engine_sync.argtypes = [vp]
        L.nf3d_engine_destroy.argtypes = [vp]
        L.nf3d_engine_destroy.restype = None
        L.nf3d_vs_signature.restype = ctypes.c_char_p
        L.nf3d_avs_signature.restype = ctypes.c_char_p
        L.nf3d_version.restype = ctypes.c_char_p
        L.nf3d_filter_create.argtypes = [ctypes.c_char_p, ctypes.POINTER(Format), ip, ip, ip,
                                         ctypes.POINTER(vp), ctypes.c_char_p, sz]
        L.nf3d_filter_request_frames.argtypes = [vp, ip, ctypes.POINTER(ctypes.c_int), ip]
        L.nf3d_filter_mt_mode.argtypes = [vp]
        L.nf3d_filter_destroy.argtypes = [vp]
        L.nf3d_filter_destroy.restype = None
        _lib = L
    return _lib


def default_params(sigma: float = 2.0, **kw) -> Params:
    p = Params()
    lib().nf3d_params_defaults(ctypes.byref(p), sigma)
    for k, v in kw.items():
        if not hasattr(p, k):
            raise TypeError(f"unknown parameter {k}")
        setattr(p, k, v)
    return p


def make_format(family: str = "yuv", bits: int = 8, is_float: bool = False, ssw: int = 1, ssh: int = 1) -> Format:
    fam = {"gray": 0, "yuv": 1, "rgb": 2}[family]
    return Format(fam, int(is_float), bits, ssw if fam == 1 else 0, ssh if fam == 1 else 0,
                  1 if fam == 0 else 3)


class Engine:
    """One per-plane engine (FFT3DEngine, src/fft3d_engine.h:22-741) on the GPU."""

    def __init__(self, params: Params, fmt: Format, width: int, height: int, frames: int, plane: int):
        L = lib()
        h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        if L.nf3d_engine_create(ctypes.byref(params), ctypes.byref(fmt), width, height, frames, plane,
                                ctypes.byref(h), err, 512) != 0:
            raise Nf3dError(err.value.decode())
        self.h = h
        self.params = params

    def geometry(self) -> dict:
        g = (ctypes.c_int * 7)()
        lib().nf3d_engine_geometry(self.h, g)
        return dict(zip(["nox", "noy", "outwidth", "mirw", "mirh", "ow", "oh"], list(g)))

    def run_device(self, first: int, count: int, src_ptrs, src_stride: int, dst_ptrs, dst_stride: int,
                   stream: int = 0):
        n = len(src_ptrs)
        s = (ctypes.c_void_p * n)(*src_ptrs)
        d = (ctypes.c_void_p * n)(*dst_ptrs)
        err = ctypes.create_string_buffer(512)
        if lib().nf3d_engine_run_device(self.h, first, count, s, src_stride, d, dst_stride,
                                        ctypes.c_void_p(stream), err, 512) != 0:
            raise Nf3dError(err.value.decode())

    def set_pattern_regions(self, rx: int, ry: int):
        """Per-region noise patterns: regions of rx x ry blocks (0, 0: one
        pattern for the plane, as the reference)."""
        L = lib()
        L.nf3d_engine_set_pattern_regions.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                                      ctypes.c_char_p, ctypes.c_size_t]
        err = ctypes.create_string_buffer(256)
        if L.nf3d_engine_set_pattern_regions(self.h, rx, ry, err, 256) != 0:
            raise Nf3dError(err.value.decode())

    def region_psd(self, src_ptr: int, src_stride: int, rx: int, ry: int, out_ptr: int, stream: int = 0):
        """Per-region pattern estimate of one device plane as a per-block PSD
        written to out_ptr ([nb][bh][bw/2+1] float32, device)."""
        L = lib()
        L.nf3d_engine_region_psd.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p,
                                             ctypes.c_size_t]
        err = ctypes.create_string_buffer(256)
        if L.nf3d_engine_region_psd(self.h, ctypes.c_void_p(src_ptr), src_stride, rx, ry, ctypes.c_void_p(out_ptr),
                                    ctypes.c_void_p(stream), err, 256) != 0:
            raise Nf3dError(err.value.decode())

    def set_dft(self, mode: int):
        """2D transform form: 0 = FFT in registers, 1 = DFT-matrix products on
        the matrix cores (32x32, 48x48 and 64x64 blocks)."""
        L = lib()
        L.nf3d_engine_set_dft.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]
        err = ctypes.create_string_buffer(256)
        if L.nf3d_engine_set_dft(self.h, mode, err, 256) != 0:
            raise Nf3dError(err.value.decode())

    def profile(self, enable: bool):
        lib().nf3d_engine_profile(self.h, int(enable))

    def profile_read(self):
        ms = ctypes.c_double()
        n = ctypes.c_int()
        if lib().nf3d_engine_profile_read(self.h, ctypes.byref(ms), ctypes.byref(n)) != 0:
            raise Nf3dError("profile read failed")
        return ms.value, n.value

    def sync(self):
        lib().nf3d_engine_sync(self.h)

    def close(self):
        if getattr(self, "h", None):
            lib().nf3d_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _run_items(items):
    """items: (engine, first, count, src_ptrs, src_stride, dst_ptrs, dst_stride) tuples."""
    arr = (RunItem * len(items))()
    keep = []
    for k, (eng, first, count, sp, ss, dp, ds) in enumerate(items):
        s = (ctypes.c_void_p * len(sp))(*sp)
        d = (ctypes.c_void_p * len(dp))(*dp)
        keep += [s, d]
        arr[k] = RunItem(eng.h, first, count, ctypes.cast(s, ctypes.POINTER(ctypes.c_void_p)), ss,
                         ctypes.cast(d, ctypes.POINTER(ctypes.c_void_p)), ds)
    return arr, keep


def run_device_batch(items, stream: int = 0):
    """nf3d_run_device_batch: several engines' device ranges, launches shared."""
    arr, keep = _run_items(items)
    err = ctypes.create_string_buffer(512)
    if lib().nf3d_run_device_batch(arr, len(items), ctypes.c_void_p(stream), err, 512) != 0:
        raise Nf3dError(err.value.decode())


def run_host_batch(items):
    """nf3d_run_host_batch: the batch with host planes (pointers to host
    memory), streamed through the device; returns when dst holds every frame."""
    arr, keep = _run_items(items)
    err = ctypes.create_string_buffer(512)
    if lib().nf3d_run_host_batch(arr, len(items), err, 512) != 0:
        raise Nf3dError(err.value.decode())


class Graph:
    """A captured nf3d_run_device_batch call (nf3d_run_device_batch_graph, or
    with ring=True nf3d_run_device_batch_graph_ring: bound to the spectra
    rings as they stand, replayed only while they hold the same frames
    shifted by whole frame steps)."""

    def __init__(self, items, stream: int, ring: bool = False):
        arr, keep = _run_items(items)
        g = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        cap = lib().nf3d_run_device_batch_graph_ring if ring else lib().nf3d_run_device_batch_graph
        if cap(arr, len(items), ctypes.c_void_p(stream), ctypes.byref(g), err, 512) != 0:
            raise Nf3dError(err.value.decode())
        self.g = g

    def launch(self, stream: int):
        if lib().nf3d_graph_launch(self.g, ctypes.c_void_p(stream)) != 0:
            raise Nf3dError("graph launch failed (a ring-bound graph needs the rings to hold the frames it was "
                            "captured against, shifted by whole frame steps)")

    def close(self):
        if getattr(self, "g", None):
            lib().nf3d_graph_destroy(self.g)
            self.g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def vs_signature() -> str:
    return lib().nf3d_vs_signature().decode()


def avs_signature() -> str:
    return lib().nf3d_avs_signature().decode()
