"""Seeded synthetic clips (libnf3d_synth.so) for the benchmark workloads.

Workload seeds are 0x4E463344 + workload index:
  gratings_clip  Gray/planar gradient + 3 drifting gratings + N(0, noise^2)
  rects_clip     gradient + moving fBm-textured rectangles + N(0, noise^2),
                 YUV 4:2:0 / 4:4:4, 8..16-bit or float
Frames are generated in parallel host threads (the C calls release the GIL).
"""
from __future__ import annotations

import ctypes
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import SYNTH_PATH, Nf3dError

SEED_BASE = 0x4E463344
_lib = None


def _synth():
    global _lib
    if _lib is None:
        if not os.path.exists(SYNTH_PATH):
            raise Nf3dError(f"{SYNTH_PATH} is missing: build it with `python -m neo_fft3d_amd.build`")
        L = ctypes.CDLL(SYNTH_PATH)
        L.nf3d_synth_gratings.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_int]
        L.nf3d_synth_rects_create.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int]
        L.nf3d_synth_rects_create.restype = ctypes.c_void_p
        L.nf3d_synth_rects_free.argtypes = [ctypes.c_void_p]
        L.nf3d_synth_rects_frame.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]
        _lib = L
    return _lib


def _dtype(bits: int, is_float: bool):
    return np.float32 if is_float else (np.uint8 if bits <= 8 else np.uint16)


def gratings_clip(width: int, height: int, frames: int, seed: int, bits: int = 8, is_float: bool = False,
                  noise: float = 4.0, threads: int = 16) -> np.ndarray:
    """[frames, height, width] single-plane clip."""
    L = _synth()
    out = np.empty((frames, height, width), dtype=_dtype(bits, is_float))

    def one(n):
        L.nf3d_synth_gratings(seed, n, width, height, bits, int(is_float), noise,
                              out[n].ctypes.data, out.strides[1])

    with ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(one, range(frames)))
    return out


def rects_clip(width: int, height: int, frames: int, seed: int, ssw: int = 1, ssh: int = 1, bits: int = 8,
               is_float: bool = False, noise_luma: float = 3.0, noise_chroma: float = 2.0, threads: int = 16):
    """(Y, U, V) arrays of shape [frames, h, w] per plane."""
    L = _synth()
    dt = _dtype(bits, is_float)
    planes = [np.empty((frames, height, width), dtype=dt),
              np.empty((frames, height >> ssh, width >> ssw), dtype=dt),
              np.empty((frames, height >> ssh, width >> ssw), dtype=dt)]
    h = L.nf3d_synth_rects_create(seed, width, height)
    try:
        def one(n):
            ptrs = (ctypes.c_void_p * 3)(*[p[n].ctypes.data for p in planes])
            strides = (ctypes.c_int * 3)(*[p.strides[1] for p in planes])
            L.nf3d_synth_rects_frame(h, seed, n, width, height, ssw, ssh, bits, int(is_float), noise_luma,
                                     noise_chroma, ptrs, strides)

        with ThreadPoolExecutor(max_workers=threads) as ex:
            list(ex.map(one, range(frames)))
    finally:
        L.nf3d_synth_rects_free(h)
    return planes
