"""Shared test helpers: small seeded clips and ctypes runners.

A "runner" wraps any library that exports the FFT3D clip interface
``<prefix>_create / <prefix>_get_frame / <prefix>_free``: the FP64 oracle
(oracle/libnf3d_oracle.so) and the product's host front end
(neo_fft3d_amd/libnf3d.so, prefix ``nf3d_clip``) both speak it, so one test
can run the same clip through each.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# NF3D_ORACLE_SO points the tests at another build of the oracle (the
# planted-mistake checks of tests/test_oracle_mutations.py use it)
ORACLE_SO = os.environ.get("NF3D_ORACLE_SO") or os.path.join(ROOT, "oracle", "libnf3d_oracle.so")

FAMILY = {"gray": 0, "yuv": 1, "rgb": 2, "yuva": 3, "rgba": 4}


@dataclass
class Clip:
    family: str
    bits: int
    is_float: bool
    ssw: int
    ssh: int
    width: int
    height: int
    frames: list  # frames[n][p] -> 2D numpy array

    @property
    def nplanes(self) -> int:
        return {"gray": 1, "yuv": 3, "rgb": 3, "yuva": 4, "rgba": 4}[self.family]

    @property
    def dtype(self):
        if self.is_float:
            return np.float32
        return np.uint8 if self.bits <= 8 else np.uint16

    def plane_shape(self, p: int):
        if self.family in ("yuv", "yuva") and p in (1, 2):
            return (self.height >> self.ssh, self.width >> self.ssw)
        return (self.height, self.width)


def make_clip(family="gray", bits=8, is_float=False, ssw=1, ssh=1, width=96, height=64,
              nframes=5, seed=1, noise=8.0, motion=1) -> Clip:
    """Seeded test content: gradient + moving gratings + Gaussian noise."""
    rng = np.random.default_rng(seed)
    maxv = 1.0 if is_float else float((1 << bits) - 1)
    scale = maxv / 255.0
    yuv = family in ("yuv", "yuva")
    clip = Clip(family, bits, is_float, ssw if yuv else 0, ssh if yuv else 0, width, height, [])
    periods = rng.uniform(6, 40, size=3)
    angles = rng.uniform(0, np.pi, size=3)
    for n in range(nframes):
        planes = []
        for p in range(clip.nplanes):
            h, w = clip.plane_shape(p)
            yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
            img = 60 + 100 * (xx / max(w - 1, 1)) * 0.5 + 60 * (yy / max(h - 1, 1))
            for k in range(3):
                ph = (xx * np.cos(angles[k]) + yy * np.sin(angles[k]) + motion * n) * 2 * np.pi / periods[k]
                img = img + 18 * np.sin(ph + p)
            img = img + rng.normal(0, noise, size=img.shape)
            img = img * scale
            if is_float:
                arr = np.clip(img, 0.0, 1.0).astype(np.float32)
            else:
                arr = np.clip(np.rint(img), 0, maxv).astype(clip.dtype)
            planes.append(np.ascontiguousarray(arr))
        clip.frames.append(planes)
    return clip


class Runner:
    """ctypes wrapper over an FFT3D clip interface (create / get_frame / free)."""

    def __init__(self, so_path: str, prefix: str):
        self.lib = ctypes.CDLL(so_path)
        self.create_fn = getattr(self.lib, prefix + "_create")
        self.create_fn.restype = ctypes.c_void_p
        self.create_fn.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_char_p, ctypes.c_int]
        self.get_fn = getattr(self.lib, prefix + "_get_frame")
        self.get_fn.restype = ctypes.c_int
        self.get_fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                ctypes.c_char_p, ctypes.c_int]
        self.free_fn = getattr(self.lib, prefix + "_free")
        self.free_fn.restype = None
        self.free_fn.argtypes = [ctypes.c_void_p]

    def open(self, args: str, clip: Clip) -> "Handle":
        return Handle(self, args, clip)


class FilterError(RuntimeError):
    pass


class Handle:
    def __init__(self, runner: Runner, args: str, clip: Clip):
        self.r = runner
        self.clip = clip
        np_ = clip.nplanes
        self._keep = clip.frames
        ptrs = [clip.frames[n][p].ctypes.data for n in range(len(clip.frames)) for p in range(np_)]
        self._ptrs = (ctypes.c_void_p * len(ptrs))(*ptrs)
        self._strides = (ctypes.c_int * np_)(*[clip.frames[0][p].strides[0] for p in range(np_)])
        err = ctypes.create_string_buffer(512)
        self.h = runner.create_fn(args.encode(), FAMILY[clip.family], int(clip.is_float), clip.bits,
                                  clip.ssw, clip.ssh, clip.width, clip.height, len(clip.frames),
                                  ctypes.cast(self._ptrs, ctypes.c_void_p),
                                  ctypes.cast(self._strides, ctypes.c_void_p), err, 512)
        if not self.h:
            raise FilterError(err.value.decode())

    def get(self, n: int):
        c = self.clip
        outs = [np.zeros(c.plane_shape(p), dtype=c.dtype) for p in range(c.nplanes)]
        optr = (ctypes.c_void_p * c.nplanes)(*[o.ctypes.data for o in outs])
        ostr = (ctypes.c_int * c.nplanes)(*[o.strides[0] for o in outs])
        err = ctypes.create_string_buffer(512)
        rc = self.r.get_fn(self.h, n, ctypes.cast(optr, ctypes.c_void_p),
                           ctypes.cast(ostr, ctypes.c_void_p), err, 512)
        if rc != 0:
            raise FilterError(err.value.decode())
        return outs

    def close(self):
        if self.h:
            self.r.free_fn(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def compare(a, b, is_float: bool):
    """Per-plane (max |diff|, fraction exact, rms) between two frame lists."""
    res = []
    for x, y in zip(a, b):
        d = np.abs(x.astype(np.float64) - y.astype(np.float64))
        if is_float:
            res.append((float(d.max()), float(np.mean(d == 0)), float(np.sqrt(np.mean(d * d)))))
        else:
            res.append((float(d.max()), float(np.mean(d == 0)), float(np.sqrt(np.mean(d * d)))))
    return res


def oracle_runner() -> Runner:
    return Runner(ORACLE_SO, "nf3d_oracle")
