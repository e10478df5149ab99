"""Launch autotuner of the frame-loop path.

The frame-loop kernel (k_run) filters a run of frames per launch.  Two launch
shapes are free, because no choice of them changes a single output bit
(engines with a learned PSD excepted: their PSD is predicted once per launch
unless a prediction period is set, so table entries are not applied to them):

  frames per launch       1..16; more frames amortise the launch and the
                          per-launch seam pass, fewer start the next launch's
                          seam pass sooner
  frames per workgroup    Wiener 2D and sharpen-only (no temporal recursion):
                          how many frames one workgroup filters in turn (more
                          workgroups in flight with 1, fewer block loads of the
                          tables with more); Wiener 3D and Kalman keep a
                          launch's frames in one workgroup

``tune`` measures every candidate on a resident synthetic clip of the given
format (all planes in one batched call, as a serving loop runs them), checks
that each candidate's output equals the default's bit for bit, and returns the
fastest; ``write_table`` writes the table that ``nf3d_tune_load`` or the
NF3D_TUNE_TABLE environment variable feeds to the library (engines created
afterwards take their entry).  The table is keyed by block shape, bit depth
(32 = float) and bt.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import Engine, Nf3dError, default_params, lib, make_format, run_device_batch

FRAMES_CANDIDATES = (4, 8, 12, 16)
FPW_CANDIDATES = (1, 2, 4)
DEFAULT = (16, 1)


def _bind():
    L = lib()
    if getattr(L, "_tune_bound", False):
        return L
    L.nf3d_tune_load.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t]
    L.nf3d_tune_lookup.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int)]
    L.nf3d_engine_set_tuning.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p,
                                         ctypes.c_size_t]
    L.nf3d_engine_get_tuning.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
    L._tune_bound = True
    return L


def set_tuning(engine: Engine, frames_per_launch: int, frames_per_workgroup: int):
    err = ctypes.create_string_buffer(256)
    if _bind().nf3d_engine_set_tuning(engine.h, frames_per_launch, frames_per_workgroup, err, 256) != 0:
        raise Nf3dError(err.value.decode())


def get_tuning(engine: Engine):
    out = (ctypes.c_int * 2)()
    _bind().nf3d_engine_get_tuning(engine.h, out)
    return int(out[0]), int(out[1])


def load_table(path: str) -> int:
    err = ctypes.create_string_buffer(256)
    n = _bind().nf3d_tune_load(path.encode(), err, 256)
    if n < 0:
        raise Nf3dError(err.value.decode())
    return n


def lookup(bw: int, bh: int, bits: int, bt: int):
    out = (ctypes.c_int * 2)()
    if _bind().nf3d_tune_lookup(bw, bh, bits, bt, out):
        return int(out[0]), int(out[1])
    return None


def write_table(path: str, entries: dict):
    """entries: {(bw, bh, bits, bt): (frames_per_launch, frames_per_workgroup)}."""
    with open(path, "w") as f:
        f.write("# bw bh bits bt frames_per_launch frames_per_workgroup (neo_fft3d_amd.autotune)\n")
        for (bw, bh, bits, bt), (fr, fpw) in sorted(entries.items()):
            f.write(f"{bw} {bh} {bits} {bt} {fr} {fpw}\n")


def candidates(bt: int):
    fpws = FPW_CANDIDATES if bt in (1, -1) else (1,)
    return [(fr, fpw) for fr in FRAMES_CANDIDATES for fpw in fpws]


def _clip(width, height, frames, bits, seed):
    from .synth import rects_clip
    is_float = bits == 32
    if is_float:
        return rects_clip(width, height, frames, seed, ssw=0, ssh=0, bits=32, is_float=True,
                          noise_luma=3.0, noise_chroma=2.0), make_format("rgb", 32, True)
    return (rects_clip(width, height, frames, seed, bits=bits, noise_luma=3.0, noise_chroma=2.0),
            make_format("yuv", bits))


def tune(bw: int, bh: int, bits: int, bt: int, width: int = 1920, height: int = 1080, frames: int = 64,
         reps: int = 3, sigma: float = 2.0, sharpen: float = 0.0, seed: int = 0x7A11, log=None) -> dict:
    """Time every candidate launch shape on a resident clip; outputs must equal
    the default shape's bit for bit.  Returns {"best", "ms", "default"}."""
    import torch
    planes, fmt = _clip(width, height, frames, bits, seed)
    ow = bw // 3 // 2 * 2
    oh = bh // 3 // 2 * 2
    params = default_params(sigma, bt=bt, bw=bw, bh=bh, ow=ow, oh=oh, sharpen=sharpen)
    src = [torch.from_numpy(np.ascontiguousarray(pl)).cuda() for pl in planes]
    dst = [torch.zeros_like(s) for s in src]
    torch.cuda.synchronize()

    def measure(fr, fpw):
        # fresh engines per candidate: the Kalman state and the spectra ring
        # carry over from pass to pass, so outputs are compared on first passes
        engines = [Engine(params, fmt, width, height, frames, p) for p in range(3)]
        items = []
        for p, e in enumerate(engines):
            set_tuning(e, fr, fpw)
            stride = src[p][0].stride(0) * src[p].element_size()
            items.append((e, 0, frames, [src[p][k].data_ptr() for k in range(frames)], stride,
                          [dst[p][k].data_ptr() for k in range(frames)], stride))

        def run_once():
            run_device_batch(items)
            for e in engines:
                e.sync()
        torch.cuda.synchronize()
        run_once()
        out = [d.cpu().numpy() for d in dst]
        best = float("inf")
        for _ in range(reps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            run_once()
            best = min(best, time.perf_counter() - t)
        for e in engines:
            e.close()
        return best * 1e3, out

    ms = {}
    ref_ms, ref = measure(*DEFAULT)
    ms[DEFAULT] = ref_ms
    for cand in candidates(bt):
        if cand == DEFAULT:
            continue
        t, out = measure(*cand)
        for a, b in zip(out, ref):
            if not np.array_equal(a, b):
                raise Nf3dError(f"launch shape {cand} changed the output of ({bw}x{bh}, {bits} bit, bt={bt})")
        ms[cand] = t
        if log:
            log(f"{bw}x{bh} bits={bits} bt={bt} frames/launch={cand[0]} frames/wg={cand[1]}: {t:.3f} ms")
    best = min(ms, key=ms.get)
    return {"best": best, "ms": ms, "default": DEFAULT}
