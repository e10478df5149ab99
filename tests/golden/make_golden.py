"""Writes the golden vectors of tests/golden/*.npz from the compiled reference.

Run in a container that has the reference tree (oracle/build_ref.sh builds
oracle/_ref/libnf3d_ref.so from it): ``python tests/golden/make_golden.py``
(all fixtures) or ``... make_golden.py <name> ...`` (the named ones).
Each fixture holds the seeded input frames, the filter arguments and the
reference's output frames (opt=1, the C kernels; ncpu=1; MKL sequential
FFTW3 interface as the FFT backend).  The GPU box never runs the reference:
the tests compare against these files.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from harness import make_clip, ref_runner  # noqa: E402

# name, family, bits, float, width, height, frames, args
CASES = [
    ("g8_bt1", "gray", 8, False, 96, 64, 5, "sigma=3;bt=1;bw=16;bh=16"),
    ("g8_bt2_sharpen", "gray", 8, False, 96, 64, 5, "sigma=3;bt=2;bw=16;bh=16;sharpen=0.5"),
    ("g8_bt3", "gray", 8, False, 96, 64, 5, "sigma=3;bt=3;bw=16;bh=16"),
    ("g8_bt4_dehalo", "gray", 8, False, 96, 64, 5, "sigma=3;bt=4;bw=32;bh=32;dehalo=1"),
    ("g8_bt5_sharpen_dehalo", "gray", 8, False, 96, 64, 6, "sigma=3;bt=5;bw=16;bh=16;sharpen=0.4;dehalo=0.5"),
    ("g8_kalman", "gray", 8, False, 96, 64, 6, "sigma=3;bt=0;bw=16;bh=16"),
    ("g8_kalman_pattern", "gray", 8, False, 96, 64, 6, "sigma=3;bt=0;bw=16;bh=16;pfactor=1"),
    ("g8_sharpen_only", "gray", 8, False, 96, 64, 4, "sigma=3;bt=-1;bw=16;bh=16;sharpen=1"),
    ("g8_pattern_bt3", "gray", 8, False, 96, 64, 5, "sigma=3;bt=3;bw=16;bh=16;pfactor=1;pframe=2"),
    ("g8_pattern_bt1_px", "gray", 8, False, 96, 64, 4, "sigma=3;bt=1;bw=16;bh=16;pfactor=1;px=3;py=2;sharpen=0.3"),
    ("g8_sigmas", "gray", 8, False, 96, 64, 5, "sigma=3;bt=3;bw=16;bh=16;sigma2=5;sigma3=1;sigma4=0.5"),
    ("g8_nodegrid_win1", "gray", 8, False, 96, 64, 5, "sigma=3;bt=3;bw=16;bh=16;degrid=0;wintype=1"),
    ("g8_win2_beta", "gray", 8, False, 96, 64, 5, "sigma=3;bt=3;bw=16;bh=16;wintype=2;beta=1.5"),
    ("g8_interlaced", "gray", 8, False, 96, 64, 5, "sigma=3;bt=3;bw=16;bh=16;interlaced=1"),
    ("g8_b24", "gray", 8, False, 96, 64, 5, "sigma=3;bt=3;bw=24;bh=24;ow=8;oh=8"),
    ("g8_b32x16", "gray", 8, False, 96, 64, 5, "sigma=3;bt=3;bw=32;bh=16;ow=8;oh=4"),
    ("g8_b48", "gray", 8, False, 160, 120, 5, "sigma=2;bt=3;bw=48;bh=48;ow=16;oh=16"),
    ("g8_b64", "gray", 8, False, 160, 128, 4, "sigma=2;bt=1;bw=64;bh=64;ow=16;oh=16"),
    ("yuv420p8_crop", "yuv", 8, False, 96, 64, 5, "sigma=3;bt=3;bw=16;bh=16;l=4;t=2;r=6;b=4"),
    ("yuv420p10_planes", "yuv", 10, False, 96, 64, 5, "sigma=3;bt=3;bw=16;bh=16;planes=0,2"),
    ("yuv420p16_kalman", "yuv", 16, False, 96, 64, 5, "sigma=3;bt=0;bw=32;bh=16;ow=8;oh=4;kratio=1.5"),
    ("rgbps_float", "rgb", 32, True, 64, 48, 4, "sigma=3;bt=3;bw=16;bh=16"),
    ("g8_pshow", "gray", 8, False, 160, 120, 3, "sigma=3;bt=3;bw=16;bh=16;pfactor=1;pshow=1"),
    ("yuv420p8_pshow_pxpy", "yuv", 8, False, 96, 64, 3, "sigma=3;bt=1;bw=16;bh=16;pfactor=1;pshow=1;px=3;py=2"),
]

# The CPU-runnable workload, 5 frames: Gray8 640x360, bt=3, sigma=2, 48/16.
CONFIG1_ARGS = "sigma=2;bt=3;bw=48;bh=48;ow=16;oh=16;sharpen=0"


def run_ref(clip, args):
    h = ref_runner().open(args + ";opt=1;ncpu=1", clip)
    return [h.get(n) for n in range(len(clip.frames))]


def save(name, clip, args, outs):
    arrays = {}
    for p in range(clip.nplanes):
        arrays[f"in{p}"] = np.stack([f[p] for f in clip.frames])
        arrays[f"out{p}"] = np.stack([o[p] for o in outs])
    meta = dict(family=clip.family, bits=clip.bits, is_float=clip.is_float, ssw=clip.ssw, ssh=clip.ssh,
                width=clip.width, height=clip.height, args=args)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), meta=json.dumps(meta), **arrays)


def main():
    os.environ.setdefault("MKL_THREADING_LAYER", "SEQUENTIAL")
    only = set(sys.argv[1:])  # optional: names of the fixtures to (re)write
    for name, fam, bits, isf, w, h, f, args in CASES:
        if only and name not in only:
            continue
        clip = make_clip(fam, bits, isf, width=w, height=h, nframes=f, seed=7)
        save(name, clip, args, run_ref(clip, args))
        print("wrote", name)
    if only:
        return
    from neo_fft3d_amd.synth import SEED_BASE, gratings_clip
    g = gratings_clip(640, 360, 5, SEED_BASE + 0, noise=4.0)
    from harness import Clip
    clip = Clip("gray", 8, False, 0, 0, 640, 360, [[np.ascontiguousarray(g[n])] for n in range(5)])
    save("config1_gray8_640x360_bt3", clip, CONFIG1_ARGS, run_ref(clip, CONFIG1_ARGS))
    print("wrote config1")
    # the plugin argument signatures the reference registers with both hosts
    import ctypes
    lib = ref_runner().lib
    lib.nf3dref_vs_params.restype = ctypes.c_char_p
    lib.nf3dref_avs_params.restype = ctypes.c_char_p
    with open(os.path.join(HERE, "signatures.json"), "w") as f:
        json.dump({"vs": lib.nf3dref_vs_params().decode(), "avs": lib.nf3dref_avs_params().decode()}, f, indent=1)
    print("wrote signatures")


if __name__ == "__main__":
    main()
