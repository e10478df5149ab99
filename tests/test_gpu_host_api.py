"""Host-side behaviour of the C ABI on the GPU: concurrent calls (the
reference's engines are driven from several threads by the hosts' MT modes,
src/fft3d.hpp:201-264) give the sequential results bit for bit, and a device
allocation failure surfaces as an ordinary error from nf3d_engine_create."""
import ctypes
import sys
import threading

import numpy as np
import pytest

from harness import ROOT, make_clip

sys.path.insert(0, ROOT)

pytestmark = pytest.mark.gpu


def engine(args_kw, w, h, frames):
    from neo_fft3d_amd import Engine, default_params, make_format
    return Engine(default_params(3.0, **args_kw), make_format("gray", 8), w, h, frames, 0)


def get_frames(eng, clip, order, pad=0):
    """nf3d_engine_get_frame (host planes) for the frames in `order`; each
    output has its own row padding (stride) `pad`."""
    from neo_fft3d_amd import FrameRef, lib
    L = lib()
    out = {}
    F = len(clip.frames)
    for n in order:
        refs = (FrameRef * F)(*[FrameRef(k, clip.frames[k][0].ctypes.data, clip.frames[k][0].strides[0])
                                for k in range(F)])
        h, w = clip.frames[n][0].shape
        dst = np.zeros((h, w + pad), dtype=np.uint8)
        err = ctypes.create_string_buffer(256)
        assert L.nf3d_engine_get_frame(eng.h, n, refs, F, dst.ctypes.data, dst.strides[0], err, 256) == 0, err.value
        out[n] = dst[:, :w].copy()
    return out


def test_concurrent_engines_match_sequential():
    F = 8
    cases = [({"bt": 3, "bw": 16, "bh": 16}, 7), ({"bt": 1, "bw": 32, "bh": 32}, 8),
             ({"bt": 0, "bw": 16, "bh": 16}, 9), ({"bt": 5, "bw": 24, "bh": 24}, 10)]
    clips = [make_clip("gray", 8, False, width=160, height=112, nframes=F, seed=s) for _, s in cases]
    seq = [get_frames(engine(kw, 160, 112, F), c, range(F)) for (kw, _), c in zip(cases, clips)]
    engines = [engine(kw, 160, 112, F) for kw, _ in cases]
    results = [None] * len(cases)
    errors = []

    def work(i):
        try:
            results[i] = get_frames(engines[i], clips[i], range(F), pad=16 * i)
        except Exception as e:  # noqa: BLE001
            errors.append(e)
    threads = [threading.Thread(target=work, args=(i,)) for i in range(len(cases))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for i in range(len(cases)):
        for n in range(F):
            assert np.array_equal(results[i][n], seq[i][n]), (i, n)


def test_shared_engine_calls_are_serialised():
    """Two threads on one engine (bt = 1: frames independent), different
    output strides: every frame as from one thread."""
    F = 12
    clip = make_clip("gray", 8, False, width=128, height=96, nframes=F, seed=4)
    want = get_frames(engine({"bt": 1, "bw": 16, "bh": 16}, 128, 96, F), clip, range(F))
    eng = engine({"bt": 1, "bw": 16, "bh": 16}, 128, 96, F)
    got = [None, None]

    def work(i):
        got[i] = get_frames(eng, clip, range(i, F, 2), pad=32 * (i + 1))
    threads = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for i in range(2):
        for n, v in got[i].items():
            assert np.array_equal(v, want[n]), n


def test_device_allocation_failure_is_an_error():
    """A plane whose spectra ring cannot fit in HBM: nf3d_engine_create fails
    with a message, no crash, and the library keeps working."""
    from neo_fft3d_amd import Nf3dError
    with pytest.raises(Nf3dError, match="hipMalloc"):
        engine({"bt": 3, "bw": 16, "bh": 16}, 180000, 180000, 4)  # 518 GB seam buffer
    clip = make_clip("gray", 8, False, width=96, height=64, nframes=3, seed=1)
    out = get_frames(engine({"bt": 1, "bw": 16, "bh": 16}, 96, 64, 3), clip, range(3))
    assert len(out) == 3
