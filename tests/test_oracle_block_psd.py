"""The oracle's per-block noise PSD input (nf3d_oracle_set_block_psd), pinned
on the CPU against the oracle's own sigma path and by block locality.

A constant per-block PSD c = sigma^2 * bw * bh is exactly the noise power the
sigma path uses (times btcur in the 3D gains), so outputs must be identical;
a huge PSD planted in one block removes that block's (non-grid) spectrum, so
only pixels inside that block's footprint may change."""
import ctypes

import numpy as np
import pytest

from harness import make_clip, oracle_runner


def runner():
    r = oracle_runner()
    r.lib.nf3d_oracle_set_block_psd.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    r.lib.nf3d_oracle_geometry.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    return r


def geometry(r, h):
    g = (ctypes.c_int * 5)()
    assert r.lib.nf3d_oracle_geometry(h.h, 0, g) == 0
    return dict(zip(["nox", "noy", "outwidth", "mirw", "mirh"], list(g)))


@pytest.mark.parametrize("bt", [3, 1, 0])
def test_constant_block_psd_equals_sigma_path(bt):
    clip = make_clip("gray", 8, False, width=96, height=64, nframes=5, seed=3)
    args = f"sigma=4;bt={bt};bw=16;bh=16"
    r = runner()
    hs = r.open(args, clip)
    want = [hs.get(n)[0] for n in range(5)]
    hp = r.open(args, clip)
    g = geometry(r, hp)
    psd = np.full(g["nox"] * g["noy"] * 16 * g["outwidth"], 4.0 ** 2 * 16 * 16, dtype=np.float32)
    assert r.lib.nf3d_oracle_set_block_psd(hp.h, 0, psd.ctypes.data) == 0
    got = [hp.get(n)[0] for n in range(5)]
    moved = max(np.abs(want[n].astype(int) - clip.frames[n][0].astype(int)).max() for n in range(5))
    assert moved > 0
    for n in range(5):
        assert np.array_equal(got[n], want[n]), n


def test_planted_block_changes_only_its_footprint():
    clip = make_clip("gray", 8, False, width=128, height=96, nframes=3, seed=8, noise=10.0)
    args = "sigma=3;bt=1;bw=16;bh=16;ow=4;oh=4"
    r = runner()
    h0 = r.open(args, clip)
    g = geometry(r, h0)
    d = 16 * g["outwidth"]
    nb = g["nox"] * g["noy"]
    base = np.full(nb * d, 9.0 * 256, dtype=np.float32)
    assert r.lib.nf3d_oracle_set_block_psd(h0.h, 0, base.ctypes.data) == 0
    ref = h0.get(1)[0]
    bx, by = 3, 2
    planted = base.copy()
    b0 = by * g["nox"] + bx
    planted[b0 * d:(b0 + 1) * d] = 1e12
    h1 = r.open(args, clip)
    assert r.lib.nf3d_oracle_set_block_psd(h1.h, 0, planted.ctypes.data) == 0
    got = h1.get(1)[0]
    diff = got.astype(int) != ref.astype(int)
    step = 16 - 4
    x0, y0 = bx * step - g["mirw"], by * step - g["mirh"]
    inside = np.zeros_like(diff)
    inside[max(y0, 0):y0 + 16, max(x0, 0):x0 + 16] = True
    print(f"planted block ({bx},{by}): {diff.sum()} pixels changed, {diff[inside].sum()} inside its footprint")
    assert diff[inside].sum() > 50
    assert not diff[~inside].any()
