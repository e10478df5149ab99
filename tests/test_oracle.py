"""The FP64 oracle, pinned before it is trusted (CPU only).

Closed forms that a plausible restatement mistake would break: the
mixed-radix DFT vs a brute-force DFT; sigma = 0 reconstructs the input exactly
(the analysis x synthesis windows form a partition of unity for every
wintype); a constant frame stays constant; a static clip under a huge sigma
goes to its floor.  The gains themselves are pinned by closed forms in
test_oracle_closed_form.py, and test_oracle_mutations.py checks that planted
mistakes fail those.  The reference cannot be compiled here (its FFT backend,
FFTW, is absent), so no reference outputs are used.
"""
import ctypes
import os

import numpy as np
import pytest

from harness import ORACLE_SO, Clip, compare, make_clip, oracle_runner


def test_dft_matches_brute_force():
    lib = ctypes.CDLL(ORACLE_SO)
    rng = np.random.default_rng(0)
    for n in (1, 2, 3, 5, 7, 8, 12, 16, 24, 25, 32, 48, 49, 64, 96):
        x = rng.normal(size=n) + 1j * rng.normal(size=n)
        out = np.zeros(2 * n)
        xi = np.ascontiguousarray(np.stack([x.real, x.imag], axis=1).ravel())
        for sign in (-1, 1):
            lib.nf3d_oracle_dft(xi.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p), n, sign)
            got = out[0::2] + 1j * out[1::2]
            k = np.arange(n)
            want = (x[None, :] * np.exp(sign * 2j * np.pi * np.outer(k, k) / n)).sum(axis=1)
            assert np.max(np.abs(got - want)) < 1e-9 * max(1.0, np.max(np.abs(want)))


@pytest.mark.parametrize("wintype", [0, 1, 2])
@pytest.mark.parametrize("bt", [1, 2, 3, 4, 5, -1])
def test_sigma_zero_reconstructs_input(wintype, bt):
    """sigma=0 makes every Wiener gain exactly 1; the analysis x synthesis
    windows sum to 1 across each overlap for all three window types, so the
    output is the input.  (Kalman is excluded: with sigma=0 its covariances
    start at 0 and an unchanged element divides 0 by 0, in the reference too,
    src/code_impl/Kalman.cpp:51-55.)"""
    clip = make_clip("gray", 8, False, width=80, height=56, nframes=5, seed=4)
    h = oracle_runner().open(f"sigma=0;bt={bt};bw=16;bh=16;wintype={wintype};degrid=0", clip)
    for n in range(5):
        (mx, exact, _), = compare(h.get(n), clip.frames[n], False)
        assert mx == 0, (n, mx)


def test_constant_frame_stays_constant():
    clip = make_clip("yuv", 8, False, width=96, height=64, nframes=4, seed=1)
    for f in clip.frames:
        f[0][:] = 117
        f[1][:] = 90
        f[2][:] = 200
    h = oracle_runner().open("sigma=4;bt=3;bw=16;bh=16;sharpen=0.5", clip)
    for n in range(4):
        out = h.get(n)
        assert (out[0] == 117).all() and (out[1] == 90).all() and (out[2] == 200).all()


def test_error_messages_match_reference_wording():
    clip = make_clip("gray", 8, False, width=64, height=48, nframes=3)
    cases = {"bw=16;ow=9": "FFT3D: ow must be less than bw / 2",
             "bh=16;oh=9": "FFT3D: oh must be less than bh / 2",
             "bt=6": "FFT3D: bt must be -1 (Sharpen), 0 (Kalman), 1..5 (Wiener)",
             "beta=0.5": "FFT3D: beta must be not less 1.0",
             "l=40;r=30": "FFT3D: Width cannot be cropped to zero or below",
             "t=30;b=30": "FFT3D: Height cannot be cropped to zero or below"}
    from harness import FilterError
    for args, msg in cases.items():
        with pytest.raises(FilterError) as e:
            oracle_runner().open(args, clip)
        assert str(e.value) == msg
