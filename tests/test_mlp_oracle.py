"""The FP64 oracle of the learned-PSD network, pinned on the CPU.

Its backward pass against central finite differences of its own loss, its
GELU against math.erf, its BF16 rounding against hand-worked bit patterns,
and the Whittle loss's minimiser against the closed form ln(mean t).
"""
import math
import os
import sys

import numpy as np

from harness import ROOT

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import nf3d_mlp_oracle as mo  # noqa: E402


def tiny(seed=1, n_in=24, nh=16, n_out=24, rows=40):
    rng = np.random.default_rng(seed)
    params = [rng.standard_normal((n_in, nh)) * 0.4, rng.standard_normal(nh) * 0.1,
              rng.standard_normal((nh, nh)) * 0.3, rng.standard_normal(nh) * 0.1,
              rng.standard_normal((nh, n_out)) * 0.3, rng.standard_normal(n_out) * 0.2]
    x = rng.standard_normal((rows, n_in))
    t = rng.exponential(1.0, (rows, n_out)) * np.exp(rng.standard_normal(n_out))[None, :]
    return params, x, t


def test_backward_matches_finite_differences():
    params, x, t = tiny()
    g = mo.backward(params, x, t)
    rng = np.random.default_rng(5)
    idx = rng.choice(g.size, 60, replace=False)
    fd = mo.fd_check(params, x, t, idx, h=1e-5)
    err = np.abs(fd - g[idx]).max() / np.abs(g[idx]).max()
    print(f"finite differences vs analytic: max rel err {err:.2e}")
    assert err < 1e-6


def test_fd_catches_a_dropped_term():
    """A backward without the GELU derivative's second term must fail the check."""
    params, x, t = tiny(seed=2)
    saved = mo.gelu_grad
    try:
        mo.gelu_grad = lambda z: 0.5 * (1.0 + mo.erf(z * mo.SQRT1_2))
        g = mo.backward(params, x, t)
    finally:
        mo.gelu_grad = saved
    idx = np.arange(0, 24 * 16, 7)
    fd = mo.fd_check(params, x, t, idx, h=1e-5)
    assert np.abs(fd - g[idx]).max() / np.abs(fd).max() > 1e-3


def test_gelu_closed_form():
    z = np.linspace(-6, 6, 101)
    want = np.array([0.5 * v * (1 + math.erf(v / math.sqrt(2))) for v in z])
    assert np.abs(mo.gelu(z) - want).max() < 1e-15
    # derivative by central differences
    h = 1e-6
    fd = (mo.gelu(z + h) - mo.gelu(z - h)) / (2 * h)
    assert np.abs(mo.gelu_grad(z) - fd).max() < 1e-8


def test_bf16_rounding():
    # 1 + 2^-8 is exactly half-way between BF16 1.0 and 1 + 2^-7: ties to even -> 1.0
    assert mo.bf16_round(np.array([1.0 + 2.0 ** -8]))[0] == 1.0
    # 1 + 3*2^-8 ties between 1 + 2^-7 (odd) and 1 + 2^-6 (even) -> 1 + 2^-6
    assert mo.bf16_round(np.array([1.0 + 3 * 2.0 ** -8]))[0] == 1.0 + 2.0 ** -6
    assert mo.bf16_round(np.array([-2.5]))[0] == -2.5
    r = np.random.default_rng(0).standard_normal(1000)
    assert np.abs(mo.bf16_round(r) - r).max() <= np.abs(r).max() * 2.0 ** -8


def test_whittle_minimiser_is_log_mean():
    rng = np.random.default_rng(3)
    S = np.exp(rng.standard_normal(8))
    t = rng.exponential(1.0, (20000, 8)) * S
    ystar = mo.whittle_minimiser(t)
    # the minimiser of mean(y + t e^-y) over a grid around it
    for k in range(8):
        grid = ystar[k] + np.linspace(-0.05, 0.05, 11)
        vals = [np.mean(g + t[:, k] * np.exp(-g)) for g in grid]
        assert int(np.argmin(vals)) == 5
    assert np.abs(ystar - np.log(S)).max() < 0.05
