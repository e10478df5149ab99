"""Closed-form checks of the FP64 oracle (CPU only).

Each test feeds the oracle content whose block spectra are known exactly and
compares its output with the value the reference's definitions fix by hand:

* a vertical-stripe cosine c(x) = M + A cos(2 pi x / P) with P dividing bw
  puts, in every block, all its energy at the DC bin and at bin (0, bw/P) with
  |X_k| = A bw bh / 2 -- with the flat analysis window (wintype=2), whose
  synthesis windows still sum to one across each overlap; the plane width is
  chosen so the reflect-101 padding continues the cosine (W - 1 a multiple
  of P/2), so edge blocks see the same pure content.  A per-bin gain G then
  turns the output into M' + A G cos(2 pi x / P) exactly;
* the gains come from the reference's formulas evaluated by hand at that one
  bin: Wiener 2D (src/code_impl/Apply.cpp:13-65), the temporal Wiener of
  Apply3D2..3D5 (:67-235) on a temporally modulated cosine whose temporal
  spectrum has bins 0 and +-1 only, sharpen / dehalo factors
  (src/code_impl/Sharpen.cpp:13-63, tables src/fft3d_engine.h:285-327), the
  Kalman recursion (src/code_impl/Kalman.cpp:13-67: gains 2/3 then 16/19 on
  a static clip, reset on a jump), the noise pattern of a given block
  (src/helper.cpp:92-132) and of sigma2..sigma4 (src/helper.cpp:14-43);
* FindPatternBlock (src/helper.cpp:46-90) on planted flat blocks: the flat
  block is chosen, the first of two identical ones wins;
* the lower Wiener limit (beta) with a huge sigma halves every bin: output
  round(x / 2) for every bt, chroma centred on 2^(bits-1)
  (src/buffer.cpp:133-135), crop borders copied (src/fft3d.hpp:292-352);
* interlaced content: each field filtered as its own image
  (src/buffer.cpp:31-63, 99-115).

A mistake of the kind a restatement makes -- a dropped normalisation, bt
missing from sigma^2, a transposed table index, a wrong temporal twiddle
sign or current-frame index, a Kalman gain off by one update, a missing
chroma centring -- moves these outputs by 1e-3 or more; the bars are 2e-5
(float) and +-1 LSB (integer).
"""
import ctypes
from fractions import Fraction

import numpy as np
import pytest

from harness import ORACLE_SO, Clip, compare, make_clip, oracle_runner

BW = 16
OW = 4
STEP = BW - OW
P = 8            # period in pixels: 2 cycles per 16-pixel block, bin kx = 2
KX = BW // P
W = 97           # W - 1 = 96 is a multiple of P / 2: the reflections continue the cosine
H = 52
N = BW * BW
TOL = 2e-5


def fclip(frames):
    """Gray float clip from a list of 2D float arrays."""
    return Clip("gray", 32, True, 0, 0, frames[0].shape[1], frames[0].shape[0],
                [[np.ascontiguousarray(f, dtype=np.float32)] for f in frames])


def stripes(M, A, w=W, h=H, period=P, phase=0.0):
    x = np.arange(w, dtype=np.float64)
    row = M + A * np.cos(2 * np.pi * x / period + phase)
    return np.tile(row, (h, 1))


def wgain(psd, sig2, lowlimit=0.0):
    """Wiener gain of one bin: max((psd - sigma^2) / psd, (beta-1)/beta)."""
    return max((psd - sig2) / psd, lowlimit)


def sig2_float(sigma):
    """sigma^2 / norm for float planes (sigma scaled by 1/255)."""
    return (sigma / 255.0) ** 2 * N


def args(**kw):
    base = dict(bw=BW, bh=BW, ow=OW, oh=OW, wintype=2)
    base.update(kw)
    return ";".join(f"{k}={v}" for k, v in base.items())


def out0(h, n):
    return h.get(n)[0].astype(np.float64)


# --------------------------------------------------------------- Wiener 2D
@pytest.mark.parametrize("degrid", [0, 1])
@pytest.mark.parametrize("sigma,beta", [(250.0, 1.0), (120.0, 1.0), (400.0, 1.5)])
def test_wiener2d_cosine_gain(degrid, sigma, beta):
    M, A = 0.5, 0.2
    clip = fclip([stripes(M, A)] * 2)
    h = oracle_runner().open(args(sigma=sigma, beta=beta, bt=1, degrid=degrid), clip)
    s2 = sig2_float(sigma)
    ll = (beta - 1) / beta
    gk = wgain((A * N / 2) ** 2, s2, ll)
    # degrid subtracts the DC exactly (flat window: the grid spectrum is DC only)
    g0 = 1.0 if degrid else wgain((M * N) ** 2, s2, ll)
    want = stripes(M * g0, A * gk)
    for n in range(2):
        assert np.max(np.abs(out0(h, n) - want)) < TOL, (n, gk)


def test_wiener2d_vertical_cosine_gain():
    """Content along y: bin (ky, 0); with ky = 6 the bins ky and bh - ky both
    carry it (full ky range of the r2c layout)."""
    M, A, ky = 0.5, 0.2, 6
    hh = 53  # H - 1 multiple of 4: the reflection continues cos(2 pi 6 y / 16)
    y = np.arange(hh, dtype=np.float64)
    img = np.tile((M + A * np.cos(2 * np.pi * ky * y / BW))[:, None], (1, 60))
    clip = fclip([img])
    sigma = 250.0
    h = oracle_runner().open(args(sigma=sigma, bt=1, degrid=1), clip)
    gk = wgain((A * N / 2) ** 2, sig2_float(sigma))
    want = np.tile((M + A * gk * np.cos(2 * np.pi * ky * y / BW))[:, None], (1, 60))
    assert np.max(np.abs(out0(h, 0) - want)) < TOL


# ------------------------------------------------------------- lower limit
@pytest.mark.parametrize("bt", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("wintype", [0, 2])
def test_wiener_floor_halves_every_bin(bt, wintype):
    """beta = 2 and a huge sigma: every bin (DC included) at the lower limit
    0.5, so the output is x / 2 (+0.5, truncated); chroma is centred on 128
    before the transform and re-based after, so chroma gives 128 + (x-128)/2."""
    clip = make_clip("yuv", 8, False, width=96, height=64, nframes=5, seed=9, noise=20)
    h = oracle_runner().open(f"sigma=1e5;beta=2;bt={bt};bw=16;bh=16;degrid=0;wintype={wintype}", clip)
    for n in range(5):
        got = h.get(n)
        for p in range(3):
            x = clip.frames[n][p].astype(np.float64)
            base = 0.0 if p == 0 else 128.0
            want = np.trunc(0.5 * (x - base) + 0.5 + base)
            d = np.abs(got[p] - want)
            # x / 2 + 0.5 lands on an integer for odd x: the FP64 round trip
            # decides it either way there; even x are half a unit away
            even = (clip.frames[n][p] % 2) == 0
            assert d.max() <= 1 and np.all(d[even] == 0), (n, p, d.max())


def test_wiener_floor_with_crop_copies_borders():
    clip = make_clip("yuv", 8, False, width=96, height=64, nframes=3, seed=3, noise=20)
    l, t, r, b = 8, 4, 12, 6
    h = oracle_runner().open(f"sigma=1e5;beta=2;bt=3;bw=16;bh=16;degrid=0;l={l};t={t};r={r};b={b}", clip)
    for n in range(3):
        got = h.get(n)
        for p in range(3):
            x = clip.frames[n][p].astype(np.float64)
            s = 0 if p == 0 else 1
            L, T, R, B = l >> s, t >> s, r >> s, b >> s
            base = 0.0 if p == 0 else 128.0
            want = x.copy()
            inner = x[T:x.shape[0] - B, L:x.shape[1] - R]
            want[T:x.shape[0] - B, L:x.shape[1] - R] = np.trunc(0.5 * (inner - base) + 0.5 + base)
            d = np.abs(got[p] - want)
            even = (clip.frames[n][p] % 2) == 0
            assert d.max() <= 1 and np.all(d[even] == 0), (n, p)
            border = np.ones_like(x, dtype=bool)
            border[T:x.shape[0] - B, L:x.shape[1] - R] = False
            assert np.array_equal(got[p][border], clip.frames[n][p][border])


# -------------------------------------------------------- temporal Wiener
def modulated(M, A, eps, bt, nframes, psi=0.3):
    """Frame T: M + A (1 + eps cos(2 pi T / bt + psi)) cos(2 pi x / P)."""
    return [stripes(M, A * (1 + eps * np.cos(2 * np.pi * T / bt + psi))) for T in range(nframes)]


@pytest.mark.parametrize("bt", [2, 3, 4, 5])
def test_temporal_wiener_closed_form(bt):
    """Per bin k the window's values are X_k m(t), m(t) = 1 + eps cos(2 pi t /
    bt + phi): temporal bin 0 is bt X_k, bins +-1 are (bt eps / 2) e^(+-i phi)
    X_k (for bt = 2 the one bin 1 is 2 eps cos(phi) X_k), the others zero.
    With sigma_3D^2 = bt sigma^2 / norm the current frame's output is
    X_k (W_0 + eps W_1 cos phi); degrid keeps the DC exactly."""
    M, A, eps, psi = 0.5, 0.2, 0.5, 0.3
    F = bt + 4
    frames = modulated(M, A, eps, bt, F, psi)
    sigma = 0.4 * 255
    h = oracle_runner().open(args(sigma=sigma, bt=bt, degrid=1), fclip(frames))
    s3 = bt * sig2_float(sigma)
    xk = A * N / 2
    for n in range(bt // 2, F - (bt - 1) // 2):
        phi = 2 * np.pi * n / bt + psi
        w0 = wgain((bt * xk) ** 2, s3)
        if bt == 2:
            f1 = 2 * eps * abs(np.cos(phi)) * xk
        else:
            f1 = bt * eps / 2 * xk
        w1 = wgain(f1 ** 2, s3)
        gain = w0 + eps * w1 * np.cos(phi)
        want = stripes(M, A * gain)
        assert np.max(np.abs(out0(h, n) - want)) < TOL, (n, gain)


@pytest.mark.parametrize("bt", [2, 3, 4, 5])
def test_static_clip_bt_equals_2d_at_scaled_sigma(bt):
    """A static clip has no temporal AC energy: bin 0 is bt X and sigma_3D^2
    = bt sigma^2, so Wiener 3D equals Wiener 2D at sigma / sqrt(bt)."""
    base = make_clip("gray", 8, False, width=80, height=64, nframes=1, seed=5, noise=12).frames[0][0]
    clip = Clip("gray", 8, False, 0, 0, 80, 64, [[base.copy()] for _ in range(bt + 2)])
    sigma = 6.0
    h3 = oracle_runner().open(f"sigma={sigma};bt={bt};bw=16;bh=16;degrid=0", clip)
    h1 = oracle_runner().open(f"sigma={sigma / np.sqrt(bt):.9f};bt=1;bw=16;bh=16;degrid=0", clip)
    n = bt // 2
    (mx, exact, _), = compare(h3.get(n), h1.get(n), False)
    assert mx <= 1 and exact >= 0.999
    # and the filter does something at this sigma
    (mx0, _, _), = compare(h3.get(n), clip.frames[n], False)
    assert mx0 >= 2


def test_static_clip_large_sigma_is_temporal_mean():
    """Huge sigma, beta = 2, static clip: the DC temporal bin is halved like
    every other, so bt = 3 gives the frame halved -- the temporal mean of the
    (identical) window halved, not a mix with neighbours."""
    base = make_clip("gray", 8, False, width=64, height=48, nframes=1, seed=2).frames[0][0]
    clip = Clip("gray", 8, False, 0, 0, 64, 48, [[base.copy()] for _ in range(5)])
    h = oracle_runner().open("sigma=1e5;beta=2;bt=3;bw=16;bh=16;degrid=0", clip)
    got = h.get(2)[0].astype(np.float64)
    want = np.trunc(0.5 * base.astype(np.float64) + 0.5)
    even = (base % 2) == 0
    assert np.abs(got - want).max() <= 1 and np.all((got == want)[even])


# ------------------------------------------------------- sharpen / dehalo
def wsharpen_at(ky, kx, bw, bh, scutoff, svr):
    """src/fft3d_engine.h:285-299 at one bin (ky in 0..bh-1, kx in 0..bw/2)."""
    dj = bh - ky if ky >= bh // 2 else ky
    d2 = (dj * svr / (bh // 2)) ** 2 + (kx / (bw // 2)) ** 2
    return 1 - np.exp(-d2 / (2 * scutoff * scutoff))


def wdehalo_at(ky, kx, bw, bh, hr, svr):
    """src/fft3d_engine.h:301-327: exp(-0.7 d^2 hr^2) - exp(-d^2 hr^2),
    normalised by its maximum over the table."""
    def raw(j, i):
        dj = bh - j if j >= bh // 2 else j
        d2 = (dj * svr / (bh // 2)) ** 2 + (i / (bw // 2)) ** 2
        return np.exp(-0.7 * d2 * hr * hr) - np.exp(-d2 * hr * hr)
    mx = max(raw(j, i) for j in range(bh) for i in range(bw // 2 + 1))
    return raw(ky, kx) / mx


@pytest.mark.parametrize("bt", [-1, 1])
@pytest.mark.parametrize("sharpen,dehalo", [(0.8, 0.0), (0.0, 1.5), (0.6, 1.0)])
def test_sharpen_dehalo_factor_closed_form(bt, sharpen, dehalo):
    """Sharpen-only (bt=-1) multiplies bin k by s_fact * d_fact; Wiener 2D
    without a pattern folds the same factor into its gain.  svr != 1 makes
    the table's (ky, kx) orientation matter."""
    M, A = 0.5, 0.15
    # ht is not scaled by the bit depth (only sigma*, smin, smax are,
    # src/fft3d_engine.h:92-103): ht = 0.05 is small next to float powers
    sigma, smin, smax, scutoff, svr, hr, ht = 60.0, 4.0, 20.0, 0.3, 2.0, 2.0, 0.05
    clip = fclip([stripes(M, A)])
    h = oracle_runner().open(args(sigma=sigma, bt=bt, degrid=1, sharpen=sharpen, dehalo=dehalo, smin=smin,
                                  smax=smax, scutoff=scutoff, svr=svr, hr=hr, ht=ht), clip)
    psd = (A * N / 2) ** 2
    smin_n, smax_n, ht2n = sig2_float(smin), sig2_float(smax), ht * ht * N
    f = 1.0
    if sharpen:
        f *= 1 + sharpen * wsharpen_at(0, KX, BW, BW, scutoff, svr) * np.sqrt(
            psd * smax_n / ((psd + smin_n) * (psd + smax_n)))
    if dehalo:
        f *= (psd + ht2n) / ((psd + ht2n) + dehalo * wdehalo_at(0, KX, BW, BW, hr, svr) * psd)
    if bt == 1:
        f *= wgain(psd, sig2_float(sigma))
    want = stripes(M, A * f)
    assert np.max(np.abs(out0(h, 0) - want)) < TOL, f


# ------------------------------------------------------------------ Kalman
def test_kalman_gains_on_static_clip_then_reset():
    """Frame 0 passes through; the state starts at outLast = 0, covariances
    sigma^2 / norm.  The DC resets on frame 1 (|X_0|^2 > sigma^2 k^2) and then
    stays exact; the cosine bin (real in every block: blocks start at
    multiples of P/2) does not reset, so its gains are K1 = 2/3 and
    K2 = 10/19 (output 16/19 of the input), K3 from the exact recursion.
    A jump larger than kratio * sigma on frame 4 resets: output = input."""
    M, A = 0.5, 0.2
    sigma, kratio = 255.0, 2.0
    A_jump = A + 0.25
    frames = [stripes(M, A)] * 4 + [stripes(M, A_jump)] * 2
    h = oracle_runner().open(args(sigma=sigma, bt=0, kratio=kratio), fclip(frames))
    assert np.max(np.abs(out0(h, 0) - frames[0])) < 1e-7
    s2 = Fraction(1)  # gains depend on ratios to sigma^2 only
    P_, Q_ = s2, s2
    last = Fraction(0)
    amps = []
    for n in range(1, 4):
        K = (P_ + Q_) / (P_ + Q_ + s2)
        out = K * 1 + (1 - K) * last  # in units of the input amplitude
        Q_, P_, last = K * K * s2, (1 - K) * (P_ + Q_), out
        amps.append(out)
    assert amps[0] == Fraction(2, 3) and amps[1] == Fraction(16, 19)
    xk = A * N / 2
    assert xk ** 2 <= sig2_float(sigma) * kratio ** 2  # no reset of the cosine bin
    for n, a in zip(range(1, 4), amps):
        want = stripes(M, A * float(a))
        assert np.max(np.abs(out0(h, n) - want)) < TOL, (n, a)
    # the jump: (A_jump - last) N / 2 exceeds kratio * sigma -> reset, output = input
    assert ((A_jump - A * float(amps[-1])) * N / 2) ** 2 > sig2_float(sigma) * kratio ** 2
    assert np.max(np.abs(out0(h, 4) - frames[4])) < TOL
    # next frame from the reset state: gain 2/3 on the unchanged input -> input
    assert np.max(np.abs(out0(h, 5) - frames[5])) < TOL


# ----------------------------------------------------------------- pattern
def _pattern(h, plane, bh, outwidth):
    lib = ctypes.CDLL(ORACLE_SO)
    lib.nf3d_oracle_pattern.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    out = np.zeros(bh * outwidth)
    pxpy = (ctypes.c_int * 2)()
    rc = lib.nf3d_oracle_pattern(ctypes.c_void_p(h.h), plane, out.ctypes.data_as(ctypes.c_void_p),
                                 ctypes.cast(pxpy, ctypes.c_void_p))
    assert rc == 0
    return out.reshape(bh, outwidth), (pxpy[0], pxpy[1])


def block_region(bx, by):
    """Plane rows / columns of block (bx, by)'s support (cover minus mirror)."""
    return slice(by * STEP - STEP, by * STEP - STEP + BW), slice(bx * STEP - STEP, bx * STEP - STEP + BW)


def test_find_pattern_block_picks_flat_block_first_of_ties():
    M, A = 0.5, 0.2
    img = stripes(M, A, w=W, h=97)
    ys, xs = block_region(5, 4)
    img[ys, xs] = M
    h = oracle_runner().open(args(sigma=2, bt=1, pfactor=1, wintype=0), fclip([img]))
    h.get(0)
    pat, (px, py) = _pattern(h, 0, BW, BW // 2 + 1)
    assert (px, py) == (5, 4)
    assert np.max(pat) < 1e-12 * (M * N) ** 2
    # two identical flat blocks: the first in (by, bx) scan order wins
    img2 = img.copy()
    ys2, xs2 = block_region(3, 6)
    img2[ys2, xs2] = M
    h2 = oracle_runner().open(args(sigma=2, bt=1, pfactor=1, wintype=0), fclip([img2]))
    h2.get(0)
    _, (px2, py2) = _pattern(h2, 0, BW, BW // 2 + 1)
    assert (px2, py2) == (5, 4)
    img3 = img.copy()
    ys3, xs3 = block_region(2, 4)
    img3[ys3, xs3] = M
    h3 = oracle_runner().open(args(sigma=2, bt=1, pfactor=1, wintype=0), fclip([img3]))
    h3.get(0)
    _, (px3, py3) = _pattern(h3, 0, BW, BW // 2 + 1)
    assert (px3, py3) == (2, 4)


def far_mask(bx, by, shape):
    """Pixels no block overlapping block (bx, by)'s support covers."""
    m = np.ones(shape, dtype=bool)
    for yy in range(by - 1, by + 2):
        for xx in range(bx - 1, bx + 2):
            ys, xs = block_region(xx, yy)
            m[max(ys.start, 0):max(ys.stop, 0), max(xs.start, 0):max(xs.stop, 0)] = False
    return m


@pytest.mark.parametrize("bt", [1, 3])
def test_pattern_from_given_block_gain(bt):
    """SetPattern at (px, py): pattern_k = |X_k - g G_k|^2 pwin_k with the
    block holding amplitude A' there, so bins elsewhere (amplitude A) get
    W = 1 - (A'/A)^2 pwin_k (2D) or 1 - (A'/A)^2 pwin_k / bt (static 3D:
    bin 0 = bt X, pattern3d = bt pattern); pwin_k = fw^2 / (fw^2 + pcutoff^2),
    fw = 2 kx / bw."""
    M, A, A2, pcut = 0.5, 0.2, 0.1, 0.1
    img = stripes(M, A, w=W, h=97)
    px, py = 4, 3
    ys, xs = block_region(px, py)
    img[ys, xs] = stripes(M, A2, w=W, h=97)[ys, xs]
    frames = [img] * (bt + 1)
    h = oracle_runner().open(args(sigma=2, bt=bt, pfactor=1, px=px, py=py, pcutoff=pcut, degrid=1), fclip(frames))
    n = bt // 2
    got = out0(h, n)
    pat, choice = _pattern(h, 0, BW, BW // 2 + 1)
    assert choice == (px, py)
    fw = 2 * KX / BW
    pwin = fw * fw / (fw * fw + pcut * pcut)
    assert abs(pat[0, KX] - (A2 * N / 2) ** 2 * pwin) < 1e-6 * (A2 * N / 2) ** 2  # pwin is a float table
    gain = 1 - (A2 / A) ** 2 * pwin / bt
    want = stripes(M, A * gain, w=W, h=97)
    m = far_mask(px, py, img.shape)
    assert m.sum() > 0.5 * m.size
    assert np.max(np.abs(got - want)[m]) < TOL


def test_sigmas_to_pattern_inner_region():
    """sigma2 = sigma3 = sigma4 = s' != sigma: the frequency profile is s' for
    f < ft2 = sqrt(0.5)/2 (src/helper.cpp:14-43); bin (0, kx = 2) has
    f = (2/9)/sqrt(2) < ft2 and the DC f = 0, so both are filtered at s'."""
    M, A = 0.5, 0.2
    sp, sigma = 200.0, 40.0
    h = oracle_runner().open(args(sigma=sigma, sigma2=sp, sigma3=sp, sigma4=sp, bt=1, degrid=0), fclip([stripes(M, A)]))
    s2 = sig2_float(sp)
    want = stripes(M * wgain((M * N) ** 2, s2), A * wgain((A * N / 2) ** 2, s2))
    assert np.max(np.abs(out0(h, 0) - want)) < TOL


def test_sigmas_to_pattern_outer_region():
    """Bin (ky = 6, 0): fy = (bh - 2|6 - bh/2|)/bh = 3/4, f = 0.75/sqrt(2)
    >= ft2, so s = sigma + (sigma2 - sigma) (1 - f) / (1 - ft2)."""
    M, A, ky = 0.5, 0.2, 6
    hh = 53
    y = np.arange(hh, dtype=np.float64)
    img = np.tile((M + A * np.cos(2 * np.pi * ky * y / BW))[:, None], (1, 60))
    sigma, s2v = 150.0, 260.0
    h = oracle_runner().open(args(sigma=sigma, sigma2=s2v, sigma3=s2v, sigma4=s2v, bt=1, degrid=1), fclip([img]))
    f = np.sqrt((0.75 ** 2) * 0.5)
    ft2 = np.sqrt(0.5) / 2
    s = sigma + (s2v - sigma) * (1 - f) / (1 - ft2)
    gk = wgain((A * N / 2) ** 2, sig2_float(s))
    want = np.tile((M + A * gk * np.cos(2 * np.pi * ky * y / BW))[:, None], (1, 60))
    assert np.max(np.abs(out0(h, 0) - want)) < 1e-4  # s enters through float tables


# -------------------------------------------------------------- interlaced
def test_interlaced_fields_filtered_separately():
    """Even rows M1 + A1 cos, odd rows M2 + A2 cos: interlaced=1 stacks the
    fields, so a block inside one field sees pure content and its pixels get
    that field's closed form; without it the rows alternate and the result
    differs."""
    M1, A1, M2, A2 = 0.4, 0.2, 0.6, 0.1
    hh = 96
    img = np.empty((hh, W))
    img[0::2] = stripes(M1, A1, h=hh // 2)
    img[1::2] = stripes(M2, A2, h=hh // 2)
    sigma = 250.0
    h = oracle_runner().open(args(sigma=sigma, bt=1, degrid=1, interlaced=1), fclip([img]))
    got = out0(h, 0)
    s2 = sig2_float(sigma)
    e1 = stripes(M1, A1 * wgain((A1 * N / 2) ** 2, s2), h=1)[0]
    e2 = stripes(M2, A2 * wgain((A2 * N / 2) ** 2, s2), h=1)[0]
    half = hh // 2
    # stacked row r < half is field 1 (output row 2r); pure where every block
    # covering it lies in the field (cover rows of those blocks < half + mirror)
    checked = 0
    for r in range(half):
        cy = r + STEP
        bys = [b for b in range(cy // STEP + 2) if b * STEP <= cy < b * STEP + BW]
        if all(b * STEP + BW - STEP <= half for b in bys):
            assert np.max(np.abs(got[2 * r] - e1)) < TOL, r
            checked += 1
        rr = hh - 1 - r  # stacked row of field 2, counted from the bottom of the stack
        cy2 = rr + STEP
        bys2 = [b for b in range(cy2 // STEP + 2) if b * STEP <= cy2 < b * STEP + BW]
        if all(b * STEP - STEP >= half for b in bys2):
            y_out = hh - 1 - 2 * (rr - half)
            assert np.max(np.abs(got[y_out] - e2)) < TOL, rr
            checked += 1
    assert checked >= hh // 3
    plain = oracle_runner().open(args(sigma=sigma, bt=1, degrid=1, interlaced=0), fclip([img]))
    assert np.max(np.abs(out0(plain, 0) - got)) > 1e-3


# -------------------------------------------------------------------- alpha
@pytest.mark.parametrize("family,ss", [("yuva", 1), ("yuva", 0), ("rgba", 0)])
def test_alpha_plane_copied_unless_named(family, ss):
    """A fourth plane is copied by default (src/fft3d.hpp:149-169: it is
    never in the default plane list); named in `planes`, it is filtered with
    the chroma geometry and centring for YUV (:186-191) -- here 4:4:4, where
    that is the whole plane -- and like luma for RGB: the lower-limit check
    gives 128 + (x - 128)/2 and x/2."""
    clip = make_clip(family, 8, False, ssw=ss, ssh=ss, width=96, height=64, nframes=3, seed=6, noise=20)
    h = oracle_runner().open("sigma=1e5;beta=2;bt=3;bw=16;bh=16;degrid=0", clip)
    for n in range(3):
        assert np.array_equal(h.get(n)[3], clip.frames[n][3])
    if family == "yuva" and ss:
        from harness import FilterError
        with pytest.raises(FilterError, match="alpha plane of a subsampled YUVA"):
            oracle_runner().open("sigma=1e5;beta=2;bt=3;bw=16;bh=16;planes=0,1,2,3", clip)
        return
    h = oracle_runner().open("sigma=1e5;beta=2;bt=3;bw=16;bh=16;degrid=0;planes=0,1,2,3", clip)
    base = 128.0 if family == "yuva" else 0.0
    for n in range(3):
        got = h.get(n)[3]
        x = clip.frames[n][3].astype(np.float64)
        want = np.trunc(0.5 * (x - base) + 0.5 + base)
        even = (clip.frames[n][3] % 2) == 0
        d = np.abs(got - want)
        assert d.max() <= 1 and np.all(d[even] == 0), (n, d.max())


# --------------------------------------------------- per-region patterns
def _set_regions(h, plane, rx, ry):
    lib = ctypes.CDLL(ORACLE_SO)
    lib.nf3d_oracle_set_pattern_regions.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    assert lib.nf3d_oracle_set_pattern_regions(ctypes.c_void_p(h.h), plane, rx, ry) == 0


def test_region_patterns_whole_grid_equal_global():
    clip = make_clip("gray", 8, False, width=160, height=128, nframes=5, seed=14, noise=10)
    args = "sigma=3;bt=3;bw=16;bh=16;pfactor=1"
    hg = oracle_runner().open(args, clip)
    hr = oracle_runner().open(args, clip)
    _set_regions(hr, 0, 100, 100)
    for n in range(5):
        (mx, fe, _), = compare(hr.get(n), hg.get(n), False)
        assert mx == 0, n


def test_region_patterns_each_region_its_own_flat_block():
    """Two regions (left / right half of the block grid), each with one
    planted low-amplitude block: blocks of each region are filtered with that
    region's pattern, W = 1 - (A'/A)^2 pwin_k at the cosine bin."""
    M, A, pcut = 0.5, 0.2, 0.1
    w, hh = 193, 97   # w - 1 multiple of P/2
    img = stripes(M, A, w=w, h=hh)
    plant = {(4, 4): 0.1, (13, 5): 0.06}
    for (bx, by), a2 in plant.items():
        ys, xs = block_region(bx, by)
        img[ys, xs] = stripes(M, a2, w=w, h=hh)[ys, xs]
    h = oracle_runner().open(args(sigma=2, bt=1, pfactor=1, pcutoff=pcut, degrid=1), fclip([img]))
    nox = (w - OW + STEP - 1) // STEP + 2
    _set_regions(h, 0, 9, 100)
    got = out0(h, 0)
    fw = 2 * KX / BW
    pwin = fw * fw / (fw * fw + pcut * pcut)
    x = np.arange(w)
    checked = 0
    for (bx, by), a2 in plant.items():
        region = range(0, 9) if bx < 9 else range(9, nox)
        want = M + A * (1 - (a2 / A) ** 2 * pwin) * np.cos(2 * np.pi * x / P)
        m = far_mask(bx, by, img.shape)
        # pixels all of whose blocks lie in this region
        cols = np.zeros(w, bool)
        for c in range(w):
            cx = c + STEP
            bxs = [b for b in range(nox) if b * STEP <= cx < b * STEP + BW]
            cols[c] = all(b in region for b in bxs)
        m &= cols[None, :]
        assert m.sum() > 1000
        assert np.max(np.abs(got - want[None, :])[m]) < TOL, (bx, by)
        checked += 1
    assert checked == 2
