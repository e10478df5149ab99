"""Learned noise PSD on the GPU: the BF16 matrix-core network against the FP64
oracle (oracle/nf3d_mlp_oracle.py), Adam and training against closed forms,
the engine's per-block pattern path against the plain sigma path, and the
PSNR the trained network gains over the best global sigma.

Bars (BF16 operands, FP32 accumulation):
  forward      RMS error <= 2e-3 and max error <= 2e-2 of the largest |y|,
               against the oracle evaluated with the same BF16-rounded
               operands (weights, inputs of each product)
  gradients    relative Frobenius error <= 3e-2 per parameter block and
               cosine >= 0.999 against the oracle with the same rounding
  Adam         one step equals the textbook update of the GPU's own gradients
               to 1e-6 relative
  training     the mean prediction reaches ln(mean t) (the Whittle loss's
               closed-form minimiser) within 0.1 for targets independent of x
  engine       a constant learned PSD c filters like sigma with c = sigma^2 *
               bw * bh (float planes: <= 2e-5 absolute)
  oracle       the engine's learned-PSD filtering against the FP64 oracle given
               the same per-block PSDs: float <= 1e-4, 8-bit +-1 LSB with
               >= 99.9% exact
  PSNR         the trained network gains >= 0.5 dB on average (and > 0 on every
               clip) over the best global sigma on coloured, spatially varying
               noise; sigma and strength tuned on held-out clips
"""
import os
import sys

import numpy as np
import pytest

from harness import ROOT

sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import nf3d_mlp_oracle as mo  # noqa: E402

pytestmark = pytest.mark.gpu

D = 48 * 25  # 48x48 blocks: 48 rows x 25 compact columns
NH = 256


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


def rand_params(n_in, nh, n_out, seed):
    rng = np.random.default_rng(seed)
    parts = [rng.standard_normal((n_in, nh)) * np.sqrt(2.0 / n_in), rng.standard_normal(nh) * 0.1,
             rng.standard_normal((nh, nh)) * np.sqrt(2.0 / nh), rng.standard_normal(nh) * 0.1,
             rng.standard_normal((nh, n_out)) * np.sqrt(1.0 / nh), rng.standard_normal(n_out) * 0.5]
    return [p.astype(np.float32).astype(np.float64) for p in parts]


def upload_bf16(torch, x):
    from neo_fft3d_amd.learned import to_bf16
    xf = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).cuda()
    xb = torch.empty(x.shape, dtype=torch.bfloat16, device="cuda")
    to_bf16(xf.data_ptr(), xb.data_ptr(), x.size)
    return xb


def make(n_in, nh, n_out, batch, params):
    from neo_fft3d_amd.learned import Mlp
    m = Mlp(n_in, nh, n_out, batch)
    m.set_params(mo.pack(params).astype(np.float32))
    return m


def predict(torch, m, xb, rows, exp_out=False, scale=1.0):
    out = torch.empty((rows, m.n_out), dtype=torch.float32, device="cuda")
    work = torch.empty(m.work_bytes(rows), dtype=torch.uint8, device="cuda")
    m.predict(xb.data_ptr(), rows, out.data_ptr(), work.data_ptr(), exp_out, scale)
    torch.cuda.synchronize()
    return out.cpu().numpy().astype(np.float64)


# the input widths pick the inference kernel: 1200 (48x48 blocks) the
# eight-phase one, 40 and the hidden widths the register-weight one, 2112
# (64x64 blocks) the four-phase one; the row counts its tile shapes
@pytest.mark.parametrize("dims,rows", [((D, NH, D), 300), ((40, 24, 56), 37), ((D, NH, D), 4096),
                                       ((64 * 33, 64, 128), 200)])
def test_forward_matches_oracle(torch, dims, rows):
    n_in, nh, n_out = dims
    params = rand_params(n_in, nh, n_out, seed=rows)
    x = np.random.default_rng(rows + 1).standard_normal((rows, n_in))
    m = make(n_in, nh, n_out, rows, params)
    xb = upload_bf16(torch, x)
    got = predict(torch, m, xb, rows)
    want = mo.forward(params, x, bf16=True)["y"]
    err = np.abs(got - want)
    ref = np.abs(want).max()
    print(f"forward {dims} rows={rows}: max {err.max() / ref:.2e} rms {np.sqrt((err ** 2).mean()) / ref:.2e} "
          f"of max|y|; vs FP64 without rounding max {np.abs(got - mo.forward(params, x)['y']).max() / ref:.2e}",
          flush=True)
    assert np.sqrt((err ** 2).mean()) <= 2e-3 * ref
    assert err.max() <= 2e-2 * ref
    # exp epilogue: exp(y) * scale
    gexp = predict(torch, m, xb, rows, exp_out=True, scale=3.0)
    rel = np.abs(gexp / (3.0 * np.exp(got)) - 1.0).max()
    print(f"exp epilogue max rel {rel:.2e}", flush=True)
    assert rel <= 1e-5


def piece_errors(got, want, dims):
    n_in, nh, n_out = dims
    sizes = [n_in * nh, nh, nh * nh, nh, nh * n_out, n_out]
    names = ["W1", "b1", "W2", "b2", "W3", "b3"]
    out, at = {}, 0
    for nm, s in zip(names, sizes):
        g, w = got[at:at + s], want[at:at + s]
        out[nm] = np.linalg.norm(g - w) / max(np.linalg.norm(w), 1e-30)
        at += s
    return out


@pytest.mark.parametrize("dims,rows", [((D, NH, D), 200), ((40, 24, 56), 70)])
def test_gradients_match_oracle(torch, dims, rows):
    n_in, nh, n_out = dims
    params = rand_params(n_in, nh, n_out, seed=11 + rows)
    rng = np.random.default_rng(rows)
    x = rng.standard_normal((rows, n_in))
    t = (rng.exponential(1.0, (rows, n_out)) * np.exp(rng.standard_normal((1, n_out)))).astype(np.float32)
    m = make(n_in, nh, n_out, rows, params)
    xb = upload_bf16(torch, x)
    tt = torch.from_numpy(t).cuda()
    loss = m.grad(xb.data_ptr(), tt.data_ptr(), rows)
    got = m.get_grads().astype(np.float64)
    want = mo.backward(params, x, t.astype(np.float64), bf16=True)
    want64 = mo.backward(params, x, t.astype(np.float64), bf16=False)
    e = piece_errors(got, want, dims)
    e64 = piece_errors(got, want64, dims)
    cos = float(got @ want / (np.linalg.norm(got) * np.linalg.norm(want)))
    want_loss = mo.whittle_loss(mo.forward(params, x, bf16=True)["y"], t.astype(np.float64))
    print(f"grads {dims}: rel Frobenius vs oracle(BF16 operands) " +
          " ".join(f"{k}={v:.2e}" for k, v in e.items()) + f"; vs FP64 " +
          " ".join(f"{k}={v:.2e}" for k, v in e64.items()) + f"; cosine {cos:.6f}; loss {loss:.6f} vs {want_loss:.6f}",
          flush=True)
    assert max(e.values()) <= 3e-2
    assert cos >= 0.999
    assert abs(loss - want_loss) <= 1e-4 * abs(want_loss) + 1e-6


def test_adam_step_is_textbook(torch):
    dims = (40, 24, 56)
    params = rand_params(*dims, seed=3)
    rows = 50
    rng = np.random.default_rng(4)
    x = rng.standard_normal((rows, dims[0]))
    t = rng.exponential(1.0, (rows, dims[2])).astype(np.float32)
    m = make(*dims, rows, params)
    xb = upload_bf16(torch, x)
    tt = torch.from_numpy(t).cuda()
    p0 = m.get_params().astype(np.float64)
    m.grad(xb.data_ptr(), tt.data_ptr(), rows)
    g = m.get_grads().astype(np.float64)
    lr = 1e-3
    m.adam(lr)
    p1 = m.get_params().astype(np.float64)
    mm = 0.1 * g
    vv = 0.001 * g * g
    want = p0 - lr * (mm / 0.1) / (np.sqrt(vv / 0.001) + 1e-8)
    err = np.abs(p1 - want).max()
    print(f"adam step max err {err:.2e} (lr {lr})", flush=True)
    assert err <= 1e-6 * max(1.0, np.abs(p0).max())


def test_training_reaches_whittle_minimiser(torch):
    """Targets independent of the input (a fresh exponential draw of the same
    PSD S each step): the network must learn y = ln S, the Whittle loss's
    minimiser, for every row.  (The draw uses torch's sampler; every product
    of the network is a library kernel.)"""
    from neo_fft3d_amd.learned import Mlp, init_params
    import math
    rows, steps = 4096, 800
    rng = np.random.default_rng(9)
    S = np.exp(rng.standard_normal(D))
    x = rng.standard_normal((rows, D))
    S_d = torch.from_numpy(S.astype(np.float32)).cuda()
    gen = torch.Generator(device="cuda")
    gen.manual_seed(9)
    tt = torch.empty((rows, D), dtype=torch.float32, device="cuda")

    def draw():  # a fresh periodogram sample t = S * Exp(1) per step: nothing to memorise
        tt.exponential_(generator=gen)
        tt.mul_(S_d)

    m = Mlp(D, NH, D, rows)
    m.set_params(init_params(D, NH, D, seed=1))
    xb = upload_bf16(torch, x)
    draw()
    first = m.train_step(xb.data_ptr(), tt.data_ptr(), rows, 5e-3, want_loss=True)
    for s in range(1, steps):
        lr = 5e-3 * (0.02 + 0.98 * 0.5 * (1 + math.cos(math.pi * s / steps)))  # cosine decay
        draw()
        last = m.train_step(xb.data_ptr(), tt.data_ptr(), rows, lr, want_loss=(s == steps - 1))
    y = predict(torch, m, xb, rows)
    t_last = tt.cpu().numpy().astype(np.float64)
    floor = mo.whittle_loss(np.broadcast_to(np.log(S), t_last.shape), t_last)
    dev = np.abs(y - np.log(S)[None, :])
    print(f"training: loss {first:.4f} -> {last:.4f} (loss at y = ln S: {floor:.4f}); "
          f"|y - ln S| mean {dev.mean():.4f}, max over outputs of the row mean {dev.mean(0).max():.3f}", flush=True)
    assert last < first
    assert dev.mean() <= 0.06
    assert dev.mean(0).max() <= 0.1
    assert last <= floor + 0.01


def filter_planes(torch, planes, fmt, sigma, bt=3, mlp=None, strength=1.0, bw=48, ow=16):
    from neo_fft3d_amd import Engine, default_params
    from neo_fft3d_amd.learned import set_learned
    F, h, w = planes[0].shape
    outs = []
    for p, pl in enumerate(planes):
        eng = Engine(default_params(sigma, bt=bt, bw=bw, bh=bw, ow=ow, oh=ow, sharpen=0.0), fmt, w, h, F, p)
        if mlp is not None:
            set_learned(eng, mlp, strength)
        src = torch.from_numpy(np.ascontiguousarray(pl)).cuda()
        dst = torch.zeros_like(src)
        stride = w * src.element_size()
        torch.cuda.synchronize()  # the engine's stream does not wait for torch's
        eng.run_device(0, F, [src[f].data_ptr() for f in range(F)], stride, [dst[f].data_ptr() for f in range(F)],
                       stride)
        eng.sync()
        outs.append(dst.cpu().numpy())
        eng.close()
    return outs


@pytest.mark.parametrize("bt", [3, 1, 0])
def test_constant_learned_psd_equals_sigma_path(torch, bt):
    """W3 = 0, b3 = ln c: every block's PSD is c, which the sigma path uses as
    sigma^2 * bw * bh (per frame; bt frames of it in the 3D gain)."""
    from neo_fft3d_amd import make_format
    from neo_fft3d_amd.learned import Mlp
    sigma = 6.0
    c = (sigma / 255.0) ** 2 * 48 * 48
    params = rand_params(D, NH, D, seed=2)
    params[4] = np.zeros_like(params[4])
    params[5] = np.full(D, np.log(np.float32(c)), dtype=np.float64)
    m = Mlp(D, NH, D, 64)
    m.set_params(mo.pack(params).astype(np.float32))
    m.set_input_norm(np.zeros(D, np.float32), np.ones(D, np.float32), 1e-6)
    from neo_fft3d_amd.learned import noisy_clip
    _, noisy, _ = noisy_clip(160, 112, 5, seed=77)
    fmt = make_format("rgb", 32, True)
    plain = filter_planes(torch, noisy, fmt, sigma, bt=bt)
    learned = filter_planes(torch, noisy, fmt, sigma, bt=bt, mlp=m)
    err = max(np.abs(a - b).max() for a, b in zip(plain, learned))
    moved = max(np.abs(a - b).max() for a, b in zip(plain, noisy))
    print(f"bt={bt}: learned(constant) vs sigma path max |d| {err:.2e} (filter moved pixels by up to {moved:.3f})",
          flush=True)
    assert moved > 1e-2
    assert err <= 2e-5


def psnr(outs, clean, skip=0):
    mse = np.mean([np.mean((o[skip:len(o) - skip].astype(np.float64) - c[skip:len(c) - skip]) ** 2)
                   for o, c in zip(outs, clean)])
    return 10 * np.log10(1.0 / mse)


def test_learned_psd_gains_psnr(torch):
    """Config: RGBPS float planes, bt=3, 48/16; coloured noise with a Perlin
    sigma field (0.005-0.04).  The global sigma and the learned-PSD strength
    are each tuned on two held-out clips; the gain is the mean over four other
    clips, each with its own noise process."""
    from neo_fft3d_amd import make_format
    from neo_fft3d_amd.learned import noisy_clip, train_model
    mlp, data, losses = train_model(clips=48, frames=2, width=384, height=384, steps=1500,
                                    log=lambda s: print(s, flush=True))
    print(f"trained on {data.rows} blocks; loss {losses[0][1]:.4f} -> {losses[-1][1]:.4f}", flush=True)
    assert losses[-1][1] < losses[0][1]
    fmt = make_format("rgb", 32, True)
    W, H, F = 384, 288, 5
    tune_clips = [noisy_clip(W, H, F, seed=2000 + k)[:2] for k in range(2)]
    eval_clips = [noisy_clip(W, H, F, seed=3000 + k)[:2] for k in range(4)]

    def mean_psnr(clips, **kw):
        return float(np.mean([psnr(filter_planes(torch, n, fmt, **kw), c) for c, n in clips]))
    sig_grid = [4.0, 5.0, 6.0, 7.0, 8.0, 9.0, 10.0, 12.0, 14.0]
    str_grid = [1.5, 2.0, 2.5, 3.0, 3.5, 4.0, 5.0, 6.0]
    tune = {s: mean_psnr(tune_clips, sigma=s) for s in sig_grid}
    tune_l = {a: mean_psnr(tune_clips, sigma=2.0, mlp=mlp, strength=a) for a in str_grid}
    best_sigma = max(tune, key=tune.get)
    best_strength = max(tune_l, key=tune_l.get)
    print("held-out sigma sweep: " + ", ".join(f"{s}:{v:.2f}" for s, v in tune.items()), flush=True)
    print("held-out strength sweep: " + ", ".join(f"{s}:{v:.2f}" for s, v in tune_l.items()), flush=True)
    gains = []
    for k, (c, n) in enumerate(eval_clips):
        p_noisy = psnr(n, c)
        p_plain = psnr(filter_planes(torch, n, fmt, best_sigma), c)
        p_learn = psnr(filter_planes(torch, n, fmt, 2.0, mlp=mlp, strength=best_strength), c)
        gains.append(p_learn - p_plain)
        print(f"clip {k}: PSNR noisy {p_noisy:.2f} dB, plain sigma={best_sigma} {p_plain:.2f} dB, learned "
              f"(strength {best_strength}) {p_learn:.2f} dB: gain {p_learn - p_plain:+.2f} dB", flush=True)
    print(f"mean gain {np.mean(gains):+.2f} dB (min {min(gains):+.2f})", flush=True)
    assert np.mean(gains) >= 0.5
    assert min(gains) > 0.0


def psd_source_frame(n, bt, F, per_launch, period=0):
    """The frame whose PSD the engine uses for frame n of an in-order device
    range 0..F-1: frames of the frame-loop range take their launch's first
    frame (launches split as the engine splits them) or, with a prediction
    period, their span's first frame (spans start at multiples of the period
    and at the range's first frame); the others their own."""
    if bt >= 2:
        lo, hi = bt // 2, F - 1 - (bt - 1) // 2
    elif bt == 0:
        lo, hi = 1, F - 1
    else:
        lo, hi = 0, F - 1
    if n < lo or n > hi:
        return n
    if period > 0:
        return max(lo, n // period * period)
    total = hi - lo + 1
    nl = (total + per_launch - 1) // per_launch
    m0 = lo
    for j in range(nl):
        nf = total // nl + (1 if j >= nl - total % nl else 0)
        if n < m0 + nf:
            return m0
        m0 += nf
    raise AssertionError("frame outside the launches")


@pytest.mark.parametrize("family,bits,is_float,bt,per_launch,period,per_call", [
    ("rgb", 32, True, 3, 1, 0, 0), ("gray", 8, False, 1, 1, 0, 0), ("gray", 8, False, 0, 1, 0, 0),
    ("rgb", 32, True, 5, 1, 0, 0), ("rgb", 32, True, 3, 16, 0, 0), ("gray", 8, False, 1, 4, 0, 0),
    ("gray", 8, False, 0, 3, 0, 0), ("rgb", 32, True, 3, 2, 4, 0), ("gray", 8, False, 1, 3, 4, 0),
    ("gray", 8, False, 0, 16, 3, 0), ("rgb", 32, True, 3, 16, 4, 2), ("gray", 8, False, 0, 16, 3, 1),
    ("gray", 8, False, 1, 2, 4, 3)])
def test_learned_path_matches_oracle(torch, family, bits, is_float, bt, per_launch, period, per_call):
    """The engine's learned-PSD filtering against the FP64 oracle given the
    same per-block PSDs (the model's exp(log-PSD) of the frame the engine
    predicts from): per-block pattern indexing in both kernels, the exp
    epilogue, the btcur scaling, the one-PSD-per-launch rule and, with a
    prediction period, the one-PSD-per-span rule (launches of a span after
    its first keep its PSD, also across calls: per_call > 0 runs the clip as
    in-order calls of that many frames)."""
    import ctypes
    from harness import Clip, compare, oracle_runner
    from neo_fft3d_amd import Engine, default_params, make_format
    from neo_fft3d_amd.learned import Mlp, engine_spectra, noisy_clip
    from neo_fft3d_amd.autotune import set_tuning
    sigma, F, W, H = 6.0, 9, 160, 112
    _, noisy, _ = noisy_clip(W, H, F, seed=91)
    nplanes = 3 if family == "rgb" else 1
    if is_float:
        planes = [np.ascontiguousarray(noisy[p]) for p in range(nplanes)]
        unit = 1.0 / 255.0
    else:
        planes = [np.clip(np.rint(noisy[p] * 255.0), 0, 255).astype(np.uint8) for p in range(nplanes)]
        unit = 1.0
    clip = Clip(family, bits, is_float, 0, 0, W, H, [[planes[p][n] for p in range(nplanes)] for n in range(F)])
    fmt = make_format(family, bits, is_float)
    args = f"sigma={sigma};bt={bt};bw=48;bh=48;ow=16;oh=16"
    # a random model with realistic PSD levels that vary per block
    params = rand_params(D, NH, D, seed=13)
    params[4] = params[4] * 0.3
    params[5] = np.full(D, np.log((sigma * unit) ** 2 * 48 * 48), dtype=np.float64)
    m = Mlp(D, NH, D, 64)
    m.set_params(mo.pack(params).astype(np.float32))
    engines = [Engine(default_params(sigma, bt=bt, bw=48, bh=48, ow=16, oh=16), fmt, W, H, F, p)
               for p in range(nplanes)]
    g = engines[0].geometry()
    nb = g["nox"] * g["noy"]
    srcs = [torch.from_numpy(planes[p]).cuda() for p in range(nplanes)]
    spec = torch.empty((nb, D, 2), dtype=torch.float32, device="cuda")
    # input normalisation from the clip's own spectra
    lp = []
    for p in range(nplanes):
        for n in range(F):
            engine_spectra(engines[p], srcs[p][n].data_ptr(), srcs[p][n].stride(0) * srcs[p].element_size(),
                           spec.data_ptr())
            s = spec.cpu().numpy().astype(np.float64)
            lp.append(np.log(s[..., 0] ** 2 + s[..., 1] ** 2 + 1e-6 * unit ** 2))
    lp = np.concatenate(lp)
    m.set_input_norm(lp.mean(0), 1.0 / np.maximum(lp.std(0), 1e-3), 1e-6 * unit ** 2)
    # per-frame PSDs as the engine's learned path computes them
    xb = torch.empty((nb, D), dtype=torch.bfloat16, device="cuda")
    psd = torch.empty((nb, D), dtype=torch.float32, device="cuda")
    work = torch.empty(m.work_bytes(nb), dtype=torch.uint8, device="cuda")
    psds = {}
    for p in range(nplanes):
        stride = srcs[p][0].stride(0) * srcs[p].element_size()
        for n in range(F):
            engine_spectra(engines[p], srcs[p][n].data_ptr(), stride, spec.data_ptr())
            m.features(spec.data_ptr(), nb, xb.data_ptr())
            m.predict(xb.data_ptr(), nb, psd.data_ptr(), work.data_ptr(), exp_out=True, scale=1.0)
            torch.cuda.synchronize()
            psds[(p, n)] = np.ascontiguousarray(psd.cpu().numpy())
    spread = np.log(psds[(0, F // 2)]).std()
    # GPU learned filtering of the whole clip
    from neo_fft3d_amd.learned import set_learned
    gpu = []
    for p in range(nplanes):
        set_learned(engines[p], m, 1.0, period=period)
        set_tuning(engines[p], per_launch, 1)
        dst = torch.zeros_like(srcs[p])
        torch.cuda.synchronize()  # the engine's stream does not wait for torch's
        stride = srcs[p][0].stride(0) * srcs[p].element_size()
        step = per_call or F
        for a in range(0, F, step):
            engines[p].run_device(a, min(step, F - a), [srcs[p][n].data_ptr() for n in range(F)], stride,
                                  [dst[n].data_ptr() for n in range(F)], stride)
        engines[p].sync()
        gpu.append(dst.cpu().numpy())
    # the oracle with the same PSDs, frame by frame
    orc = oracle_runner()
    orc.lib.nf3d_oracle_set_block_psd.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    h = orc.open(args, clip)
    for n in range(F):
        src_n = psd_source_frame(n, bt, F, per_launch, period)
        for p in range(nplanes):
            assert orc.lib.nf3d_oracle_set_block_psd(h.h, p, psds[(p, src_n)].ctypes.data) == 0
        want = h.get(n)
        got = [gpu[p][n] for p in range(nplanes)]
        for p, (mx, fe, rms) in enumerate(compare(got, want, is_float)):
            print(f"learned {family}{bits} bt={bt} per_launch={per_launch} period={period} per_call={per_call} n={n} "
                  f"(PSD of frame {src_n}) p={p}: "
                  f"max {mx:.3g} exact {fe:.5f} rms {rms:.3g} "
                  f"(log-PSD spread {spread:.2f})", flush=True)
            if is_float:
                assert mx <= 1e-4
            else:
                assert mx <= 1 and fe >= 0.999
    for e in engines:
        e.close()


@pytest.mark.parametrize("period", [0, 4])
def test_graph_replays_with_learned_psd(torch, period):
    """A batched call over engines with the learned PSD, captured as a HIP
    graph: every replay (with eager work between replays) equals the eager
    call bit for bit -- the per-launch prediction (analysis, features, GEMMs)
    is part of the graph.  With a prediction period the planes' shared
    prediction is copied to each engine's own PSD, which the span's later
    launches reuse: the batched call equals the planes run one by one."""
    from neo_fft3d_amd import Engine, Graph, default_params, make_format, run_device_batch
    from neo_fft3d_amd.learned import Mlp, noisy_clip, set_learned
    F, W, H = 10, 160, 112
    _, noisy, _ = noisy_clip(W, H, F, seed=44)
    params = rand_params(D, NH, D, seed=5)
    params[4] = params[4] * 0.3
    params[5] = np.full(D, np.log((6.0 / 255.0) ** 2 * 48 * 48), dtype=np.float64)
    m = Mlp(D, NH, D, 64)
    m.set_params(mo.pack(params).astype(np.float32))
    m.set_input_norm(np.full(D, -6.0, np.float32), np.full(D, 0.3, np.float32), 1e-6)
    fmt = make_format("rgb", 32, True)
    engs = [Engine(default_params(6.0, bt=3, bw=48, bh=48, ow=16, oh=16), fmt, W, H, F, p) for p in range(3)]
    for e in engs:
        set_learned(e, m, 1.0, period=period)
    if period:
        from neo_fft3d_amd.autotune import set_tuning
        for e in engs:
            set_tuning(e, 2, 1)  # spans of 4 frames in launches of 2: every other launch reuses
    src = [torch.from_numpy(np.ascontiguousarray(noisy[p])).cuda() for p in range(3)]
    dst = [torch.zeros_like(x) for x in src]
    st = W * 4
    items = [(e, 0, F, [src[p][k].data_ptr() for k in range(F)], st, [dst[p][k].data_ptr() for k in range(F)], st)
             for p, e in enumerate(engs)]
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        run_device_batch(items, stream.cuda_stream)
    stream.synchronize()
    want = [d.cpu().numpy().copy() for d in dst]
    if period:
        for p, e in enumerate(engs):
            one = torch.zeros_like(src[p])
            with torch.cuda.stream(stream):
                run_device_batch([(e, 0, F, items[p][3], st, [one[k].data_ptr() for k in range(F)], st)],
                                 stream.cuda_stream)
            stream.synchronize()
            print(f"period {period} plane {p} alone vs batched: max diff "
                  f"{np.abs(one.cpu().numpy() - want[p]).max():.3g}", flush=True)
            assert np.array_equal(one.cpu().numpy(), want[p]), p
    g = Graph(items, stream.cuda_stream)
    for r in range(3):
        with torch.cuda.stream(stream):
            for d in dst:
                d.zero_()
            x = torch.randn(1 << 20, device="cuda")
            float(x.sum())  # eager work and a synchronisation between replays
        g.launch(stream.cuda_stream)
        stream.synchronize()
        for p in range(3):
            got = dst[p].cpu().numpy()
            print(f"learned graph replay {r} plane {p}: max diff {np.abs(got - want[p]).max():.3g}", flush=True)
            assert np.array_equal(got, want[p]), (r, p)
    g.close()
    for e in engs:
        e.close()


def test_ring_bound_graph_checks_the_kept_span(torch):
    """A ring-bound graph whose captured launches reuse a learned PSD kept
    from an earlier call (bt=1: no ring, the span is what it relies on): its
    replay right after the capture equals one eager call over the same
    frames bit for bit; once the engine has moved on by a frame count that is
    not a whole number of spans, the replay is refused and nothing runs."""
    from neo_fft3d_amd import Engine, Graph, Nf3dError, default_params, make_format, run_device_batch
    from neo_fft3d_amd.learned import Mlp, noisy_clip, set_learned
    F, W, H = 12, 160, 112
    _, noisy, _ = noisy_clip(W, H, F, seed=61)
    params = rand_params(D, NH, D, seed=8)
    params[4] = params[4] * 0.3
    params[5] = np.full(D, np.log((6.0 / 255.0) ** 2 * 48 * 48), dtype=np.float64)
    m = Mlp(D, NH, D, 64)
    m.set_params(mo.pack(params).astype(np.float32))
    m.set_input_norm(np.full(D, -6.0, np.float32), np.full(D, 0.3, np.float32), 1e-6)
    fmt = make_format("gray", 32, True)
    src = torch.from_numpy(np.ascontiguousarray(noisy[0])).cuda()
    st = W * 4

    def engine():
        e = Engine(default_params(6.0, bt=1, bw=48, bh=48, ow=16, oh=16), fmt, W, H, F, 0)
        set_learned(e, m, 1.0, period=4)
        return e

    def item(e, dst, first, count):
        return (e, first, count, [src[k].data_ptr() for k in range(F)], st, [dst[k].data_ptr() for k in range(F)], st)
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    want = torch.zeros_like(src)
    ref = engine()
    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        run_device_batch([item(ref, want, 0, 4)], sp)       # frames 0..3: the PSD of frame 0
    stream.synchronize()
    got = torch.zeros_like(src)
    e = engine()
    with torch.cuda.stream(stream):
        run_device_batch([item(e, got, 0, 2)], sp)           # span 0 predicted, kept
        g = Graph([item(e, got, 2, 2)], sp, ring=True)      # frames 2, 3 reuse it
        g.launch(sp)
    stream.synchronize()
    for n in range(4):
        assert torch.equal(got[n], want[n]), n
    with torch.cuda.stream(stream):
        run_device_batch([item(e, got, 4, 5)], sp)           # now 5 frames on: not whole spans
    stream.synchronize()
    before = got.cpu().numpy().copy()
    with pytest.raises(Nf3dError):
        g.launch(sp)
    stream.synchronize()
    assert np.array_equal(got.cpu().numpy(), before)
    g.close()
    e.close()
    ref.close()


def test_learned_period_rejects_negative(torch):
    """nf3d_engine_set_learned_period takes 0 (one prediction per launch) or
    a frame count; a negative period is refused with the library's message."""
    from neo_fft3d_amd import Engine, Nf3dError, default_params, make_format
    from neo_fft3d_amd.learned import set_learned_period
    e = Engine(default_params(2.0, bt=3, bw=48, bh=48, ow=16, oh=16), make_format("gray", 8), 160, 112, 4, 0)
    set_learned_period(e, 0)
    set_learned_period(e, 64)
    with pytest.raises(Nf3dError, match="prediction period"):
        set_learned_period(e, -1)
    e.close()


def test_trainer_block_shape_and_sample_scale(torch):
    """A network for 32x32 blocks (overlap 10) in 8-bit sample units: trained
    on the GPU, calibrated on a held-out clip -- the mean of ln t - y over its
    blocks is -0.577 (Euler's gamma) when y = ln E[t]."""
    from neo_fft3d_amd.learned import analysis_engine, engine_spectra, noisy_clip, train_model
    mlp, data, losses = train_model(clips=12, frames=2, width=256, height=256, steps=500, bw=32, ow=10,
                                    scale=255.0)
    assert mlp.n_in == 32 * 17
    assert losses[-1][1] < losses[0][1]
    W = H = 256
    _, noisy, noise = noisy_clip(W, H, 1, seed=777, scale=255.0)
    eng = analysis_engine(W, H, bw=32, ow=10)
    g = eng.geometry()
    nb = g["nox"] * g["noy"]
    spec = torch.empty((nb, 544, 2), dtype=torch.float32, device="cuda")
    xb = torch.empty((nb, 544), dtype=torch.bfloat16, device="cuda")
    y = torch.empty((nb, 544), dtype=torch.float32, device="cuda")
    work = torch.empty(mlp.work_bytes(nb), dtype=torch.uint8, device="cuda")
    r = []
    for p in range(3):
        src = torch.from_numpy(np.ascontiguousarray(noisy[p][0])).cuda()
        nz = torch.from_numpy(np.ascontiguousarray(noise[p][0])).cuda()
        engine_spectra(eng, src.data_ptr(), W * 4, spec.data_ptr())
        mlp.features(spec.data_ptr(), nb, xb.data_ptr())
        mlp.predict(xb.data_ptr(), nb, y.data_ptr(), work.data_ptr())
        engine_spectra(eng, nz.data_ptr(), W * 4, spec.data_ptr())
        s = spec.cpu().numpy().astype(np.float64)
        t = s[..., 0] ** 2 + s[..., 1] ** 2
        r.append(np.log(t + 1e-30) - y.cpu().numpy().astype(np.float64))
    r = np.concatenate(r)
    print(f"32x32 8-bit network: loss {losses[0][1]:.3f} -> {losses[-1][1]:.3f}; held-out mean(ln t - y) "
          f"{r.mean():.3f} (calibrated: -0.577)", flush=True)
    assert abs(r.mean() + 0.5772) < 0.3
    eng.close()
