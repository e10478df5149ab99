"""The device-range path (nf3d_engine_run_device) on the GPU.

run_device filters a device-resident clip in order: frames whose filter
window lies inside the clip go through the frame-loop kernel (k_run, up to 16
frames per launch, the Wiener-3D spectra ring and the Kalman state carried
between launches), the clip ends through the per-frame kernel.  Each case is
checked against the FP64 oracle at the parity bar (integer formats within
+-1 LSB with >= 99.9% of pixels exact, float within 1e-4) and against the
per-frame path of the same library (nf3d_clip), to the same bar (the two
kernels associate their transforms differently).  Clips are 40 frames
so every case spans several launches; a second run splits the clip in two
calls to check that the ring and the Kalman state carry across calls.
"""
import os

import numpy as np
import pytest

from harness import ROOT, Runner, compare, make_clip, oracle_runner

pytestmark = pytest.mark.gpu

PRODUCT_SO = os.path.join(ROOT, "neo_fft3d_amd", "libnf3d.so")

# (args, family, bits, float, width, height, frames)
CASES = [
    ("sigma=3;bt=3;bw=16;bh=16;sharpen=0.3", "gray", 8, False, 160, 128, 40),
    ("sigma=2;bt=3;bw=48;bh=48;ow=16;oh=16;sharpen=0.3", "gray", 8, False, 320, 200, 36),
    ("sigma=2;bt=3;bw=48;bh=48;ow=16;oh=16", "gray", 8, False, 320, 200, 20),
    ("sigma=3;bt=2;bw=16;bh=16;sharpen=0.5", "gray", 8, False, 160, 128, 40),
    ("sigma=3;bt=4;bw=32;bh=32;dehalo=1", "gray", 8, False, 160, 128, 40),
    ("sigma=3;bt=5;bw=16;bh=16;sharpen=0.4;dehalo=0.5", "gray", 8, False, 160, 128, 40),
    ("sigma=3;bt=5;bw=16;bh=16", "gray", 8, False, 160, 128, 24),
    ("sigma=3;bt=1;bw=16;bh=16;sharpen=0.3", "gray", 8, False, 160, 128, 40),
    ("sigma=3;bt=1;bw=64;bh=64;ow=16;oh=16", "gray", 8, False, 200, 160, 20),
    ("sigma=3;bt=0;bw=16;bh=16;sharpen=0.3", "gray", 8, False, 160, 128, 40),
    ("sigma=3;bt=0;bw=16;bh=16", "gray", 8, False, 160, 128, 24),
    ("sigma=3;bt=-1;bw=16;bh=16;sharpen=1", "gray", 8, False, 160, 128, 24),
    ("sigma=3;bt=-1;bw=16;bh=16;dehalo=1;sharpen=0.5", "gray", 8, False, 160, 128, 24),
    ("sigma=3;bt=3;bw=16;bh=16;pfactor=1", "gray", 8, False, 160, 128, 24),
    ("sigma=3;bt=1;bw=16;bh=16;pfactor=1;sharpen=0.3", "gray", 8, False, 160, 128, 24),
    ("sigma=3;bt=0;bw=16;bh=16;pfactor=1", "gray", 8, False, 160, 128, 24),
    ("sigma=3;bt=3;bw=16;bh=16;sigma2=5;sigma3=1;sigma4=0.5", "gray", 8, False, 160, 128, 24),
    ("sigma=3;bt=3;bw=16;bh=16;degrid=0;wintype=1", "gray", 8, False, 160, 128, 24),
    ("sigma=3;bt=3;bw=16;bh=16;interlaced=1", "gray", 8, False, 160, 128, 24),
    ("sigma=3;bt=3;bw=8;bh=8;ow=2;oh=2", "gray", 8, False, 96, 64, 24),
    ("sigma=3;bt=3;bw=24;bh=24;ow=8;oh=8", "gray", 8, False, 160, 128, 24),
    ("sigma=3;bt=3;bw=32;bh=16;ow=8;oh=4", "gray", 8, False, 160, 128, 24),
    ("sigma=3;bt=3;bw=16;bh=32;ow=4;oh=8", "gray", 8, False, 160, 128, 24),
    ("sigma=3;bt=3;bw=16;bh=16;sharpen=0.3", "gray", 10, False, 160, 128, 24),
    ("sigma=3;bt=0;bw=32;bh=16;ow=8;oh=4;kratio=1.5", "gray", 16, False, 160, 128, 24),
    ("sigma=3;bt=3;bw=16;bh=16;sharpen=0.3", "rgb", 32, True, 96, 64, 24),
    # the matrix-core DFT form in the frame loop (32x32 blocks)
    ("sigma=3;bt=3;bw=32;bh=32;ow=16;oh=16;sharpen=0.3;dft=1", "gray", 8, False, 160, 128, 40),
    ("sigma=3;bt=5;bw=32;bh=32;dft=1", "gray", 8, False, 160, 128, 24),
    ("sigma=3;bt=1;bw=32;bh=32;dehalo=1;dft=1", "gray", 8, False, 160, 128, 24),
    ("sigma=3;bt=0;bw=32;bh=32;ow=16;oh=16;kratio=2;dft=1", "gray", 10, False, 160, 128, 24),
    ("sigma=3;bt=-1;bw=32;bh=32;sharpen=1;dft=1", "gray", 8, False, 160, 128, 24),
    ("sigma=3;bt=3;bw=32;bh=32;pfactor=1;dft=1", "gray", 8, False, 160, 128, 24),
]

INT_FIELDS = {"bw", "bh", "bt", "ow", "oh", "measure", "interlaced", "wintype", "pframe", "px", "py",
              "pshow", "l", "t", "r", "b", "opt"}


def engine_params(args):
    from neo_fft3d_amd import default_params
    kv = dict(a.split("=") for a in args.split(";"))
    kv.pop("dft", None)  # engine settings, not parameters: make_engine applies them
    kv.pop("pregion", None)
    p = default_params(float(kv.pop("sigma")))
    for k, v in kv.items():
        setattr(p, k, int(v) if k in INT_FIELDS else float(v))
    return p


def make_engine(args, fmt, width, height, frames, plane):
    from neo_fft3d_amd import Engine
    e = Engine(engine_params(args), fmt, width, height, frames, plane)
    if "dft=1" in args.split(";"):
        e.set_dft(1)
    kv = dict(a.split("=") for a in args.split(";"))
    if "pregion" in kv:
        r = [int(v) for v in kv["pregion"].split(",")]
        e.set_pattern_regions(r[0], r[-1])
    return e


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available()
    return torch


@pytest.fixture(scope="module")
def product():
    return Runner(PRODUCT_SO, "nf3d_clip")


@pytest.fixture(scope="module")
def oracle():
    return oracle_runner()


def run_range(torch, clip, args, splits):
    """Plane 0 of `clip` through Engine.run_device, called once per split."""
    from neo_fft3d_amd import make_format
    F = len(clip.frames)
    fmt = make_format(clip.family, clip.bits, clip.is_float)
    eng = make_engine(args, fmt, clip.width, clip.height, F, 0)
    host = np.stack([clip.frames[n][0] for n in range(F)])
    src = torch.from_numpy(host).cuda()
    dst = torch.zeros_like(src)
    sp = [src[n].data_ptr() for n in range(F)]
    dp = [dst[n].data_ptr() for n in range(F)]
    stride = src.stride(1) * src.element_size()
    start = 0
    for cnt in splits:
        print(f"run_device {args} frames {start}..{start + cnt - 1}", flush=True)
        eng.run_device(start, cnt, sp, stride, dp, stride, 0)
        start += cnt
    eng.sync()
    torch.cuda.synchronize()
    out = dst.cpu().numpy()
    eng.close()
    return out


@pytest.mark.parametrize("case", CASES, ids=[c[0] + f"-{c[1]}{c[2]}" for c in CASES])
def test_run_device(torch_cuda, product, oracle, case):
    args, fam, bits, isf, W, H, F = case
    clip = make_clip(fam, bits, isf, width=W, height=H, nframes=F, seed=11)
    got = run_range(torch_cuda, clip, args, [F])
    ho = oracle.open(args, clip)
    hp = product.open(args, clip)
    for n in range(F):
        want = ho.get(n)[0]
        per_frame = hp.get(n)[0]
        (mx, fe, rms), = compare([got[n]], [want], isf)
        (mx2, fe2, rms2), = compare([got[n]], [per_frame], isf)
        print(f"{fam}{bits} {args} n={n} vs oracle max={mx:.3g} exact={fe:.5f} rms={rms:.3g}; "
              f"vs per-frame max={mx2:.3g} exact={fe2:.6f}", flush=True)
        if isf:
            assert mx <= 1e-4 and mx2 <= 1e-4, (n, mx, mx2)
        else:
            exact = 0.99 if bits == 16 else 0.999
            assert mx <= 1 and fe >= exact, (n, mx, fe)
            # the two kernels associate the transforms differently (the frame
            # loop splits each one over a lane pair), so they agree to the
            # oracle bar, not bit for bit
            assert mx2 <= 1 and fe2 >= exact, (n, mx2, fe2)


SPLIT_CASES = [
    ("sigma=3;bt=3;bw=16;bh=16;sharpen=0.3", [7, 20, 13]),
    ("sigma=3;bt=5;bw=16;bh=16", [3, 18, 19]),
    ("sigma=3;bt=0;bw=16;bh=16;sharpen=0.3", [1, 22, 17]),
    # one frame per call (a serving loop): the window's older spectra stay in the ring
    ("sigma=3;bt=3;bw=48;bh=48;ow=16;oh=16;sharpen=0.3", [1] * 40),
    ("sigma=3;bt=5;bw=32;bh=32", [2, 1, 1] + [1] * 36),
    ("sigma=3;bt=3;bw=32;bh=32;sharpen=0.3;dft=1", [5, 1, 1, 20, 13]),
    ("sigma=3;bt=0;bw=32;bh=32;dft=1", [1, 22, 17]),
]


@pytest.mark.parametrize("args,splits", SPLIT_CASES, ids=[c[0] for c in SPLIT_CASES])
def test_run_device_split_calls(torch_cuda, args, splits):
    """The same clip in one call and in several: identical outputs."""
    clip = make_clip("gray", 8, False, width=160, height=128, nframes=40, seed=5)
    whole = run_range(torch_cuda, clip, args, [40])
    parts = run_range(torch_cuda, clip, args, splits)
    diff = np.abs(whole.astype(np.int32) - parts.astype(np.int32))
    print(f"{args} splits={splits} max diff {diff.max()}", flush=True)
    assert diff.max() == 0


# ------------------------------------------------------ batched launches
def _device_clip(torch, clip, plane):
    host = np.stack([clip.frames[n][plane] for n in range(len(clip.frames))])
    src = torch.from_numpy(host).cuda()
    dst = torch.zeros_like(src)
    F = len(clip.frames)
    stride = src.stride(1) * src.element_size()
    return src, dst, [src[n].data_ptr() for n in range(F)], [dst[n].data_ptr() for n in range(F)], stride


BATCH_CLIPS = [
    # (args, family, bits, width, height, frames, seed): the first two share
    # launches (same shape, bt, sharpen), the third does not, the YUV clip
    # adds three planes of another shape
    ("sigma=3;bt=3;bw=16;bh=16;sharpen=0.3", "gray", 8, 160, 128, 24, 1),
    ("sigma=3;bt=3;bw=16;bh=16;sharpen=0.3", "gray", 8, 96, 64, 20, 2),
    ("sigma=3;bt=0;bw=16;bh=16", "gray", 8, 128, 96, 24, 3),
    ("sigma=2;bt=3;bw=48;bh=48;ow=16;oh=16;sharpen=0.3", "yuv", 8, 320, 200, 22, 4),
    # same shape, mode and sharpen case, the two transform forms: separate launches
    ("sigma=3;bt=3;bw=32;bh=32;sharpen=0.3", "gray", 8, 160, 128, 20, 6),
    ("sigma=3;bt=3;bw=32;bh=32;sharpen=0.3;dft=1", "gray", 8, 160, 128, 20, 7),
]


def _batch_setup(torch):
    from neo_fft3d_amd import make_format
    jobs = []
    for args, fam, bits, W, H, F, seed in BATCH_CLIPS:
        clip = make_clip(fam, bits, False, width=W, height=H, nframes=F, seed=seed)
        for p in range(clip.nplanes):
            jobs.append((args, clip, p, make_format(fam, bits)))
    def engines():
        return [make_engine(a, fmt, c.width, c.height, len(c.frames), p) for a, c, p, fmt in jobs]
    return jobs, engines


def test_batch_matches_separate_calls(torch_cuda):
    """N engines in one batched call == N separate run_device calls, bit for bit."""
    from neo_fft3d_amd import run_device_batch
    torch = torch_cuda
    jobs, engines = _batch_setup(torch)
    bufs = [_device_clip(torch, c, p) for _, c, p, _ in jobs]
    sep = engines()
    for e, (src, dst, sp, dp, st), (a, c, p, _) in zip(sep, bufs, jobs):
        e.run_device(0, len(c.frames), sp, st, dp, st, 0)
        e.sync()
    torch.cuda.synchronize()
    want = [b[1].cpu().numpy().copy() for b in bufs]
    for b in bufs:
        b[1].zero_()
    torch.cuda.synchronize()
    bat = engines()
    items = [(e, 0, len(c.frames), sp, st, dp, st) for e, (src, dst, sp, dp, st), (a, c, p, _) in zip(bat, bufs, jobs)]
    print(f"batch of {len(items)} engines", flush=True)
    run_device_batch(items, 0)
    bat[0].sync()
    torch.cuda.synchronize()
    for k, b in enumerate(bufs):
        got = b[1].cpu().numpy()
        print(f"job {k} {jobs[k][0]} plane {jobs[k][2]}: max diff {np.abs(got.astype(int) - want[k].astype(int)).max()}",
              flush=True)
        assert np.array_equal(got, want[k]), k


@pytest.mark.parametrize("args", ["sigma=2;bt=3;bw=48;bh=48;ow=16;oh=16;sharpen=0.3",
                                  "sigma=2;bt=3;bw=32;bh=32;ow=16;oh=16;sharpen=0.3;dft=1"])
def test_graph_replays_match_eager(torch_cuda, args):
    """The batch captured as a HIP graph: every replay equals the eager call
    (FFT form, and the matrix-core DFT form)."""
    from neo_fft3d_amd import Graph, make_format, run_device_batch
    torch = torch_cuda
    clip = make_clip("yuv", 8, False, width=320, height=200, nframes=22, seed=9)
    fmt = make_format("yuv", 8)
    engs = [make_engine(args, fmt, clip.width, clip.height, 22, p) for p in range(3)]
    bufs = [_device_clip(torch, clip, p) for p in range(3)]
    items = [(e, 0, 22, sp, st, dp, st) for e, (src, dst, sp, dp, st) in zip(engs, bufs)]
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        run_device_batch(items, stream.cuda_stream)
    stream.synchronize()
    want = [b[1].cpu().numpy().copy() for b in bufs]
    g = Graph(items, stream.cuda_stream)
    for r in range(3):
        with torch.cuda.stream(stream):
            for b in bufs:
                b[1].zero_()
            x = torch.randn(1 << 20, device="cuda")
            float(x.sum())  # eager work and a synchronisation between replays
        g.launch(stream.cuda_stream)
        stream.synchronize()
        for p, b in enumerate(bufs):
            got = b[1].cpu().numpy()
            print(f"replay {r} plane {p}: max diff {np.abs(got.astype(int) - want[p].astype(int)).max()}", flush=True)
            assert np.array_equal(got, want[p]), (r, p)
    g.close()


@pytest.mark.parametrize("bw,ow,rx,ry", [(16, 4, 3, 3), (48, 16, 2, 1), (32, 10, 100, 100)])
def test_region_psd_matches_oracle(torch_cuda, oracle, bw, ow, rx, ry):
    """nf3d_engine_region_psd (the learned PSD's second label source) against
    the oracle's per-region patterns of the same plane (estimated on frame
    pframe = 0): FP32 spectra against FP64, <= 1e-4 of the largest value; one
    region spanning the grid gives the plane's single pattern everywhere."""
    import ctypes
    from harness import ORACLE_SO
    from neo_fft3d_amd import make_format
    clip = make_clip("gray", 8, False, width=320, height=200, nframes=3, seed=17)
    args = f"sigma=2;bt=1;bw={bw};bh={bw};ow={ow};oh={ow};pfactor=1;pregion={rx},{ry}"
    ho = oracle.open(args, clip)
    ho.get(0)
    lib = ctypes.CDLL(ORACLE_SO)
    geo = (ctypes.c_int * 5)()
    lib.nf3d_oracle_geometry.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    assert lib.nf3d_oracle_geometry(ctypes.c_void_p(ho.h), 0, geo) == 0
    nb, n = geo[0] * geo[1], bw * geo[2]
    want = np.zeros(nb * n)
    lib.nf3d_oracle_block_psd_get.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    assert lib.nf3d_oracle_block_psd_get(ctypes.c_void_p(ho.h), 0, want.ctypes.data_as(ctypes.c_void_p)) == 0
    e = make_engine(args, make_format("gray", 8, False), clip.width, clip.height, 3, 0)
    src = torch_cuda.from_numpy(np.ascontiguousarray(clip.frames[0][0])).cuda()
    out = torch_cuda.full((nb * n,), float("nan"), dtype=torch_cuda.float32, device="cuda")
    torch_cuda.cuda.synchronize()
    print(f"region_psd bw={bw} regions {rx}x{ry} blocks={nb}", flush=True)
    e.region_psd(src.data_ptr(), src.stride(0), rx, ry, out.data_ptr())
    e.sync()
    got = out.cpu().numpy().astype(np.float64)
    err = np.max(np.abs(got - want)) / np.max(want)
    print(f"region_psd bw={bw} max rel err {err:.3g}", flush=True)
    assert np.isfinite(got).all() and err <= 1e-4
    if rx >= geo[0] and ry >= geo[1]:
        assert np.array_equal(got.reshape(nb, n), np.broadcast_to(got[:n], (nb, n)))
    e.close()


@pytest.mark.parametrize("args", ["sigma=3;bt=3;bw=16;bh=16;pfactor=1;pregion=3",
                                  "sigma=2;bt=3;bw=48;bh=48;ow=16;oh=16;pfactor=1;pregion=2,1;sharpen=0.3"])
def test_run_device_region_patterns(torch_cuda, oracle, args):
    """Per-region patterns through the frame-loop kernel (one pattern per
    block) against the oracle."""
    clip = make_clip("gray", 8, False, width=320, height=200, nframes=24, seed=17)
    from harness import Runner
    got = run_range(torch_cuda, clip, args, [24])
    ho = oracle.open(args, clip)
    for n in range(24):
        (mx, fe, rms), = compare([got[n]], [ho.get(n)[0]], False)
        print(f"{args} n={n} max={mx:.3g} exact={fe:.5f} rms={rms:.3g}", flush=True)
        assert mx <= 1 and fe >= 0.999, (n, mx, fe)
