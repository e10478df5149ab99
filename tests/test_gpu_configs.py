"""The workloads' configurations through the frame-loop path, against the
FP64 oracle (GPU).

Every clip runs through nf3d_run_device_batch (all planes, the frame-loop
kernel for whole windows, the per-frame kernel at the clip ends) and every
plane of every frame is compared with the oracle: integer formats within
+-1 LSB with >= 99.9% of pixels exact (>= 99% for full-range 16-bit samples),
each comparison printing max and RMS error.  Content comes from the
workloads' own seeded generators at reduced sizes:

* 4:2:0 8-bit moving textured rectangles, bt = 1 / 3 / 5, 48/16, sharpen 0.3
  (the scale workload; chroma planes centred on 128, src/buffer.cpp:133-135);
* 4:2:0 10-bit, noise 12, 64/16, 48/16 and 32/16 blocks, Kalman (kratio 2) and
  Wiener bt = 3 (the precision workload's shapes in the FFT form);
* 4:4:4 16-bit, 32/10 (ow = nearest even to bw/3), bt = 5;
* the Gray8 gratings workload at its full 640x360 size;
* a mixed set like the 16-clip serving workload -- 4:2:0 8/10-bit and 4:4:4
  16-bit clips, bw 32/48/64 with ow 10/16/22, bt 0/1/3/5 -- in one batched
  call; half of the clips use a learned PSD whose network outputs the
  constant sigma^2 * bw * bh (zero weights, output bias ln(sigma^2 bw bh)),
  which must reproduce the oracle's sigma path.
"""
import numpy as np
import pytest

from harness import Clip, compare, oracle_runner

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available()
    return torch


def rects(w, h, frames, seed, ssw=1, ssh=1, bits=8, noise_luma=3.0, noise_chroma=2.0):
    from neo_fft3d_amd.synth import rects_clip
    planes = rects_clip(w, h, frames, seed, ssw=ssw, ssh=ssh, bits=bits, noise_luma=noise_luma,
                        noise_chroma=noise_chroma)
    return Clip("yuv", bits, False, ssw, ssh, w, h,
                [[np.ascontiguousarray(planes[p][n]) for p in range(3)] for n in range(frames)])


def engine_params(args):
    from test_gpu_run import engine_params as ep
    return ep(args)


class DeviceClip:
    """A clip resident on the GPU with one engine per plane."""

    def __init__(self, torch, args, clip, learned_sigma=None):
        from neo_fft3d_amd import Engine, make_format
        F = len(clip.frames)
        self.clip, self.args, self.F = clip, args, F
        self.fmt = make_format("yuv", clip.bits, False, clip.ssw, clip.ssh)
        self.src, self.dst, self.items, self.engines, self.mlps = [], [], [], [], []
        for p in range(3):
            host = np.stack([clip.frames[n][p] for n in range(F)])
            s = torch.from_numpy(host).cuda()
            d = torch.zeros_like(s)
            st = s.stride(1) * s.element_size()
            e = Engine(engine_params(args), self.fmt, clip.width, clip.height, F, p)
            if learned_sigma is not None:
                self.mlps.append(constant_psd_model(e, args, clip, learned_sigma))
            self.src.append(s)
            self.dst.append(d)
            self.engines.append(e)
            self.items.append((e, 0, F, [s[n].data_ptr() for n in range(F)], st,
                               [d[n].data_ptr() for n in range(F)], st))

    def outputs(self):
        return [[self.dst[p][n].cpu().numpy() for p in range(3)] for n in range(self.F)]


def constant_psd_model(engine, args, clip, sigma):
    """A network whose output is ln(sigma_s^2 * bw * bh) for every input:
    zero weights and hidden biases, that constant as the output bias (sigma_s
    is sigma scaled to the sample range, as the engine scales it)."""
    from neo_fft3d_amd.learned import Mlp, set_learned
    kv = dict(a.split("=") for a in args.split(";"))
    bw, bh = int(kv["bw"]), int(kv["bh"])
    n = bh * (bw // 2 + 1)
    m = Mlp(n, 64, n, 64)
    scale = 2.0 ** (clip.bits - 8)
    params = np.zeros(m.nparam, dtype=np.float32)
    params[-n:] = np.log((sigma * scale) ** 2 * bw * bh)
    m.set_params(params)
    m.set_input_norm(np.zeros(n, np.float32), np.ones(n, np.float32), 1e-10)
    set_learned(engine, m, 1.0)
    return m


def compare_with_oracle(label, clip, args, got, exact=0.999):
    """+-1 LSB everywhere; `exact` of the clip's pixels exact per plane, and
    no frame below 99.5% (a Kalman reset decided on the other side of its
    threshold by FP32 and FP64 moves a whole block by one LSB)."""
    ho = oracle_runner().open(args, clip)
    same = np.zeros(3)
    total = np.zeros(3)
    for n in range(len(clip.frames)):
        want = ho.get(n)
        for p, (mx, fe, rms) in enumerate(compare(got[n], want, False)):
            print(f"{label} n={n} p={p} max={mx:.3g} exact={fe:.5f} rms={rms:.3g}", flush=True)
            assert mx <= 1 and fe >= 0.995, (label, n, p, mx, fe)
            same[p] += fe * want[p].size
            total[p] += want[p].size
    for p in range(3):
        print(f"{label} plane {p}: exact over the clip {same[p] / total[p]:.5f}", flush=True)
        assert same[p] / total[p] >= exact, (label, p)


def run_batch(torch, clips):
    from neo_fft3d_amd import run_device_batch
    items = [it for c in clips for it in c.items]
    print(f"batch of {len(items)} engines", flush=True)
    run_device_batch(items, 0)
    for c in clips:
        for e in c.engines:
            e.sync()
    torch.cuda.synchronize()


@pytest.mark.parametrize("bt", [1, 3, 5])
def test_scale_workload_yuv420p8_48_sharpen(torch_cuda, bt):
    from neo_fft3d_amd.synth import SEED_BASE
    clip = rects(480, 270, 20, SEED_BASE + 1)
    args = f"sigma=2;bt={bt};bw=48;bh=48;ow=16;oh=16;sharpen=0.3"
    dc = DeviceClip(torch_cuda, args, clip)
    run_batch(torch_cuda, [dc])
    compare_with_oracle(f"420p8 bt={bt}", clip, args, dc.outputs())


@pytest.mark.parametrize("bw,ow", [(64, 16), (48, 16), (32, 16)])
@pytest.mark.parametrize("bt", [0, 3])
def test_precision_workload_shapes_yuv420p10(torch_cuda, bw, ow, bt):
    from neo_fft3d_amd.synth import SEED_BASE
    clip = rects(480, 272, 12, SEED_BASE + 2, bits=10, noise_luma=12.0, noise_chroma=8.0)
    args = f"sigma=3;bt={bt};bw={bw};bh={bw};ow={ow};oh={ow};kratio=2"
    dc = DeviceClip(torch_cuda, args, clip)
    run_batch(torch_cuda, [dc])
    compare_with_oracle(f"420p10 {bw}/{ow} bt={bt}", clip, args, dc.outputs())


def test_yuv444p16_bt5(torch_cuda):
    from neo_fft3d_amd.synth import SEED_BASE
    clip = rects(200, 136, 10, SEED_BASE + 4, ssw=0, ssh=0, bits=16, noise_luma=768.0, noise_chroma=512.0)
    args = "sigma=3;bt=5;bw=32;bh=32;ow=10;oh=10"
    dc = DeviceClip(torch_cuda, args, clip)
    run_batch(torch_cuda, [dc])
    compare_with_oracle("444p16 bt=5", clip, args, dc.outputs(), exact=0.99)


def test_gratings_workload_gray8_640x360():
    """The CPU-runnable workload with its own generator and seed, through the
    front end (nf3d_clip) at full size."""
    import os
    from harness import ROOT, Runner
    from neo_fft3d_amd.synth import SEED_BASE, gratings_clip
    frames = gratings_clip(640, 360, 30, SEED_BASE + 0, noise=4.0)
    clip = Clip("gray", 8, False, 0, 0, 640, 360, [[np.ascontiguousarray(frames[n])] for n in range(30)])
    args = "sigma=2;bt=3;bw=48;bh=48;ow=16;oh=16;sharpen=0"
    hp = Runner(os.path.join(ROOT, "neo_fft3d_amd", "libnf3d.so"), "nf3d_clip").open(args, clip)
    ho = oracle_runner().open(args, clip)
    for n in (0, 1, 2, 14, 15, 28, 29):
        (mx, fe, rms), = compare(hp.get(n), ho.get(n), False)
        print(f"gratings n={n} max={mx:.3g} exact={fe:.5f} rms={rms:.3g}", flush=True)
        assert mx <= 1 and fe >= 0.999, (n, mx, fe)


MIX = [
    # (width, height, ssw/ssh, bits, bw, ow, bt, sigma, learned)
    (212, 120, 1, 8, 32, 10, 0, 2.5, False),
    (212, 120, 1, 8, 48, 16, 1, 1.5, True),
    (320, 180, 1, 8, 64, 22, 3, 3.0, False),
    (320, 180, 1, 8, 32, 10, 5, 2.0, True),
    (480, 272, 1, 10, 48, 16, 3, 4.0, False),
    (480, 272, 1, 10, 64, 22, 0, 2.0, True),
    (240, 136, 0, 16, 32, 10, 1, 3.5, False),
    (240, 136, 0, 16, 48, 16, 3, 1.8, True),
]


def test_mixed_clip_set_one_batch(torch_cuda):
    from neo_fft3d_amd.synth import SEED_BASE
    clips = []
    for k, (w, h, ss, bits, bw, ow, bt, sigma, learned) in enumerate(MIX):
        scale = 2.0 ** (bits - 8)
        clip = rects(w, h, 12, SEED_BASE + 5 + k, ssw=ss, ssh=ss, bits=bits,
                     noise_luma=3.0 * scale, noise_chroma=2.0 * scale)
        args = f"sigma={sigma};bt={bt};bw={bw};bh={bw};ow={ow};oh={ow}"
        clips.append(DeviceClip(torch_cuda, args, clip, learned_sigma=sigma if learned else None))
    run_batch(torch_cuda, clips)
    for k, c in enumerate(clips):
        exact = 0.99 if c.clip.bits == 16 else 0.999
        compare_with_oracle(f"mix {k} {c.args} learned={MIX[k][-1]}", c.clip, c.args, c.outputs(), exact=exact)


# ------------------------------------------------------------ serving form
S = 15  # staging slots per plane: a multiple of every bt, longer than any window


def random_psd_model(bw, bits, sigma, seed):
    """A network with random weights around ln(sigma_s^2 bw bh): a PSD that
    varies from block to block."""
    from neo_fft3d_amd.learned import Mlp, init_params
    n = bw * (bw // 2 + 1)
    m = Mlp(n, 64, n, 64)
    bias = np.full(n, np.log((sigma * 2.0 ** (bits - 8)) ** 2 * bw * bw), np.float32)
    m.set_params(init_params(n, 64, n, seed, out_bias=bias))
    m.set_input_norm(np.full(n, 8.0, np.float32), np.full(n, 0.3, np.float32), 1e-6)
    return m


class StagedClip:
    """One clip of the mixed set served frame by frame: frame j of each plane
    is copied into staging slot j % S when it arrives and filtered into
    output slot j % S, so the engines see the clip's frames at fixed
    addresses that repeat every S frames."""

    def __init__(self, torch, args, clip, model, period=0):
        from neo_fft3d_amd import Engine, make_format
        from neo_fft3d_amd.learned import set_learned
        F = len(clip.frames)
        fmt = make_format("yuv", clip.bits, False, clip.ssw, clip.ssh)
        self.res, self.stg, self.out, self.engines = [], [], [], []
        for p in range(3):
            host = np.stack([clip.frames[n][p] for n in range(F)])
            r = torch.from_numpy(host).cuda()
            self.res.append(r)
            self.stg.append(torch.zeros((S,) + tuple(r.shape[1:]), dtype=r.dtype, device="cuda"))
            self.out.append(torch.zeros_like(self.stg[-1]))
            e = Engine(engine_params(args), fmt, clip.width, clip.height, F, p)
            if model is not None:
                set_learned(e, model, 1.0, period=period)
            self.engines.append(e)
        self.F = F

    def arrive(self, j):
        for p in range(3):
            self.stg[p][j % S].copy_(self.res[p][j])

    def items(self, k):
        out = []
        for p, e in enumerate(self.engines):
            st = self.stg[p][0].stride(0) * self.stg[p].element_size()
            out.append((e, k, 1, [self.stg[p][j % S].data_ptr() for j in range(self.F)], st,
                        [self.out[p][j % S].data_ptr() for j in range(self.F)], st))
        return out


@pytest.mark.parametrize("period", [0, S])
def test_mixed_set_serving_graphs_bit_exact(torch_cuda, period):
    """The mixed clip set served one frame step at a time: (a) one batched
    call per step, (b) per-clip calls per step, (c) after two eager steps, one
    HIP graph per phase k mod S captured from the batched call of that step
    (each replayed once at capture to advance the ring), then replayed for
    the following periods, (d) as (c) with graphs bound to the spectra rings
    (nf3d_run_device_batch_graph_ring: the captured steps read the frames the
    previous step left in the rings instead of transforming them again).
    Every output frame of (b), (c) and (d) equals (a) bit for bit; the
    learned clips use networks whose PSD varies per block.  With a prediction
    period of S frames the learned PSD is predicted once per S steps and
    reused by the steps between, eager and ring-bound alike (a graph not
    bound to the rings predicts at every replay, so (c) is left out)."""
    from neo_fft3d_amd import Graph, run_device_batch
    from neo_fft3d_amd.synth import SEED_BASE
    torch = torch_cuda
    F = 40
    sets = {"batch": [], "single": [], "ring": []} if period else {"batch": [], "single": [], "graph": [], "ring": []}
    models = []
    for k, (w, h, ss, bits, bw, ow, bt, sigma, learned) in enumerate(MIX):
        scale = 2.0 ** (bits - 8)
        clip = rects(w, h, F, SEED_BASE + 30 + k, ssw=ss, ssh=ss, bits=bits, noise_luma=3.0 * scale,
                     noise_chroma=2.0 * scale)
        args = f"sigma={sigma};bt={bt};bw={bw};bh={bw};ow={ow};oh={ow}"
        m = random_psd_model(bw, bits, sigma, 100 + k) if learned else None
        models.append(m)
        for key in sets:
            sets[key].append(StagedClip(torch, args, clip, m, period))
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    graphs, rgraphs = {}, {}
    warm = 3  # the first frame-loop step (bt=5: frame 2) fills the ring eagerly
    checked = 0
    with torch.cuda.stream(stream):
        for j in range(3):
            for key in sets:
                for c in sets[key]:
                    c.arrive(j)
        for k in range(F - 2):
            if k > 0:
                for key in sets:
                    for c in sets[key]:
                        c.arrive(k + 2)
            run_device_batch([it for c in sets["batch"] for it in c.items(k)], sp)
            for c in sets["single"]:
                for it in c.items(k):
                    run_device_batch([it], sp)
            for key, gs in (("graph", graphs), ("ring", rgraphs)):
                if key not in sets:
                    continue
                items = [it for c in sets[key] for it in c.items(k)]
                if k < warm:
                    run_device_batch(items, sp)
                else:
                    if k < warm + S:
                        print(f"capture {key} phase {k % S} at frame {k}", flush=True)
                        gs[k % S] = Graph(items, sp, ring=key == "ring")
                    gs[k % S].launch(sp)
            stream.synchronize()
            for i, a in enumerate(sets["batch"]):
                for p in range(3):
                    want = a.out[p][k % S]
                    for key in sets:
                        assert torch.equal(sets[key][i].out[p][k % S], want), (key, k, p)
                    checked += 1
    print(f"{checked} plane-frames equal; {len(graphs)} + {len(rgraphs)} graphs, replayed over {F - 2 - warm} steps",
          flush=True)
    for g in list(graphs.values()) + list(rgraphs.values()):
        g.close()
    for key in sets:
        for c in sets[key]:
            for e in c.engines:
                e.close()
