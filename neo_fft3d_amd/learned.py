"""Learned noise PSD: host side of the nf3d_mlp_* C ABI, the training data
generator and the trainer.

The network (n_in = bh*(bw/2+1) log-power values of a block spectrum -> 2
hidden GELU layers -> n_out log-PSD values) replaces the single noise
pattern of the reference (src/helper.cpp:57-132, one [bh][outwidth] PSD for
the whole plane) with one PSD per block.  Every product runs in the library's
BF16 matrix-core kernels (forward, backward, Adam); this module only drives
them and generates the training data:

  clean planes    the seeded moving-rectangles generator (float, [0, 1])
  noise           coloured Gaussian: power law |f|^a (a ~ U[-1, 1]) times a
                  low-order anisotropic term exp(c1 cos 2t + c2 sin 2t), unit
                  variance, times a Perlin sigma field in [0.005, 0.04]
  features        the filter's own analysis (nf3d_engine_spectra: cover,
                  window, r2c) of the noisy plane, ln(|c|^2 + eps) normalised
  targets         labels="noise": the same analysis of the noise alone, its
                  periodogram, whose expectation is the block's noise PSD;
                  labels="regions": the method's own pattern estimator on the
                  noisy plane (nf3d_engine_region_psd: each region's flattest
                  block, |X - degrid correction|^2 x pattern window), which
                  needs no knowledge of the noise
  loss            Whittle: mean(y + t e^-y), minimised by y = ln E[t]

Device memory is held in torch tensors (allocation only); every computation on
them is a library kernel.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import Engine, Nf3dError, default_params, lib, make_format

EULER_GAMMA = 0.5772156649015329


def _bind():
    L = lib()
    if getattr(L, "_mlp_bound", False):
        return L
    vp, ip, sz, fp = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, ctypes.c_float
    cp = ctypes.c_char_p
    f32p = ctypes.POINTER(ctypes.c_float)
    L.nf3d_mlp_create.argtypes = [ip, ip, ip, ip, ctypes.POINTER(vp), cp, sz]
    L.nf3d_mlp_destroy.argtypes = [vp]
    L.nf3d_mlp_destroy.restype = None
    L.nf3d_mlp_param_count.argtypes = [vp]
    L.nf3d_mlp_param_count.restype = ctypes.c_longlong
    L.nf3d_mlp_dims.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    L.nf3d_mlp_dims.restype = None
    L.nf3d_mlp_set_params.argtypes = [vp, f32p, cp, sz]
    L.nf3d_mlp_get_params.argtypes = [vp, f32p, cp, sz]
    L.nf3d_mlp_get_grads.argtypes = [vp, f32p, cp, sz]
    L.nf3d_mlp_set_input_norm.argtypes = [vp, f32p, f32p, fp, cp, sz]
    L.nf3d_mlp_features.argtypes = [vp, vp, ip, vp, vp]
    L.nf3d_to_bf16.argtypes = [vp, vp, ctypes.c_longlong, vp]
    L.nf3d_periodogram.argtypes = [vp, vp, ctypes.c_longlong, vp]
    L.nf3d_gather_rows.argtypes = [vp, vp, vp, ip, ip, vp]
    L.nf3d_mlp_work_bytes.argtypes = [vp, ip]
    L.nf3d_mlp_work_bytes.restype = ctypes.c_longlong
    L.nf3d_mlp_predict.argtypes = [vp, vp, ip, vp, ip, fp, vp, vp, cp, sz]
    L.nf3d_mlp_grad.argtypes = [vp, vp, vp, ip, ctypes.POINTER(ctypes.c_double), vp, cp, sz]
    L.nf3d_mlp_adam.argtypes = [vp, fp, fp, fp, fp, vp, cp, sz]
    L.nf3d_mlp_train_step.argtypes = [vp, vp, vp, ip, fp, ctypes.POINTER(ctypes.c_double), vp, cp, sz]
    L.nf3d_engine_spectra.argtypes = [vp, vp, ip, vp, vp, cp, sz]
    L.nf3d_engine_set_learned.argtypes = [vp, vp, fp, cp, sz]
    L.nf3d_engine_set_learned_period.argtypes = [vp, ip, cp, sz]
    L._mlp_bound = True
    return L


def _f32(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def init_params(n_in: int, nh: int, n_out: int, seed: int, out_bias=None) -> np.ndarray:
    """Seeded initial parameters in host order (W1, b1, W2, b2, W3, b3), float32."""
    rng = np.random.default_rng(seed)
    W1 = rng.standard_normal((n_in, nh)) * math.sqrt(2.0 / n_in)
    W2 = rng.standard_normal((nh, nh)) * math.sqrt(2.0 / nh)
    W3 = rng.standard_normal((nh, n_out)) * math.sqrt(1.0 / nh) * 0.1
    b3 = np.zeros(n_out) if out_bias is None else np.asarray(out_bias, dtype=np.float64)
    return np.concatenate([W1.ravel(), np.zeros(nh), W2.ravel(), np.zeros(nh), W3.ravel(), b3]).astype(np.float32)


class Mlp:
    """nf3d_mlp: the network's parameters, Adam state and training workspace on the GPU."""

    def __init__(self, n_in: int, n_hidden: int, n_out: int, max_batch: int):
        L = _bind()
        h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        if L.nf3d_mlp_create(n_in, n_hidden, n_out, max_batch, ctypes.byref(h), err, 512) != 0:
            raise Nf3dError(err.value.decode())
        self.h = h
        self.n_in, self.nh, self.n_out, self.max_batch = n_in, n_hidden, n_out, max_batch
        self.nparam = int(L.nf3d_mlp_param_count(h))

    def _call(self, fn, *args):
        err = ctypes.create_string_buffer(512)
        if fn(self.h, *args, err, 512) != 0:
            raise Nf3dError(err.value.decode())

    def set_params(self, flat: np.ndarray):
        flat = np.ascontiguousarray(flat, dtype=np.float32)
        if flat.size != self.nparam:
            raise ValueError(f"expected {self.nparam} parameters, got {flat.size}")
        self._call(_bind().nf3d_mlp_set_params, _f32(flat))

    def get_params(self) -> np.ndarray:
        out = np.empty(self.nparam, dtype=np.float32)
        self._call(_bind().nf3d_mlp_get_params, _f32(out))
        return out

    def get_grads(self) -> np.ndarray:
        out = np.empty(self.nparam, dtype=np.float32)
        self._call(_bind().nf3d_mlp_get_grads, _f32(out))
        return out

    def set_input_norm(self, mu: np.ndarray, isd: np.ndarray, eps: float):
        mu = np.ascontiguousarray(mu, dtype=np.float32)
        isd = np.ascontiguousarray(isd, dtype=np.float32)
        self.norm = (mu.copy(), isd.copy(), float(eps))
        self._call(_bind().nf3d_mlp_set_input_norm, _f32(mu), _f32(isd), ctypes.c_float(eps))

    def features(self, spec_ptr: int, n: int, x_ptr: int, stream: int = 0):
        if _bind().nf3d_mlp_features(self.h, ctypes.c_void_p(spec_ptr), n, ctypes.c_void_p(x_ptr),
                                     ctypes.c_void_p(stream)) != 0:
            raise Nf3dError("feature kernel failed")

    def work_bytes(self, n: int) -> int:
        return int(_bind().nf3d_mlp_work_bytes(self.h, n))

    def predict(self, x_ptr: int, n: int, out_ptr: int, work_ptr: int, exp_out: bool = False, scale: float = 1.0,
                stream: int = 0):
        self._call(_bind().nf3d_mlp_predict, ctypes.c_void_p(x_ptr), n, ctypes.c_void_p(out_ptr), int(exp_out),
                   ctypes.c_float(scale), ctypes.c_void_p(work_ptr), ctypes.c_void_p(stream))

    def grad(self, x_ptr: int, t_ptr: int, n: int, stream: int = 0) -> float:
        loss = ctypes.c_double()
        self._call(_bind().nf3d_mlp_grad, ctypes.c_void_p(x_ptr), ctypes.c_void_p(t_ptr), n, ctypes.byref(loss),
                   ctypes.c_void_p(stream))
        return loss.value

    def adam(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8, stream: int = 0):
        self._call(_bind().nf3d_mlp_adam, ctypes.c_float(lr), ctypes.c_float(beta1), ctypes.c_float(beta2),
                   ctypes.c_float(eps), ctypes.c_void_p(stream))

    def train_step(self, x_ptr: int, t_ptr: int, n: int, lr: float, want_loss: bool = False,
                   stream: int = 0):
        loss = ctypes.c_double()
        self._call(_bind().nf3d_mlp_train_step, ctypes.c_void_p(x_ptr), ctypes.c_void_p(t_ptr), n,
                   ctypes.c_float(lr), ctypes.byref(loss) if want_loss else None, ctypes.c_void_p(stream))
        return loss.value if want_loss else None

    def close(self):
        if getattr(self, "h", None):
            _bind().nf3d_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def to_bf16(x_ptr: int, out_ptr: int, count: int, stream: int = 0):
    if _bind().nf3d_to_bf16(ctypes.c_void_p(x_ptr), ctypes.c_void_p(out_ptr), count, ctypes.c_void_p(stream)):
        raise Nf3dError("nf3d_to_bf16 failed")


def periodogram(spec_ptr: int, out_ptr: int, count: int, stream: int = 0):
    if _bind().nf3d_periodogram(ctypes.c_void_p(spec_ptr), ctypes.c_void_p(out_ptr), count, ctypes.c_void_p(stream)):
        raise Nf3dError("nf3d_periodogram failed")


def engine_spectra(engine: Engine, src_ptr: int, stride: int, spec_ptr: int, stream: int = 0):
    """The engine's own analysis of one device plane: compact spectra [nb][bh][bw/2+1].

    With stream 0 the analysis runs on the engine's own (non-blocking) stream:
    the call first waits for the device (the plane may have been written on
    another stream) and returns once the spectra are written, so any stream may
    read them.  A non-zero stream enqueues without waiting."""
    err = ctypes.create_string_buffer(512)
    if not stream:
        import torch
        torch.cuda.synchronize()
    if _bind().nf3d_engine_spectra(engine.h, ctypes.c_void_p(src_ptr), stride, ctypes.c_void_p(spec_ptr),
                                   ctypes.c_void_p(stream), err, 512) != 0:
        raise Nf3dError(err.value.decode())
    if not stream:
        engine.sync()


def region_labels(engine: Engine, src_ptr: int, stride: int, rx: int, ry: int, out_ptr: int):
    """The engine's per-region pattern estimate of one device plane as
    per-block PSD targets [nb][bh*(bw/2+1)] (nf3d_engine_region_psd), with
    the same waits as engine_spectra on stream 0."""
    import torch
    torch.cuda.synchronize()
    engine.region_psd(src_ptr, stride, rx, ry, out_ptr)
    engine.sync()


def set_learned(engine: Engine, mlp: Mlp | None, strength: float = 1.0, period: int | None = None):
    """Attach (or with None detach) a learned PSD; `period` (frames) sets
    how often the frame-loop path predicts it: 0 once per launch, n > 0 once
    per span of n frames (None leaves the engine's setting)."""
    err = ctypes.create_string_buffer(512)
    if _bind().nf3d_engine_set_learned(engine.h, mlp.h if mlp is not None else None, ctypes.c_float(strength),
                                       err, 512) != 0:
        raise Nf3dError(err.value.decode())
    if period is not None:
        set_learned_period(engine, period)


def set_learned_period(engine: Engine, period: int):
    """nf3d_engine_set_learned_period: one frame-loop prediction per launch
    (0) or per span of `period` frames; also forgets the kept span."""
    err = ctypes.create_string_buffer(512)
    if _bind().nf3d_engine_set_learned_period(engine.h, int(period), err, 512) != 0:
        raise Nf3dError(err.value.decode())


# ------------------------------------------------------------------ noise
def perlin(h: int, w: int, cell: float, rng: np.random.Generator) -> np.ndarray:
    """2D Perlin gradient noise (quintic fade) with the given cell size in pixels."""
    gy, gx = int(h / cell) + 2, int(w / cell) + 2
    ang = rng.uniform(0.0, 2.0 * np.pi, (gy, gx))
    gxv, gyv = np.cos(ang), np.sin(ang)
    y = np.arange(h) / cell
    x = np.arange(w) / cell
    y0 = np.floor(y).astype(int)
    x0 = np.floor(x).astype(int)
    fy = (y - y0)[:, None]
    fx = (x - x0)[None, :]

    def corner(dy, dx):
        iy = (y0 + dy)[:, None]
        ix = (x0 + dx)[None, :]
        return gxv[iy, ix] * (fx - dx) + gyv[iy, ix] * (fy - dy)

    def fade(t):
        return t * t * t * (t * (t * 6 - 15) + 10)

    u, v = fade(fx), fade(fy)
    top = corner(0, 0) * (1 - u) + corner(0, 1) * u
    bot = corner(1, 0) * (1 - u) + corner(1, 1) * u
    return top * (1 - v) + bot * v


class NoiseModel:
    """One coloured-noise process: PSD shape and sigma field of a plane (fixed
    over a clip; each frame draws a new realisation)."""

    def __init__(self, h: int, w: int, rng: np.random.Generator, smin: float = 0.005, smax: float = 0.04):
        self.h, self.w = h, w
        self.exponent = float(rng.uniform(-1.0, 1.0))
        c1, c2 = rng.uniform(-0.6, 0.6, 2)
        fy = np.fft.fftfreq(h)[:, None]
        fx = np.fft.fftfreq(w)[None, :]
        r = np.sqrt(fx * fx + fy * fy)
        r[0, 0] = 1.0
        theta = np.arctan2(fy, fx)
        S = (r / 0.25) ** self.exponent * np.exp(c1 * np.cos(2 * theta) + c2 * np.sin(2 * theta))
        S[0, 0] = 0.0
        S *= S.size / S.sum()  # unit variance after shaping
        self.amp = np.sqrt(S)
        p = perlin(h, w, max(h, w) / 3.0, rng) + 0.5 * perlin(h, w, max(h, w) / 7.0, rng)
        p = (p - p.min()) / max(p.max() - p.min(), 1e-12)
        self.sigma = smin + (smax - smin) * p

    def sample(self, rng: np.random.Generator) -> np.ndarray:
        white = rng.standard_normal((self.h, self.w))
        field = np.fft.ifft2(np.fft.fft2(white) * self.amp).real
        return (field * self.sigma).astype(np.float32)


def clean_planes(width: int, height: int, frames: int, seed: int) -> list:
    """Three float planes in [0, 1] (4:4:4) of the moving-rectangles generator, no noise."""
    from .synth import rects_clip
    return rects_clip(width, height, frames, seed, ssw=0, ssh=0, bits=32, is_float=True, noise_luma=0.0,
                      noise_chroma=0.0)


def noisy_clip(width: int, height: int, frames: int, seed: int, scale: float = 1.0):
    """(clean, noisy, noise) lists of three [frames, h, w] float32 planes; one
    noise process per plane.  `scale` multiplies content and noise (255 for
    8-bit sample units, 1023 for 10-bit, ...)."""
    rng = np.random.default_rng(seed)
    clean = clean_planes(width, height, frames, seed)
    noisy, noise = [], []
    for p in range(3):
        model = NoiseModel(height, width, rng)
        n = np.stack([model.sample(rng) for _ in range(frames)])
        if scale != 1.0:
            clean[p] = (clean[p] * scale).astype(np.float32)
            n = (n * scale).astype(np.float32)
        noise.append(n)
        noisy.append((clean[p] + n).astype(np.float32))
    return clean, noisy, noise


# ------------------------------------------------------------------ trainer
BW = BH = 48
OW = OH = 16


def analysis_engine(width: int, height: int, frames: int = 1, bw: int = BW, ow: int = OW) -> Engine:
    p = default_params(2.0, bt=3, bw=bw, bh=bw, ow=ow, oh=ow, sharpen=0.0)
    return Engine(p, make_format("gray", 32, True), width, height, max(frames, 1), 0)


class Dataset:
    """Features (BF16) and periodogram targets (FP32) of training blocks, in HBM."""

    def __init__(self, X, T, n_in: int, mu: np.ndarray, isd: np.ndarray, eps: float, out_bias: np.ndarray):
        self.X, self.T, self.n_in = X, T, n_in
        self.mu, self.isd, self.eps, self.out_bias = mu, isd, eps, out_bias

    @property
    def rows(self) -> int:
        return self.X.shape[0]


def build_dataset(mlp: Mlp, clips: int, width: int, height: int, frames: int, seed: int, eps: float = 1e-6,
                  log=None, bw: int = BW, ow: int = OW, scale: float = 1.0, labels: str = "noise",
                  regions: tuple = (4, 4)) -> Dataset:
    """Blocks of `clips` seeded clips (3 planes x `frames` frames each) through
    the filter's analysis (bw x bw blocks, overlap ow), samples in units of
    `scale`, with targets from `labels` ("noise" or "regions" of
    regions[0] x regions[1] blocks).  Sets the model's input normalisation
    from the first clip's features."""
    if labels not in ("noise", "regions"):
        raise ValueError("labels: 'noise' or 'regions'")
    import torch
    eng = analysis_engine(width, height, bw=bw, ow=ow)
    g = eng.geometry()
    nb = g["nox"] * g["noy"]
    d = bw * g["outwidth"]
    eps = eps * scale * scale
    if d != mlp.n_in or d != mlp.n_out:
        raise ValueError("model width does not match the block shape")
    rows = clips * 3 * frames * nb
    X = torch.empty((rows, d), dtype=torch.bfloat16, device="cuda")
    T = torch.empty((rows, d), dtype=torch.float32, device="cuda")
    spec = torch.empty((nb, d, 2), dtype=torch.float32, device="cuda")
    stride = width * 4
    row = 0
    for c in range(clips):
        _, noisy, noise = noisy_clip(width, height, frames, seed + c, scale)
        planes_noisy = torch.from_numpy(np.stack(noisy)).cuda()  # [3][frames][h][w]
        planes_noise = torch.from_numpy(np.stack(noise)).cuda()
        if c == 0:
            # normalisation and the output bias from the first clip
            lp, lt = [], []
            for p in range(3):
                for f in range(frames):
                    engine_spectra(eng, planes_noisy[p, f].data_ptr(), stride, spec.data_ptr())
                    s = spec.cpu().numpy().astype(np.float64)
                    lp.append(np.log(s[..., 0] ** 2 + s[..., 1] ** 2 + eps))
                    if labels == "noise":
                        engine_spectra(eng, planes_noise[p, f].data_ptr(), stride, spec.data_ptr())
                        s = spec.cpu().numpy().astype(np.float64)
                        lt.append(np.log(s[..., 0] ** 2 + s[..., 1] ** 2 + 1e-30))
                    else:
                        tr = torch.empty((nb, d), dtype=torch.float32, device="cuda")
                        region_labels(eng, planes_noisy[p, f].data_ptr(), stride, regions[0], regions[1],
                                      tr.data_ptr())
                        lt.append(np.log(tr.cpu().numpy().astype(np.float64) + 1e-30))
            lp = np.concatenate(lp)
            lt = np.concatenate(lt)
            mu = lp.mean(0)
            isd = 1.0 / np.maximum(lp.std(0), 1e-3)
            out_bias = lt.mean(0) + EULER_GAMMA
            mlp.set_input_norm(mu, isd, eps)
        for p in range(3):
            for f in range(frames):
                engine_spectra(eng, planes_noisy[p, f].data_ptr(), stride, spec.data_ptr())
                mlp.features(spec.data_ptr(), nb, X[row].data_ptr())
                if labels == "noise":
                    engine_spectra(eng, planes_noise[p, f].data_ptr(), stride, spec.data_ptr())
                    periodogram(spec.data_ptr(), T[row].data_ptr(), nb * d)
                else:
                    region_labels(eng, planes_noisy[p, f].data_ptr(), stride, regions[0], regions[1],
                                  T[row].data_ptr())
                row += nb
        if log and (c + 1) % max(1, clips // 4) == 0:
            log(f"dataset: {c + 1}/{clips} clips, {row} blocks")
    torch.cuda.synchronize()
    eng.close()
    return Dataset(X, T, d, mu.astype(np.float32), isd.astype(np.float32), eps, out_bias.astype(np.float32))


def train(mlp: Mlp, data: Dataset, steps: int, batch: int = 4096, lr: float = 1e-3, seed: int = 0,
          log=None, log_every: int = 100) -> list:
    """Adam on shuffled batches of the dataset; returns the logged losses."""
    import torch
    gen = torch.Generator(device="cpu")
    gen.manual_seed(seed)
    losses = []
    rows = data.rows
    Xb = torch.empty((batch, data.n_in), dtype=torch.bfloat16, device="cuda")
    Tb = torch.empty((batch, data.n_in), dtype=torch.float32, device="cuda")
    L = _bind()
    perm = torch.randperm(rows, generator=gen).to(torch.int32).cuda()
    at = 0
    for s in range(steps):
        if at + batch > rows:
            perm = torch.randperm(rows, generator=gen).to(torch.int32).cuda()
            at = 0
        idx = perm[at:at + batch]
        at += batch
        # the batch's rows, by the library's gather kernel (same stream as the step)
        if (L.nf3d_gather_rows(ctypes.c_void_p(data.X.data_ptr()), ctypes.c_void_p(idx.data_ptr()),
                               ctypes.c_void_p(Xb.data_ptr()), batch, data.n_in * 2, None) or
                L.nf3d_gather_rows(ctypes.c_void_p(data.T.data_ptr()), ctypes.c_void_p(idx.data_ptr()),
                                   ctypes.c_void_p(Tb.data_ptr()), batch, data.n_in * 4, None)):
            raise Nf3dError("batch gather failed")
        # cosine decay to lr / 20
        lr_s = lr * (0.05 + 0.95 * 0.5 * (1 + math.cos(math.pi * s / max(steps, 1))))
        want = (s % log_every == 0) or s == steps - 1
        loss = mlp.train_step(Xb.data_ptr(), Tb.data_ptr(), batch, lr_s, want_loss=want)
        if want:
            losses.append((s, loss))
            if log:
                log(f"step {s}: whittle loss {loss:.5f}")
    torch.cuda.synchronize()
    return losses


def train_model(clips: int = 48, frames: int = 2, width: int = 384, height: int = 384, steps: int = 1500,
                batch: int = 4096, hidden: int = 256, lr: float = 1e-3, seed: int = 0x4C50, log=None,
                bw: int = BW, ow: int = OW, scale: float = 1.0, labels: str = "noise", regions: tuple = (4, 4)):
    """Model trained from scratch on seeded data for bw x bw blocks with
    overlap ow, samples in units of `scale` (1: float planes in [0, 1]; 255:
    8-bit; 1023: 10-bit; 65535: 16-bit), targets from `labels` (see
    build_dataset); returns (mlp, dataset, losses)."""
    d = bw * (bw // 2 + 1)
    mlp = Mlp(d, hidden, d, batch)
    data = build_dataset(mlp, clips, width, height, frames, seed, log=log, bw=bw, ow=ow, scale=scale,
                         labels=labels, regions=regions)
    mlp.set_params(init_params(d, hidden, d, seed, out_bias=data.out_bias))
    losses = train(mlp, data, steps, batch, lr, seed, log=log)
    return mlp, data, losses
