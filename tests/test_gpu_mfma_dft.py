"""DFT-as-GEMM precision form (nf3d_dft2d_mfma) against the FP64 DFT.

The 32x32, 48x48 and 64x64 block transforms run as dense DFT-matrix products on the
matrix cores with FP16 hi/lo split operands and FP32 accumulation.  The closed form
is numpy's float64 FFT (an FP64 DFT); the bar is a maximum error of 2e-6 of
the largest spectral magnitude (forward) or of the largest sample (inverse),
the same order as an FP32 FFT's rounding, with max and RMS reported.
"""
import ctypes
import os

import numpy as np
import pytest

from harness import ROOT

pytestmark = pytest.mark.gpu

LIB = os.path.join(ROOT, "neo_fft3d_amd", "libnf3d.so")


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available()
    lib = ctypes.CDLL(LIB)
    lib.nf3d_dft2d_mfma.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_float, ctypes.c_void_p]
    return torch, lib


def windowed_blocks(nb, maxv, seed, n=32):
    """Blocks of [0, maxv] samples times a raised-cosine window, as the filter sees them."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, maxv, size=(nb, n, n))
    w = 0.5 - 0.5 * np.cos(2 * np.pi * (np.arange(n) + 0.5) / n)
    return (x * w[None, :, None] * w[None, None, :]).astype(np.float32)


def run(torch, lib, host_in, out_shape, out_dtype, inverse, scale, n=32):
    d_in = torch.from_numpy(np.ascontiguousarray(host_in)).cuda()
    d_out = torch.empty(out_shape, dtype=out_dtype, device="cuda")
    nb = out_shape[0]
    rc = lib.nf3d_dft2d_mfma(ctypes.c_void_p(d_in.data_ptr()), ctypes.c_void_p(d_out.data_ptr()), nb, n,
                             int(inverse), ctypes.c_float(scale), None)
    assert rc == 0
    torch.cuda.synchronize()
    return d_out.cpu().numpy()


@pytest.mark.parametrize("n", [32, 48, 64])
@pytest.mark.parametrize("maxv,scale", [(255.0, 1.0), (1023.0, 1.0), (65535.0, 0.25)])
def test_forward_matches_fp64(env, maxv, scale, n):
    torch, lib = env
    nb = 2000 if n == 32 else 500
    x = windowed_blocks(nb, maxv, seed=int(maxv), n=n)
    got = run(torch, lib, x, (nb, n, n // 2 + 1, 2), torch.float32, False, scale, n)
    got = got[..., 0] + 1j * got[..., 1]
    want = np.fft.rfft2(x.astype(np.float64))
    err = np.abs(got - want)
    ref = np.abs(want).max()
    print(f"forward n={n} maxv={maxv}: max err {err.max() / ref:.3e} of max|Z|, rms {np.sqrt((err ** 2).mean()) / ref:.3e}",
          flush=True)
    assert err.max() <= 2e-6 * ref


@pytest.mark.parametrize("n", [32, 48, 64])
@pytest.mark.parametrize("maxv,scale", [(255.0, 1.0 / 32), (1023.0, 1.0 / 32), (65535.0, 1.0 / 4096)])
def test_inverse_matches_fp64(env, maxv, scale, n):
    torch, lib = env
    nb = 2000 if n == 32 else 500
    x = windowed_blocks(nb, maxv, seed=7 + int(maxv), n=n)
    Z = np.fft.rfft2(x.astype(np.float64))
    Zf = np.stack([Z.real, Z.imag], axis=-1).astype(np.float32)
    got = run(torch, lib, Zf, (nb, n, n), torch.float32, True, scale, n)
    Z32 = Zf[..., 0].astype(np.float64) + 1j * Zf[..., 1].astype(np.float64)
    want = np.fft.irfft2(Z32, s=(n, n)) * float(n * n)  # FFTW c2r is unnormalised
    err = np.abs(got - want)
    ref = np.abs(want).max()
    print(f"inverse n={n} maxv={maxv}: max err {err.max() / ref:.3e} of max|x|, rms {np.sqrt((err ** 2).mean()) / ref:.3e}",
          flush=True)
    assert err.max() <= 2e-6 * ref


def test_round_trip_recovers_blocks(env):
    torch, lib = env
    nb = 512
    x = windowed_blocks(nb, 1023.0, seed=3)
    Z = run(torch, lib, x, (nb, 32, 17, 2), torch.float32, False, 1.0)
    back = run(torch, lib, Z, (nb, 32, 32), torch.float32, True, 1.0 / 32) / 1024.0
    err = np.abs(back - x).max()
    print(f"round trip max err {err:.3e} (samples up to 1023)", flush=True)
    assert err <= 2e-3


def test_rejects_unsupported_sizes(env):
    """Argument checks on the host: no kernel is launched for any of these
    (a block size without tables, a scale that is not a power of two, null
    buffers with blocks to transform)."""
    torch, lib = env
    buf = torch.zeros(64 * 64 * 2, dtype=torch.float32, device="cuda")
    p = ctypes.c_void_p(buf.data_ptr())
    assert lib.nf3d_dft2d_mfma(p, p, 1, 24, 0, ctypes.c_float(1.0), None) != 0
    assert lib.nf3d_dft2d_mfma(p, p, 1, 32, 0, ctypes.c_float(0.3), None) != 0
    assert lib.nf3d_dft2d_mfma(None, None, 1, 48, 0, ctypes.c_float(1.0), None) != 0
    assert lib.nf3d_dft2d_mfma(p, None, 1, 32, 1, ctypes.c_float(1.0), None) != 0
    torch.cuda.synchronize()
