"""nf3d_mlp_predict_noalloc against nf3d_mlp_predict (GPU, through the C ABI).

The no-allocation entry point keeps the hidden activations in the library's
fixed per-GPU scratch array instead of a caller's workspace.  Two calls in a
row -- the second with fewer rows, so stale rows of the first call sit in
the scratch array -- must equal the workspace entry point bit for bit (same
kernels, same operands), for both output forms (log-PSD and exp(log-PSD) *
scale).  Sizes beyond the scratch array are refused with an error.
"""
import ctypes
import os

import numpy as np
import pytest

from harness import ROOT

pytestmark = pytest.mark.gpu

LIB = os.path.join(ROOT, "neo_fft3d_amd", "libnf3d.so")


def test_noalloc_predict_matches_workspace_predict():
    import torch
    assert torch.cuda.is_available()
    torch.cuda.set_device(0)
    L = ctypes.CDLL(LIB)
    vp, ip, fp, cp, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_char_p, ctypes.c_size_t
    L.nf3d_mlp_create.argtypes = [ip, ip, ip, ip, ctypes.POINTER(vp), cp, sz]
    L.nf3d_mlp_set_params.argtypes = [vp, ctypes.POINTER(ctypes.c_float), cp, sz]
    L.nf3d_mlp_param_count.argtypes = [vp]
    L.nf3d_mlp_param_count.restype = ctypes.c_longlong
    L.nf3d_mlp_work_bytes.argtypes = [vp, ip]
    L.nf3d_mlp_work_bytes.restype = ctypes.c_longlong
    L.nf3d_mlp_predict.argtypes = [vp, vp, ip, vp, ip, fp, vp, vp, cp, sz]
    L.nf3d_mlp_predict_noalloc.argtypes = [vp, vp, ip, vp, ip, fp, vp, cp, sz]
    L.nf3d_mlp_destroy.argtypes = [vp]
    err = ctypes.create_string_buffer(512)
    n_in, nh, n_out = 48 * 25, 256, 48 * 25
    m = vp()
    assert L.nf3d_mlp_create(n_in, nh, n_out, 64, ctypes.byref(m), err, 512) == 0, err.value
    rng = np.random.default_rng(7)
    npar = L.nf3d_mlp_param_count(m)
    params = (rng.standard_normal(npar) * 0.05).astype(np.float32)
    assert L.nf3d_mlp_set_params(m, params.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), err, 512) == 0
    stream = torch.cuda.current_stream().cuda_stream
    for exp_out, scale in ((0, 1.0), (1, 3.0)):
        for n in (2232, 608):  # a 1080p luma plane's blocks, then a chroma plane's
            print(f"noalloc exp_out={exp_out} rows={n}", flush=True)
            x = torch.randn(n, n_in, device="cuda").to(torch.bfloat16)
            want = torch.empty(n, n_out, device="cuda", dtype=torch.float32)
            got = torch.full((n, n_out), float("nan"), device="cuda", dtype=torch.float32)
            work = torch.empty(L.nf3d_mlp_work_bytes(m, n), device="cuda", dtype=torch.uint8)
            assert L.nf3d_mlp_predict(m, x.data_ptr(), n, want.data_ptr(), exp_out, scale, work.data_ptr(),
                                      stream, err, 512) == 0, err.value
            assert L.nf3d_mlp_predict_noalloc(m, x.data_ptr(), n, got.data_ptr(), exp_out, scale, stream,
                                              err, 512) == 0, err.value
            torch.cuda.synchronize()
            assert torch.isfinite(want).all()
            assert torch.equal(got, want), (exp_out, n, float((got - want).abs().max()))
    # beyond the scratch array: refused before any launch
    rc = L.nf3d_mlp_predict_noalloc(m, ctypes.c_void_p(1), 16385, ctypes.c_void_p(1), 0, 1.0, stream, err, 512)
    assert rc == -1 and b"at most 16384 rows" in err.value
    L.nf3d_mlp_destroy(m)
