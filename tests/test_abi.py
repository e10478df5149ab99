"""C ABI of the product library (CPU: no kernel is launched here).

The library loads, exports every function include/nf3d.h declares, carries
the reference's plugin signatures and parameter defaults, and rejects bad
parameters with the reference's messages before touching a GPU.
"""
import ctypes
import json
import os
import re

import pytest

from harness import ROOT

LIB = os.path.join(ROOT, "neo_fft3d_amd", "libnf3d.so")
HEADER = os.path.join(ROOT, "include", "nf3d.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(nf3d_[a-z_0-9]+)\s*\(", src)))


@pytest.fixture(scope="module")
def lib():
    return ctypes.CDLL(LIB)


def test_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n


def test_signatures_match_reference(lib):
    lib.nf3d_vs_signature.restype = ctypes.c_char_p
    lib.nf3d_avs_signature.restype = ctypes.c_char_p
    ref = json.load(open(os.path.join(ROOT, "tests", "golden", "signatures.json")))
    assert lib.nf3d_vs_signature().decode() == ref["vs"]
    assert lib.nf3d_avs_signature().decode() == ref["avs"]


def test_defaults(lib):
    import sys
    sys.path.insert(0, ROOT)
    from neo_fft3d_amd import Params
    p = Params()
    lib.nf3d_params_defaults(ctypes.byref(p), ctypes.c_float(2.5))
    # src/fft3d.hpp:77-106
    assert (p.sigma, p.beta, p.bw, p.bh, p.bt, p.ow, p.oh) == (2.5, 1.0, 32, 32, 3, -1, -1)
    assert abs(p.kratio - 2.0) < 1e-7 and p.sharpen == 0 and abs(p.scutoff - 0.3) < 1e-7
    assert (p.svr, p.smin, p.smax) == (1.0, 4.0, 20.0)
    assert (p.measure, p.interlaced, p.wintype, p.pframe, p.px, p.py, p.pshow) == (1, 0, 0, 0, 0, 0, 0)
    assert abs(p.pcutoff - 0.1) < 1e-7 and p.pfactor == 0
    assert (p.sigma2, p.sigma3, p.sigma4) == (2.5, 2.5, 2.5)
    assert (p.degrid, p.dehalo, p.hr, p.ht) == (1.0, 0.0, 2.0, 50.0)
    assert (p.l, p.t, p.r, p.b, p.opt) == (0, 0, 0, 0, 0)


@pytest.mark.parametrize("args,msg", [
    ("bw=16;ow=9", "ow must be less than bw / 2"),
    ("bh=16;oh=9", "oh must be less than bh / 2"),
    ("bt=6", "bt must be -1 (Sharpen), 0 (Kalman), 1..5 (Wiener)"),
    ("bt=-2", "bt must be -1 (Sharpen), 0 (Kalman), 1..5 (Wiener)"),
    ("beta=0.5", "beta must be not less 1.0"),
    ("l=40;r=30", "Width cannot be cropped to zero or below"),
    ("t=30;b=30", "Height cannot be cropped to zero or below"),
])
def test_reference_error_messages(lib, args, msg):
    import sys
    sys.path.insert(0, ROOT)
    from neo_fft3d_amd import Format
    fmt = Format(0, 0, 8, 0, 0, 1)
    out = ctypes.c_void_p()
    err = ctypes.create_string_buffer(256)
    rc = lib.nf3d_filter_create(args.encode(), ctypes.byref(fmt), 64, 48, 10, ctypes.byref(out), err,
                                ctypes.c_size_t(256))
    assert rc == -1
    assert err.value.decode() == msg
