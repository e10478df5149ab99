"""HIP path vs the FP64 oracle.

Bar: integer formats within +-1 LSB with >= 99.9% of pixels exact, float
formats within 1e-4 absolute (tolerances from the parity specification of this
port); every comparison also reports max and RMS error.  16-bit containers
with full 16-bit content are checked at >= 99% exact: FP32 transforms cannot
resolve the 1/65535 rounding boundary more finely.
"""
import os

import numpy as np
import pytest

from harness import ROOT, Runner, compare, make_clip, oracle_runner

pytestmark = pytest.mark.gpu

PRODUCT_SO = os.path.join(ROOT, "neo_fft3d_amd", "libnf3d.so")


@pytest.fixture(scope="module")
def product():
    return Runner(PRODUCT_SO, "nf3d_clip")


@pytest.fixture(scope="module")
def oracle():
    return oracle_runner()


def check_clip(product, oracle, fam, bits, isf, args, W=96, H=64, F=6, seed=3, exact=0.999, frames=None):
    clip = make_clip(fam, bits, isf, width=W, height=H, nframes=F, seed=seed)
    hp = product.open(args, clip)
    ho = oracle.open(args, clip)
    report = []
    for n in (frames if frames is not None else range(F)):
        a = hp.get(n)
        b = ho.get(n)
        for p, (mx, fe, rms) in enumerate(compare(a, b, isf)):
            report.append((n, p, mx, fe, rms))
            print(f"{fam}{bits} {args} n={n} p={p} max={mx:.3g} exact={fe:.5f} rms={rms:.3g}", flush=True)
            if isf:
                assert mx <= 1e-4, (n, p, mx)
            else:
                assert mx <= 1, (n, p, mx)
                assert fe >= exact, (n, p, fe)
    return report


MODES = [
    "sigma=3;bt=1;bw=16;bh=16",
    "sigma=3;bt=2;bw=16;bh=16;sharpen=0.5",
    "sigma=3;bt=3;bw=16;bh=16",
    "sigma=3;bt=4;bw=32;bh=32;dehalo=1",
    "sigma=3;bt=5;bw=16;bh=16;sharpen=0.4;dehalo=0.5",
    "sigma=3;bt=0;bw=16;bh=16",
    "sigma=3;bt=-1;bw=16;bh=16;sharpen=1",
    "sigma=3;bt=3;bw=16;bh=16;pfactor=1",
    "sigma=3;bt=3;bw=16;bh=16;pfactor=1;pframe=2",
    "sigma=3;bt=1;bw=16;bh=16;pfactor=1;sharpen=0.3",
    "sigma=3;bt=0;bw=16;bh=16;pfactor=1",
    "sigma=3;bt=3;bw=16;bh=16;sigma2=5;sigma3=1;sigma4=0.5",
    "sigma=3;bt=3;bw=16;bh=16;degrid=0;wintype=1",
    "sigma=3;bt=3;bw=16;bh=16;wintype=2;beta=1.5",
    "sigma=3;bt=3;bw=16;bh=16;interlaced=1",
    "sigma=3;bt=3;bw=8;bh=8;ow=2;oh=2",
    "sigma=3;bt=3;bw=24;bh=24;ow=8;oh=8",
    "sigma=3;bt=3;bw=32;bh=16;ow=8;oh=4",
    "sigma=3;bt=3;bw=16;bh=32;ow=4;oh=8",
    "sigma=2;bt=3;bw=48;bh=48;ow=16;oh=16",
    "sigma=2;bt=1;bw=64;bh=64;ow=16;oh=16",
    "sigma=3;bt=3;bw=16;bh=16;pfactor=1;pshow=1",
    "sigma=3;bt=1;bw=16;bh=16;pfactor=1;pshow=1;px=4;py=3",
]


@pytest.mark.parametrize("args", MODES)
def test_gray8_modes(product, oracle, args):
    check_clip(product, oracle, "gray", 8, False, args, W=160, H=128)


def test_yuv420p8_crop(product, oracle):
    check_clip(product, oracle, "yuv", 8, False, "sigma=3;bt=3;bw=16;bh=16;l=4;t=2;r=6;b=4")


def test_yuv420p10_planes(product, oracle):
    check_clip(product, oracle, "yuv", 10, False, "sigma=3;bt=3;bw=16;bh=16;planes=0,2")


def test_rgbps_float(product, oracle):
    check_clip(product, oracle, "rgb", 32, True, "sigma=3;bt=3;bw=16;bh=16")


def test_yuv_p16_kalman(product, oracle):
    check_clip(product, oracle, "yuv", 16, False, "sigma=3;bt=0;bw=32;bh=16;ow=8;oh=4;kratio=1.5",
               exact=0.99)


def test_yuv420p8_pshow_pxpy(product, oracle):
    check_clip(product, oracle, "yuv", 8, False, "sigma=3;bt=1;bw=16;bh=16;pfactor=1;pshow=1;px=3;py=2", F=3)


@pytest.mark.parametrize("family,ss,args", [
    ("yuva", 1, "sigma=3;bt=3;bw=16;bh=16"),                 # alpha copied
    ("yuva", 0, "sigma=3;bt=3;bw=16;bh=16;planes=0,1,2,3"),  # alpha filtered as chroma
    ("rgba", 0, "sigma=3;bt=1;bw=16;bh=16;planes=0,3"),      # alpha filtered like luma
    ("yuva", 0, "sigma=3;bt=0;bw=16;bh=16;planes=3"),
])
def test_alpha_planes(product, oracle, family, ss, args):
    """Four-plane formats: the alpha plane copied, or filtered when named."""
    clip = make_clip(family, 8, False, ssw=ss, ssh=ss, width=96, height=64, nframes=6, seed=8)
    hp = product.open(args, clip)
    ho = oracle.open(args, clip)
    for n in range(6):
        for p, (mx, fe, rms) in enumerate(compare(hp.get(n), ho.get(n), False)):
            print(f"{family} {args} n={n} p={p} max={mx:.3g} exact={fe:.5f} rms={rms:.3g}", flush=True)
            assert mx <= 1 and fe >= 0.999, (n, p, mx, fe)


def test_alpha_of_subsampled_yuva_refused(product):
    from harness import FilterError
    clip = make_clip("yuva", 8, False, ssw=1, ssh=1, width=96, height=64, nframes=3, seed=8)
    with pytest.raises(FilterError, match="alpha plane of a subsampled YUVA"):
        product.open("sigma=3;bt=3;bw=16;bh=16;planes=0,3", clip)


@pytest.mark.parametrize("args", ["sigma=3;bt=3;bw=16;bh=16;pfactor=1;pregion=3",
                                  "sigma=3;bt=1;bw=16;bh=16;pfactor=1;pregion=4,2;sharpen=0.3",
                                  "sigma=3;bt=0;bw=16;bh=16;pfactor=1;pregion=5,3"])
def test_region_patterns(product, oracle, args):
    """Per-region noise patterns through the per-frame path and the oracle."""
    check_clip(product, oracle, "gray", 8, False, args, W=160, H=128, F=8)
