"""Launch autotuner on the GPU: every launch shape gives the same output bit
for bit (the tuner itself refuses a shape that changes one), and a loaded
table reaches engines created afterwards."""
import os
import sys

import pytest

from harness import ROOT

sys.path.insert(0, ROOT)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bw,bits,bt,sharpen", [(48, 8, 3, 0.3), (16, 10, 1, 0.3), (32, 8, 0, 0.0),
                                                 (16, 8, -1, 0.5), (32, 32, 3, 0.0), (48, 16, 5, 0.0)])
def test_launch_shapes_are_results_invariant(bw, bits, bt, sharpen):
    from neo_fft3d_amd import autotune
    r = autotune.tune(bw, bw, bits, bt, width=320, height=240, frames=40, reps=1, sharpen=sharpen,
                      log=lambda s: print(s, flush=True))
    print(f"{bw}x{bw} bits={bits} bt={bt}: best {r['best']} of {len(r['ms'])} shapes, "
          f"default {r['ms'][r['default']]:.3f} ms, best {r['ms'][r['best']]:.3f} ms", flush=True)
    assert r["best"] in r["ms"]
    assert len(r["ms"]) == len(autotune.candidates(bt))


def test_table_reaches_new_engines(tmp_path):
    from neo_fft3d_amd import Engine, default_params, make_format
    from neo_fft3d_amd import autotune
    path = str(tmp_path / "tune.txt")
    autotune.write_table(path, {(48, 48, 8, 3): (8, 1), (16, 16, 10, 1): (4, 2)})
    try:
        assert autotune.load_table(path) == 2
        e = Engine(default_params(2.0, bt=3, bw=48, bh=48, ow=16, oh=16), make_format("yuv", 8), 320, 240, 10, 0)
        assert autotune.get_tuning(e) == (8, 1)
        e2 = Engine(default_params(2.0, bt=1, bw=16, bh=16, ow=4, oh=4), make_format("yuv", 10), 320, 240, 10, 0)
        assert autotune.get_tuning(e2) == (4, 2)
        e3 = Engine(default_params(2.0, bt=3, bw=32, bh=32), make_format("yuv", 8), 320, 240, 10, 0)
        assert autotune.get_tuning(e3) == autotune.DEFAULT
        for x in (e, e2, e3):
            x.close()
    finally:
        empty = str(tmp_path / "empty.txt")
        open(empty, "w").write("# empty\n")
        autotune.load_table(empty)
