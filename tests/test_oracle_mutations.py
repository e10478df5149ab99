"""Planted mistakes in the FP64 oracle must fail its closed-form checks (CPU).

Each case copies oracle/nf3d_oracle.c, changes one expression the way a
restatement could plausibly get it wrong, builds the copy and runs
tests/test_oracle_closed_form.py against it (NF3D_ORACLE_SO): the named test
must fail.  The unchanged source passes the same tests (that file's own run).
"""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(os.path.dirname(HERE), "oracle", "nf3d_oracle.c")

# (name, original text, planted text, -k expression of the test that must fail)
MUTATIONS = [
    ("3D sigma without bt",
     "float sig2f = (float)btcur * e->ep.sigma * e->ep.sigma / e->norm;",
     "float sig2f = e->ep.sigma * e->ep.sigma / e->norm;",
     "temporal_wiener_closed_form or static_clip_bt_equals"),
    ("temporal output not normalised by bt",
     "out[off + k].re = (sr + g0) / btc;",
     "out[off + k].re = (sr + g0) / (btc + 1);",
     "temporal_wiener_closed_form"),
    ("temporal DFT indexed from the oldest frame instead of the current one",
     "int tt = from + t;",
     "int tt = t;",
     "temporal_wiener_closed_form"),
    ("Kalman process covariance with K instead of K^2",
     "e->covarProcess[o].re = gr * gr * sigma;",
     "e->covarProcess[o].re = gr * sigma;",
     "kalman"),
    ("sharpen window transposed",
     "e->wsharpen[j * ow + i] = (float)(1 - exp((double)(-d2 / (2 * e->ep.scutoff * e->ep.scutoff))));",
     "e->wsharpen[(i % bh) * ow + (j % ow)] = (float)(1 - exp((double)(-d2 / (2 * e->ep.scutoff * e->ep.scutoff))));",
     "sharpen_dehalo"),
    ("chroma not centred in the analysis",
     "double base = (!e->fmt.is_float && chroma_centre) ? (double)(1 << (e->fmt.bits - 1)) : 0.0;\n  for (int by",
     "double base = 0.0;\n  for (int by",
     "floor_halves"),
    ("pattern search keeps the last minimum",
     "if (s < best) { *px = bx; *py = by; best = s; }",
     "if (s <= best) { *px = bx; *py = by; best = s; }",
     "find_pattern_block"),
    ("pattern window over the wrong axis",
     "float fw = (float)i * 2.0f / (float)bw;",
     "float fw = (float)(i + 1) * 2.0f / (float)bw;",
     "pattern_from_given_block"),
    ("interlaced fields not un-reordered",
     "else ry = (y % 2 == 0) ? y / 2 : e->Hc / 2 + (e->Hc - 1 - y) / 2;",
     "else ry = (y % 2 == 0) ? y / 2 : e->Hc / 2 + (y - 1) / 2;",
     "interlaced"),
]


@pytest.mark.parametrize("name,old,new,select", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_planted_mistake_is_caught(tmp_path, name, old, new, select):
    src = open(SRC).read()
    assert src.count(old) == 1, f"mutation point not found once: {old!r}"
    mutated = tmp_path / "mut.c"
    mutated.write_text(src.replace(old, new))
    so = tmp_path / "libmut.so"
    subprocess.run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-o", str(so), str(mutated), "-lm"], check=True)
    env = dict(os.environ, NF3D_ORACLE_SO=str(so))
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        os.path.join(HERE, "test_oracle_closed_form.py"), "-k", select],
                       cwd=HERE, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 1 and " failed" in r.stdout, f"{name}: not caught\n{r.stdout[-2000:]}"
