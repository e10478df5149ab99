"""Launch tuning table of the library (CPU: parsing and lookup only)."""
import ctypes
import os

import pytest

from harness import ROOT

LIB = os.path.join(ROOT, "neo_fft3d_amd", "libnf3d.so")


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    L = ctypes.CDLL(LIB)
    L.nf3d_tune_load.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t]
    L.nf3d_tune_lookup.argtypes = [ctypes.c_int] * 4 + [ctypes.POINTER(ctypes.c_int)]
    yield L
    # leave an empty table behind for the rest of the process
    empty = tmp_path_factory.mktemp("tune") / "empty.txt"
    empty.write_text("# empty\n")
    L.nf3d_tune_load(str(empty).encode(), None, 0)


def test_load_and_lookup(lib, tmp_path):
    p = tmp_path / "tune.txt"
    p.write_text("# bw bh bits bt frames fpw\n48 48 8 3 8 1\n\n  16 16 10 1 4 2\n32 32 32 0 12 1\n")
    err = ctypes.create_string_buffer(256)
    assert lib.nf3d_tune_load(str(p).encode(), err, 256) == 3
    out = (ctypes.c_int * 2)()
    assert lib.nf3d_tune_lookup(48, 48, 8, 3, out) == 1 and tuple(out) == (8, 1)
    assert lib.nf3d_tune_lookup(16, 16, 10, 1, out) == 1 and tuple(out) == (4, 2)
    assert lib.nf3d_tune_lookup(32, 32, 32, 0, out) == 1 and tuple(out) == (12, 1)
    assert lib.nf3d_tune_lookup(48, 48, 10, 3, out) == 0
    assert lib.nf3d_tune_lookup(48, 48, 8, 5, out) == 0


@pytest.mark.parametrize("text", ["48 48 8 3 17 1\n", "48 48 8 3 8\n", "48 48 8 3 0 1\n", "x y\n"])
def test_rejects_bad_lines(lib, tmp_path, text):
    p = tmp_path / "bad.txt"
    p.write_text(text)
    err = ctypes.create_string_buffer(256)
    assert lib.nf3d_tune_load(str(p).encode(), err, 256) == -1
    assert b"bad line 1" in err.value


def test_missing_file(lib):
    err = ctypes.create_string_buffer(256)
    assert lib.nf3d_tune_load(b"/nonexistent/tune.txt", err, 256) == -1
    assert b"cannot open" in err.value
