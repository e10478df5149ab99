"""The engine-level binding shown in INTEGRATION.md compiles (CPU only).

The C++ block of INTEGRATION.md -- the FFT3DEngine replacement class a
maintainer puts in place of src/fft3d_engine.h -- is extracted and compiled
(g++ -fsyntax-only) against the reference's own headers (EngineParams,
DSFrame, FetchFrameFunctor, FFTFunctionPointers) and include/nf3d.h, so the
documented binding cannot drift from either side.  Skipped when the reference
tree is absent.
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src")), reason="reference tree not present")
def test_integration_wrapper_compiles(tmp_path):
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```cpp\n(.*?)```", doc, re.S)
    wrapper = [b for b in blocks if "class FFT3DEngine" in b]
    assert wrapper, "INTEGRATION.md has no FFT3DEngine block"
    src = tmp_path / "wrapper.cpp"
    src.write_text('#include "fft3d_common.h"\n#include <unordered_map>\n#include <vector>\n'
                   + wrapper[0].replace('#include "nf3d.h"', '#include "nf3d.h"  // include/')
                   + "\nint main() { return 0; }\n")
    cmd = ["g++", "-std=gnu++17", "-fsyntax-only", "-DHAS_EXECUTION",
           f"-I{REF}/src", f"-I{REF}/include/avisynth", f"-I{REF}/include/vapoursynth",
           f"-I{REF}/include/dualsynth", f"-I{ROOT}/include", str(src)]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-3000:]
