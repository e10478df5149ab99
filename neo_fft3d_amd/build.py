"""Builds the package's native libraries in-tree.

  neo_fft3d_amd/libnf3d.so        the product: gfx950 kernels + C ABI (hipcc)
  neo_fft3d_amd/libnf3d_synth.so  seeded synthetic clips (g++, host only)

Run as ``python -m neo_fft3d_amd.build`` or through ``__graft_entry__.build()``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(ROOT, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PYTORCH_ROCM_ARCH", "gfx950")
# -ffp-contract=off: the spectral pass and the overlap-add follow the
# reference's float rounding (the FFT opts into FMA explicitly);
# -packed-fp32-ops: no v_pk_*_f32 generation.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", f"-I{os.path.join(ROOT, 'include')}",
]

KERNEL_UNITS = ["nf3d_kernels.hip", "nf3d_mfma_dft.hip", "nf3d_mlp.hip"]
HOST_UNITS = ["nf3d_engine.cpp", "nf3d_frontend.cpp"]


def block_shapes():
    """(bw, bh) pairs compiled for k_block, read from NF3D_SHAPES."""
    import re
    src = open(os.path.join(CSRC, "nf3d_kernels.hip")).read()
    line = re.search(r"#define NF3D_SHAPES\(X\)(.*)", src).group(1)
    return [(int(a), int(b)) for a, b in re.findall(r"X\((\d+),\s*(\d+)\)", line)]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}")
    return r.stdout


def _stale(obj: str, src: str) -> bool:
    """An object is stale when older than its source or any header of csrc/
    (every unit may include any of them), or, for host units, of include/."""
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [src] + [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    if src.endswith(".cpp"):
        deps += [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))]
    return any(os.path.getmtime(d) > t for d in deps)


def build_product(verbose: bool = False, incremental: bool = False) -> str:
    """Compiles every unit (incremental: only units older than their sources)."""
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    objs = []
    for u in KERNEL_UNITS:
        o = os.path.join(BUILD, u + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, u), "-o", o])
        objs.append(o)
    # each block shape in two units (per-frame kernels, frame-loop kernels)
    # so the slowest compiles run side by side
    for bw, bh in sorted(block_shapes(), key=lambda s: -s[0] * s[1]):
        for part in ("RUN", "TILE"):
            o = os.path.join(BUILD, f"nf3d_shape_{bw}x{bh}_{part.lower()}.o")
            jobs.append([HIPCC, *HIP_FLAGS, f"-DNF3D_BW={bw}", f"-DNF3D_BH={bh}", f"-DNF3D_PART_{part}", "-c",
                         os.path.join(CSRC, "nf3d_shape.hip"), "-o", o])
            objs.append(o)
    for u in HOST_UNITS:
        o = os.path.join(BUILD, u + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, u), "-o", o])
        objs.append(o)
    if incremental:
        jobs = [j for j in jobs if _stale(j[-1], j[-3])]
    with ThreadPoolExecutor(max_workers=max(1, min(os.cpu_count() or 8, 16, len(jobs)))) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    lib = os.path.join(PKG, "libnf3d.so")
    tmp = lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, lib)
    return lib


def build_synth() -> str:
    lib = os.path.join(PKG, "libnf3d_synth.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp",
          os.path.join(CSRC, "synth.cpp"), "-o", lib + ".tmp"])
    os.replace(lib + ".tmp", lib)
    return lib


def main(argv=None) -> int:
    args = argv or sys.argv[1:]
    verbose = "-v" in args
    build_synth()
    lib = build_product(verbose, incremental="--incremental" in args)
    print(f"built {lib}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
