"""Compile libedhip.so (all HIP sources under csrc/) for gfx950, in-tree.

    python -m exact_diagonalization_for_fermionic_systems_with_gpu_amd.build

hipcc cross-compiles without a GPU, so this runs on any ROCm host.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
LIB = os.path.join(PKG, "libedhip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3",
    "-std=c++17",
    "--offload-arch=gfx950",
    "-fPIC",
    "-shared",
    # keep hipcc from emitting packed FP32 instructions (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def sources():
    return sorted(glob.glob(os.path.join(PKG, "csrc", "*.hip")))


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed: {' '.join(cmd[-3:])}")


def build(verbose=False, jobs=None):
    """Compile every csrc/*.hip to an object (in parallel), then link libedhip.so."""
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(REPO, "_build", "edhip")
    os.makedirs(objdir, exist_ok=True)
    cflags = [f for f in FLAGS if f != "-shared"]
    objs, cmds = [], []
    for src in sources():
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *cflags, "-I", os.path.join(REPO, "include"), "-c", "-o", obj, src])
    jobs = jobs or min(8, os.cpu_count() or 1, len(cmds))
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-L/opt/rocm/lib", "-lrocsolver",
          "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_drivers(verbose=False):
    """Compile the C++ drivers under tests/cpp (written against include/ed/ed.hpp) into
    tests/cpp/bin, linked to the in-tree libedhip.so by a relative rpath."""
    out = []
    srcdir = os.path.join(REPO, "tests", "cpp")
    bindir = os.path.join(srcdir, "bin")
    os.makedirs(bindir, exist_ok=True)
    for src in sorted(glob.glob(os.path.join(srcdir, "*.cpp"))):
        exe = os.path.join(bindir, os.path.splitext(os.path.basename(src))[0])
        _run([HIPCC, "-std=c++17", "-O2", "-I", os.path.join(REPO, "include"), "-o", exe, src, "-L", PKG, "-ledhip",
              "-Wl,-rpath,$ORIGIN/../../../" + os.path.basename(PKG)], verbose)
        out.append(exe)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_drivers(verbose=True))
