"""Builds the HIP library ``partgs_amd/lib/libsurfel_rast.so`` for gfx950.

Usage: ``python -m partgs_amd.build`` (or ``build()`` from Python).

Compiler policy, shared by every kernel in ``partgs_amd/csrc``:

* ``-ffp-contract=off``: no implicit FMA contraction; every FP32 product and
  sum is rounded on its own in source order, so device results equal a host
  evaluation of the same expressions (integer outputs bit-exact).
* ``-fhip-fp32-correctly-rounded-divide-sqrt``: IEEE division and sqrt.
* ``-munsafe-fp-atomics``: FP32 atomics on device memory use the hardware
  ``global_atomic_add_f32`` instead of a compare-and-swap loop.
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32
  instructions are generated.
"""
from __future__ import annotations

import concurrent.futures as _cf
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
LIBDIR = os.path.join(ROOT, "lib")
LIBNAME = "libsurfel_rast.so"
SOURCES = ["sr_api.hip", "sr_forward.hip", "sr_backward.hip", "sr_part.hip", "sr_block.hip", "sp_mlp.hip",
           "sk_knn.hip"]
ARCH = os.environ.get("PARTGS_AMD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I", os.path.join(REPO, "include"),
    "-I", CSRC,
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + p.stdout)
    return p.stdout


def build(verbose: bool = False) -> str:
    """Compiles every source and links the shared library; returns its path."""
    hipcc = _hipcc()
    objdir = os.path.join(ROOT, "_build")
    os.makedirs(objdir, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = []
    cmds = []
    for s in srcs:
        o = os.path.join(objdir, os.path.splitext(s)[0] + ".o")
        objs.append(o)
        cmds.append([hipcc, *HIPCC_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
    with _cf.ThreadPoolExecutor(max_workers=min(8, len(cmds))) as ex:
        outs = list(ex.map(_run, cmds))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    tmp = os.path.join(LIBDIR, LIBNAME + ".tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    out = os.path.join(LIBDIR, LIBNAME)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
