"""Build recipe of the HIP library libmdffgpu.so (gfx950), in-tree.

    python -m mdff_amd.build

compiles every .hip file of mdff_amd/csrc with hipcc and links them into
mdff_amd/libmdffgpu.so.  Packed FP32 instructions are disabled for all
device code (-packed-fp32-ops).
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmdffgpu.so")
ARCH = os.environ.get("MDFF_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["kernels64.hip", "nlist.hip", "md.hip", "quad.hip", "realspace.hip", "mixed.hip", "batch.hip", "efg.hip", "mdffgpu.hip"]

DEVICE_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError(f"command failed: {' '.join(cmd)}")
    if out.strip():
        sys.stderr.write(out + "\n")


FORTRAN = os.path.join(HERE, "fortran")
FC = os.environ.get("MDFF_FC", "/opt/rocm/bin/amdflang")
DRIVER = os.path.join(FORTRAN, "_build", "engforce_drive")


def build_fortran(verbose: bool = False) -> str:
    """The Fortran host: the ISO_C_BINDING module and the driver program,
    linked against libmdffgpu.so (rpath relative to the executable)."""
    bdir = os.path.join(FORTRAN, "_build")
    os.makedirs(bdir, exist_ok=True)
    mod = os.path.join(FORTRAN, "mdffgpu_mod.f90")
    prog = os.path.join(FORTRAN, "engforce_drive.f90")
    cmd = [FC, "-O2", "-J", bdir, "-o", DRIVER + ".tmp", mod, prog, "-L", HERE, "-lmdffgpu",
           "-Wl,-rpath,$ORIGIN/../.."]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


def build(verbose: bool = False) -> str:
    srcs = list(SOURCES)
    with tempfile.TemporaryDirectory() as td:
        objs = []
        for s in srcs:
            o = os.path.join(td, s.replace(".hip", ".o"))
            cmd = [HIPCC] + DEVICE_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o]
            if verbose:
                print(" ".join(cmd))
            _run(cmd)
            objs.append(o)
        tmp = OUT + ".tmp"
        _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
        os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    if os.path.exists(FC):
        print(build_fortran(verbose="-v" in sys.argv))
