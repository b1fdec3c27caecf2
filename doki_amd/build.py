"""Builds the native parts of doki_amd in-tree.

  libdoki_hip.so   HIP kernels + C ABI (hipcc, gfx950 only)
  doki<EXT>        the `doki` CPython extension (gcc, C11), linked to the
                   library above with an $ORIGIN rpath

Run ``python -m doki_amd.build`` (or ``__graft_entry__.build()``).
"""
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
EXT = sysconfig.get_config_var("EXT_SUFFIX")

HIP_SOURCES = ["dk_runtime.hip", "dk_kernels.hip", "dk_dense.hip",
               "dk_batch.hip", "dk_pass.hip", "dk_circuit.cpp", "dk_memory.cpp",
               "dk_serve.cpp"]
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions are kept out of the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    f"-I{INC}", f"-I{CSRC}",
]

LIB = os.path.join(HERE, "libdoki_hip.so")
EXTMOD = os.path.join(HERE, "doki" + EXT)


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    noise = ("packed-fp32-ops", "not a recognized feature")
    msg = "\n".join(ln for ln in (r.stdout + r.stderr).splitlines()
                    if not any(x in ln for x in noise))
    if verbose and msg.strip():
        print(msg)


def build_library(verbose=False):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_extension(verbose=False):
    import numpy
    pyinc = sysconfig.get_paths()["include"]
    tmp = EXTMOD + ".tmp"
    _run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-Wall",
          "-Wno-missing-braces", f"-I{INC}", f"-I{pyinc}",
          f"-I{numpy.get_include()}",
          "-DNPY_NO_DEPRECATED_API=NPY_1_7_API_VERSION",
          os.path.join(CSRC, "dokimodule.c"), f"-L{HERE}", "-ldoki_hip",
          "-Wl,-rpath,$ORIGIN", "-o", tmp], verbose)
    os.replace(tmp, EXTMOD)
    return EXTMOD


def build(verbose=False):
    build_library(verbose)
    build_extension(verbose)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", LIB, EXTMOD)
