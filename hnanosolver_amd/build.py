"""Builds libhns.so (HIP, gfx950) in-tree, the hns_torch autograd extension over it, and the CPU
oracle used by the tests.

    python -m hnanosolver_amd.build            # library + torch extension + oracle
    python -m hnanosolver_amd.build --lib      # library only
    python -m hnanosolver_amd.build --torch    # torch extension only (needs libhns.so)
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libhns.so")
ORACLE_DIR = os.path.join(ROOT, "oracle")

SOURCES = ["topology.hip", "kernels_step.hip", "kernels_advect.hip", "kernels_pressure.hip", "kernels_cg.hip", "solvers.hip",
           "kernels_precond.hip", "precond.hip", "kernels_adjoint.hip", "adjoint.hip", "kernels_mg.hip", "ingest.hip", "api.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("HNS_OFFLOAD_ARCH", "gfx950")
# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 in device code.
# -ffp-contract=off: the kernels execute exactly the multiplies and adds they write (fused only
# where they call fmaf, as the reference does with __fmaf_rn / fmaf), the same rule the CPU oracle
# is compiled with, so FP32 results are reproducible operation for operation.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-ffp-contract=off"]
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-result",
          "-I", os.path.join(ROOT, "include")]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s" % (r.returncode, " ".join(cmd), r.stdout))
    return r.stdout


def build_lib(verbose: bool = False) -> str:
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)

    def compile_one(src):
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        out = _run([HIPCC, *COMMON, *DEVICE_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        if verbose and out.strip():
            print(out)
        return obj

    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


TORCH_EXT = os.path.join(PKG, "hns_torch.so")


def build_torch_ext() -> str:
    """hns_torch.so: the torch::autograd::Function of the differentiable step (host C++ against
    LibTorch, linked to the in-tree libhns.so through an $ORIGIN rpath)."""
    import sysconfig

    import torch
    from torch.utils import cpp_extension as ce
    inc = []
    for d in ce.include_paths() + [sysconfig.get_paths()["include"]]:
        inc += ["-isystem", d]
    libdir = ce.library_paths()[0]
    abi = "1" if torch.compiled_with_cxx11_abi() else "0"
    tmp = TORCH_EXT + ".tmp"
    _run([os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-shared", "-fPIC", "-w",
          "-D_GLIBCXX_USE_CXX11_ABI=" + abi, "-DTORCH_EXTENSION_NAME=hns_torch", "-DTORCH_API_INCLUDE_EXTENSION_H",
          *inc, os.path.join(CSRC, "torch_ext.cpp"), "-o", tmp, "-L" + libdir, "-ltorch", "-ltorch_cpu", "-lc10",
          "-ltorch_python", "-L" + PKG, "-l:libhns.so", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + libdir])
    os.replace(tmp, TORCH_EXT)
    return TORCH_EXT


TEST_PROGRAMS = os.path.join(ROOT, "tests", "cpp", "_build")


def build_test_programs() -> str:
    """tests/cpp/adapter_cook.cpp (a SOP-shaped cook through include/hns.hpp) against the in-tree
    libhns.so, into tests/cpp/_build/ (the GPU tests run it; nothing is compiled on the GPU box)."""
    os.makedirs(TEST_PROGRAMS, exist_ok=True)
    exe = os.path.join(TEST_PROGRAMS, "adapter_cook")
    _run([os.environ.get("CXX", "g++"), "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
          os.path.join(ROOT, "tests", "cpp", "adapter_cook.cpp"), "-o", exe + ".tmp", "-L", PKG, "-l:libhns.so",
          "-Wl,-rpath,$ORIGIN/../../../hnanosolver_amd"])
    os.replace(exe + ".tmp", exe)
    return exe


def build_oracle() -> str:
    _run(["make", "-s", "-C", ORACLE_DIR])
    return os.path.join(ORACLE_DIR, "liboracle_hns.so")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true")
    ap.add_argument("--oracle", action="store_true")
    ap.add_argument("--torch", action="store_true")
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args(argv)
    every = not (a.lib or a.oracle or a.torch)
    if a.lib or every:
        print("built", build_lib(a.v))
    if a.torch or every:
        print("built", build_torch_ext())
    if a.oracle or every:
        print("built", build_oracle())
    if every:
        print("built", build_test_programs())


if __name__ == "__main__":
    sys.exit(main())
