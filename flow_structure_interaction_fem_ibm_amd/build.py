"""Build steps: the HIP library (gfx950), the oracle's C restatement and,
where the reference sources are present, nothing else (the reference is
Fortran that does not build as shipped).

Usage: python -m flow_structure_interaction_fem_ibm_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["csrc/pp_api.hip", "csrc/pp_setup.hip", "csrc/pp_geom.cpp", "csrc/pp_geom_tri.cpp",
               "csrc/pp_geom_dev.hip"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value", "-pthread",
]
# rocSOLVER factors the coarsest-level bordered matrix at set-up
HIP_LIBS = ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]


def _run(cmd, cwd=None, verbose=False):
    if verbose:
        print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out)
    return out


def build_hip(verbose=False):
    out = os.path.join(HERE, "libpp_hip.so")
    srcs = [os.path.join(HERE, s) for s in HIP_SOURCES if os.path.exists(os.path.join(HERE, s))]
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, f"-I{ROOT}/include", "-o", tmp, *srcs, *HIP_LIBS], verbose=verbose)
    os.replace(tmp, out)
    return out


def build_oracle(verbose=False):
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], verbose=verbose)
    return os.path.join(ROOT, "oracle", "_build", "liboracle.so")


def build_all(verbose=False):
    paths = [build_hip(verbose), build_oracle(verbose)]
    return paths


if __name__ == "__main__":
    for p in build_all(verbose="-v" in sys.argv):
        print(p)
