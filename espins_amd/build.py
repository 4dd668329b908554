"""Build the in-tree native artefacts.

* ``espins_amd/libespins_hip.so``  -- the HIP engine (gfx950) behind ``include/espins_hip.h``
* ``espins_amd/driver/mc_gpu.x``   -- the Fortran driver (amdflang), when amdflang exists
* ``oracle/liboracle.so``          -- the CPU restatement used by tests (test infrastructure)

Run ``python -m espins_amd.build`` (or ``__graft_entry__.build()``).  Everything is compiled
every time it is called.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libespins_hip.so")
ARCH = os.environ.get("ESPINS_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # no packed-FP32 instructions in any HIP code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} (exit {r.returncode})")
    return r


def build_hip(verbose: bool = False) -> str:
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    objdir = os.path.join(PKG, "_obj")
    os.makedirs(objdir, exist_ok=True)
    jobs = []
    for src in ("esp_model.cpp", "esp_setup.cpp", "esp_hip.hip", "esp_batch.hip", "esp_learn.hip", "esp_obs.hip"):
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(objdir, src + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        jobs.append(([hipcc, *HIP_FLAGS, *lang, "-c", path, "-o", obj], obj))
    # the translation units are independent: compile them side by side
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        list(ex.map(lambda j: _run(j[0]), jobs))
    objs = [o for _, o in jobs]
    tmp = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_oracle(verbose: bool = False) -> str:
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-C", odir, "liboracle.so"])
    if os.path.isdir("/root/reference/src") and shutil.which("amdflang"):
        try:
            _run(["make", "-C", odir, "ref"])
        except RuntimeError:
            pass  # the MT19937 cross-check is optional
    return os.path.join(odir, "liboracle.so")


def build_driver(verbose: bool = False):
    ddir = os.path.join(PKG, "driver")
    if not os.path.exists(os.path.join(ddir, "Makefile")):
        return None
    if not (shutil.which("amdflang") or os.path.exists("/opt/rocm/bin/amdflang")):
        return None
    _run(["make", "-C", ddir, "-B"])
    return os.path.join(ddir, "mc_gpu.x")


def build_all(verbose: bool = False) -> None:
    build_hip(verbose)
    build_oracle(verbose)
    build_driver(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
