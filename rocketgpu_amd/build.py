"""Builds the native libraries in-tree.

  rocketgpu_amd/_build/librocketgpu_amd.so   the product: gfx950 kernels + BatchSim + C-ABI (hipcc)
  rocketgpu_amd/_build/librg_twin.so         test-only host compile of the same step core (g++)
  rocketgpu_amd/_build/librg_phases.so       test-only gfx950 kernels of single car phases (hipcc)
  rocketgpu_amd/_build/batch_sim_client      test-only C++ caller of include/rocketgpu_amd/batch_sim.hpp (g++)
  oracle/_ref/librgref.so                    test-only reference CPU build (make -C oracle), when the reference
                                             sources are present

Floating-point contract for every build: no contraction (-ffp-contract=off), no fast-math, correctly rounded
division and square root, IEEE denormals; packed FP32 instructions disabled on gfx950.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
INCLUDE = os.path.join(REPO, "include")

PRODUCT = os.path.join(BUILD, "librocketgpu_amd.so")
TWIN = os.path.join(BUILD, "librg_twin.so")
PHASES = os.path.join(BUILD, "librg_phases.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RG_OFFLOAD_ARCH", "gfx950")

COMMON = ["-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", "-fno-math-errno", f"-I{INCLUDE}", f"-I{CSRC}"]


def _run(cmd, **kw):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd, **kw)


def build_product():
    os.makedirs(BUILD, exist_ok=True)
    tmp = PRODUCT + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", *COMMON,
          "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-shared", "-o", tmp,
          os.path.join(CSRC, "rg_batch.hip"), os.path.join(CSRC, "mesh_build.cpp"),
          os.path.join(CSRC, "rg_serve.cpp")])
    os.replace(tmp, PRODUCT)


def build_twin():
    os.makedirs(BUILD, exist_ok=True)
    tmp = TWIN + ".tmp"
    _run(["g++", *COMMON, "-shared", "-o", tmp,
          os.path.join(CSRC, "host_twin.cpp"), os.path.join(CSRC, "mesh_build.cpp")])
    os.replace(tmp, TWIN)


def build_phases():
    """Test-only gfx950 kernels of single car phases (checked against the reference's phase functions)."""
    os.makedirs(BUILD, exist_ok=True)
    tmp = PHASES + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", *COMMON,
          "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-shared", "-o", tmp, os.path.join(CSRC, "phase_kernels.hip")])
    os.replace(tmp, PHASES)


def build_client():
    """Test-only C++ program driving the library through the BatchSim header (tests/cpp/batch_sim_client.cpp)."""
    out = os.path.join(BUILD, "batch_sim_client")
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    _run(["g++", "-std=c++17", "-O2", "-D__HIP_PLATFORM_AMD__", f"-I{INCLUDE}", f"-I{rocm}/include", "-o", out + ".tmp",
          os.path.join(REPO, "tests", "cpp", "batch_sim_client.cpp"), f"-L{BUILD}", "-lrocketgpu_amd",
          "-Wl,-rpath,$ORIGIN", f"-L{rocm}/lib", "-lamdhip64", f"-Wl,-rpath,{rocm}/lib"])
    os.replace(out + ".tmp", out)


def build_oracle():
    ref = os.environ.get("REFERENCE_DIR", "/root/reference")
    if not os.path.isdir(ref) or shutil.which("make") is None:
        print("oracle: reference sources not present; using the prebuilt oracle/_ref if any", flush=True)
        return
    _run(["make", "-C", os.path.join(REPO, "oracle"), f"-j{min(8, os.cpu_count() or 1)}", f"REFERENCE_DIR={ref}"])


def build_all():
    build_product()
    build_twin()
    build_phases()
    build_client()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "product": build_product, "twin": build_twin, "phases": build_phases, "client": build_client,
         "oracle": build_oracle}[w]()
