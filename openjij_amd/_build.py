"""Native build of the package: the HIP engine ``libojmi.so`` (gfx950) and the
``cxxjij`` pybind11 module that sits on its C ABI.

Everything is built in-tree so the shared objects travel with the source tree.
"""
from __future__ import annotations

import concurrent.futures as _cf
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libojmi.so")

HIP_SOURCES = [
    "ojmi_api.hip", "k_wave.hip", "k_wave_pipe.hip", "k_batch.hip", "k_sqa.hip", "k_sqa_f32_small.hip",
    "k_sqa_f32_large.hip", "k_sqa_f64_small.hip", "k_sqa_f64_large.hip", "k_shared.hip", "k_colown.hip",
    "k_colown_i8.hip", "k_colown_i32.hip", "k_lds.hip", "k_csr.hip", "k_energy.hip",
    "k_spin_gemm.hip", "k_beta_stats.hip", "k_batch_persist.hip", "k_mfma_sweep.hip", "k_sqa_csr.hip",
    "k_sqa_wide.hip",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        ln for ln in p.stdout.splitlines() if "packed-fp32-ops" not in ln
    )
    if p.returncode != 0:
        raise RuntimeError(f"command failed ({p.returncode}): {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out)


def build_engine(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    jobs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        jobs.append([HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-c",
                     os.path.join(CSRC, src), "-o", obj])
    with _cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(lambda c: _run(c, verbose), jobs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-Wl,-z,defs", "-o", tmp, *objs], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_cxxjij(verbose: bool = False) -> str:
    import pybind11

    src = os.path.join(CSRC, "cxxjij_module.cpp")
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = os.path.join(HERE, "cxxjij" + suffix)
    inc = sysconfig.get_paths()["include"]
    cmd = [
        "g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden",
        "-I", pybind11.get_include(), "-I", inc, "-I", os.path.join(ROOT, "include"),
        src, "-o", out + ".tmp", "-L", HERE, "-lojmi", "-Wl,-rpath,$ORIGIN",
    ]
    _run(cmd, verbose)
    os.replace(out + ".tmp", out)
    return out


def build_oracle(verbose: bool = False) -> str:
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], verbose)
    return os.path.join(ROOT, "oracle", "liboj_oracle.so")


def build_all(verbose: bool = False) -> None:
    build_engine(verbose)
    if os.path.exists(os.path.join(CSRC, "cxxjij_module.cpp")):
        build_cxxjij(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    build_oracle(verbose="-v" in sys.argv)
    print("built", LIB)
