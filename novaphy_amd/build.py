"""Build recipe for the native parts of novaphy_amd.

* ``libnpx_hip.so``  -- HIP kernels + C ABI (include/novaphy_hip.h), built with
  hipcc for gfx950.  The default kernels reproduce the reference's float
  operation order, so the build keeps FMA contraction off and division /
  square root correctly rounded.
* ``_core*.so``      -- the C++20 host (FluidWorld, PBFSolver, ModelBuilder,
  ...) as a pybind11 module; it links libnpx_hip.so and calls only its C ABI.

Run ``python -m novaphy_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("NPX_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Wall", "-Wextra", "-Wno-unused-parameter",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

LIB_NAME = "libnpx_hip.so"


def lib_path() -> str:
    return os.path.join(HERE, LIB_NAME)


def core_path() -> str:
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(HERE, "_core" + suffix)


def _run(cmd, verbose):
    if verbose:
        print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = ("packed-fp32-ops",)
    err = "\n".join(l for l in (r.stderr or "").splitlines()
                    if l.strip() and not any(n in l for n in noise))
    if r.returncode != 0:
        raise RuntimeError(f"build failed ({r.returncode}): {' '.join(cmd)}\n{err}")
    if err and verbose:
        print(err)


def build_hip(verbose: bool = True) -> str:
    out = lib_path()
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, os.path.join(CSRC, "npx.hip")], verbose)
    os.replace(tmp, out)
    return out


def build_core(verbose: bool = True) -> str:
    import pybind11
    host = os.path.join(CSRC, "host")
    srcs = sorted(os.path.join(host, f) for f in os.listdir(host) if f.endswith(".cpp"))
    out = core_path()
    tmp = out + ".tmp"
    cmd = ["g++", "-O2", "-std=c++20", "-fPIC", "-shared", "-Wall", "-Wextra",
           "-Wno-unused-parameter", "-fvisibility=hidden",
           "-I", sysconfig.get_paths()["include"], "-I", pybind11.get_include(),
           "-I", os.path.join(ROOT, "include"), "-I", host,
           *srcs, "-o", tmp, "-L", HERE, "-l:" + LIB_NAME, "-Wl,-rpath,$ORIGIN"]
    _run(cmd, verbose)
    os.replace(tmp, out)
    return out


def arith_check_path() -> str:
    return os.path.join(ROOT, "tests", "hip", "libnpx_arith_check.so")


def build_arith_check(verbose: bool = True) -> str:
    """The device check of the in-range sqrt / division helpers (a test library)."""
    src = os.path.join(ROOT, "tests", "hip", "arith_check.hip")
    out = arith_check_path()
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, src], verbose)
    os.replace(tmp, out)
    return out


def build(verbose: bool = True) -> None:
    """Compiles every native part.  The two hipcc compiles run side by side;
    the host module links libnpx_hip.so, so it comes after."""
    import concurrent.futures as cf
    jobs = [build_hip]
    if os.path.isfile(os.path.join(ROOT, "tests", "hip", "arith_check.hip")):
        jobs.append(build_arith_check)
    with cf.ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        for f in [ex.submit(j, verbose) for j in jobs]:
            f.result()
    if os.path.isdir(os.path.join(CSRC, "host")):
        build_core(verbose)


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
