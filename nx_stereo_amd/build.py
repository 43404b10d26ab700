"""Builds libnxs_hip.so, the gfx950 library behind include/nxs_hip.h, in-tree,
and stereo_match_hip, the Port executable over it.

Each library source compiles to an object with hipcc for gfx950, in parallel,
and the objects link into nx_stereo_amd/libnxs_hip.so; port_main.cpp (host
code) links against it into nx_stereo_amd/stereo_match_hip. Usage:
``python -m nx_stereo_amd.build``.
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIBRARY = os.path.join(HERE, "libnxs_hip.so")
PORT = os.path.join(HERE, "stereo_match_hip")  # the Port executable over the library
SOURCES = ["sgm.hip", "band.hip", "stages.hip", "runtime.cpp"]
ARCH = "gfx950"

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    # One rounding per f32 operation, and correctly rounded division: the
    # reference's f32 results are reproduced bit for bit.
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    # No packed-f32 instructions.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT}/include",
]


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return os.path.join(rocm, "bin", "hipcc")


def _object(source: str) -> str:
    return os.path.join(CSRC, "_build", source + ".o")


def _compile(source: str) -> None:
    os.makedirs(os.path.join(CSRC, "_build"), exist_ok=True)
    cmd = [hipcc(), *FLAGS, "-c", os.path.join(CSRC, source), "-o", _object(source)]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"hipcc failed on {source}:\n{result.stderr}")
    noisy = [line for line in result.stderr.splitlines()
             if "warning" in line and "packed-fp32-ops" not in line]
    if noisy:
        sys.stderr.write("\n".join(noisy) + "\n")


def build(verbose: bool = False) -> str:
    """Compiles every source and links the library; returns its path."""
    with concurrent.futures.ThreadPoolExecutor(len(SOURCES)) as pool:
        list(pool.map(_compile, SOURCES))
    tmp = LIBRARY + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp,
           *[_object(s) for s in SOURCES], "-lpthread"]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"link failed:\n{result.stderr}")
    os.replace(tmp, LIBRARY)
    # The Port executable: host code only, linked against the library beside it.
    tmp = PORT + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", f"-I{ROOT}/include",
           os.path.join(CSRC, "port_main.cpp"), "-o", tmp, f"-L{HERE}", "-lnxs_hip",
           "-Wl,-rpath,$ORIGIN", "-lpthread"]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"building stereo_match_hip failed:\n{result.stderr}")
    os.replace(tmp, PORT)
    if verbose:
        print(f"built {LIBRARY} and {PORT}")
    return LIBRARY


if __name__ == "__main__":
    build(verbose=True)
