"""Builds graph500_amd/lib/libg5.so (HIP, gfx950) from graph500_amd/csrc and
the C++ benchmark driver graph500_amd/lib/g5bench over it.

    python -m graph500_amd.build          # or build() from Python

Every .hip source is compiled for gfx950 only, objects go to
graph500_amd/_build/, the shared library and the driver to graph500_amd/lib/.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libg5.so")
DRIVER = os.path.join(LIBDIR, "g5bench")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

SOURCES = ["g5_gen.hip", "g5_prims.hip", "g5_csr.hip", "g5_bfs.hip", "g5_msbfs.hip", "g5_validate.hip",
           "g5_serve.hip", "g5_api.hip", "g5_measure.hip"]

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    # packed FP32 instructions stay off for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # host double arithmetic of find_roots is evaluated as written (no FMA)
    "-ffp-contract=off",
    "-I" + INCLUDE,
]


def _compile(src):
    obj = os.path.join(BUILD, src.replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr))
    return obj, r.stderr


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        results = list(ex.map(_compile, SOURCES))
    for obj, err in results:
        if verbose and err.strip():
            sys.stderr.write(err)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *[o for o, _ in results]]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, LIB)
    _build_driver()
    return LIB


def _build_driver():
    """g5bench: host-only C++ over the C ABI (include/g5_bfs.hpp)."""
    tmp = DRIVER + ".tmp"
    cmd = [CXX, "-O2", "-std=c++17", "-Wall", "-I" + INCLUDE, os.path.join(HERE, "driver", "g5bench.cc"),
           "-o", tmp, "-L" + LIBDIR, "-lg5", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("driver build failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, DRIVER)


if __name__ == "__main__":
    print(build(verbose=True))
