"""Build recipe for the native libraries (in-tree, no JIT cache).

* ``cardillocxx_amd/libcdx.so``       -- HIP device library for gfx950 (hipcc), C ABI in include/cdx/cdx.h
* ``cardillocxx_amd/libcdx_scene.so`` -- CPU contact generation (g++ -fopenmp)
* ``cardillocxx_amd/cdx_adapter_demo`` -- C++ SolverBase adapter driver (host-only, links libcdx.so)

Run ``python -m cardillocxx_amd.build`` or call :func:`build`.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")

HIP_SOURCES = ["cdx_api.hip", "cdx_pj.hip", "cdx_gs.hip", "cdx_batch.hip", "cdx_adjoint.hip", "cdx_joints.hip"]
HIP_FLAGS = [
    "-O3",
    "--offload-arch=gfx950",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Wno-unused-result",
    "-Wno-unused-value",
    # keep the compiler away from packed FP32 instructions
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]
HOST_FLAGS = ["-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-std=c++17", "-Wall", "-Wextra"]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


def build_device(verbose: bool = True) -> str:
    """Each source compiled on its own (in parallel), then one link into libcdx.so."""
    from concurrent.futures import ThreadPoolExecutor

    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    out = os.path.join(PKG, "libcdx.so")
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    cflags = [f for f in HIP_FLAGS if f != "-shared"]
    objs = [os.path.join(objdir, os.path.splitext(s)[0] + ".o") for s in HIP_SOURCES]
    jobs = [[hipcc, *cflags, "-c", "-o", o, os.path.join(CSRC, s)] for s, o in zip(HIP_SOURCES, objs)]
    with ThreadPoolExecutor(max_workers=min(len(jobs), max(1, min(8, os.cpu_count() or 1)))) as ex:
        list(ex.map(_run, jobs))
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out + ".tmp", *objs])
    os.replace(out + ".tmp", out)
    return out


def build_scene() -> str:
    out = os.path.join(PKG, "libcdx_scene.so")
    # no contraction: the angular rows it writes are rebuilt bit for bit on the device from the
    # contact points (cdx_set_contact_points), with the same operations
    _run(["g++", *HOST_FLAGS, "-ffp-contract=off", "-shared", "-o", out + ".tmp",
          os.path.join(CSRC, "host", "cdx_scene.cpp")])
    os.replace(out + ".tmp", out)
    return out


def build_adapter_demo() -> str:
    out = os.path.join(PKG, "cdx_adapter_demo")
    _run(
        [
            "g++",
            "-O2",
            "-std=c++17",
            "-Wall",
            "-I",
            os.path.join(ROOT, "include"),
            "-I",
            os.path.join(CSRC, "host"),
            "-o",
            out + ".tmp",
            os.path.join(CSRC, "host", "cdx_adapter_demo.cpp"),
            "-L",
            PKG,
            "-lcdx",
            "-lcdx_scene",
            f"-Wl,-rpath,{PKG}",
            "-Wl,-rpath,$ORIGIN",
        ]
    )
    os.replace(out + ".tmp", out)
    return out


def build_oracle() -> str:
    """The CPU oracle is test infrastructure; it is built beside the product for the tests."""
    _run(["make", "-C", os.path.join(ROOT, "oracle"), "-s"])
    return os.path.join(ROOT, "oracle", "libcdx_oracle.so")


def build(oracle: bool = True) -> None:
    build_device()
    build_scene()
    if os.path.exists(os.path.join(CSRC, "host", "cdx_adapter_demo.cpp")):
        build_adapter_demo()
    if oracle:
        build_oracle()


if __name__ == "__main__":
    build(oracle="--no-oracle" not in sys.argv)
