"""Build in-tree: librrh.so (HIP, gfx950, hipcc), librrh_host.so (the C++ pass classes and
Application over the C ABI, g++) and the Python module _app (pybind11, g++).

Every HIP translation unit is compiled with -ffp-contract=off so the shading arithmetic
rounds exactly as written, and with packed-FP32 instructions disabled.
Usage: python -m research_amd.build [--jobs N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
HOST = HERE / "host"
INCLUDE = HERE.parent / "include"
OUT = HERE / "librrh.so"
HOST_OUT = HERE / "librrh_host.so"
OBJ = HERE / "_build"
HOST_SOURCES = ["render.cpp", "passes.cpp", "application.cpp", "gltf.cpp"]
CXX = os.environ.get("CXX", "g++")

SOURCES = [
    "rrh_core.hip",
    "rrh_pt.hip",
    "rrh_di.hip",
    "rrh_atrous.hip",
    "rrh_guide.hip",
    "rrh_mlp.hip",
    "rrh_bvh.cpp",
    "rrh_bvh_gpu.hip",
    "rrh_gbuffer.hip",
    "rrh_restir_pt.hip",
    "rrh_restir_gi.hip",
    "rrh_scenegen.cpp",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
DEVICE = [
    "--offload-arch=gfx950", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _compile(src: str) -> Path:
    path = CSRC / src
    obj = OBJ / (src + ".o")
    cmd = [HIPCC] + COMMON
    if src.endswith(".hip"):
        cmd += ["-x", "hip"] + DEVICE
    cmd += ["-c", str(path), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int = 8, verbose: bool = False) -> Path:
    OBJ.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", str(tmp)] + [str(o) for o in objs] + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


def _run(cmd: list[str], what: str) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{what} failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def build_host(verbose: bool = False) -> Path:
    """C++ host layer over librrh.so (no device code: g++), then the pybind11 module."""
    import sysconfig

    import pybind11

    flags = ["-O2", "-std=c++17", "-fPIC", "-Wall", f"-I{INCLUDE}", f"-I{HOST}"]
    rpath = "-Wl,-rpath,$ORIGIN"
    tmp = HOST_OUT.with_suffix(".so.tmp")
    _run([CXX, "-shared"] + flags + [str(HOST / f) for f in HOST_SOURCES] +
         [f"-L{HERE}", "-lrrh", rpath, "-o", str(tmp)], "host library build")
    os.replace(tmp, HOST_OUT)
    ext = HERE / ("_app" + sysconfig.get_config_var("EXT_SUFFIX"))
    tmp = ext.with_suffix(".tmp")
    _run([CXX, "-shared"] + flags + [f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}",
                                     str(HOST / "bindings.cpp"), f"-L{HERE}", "-lrrh_host", "-lrrh", rpath,
                                     "-o", str(tmp)], "python module build")
    os.replace(tmp, ext)
    if verbose:
        print(f"built {HOST_OUT} and {ext}")
    return ext


def clean() -> None:
    shutil.rmtree(OBJ, ignore_errors=True)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    a = ap.parse_args()
    try:
        build(a.jobs, verbose=True)
        build_host(verbose=True)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
