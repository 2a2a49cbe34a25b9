"""Build the native libraries in-tree.

libradix_hip.so   the HIP kernels and the C ABI (include/radix_hip.h), gfx950
libradix_synth.so host-only synthetic splat clouds and cameras (seeded)

`python -m radix_amd.build` compiles both; translation units compile in
parallel into radix_amd/_build/ and link next to this file.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
ARCH = os.environ.get("RDX_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["api.hip", "project.hip", "project_f32.hip", "project_ext.hip", "project_packed.hip", "batch.hip", "raster.hip",
               "draw.hip", "ray.hip", "ray_uniform.hip", "ray_nonuniform.hip", "io.hip"]
SYNTH_SOURCES = ["synth.cpp"]

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# Packed fp32 ops stay off: ordinary float code only.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-ffp-contract=on",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")


def _newer(target: str, deps: list[str]) -> bool:
    if not os.path.exists(target):
        return False
    t = os.path.getmtime(target)
    return all(os.path.getmtime(d) <= t for d in deps)


def build_hip(force: bool = True, jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    headers = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h"))]
    headers.append(os.path.join(ROOT, "include", "radix_hip.h"))
    objs = []
    jobs_list = []
    for src in HIP_SOURCES:
        s = os.path.join(CSRC, src)
        o = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(o)
        if force or not _newer(o, [s] + headers):
            jobs_list.append([HIPCC, *HIP_FLAGS, "-c", s, "-o", o])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(_run, c) for c in jobs_list]:
            f.result()
    lib = os.path.join(HERE, "libradix_hip.so")
    tmp = lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, lib)
    return lib


def build_synth(force: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    lib = os.path.join(HERE, "libradix_synth.so")
    srcs = [os.path.join(CSRC, s) for s in SYNTH_SOURCES]
    if force or not _newer(lib, srcs):
        tmp = lib + ".tmp"
        _run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-fopenmp", "-o", tmp, *srcs])
        os.replace(tmp, lib)
    return lib


NODE_INCLUDE = "/usr/include/node"


def build_napi(force: bool = True) -> str | None:
    """The Node N-API addon (radix_amd/napi/radix_napi.node), when Node's headers are present."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        return None
    src = os.path.join(HERE, "napi", "radix_napi.cc")
    out = os.path.join(HERE, "napi", "radix_napi.node")
    if force or not _newer(out, [src, os.path.join(ROOT, "include", "radix_hip.h")]):
        tmp = out + ".tmp"
        _run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-DNODE_GYP_MODULE_NAME=radix_napi",
              "-I" + NODE_INCLUDE, "-o", tmp, src, "-L" + HERE, "-lradix_hip", "-Wl,-rpath,$ORIGIN/.."])
        os.replace(tmp, out)
    return out


def build(force: bool = True) -> None:
    build_synth(force)
    build_hip(force)
    build_napi(force)


if __name__ == "__main__":
    build(force="--no-force" not in sys.argv)
    print("built", os.path.join(HERE, "libradix_hip.so"))
