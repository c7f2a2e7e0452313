"""Builds the in-tree native pieces.

  winding_amd/libwinding_amd.so   gfx950 kernels + the C ABI (hipcc)
  winding_amd/winding_amd.node    the N-API addon over it (gcc, Node headers)
  oracle/_build/liboracle.so      the CPU oracle (gcc), test infrastructure

Run as ``python -m winding_amd.build``. Every call compiles from source.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "winding_amd")
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(ROOT, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
NODE_INCLUDE = "/usr/include/node"

LIB = os.path.join(PKG, "libwinding_amd.so")
ADDON = os.path.join(PKG, "winding_amd.node")
ORACLE_LIB = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

# Compiled for gfx950 only. Packed f32 forms stay off, and no multiply-add is
# contracted: the FP64 path's rounding is part of its definition.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Werror", "-Wno-unused-function", "-Wno-unknown-pragmas",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Werror", "-Wno-unknown-pragmas",
              "-D__HIP_PLATFORM_AMD__"]

HIP_SOURCES = ["splat_sort.hip", "sort_passes.hip", "batch.hip", "mlp.hip", "learned.hip", "decode.hip"]
# Sources built once more with other defines: the single sort's passes at 4 and 8 keys a thread
# (tiles of 1024 and 2048, for smaller clouds) and the batch kernels at 8, beside the default 16 (4096).
HIP_VARIANTS = [("sort_passes.hip", "sort_passes.kpt4.hip.o", ["-DWD_KEYS_PER_THREAD=4"]),
                ("sort_passes.hip", "sort_passes.kpt8.hip.o", ["-DWD_KEYS_PER_THREAD=8"]),
                ("batch.hip", "batch.kpt8.hip.o", ["-DWD_KEYS_PER_THREAD=8"])]
HOST_SOURCES = ["api.cpp", "synth.cpp", "batch_host.cpp", "mlp_host.cpp", "learned_host.cpp", "decode_host.cpp"]


def _run(cmd, cwd=ROOT):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n  %s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def _objects():
    os.makedirs(BUILD, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    jobs = []
    for s in HIP_SOURCES:
        src = os.path.join(CSRC, s)
        if os.path.exists(src):
            obj = os.path.join(BUILD, s + ".o")
            jobs.append(([HIPCC] + HIP_FLAGS + inc + ["-c", src, "-o", obj], obj))
    for s, obj_name, defines in HIP_VARIANTS:
        src = os.path.join(CSRC, s)
        obj = os.path.join(BUILD, obj_name)
        jobs.append(([HIPCC] + HIP_FLAGS + defines + inc + ["-c", src, "-o", obj], obj))
    for s in HOST_SOURCES:
        src = os.path.join(CSRC, s)
        if os.path.exists(src):
            obj = os.path.join(BUILD, s + ".o")
            jobs.append((["g++"] + HOST_FLAGS + inc + ["-I", "/opt/rocm/include", "-c", src, "-o", obj], obj))
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(lambda j: _run(j[0]), jobs))
    return [o for _, o in jobs]


def build_library():
    objs = _objects()
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs +
         ["-lpthread", "-lz", "-Wl,-soname,libwinding_amd.so"])
    os.replace(tmp, LIB)
    return LIB


def build_addon():
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        raise RuntimeError("build failed: Node's N-API headers are not at %s" % NODE_INCLUDE)
    src = os.path.join(PKG, "napi", "addon.c")
    tmp = ADDON + ".tmp"
    _run(["gcc", "-O2", "-fPIC", "-shared", "-Wall", "-Werror", "-I", NODE_INCLUDE, "-I", os.path.join(ROOT, "include"),
          "-DNODE_GYP_MODULE_NAME=winding_amd", src, "-o", tmp, "-L", PKG, "-lwinding_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, ADDON)
    return ADDON


def build_oracle():
    os.makedirs(os.path.dirname(ORACLE_LIB), exist_ok=True)
    srcs = [os.path.join(ROOT, "oracle", f) for f in ("splat_oracle.c", "decode_oracle.c")]
    tmp = ORACLE_LIB + ".tmp"
    _run(["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Werror"] + srcs + ["-o", tmp, "-lm"])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def build_all(verbose=False):
    out = [build_library(), build_addon(), build_oracle()]
    if verbose:
        for p in out:
            print("built", os.path.relpath(p, ROOT))
    return out


if __name__ == "__main__":
    build_all(verbose=True)
    sys.exit(0)
