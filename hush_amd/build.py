"""Build recipe of the in-tree native libraries (run: python hush_amd/build.py).

  hush_amd/libhush.so        HIP kernels + C ABI (hipcc, gfx950)
  hush_amd/libhush_scene.so  procedural G-buffer generator (gcc, OpenMP)
  hush_amd/node/hush_napi.node  Node.js N-API addon over libhush.so (g++, when node headers exist)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else "hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",                      # the oracle's operation order, no implicit FMA
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mllvm", "-amdgpu-mfma-vgpr-form",       # MFMA results in VGPRs (no AGPR round trip); MFMA kernels only
    "-Wno-pass-failed", "-Wno-unused-command-line-argument",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]
HIP_SOURCES = ["hush_kernels.hip", "hush_api.hip", "svgf.hip", "hush_train.hip", "scene_gpu.hip", "post.hip"]


def run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if out.strip():
        print(out)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


def build_hip():
    out = os.path.join(HERE, "libhush.so")
    run([HIPCC] + HIP_FLAGS + ["-o", out + ".tmp"] + [os.path.join(CSRC, s) for s in HIP_SOURCES])
    os.replace(out + ".tmp", out)
    return out


def build_scene():
    out = os.path.join(HERE, "libhush_scene.so")
    run(["gcc", "-O2", "-fPIC", "-shared", "-std=gnu11", "-fopenmp", "-fvisibility=hidden", "-Wall",
         os.path.join(CSRC, "scene.c"), "-o", out + ".tmp", "-lm"])
    os.replace(out + ".tmp", out)
    return out


def build_node():
    """The N-API addon links libhush.so by rpath; skipped when no node headers are installed."""
    inc = None
    for c in ("/usr/include/node", "/usr/local/include/node"):
        if os.path.exists(os.path.join(c, "node_api.h")):
            inc = c
            break
    src = os.path.join(HERE, "node", "hush_napi.cc")
    if inc is None or not os.path.exists(src):
        print("node addon: skipped (no node_api.h or source)")
        return None
    out = os.path.join(HERE, "node", "hush_napi.node")
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I" + inc, "-I" + os.path.join(ROOT, "include"),
         "-I" + os.path.join(rocm, "include"), "-D__HIP_PLATFORM_AMD__", "-DNODE_GYP_MODULE_NAME=hush_napi",
         src, "-o", out + ".tmp", "-L" + HERE, "-lhush", "-L" + os.path.join(rocm, "lib"), "-lamdhip64",
         "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath," + os.path.join(rocm, "lib"), "-Wl,--allow-shlib-undefined"])
    os.replace(out + ".tmp", out)
    return out


def build_all():
    build_hip()
    build_scene()
    build_node()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "hip": build_hip, "scene": build_scene, "node": build_node}[w]()
