"""Builds the native libraries in-tree (they travel with the source tree; no JIT cache).

  lib/libspecrel_scene.so      host scene producer (g++)
  lib/libspecrel_hip.so        gfx950 renderer behind the C ABI of include/specrel.h (hipcc)
  lib/libspecrel_mathcheck.so  device evaluation of the renderer's f32 transcendental helpers,
                               for the test that sweeps them against the host definitions (hipcc)

Both are compiled with -ffp-contract=off and without fast-math: the renderer reproduces the
reference's f32 arithmetic bit for bit. HIP code additionally disables the packed-FP32 target
feature.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INC = os.path.join(ROOT, "include")
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")

FP_FLAGS = ["-ffp-contract=off", "-fno-fast-math"]
HIP_ARCH = os.environ.get("SPECREL_HIP_ARCH", "gfx950")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def scene_cmd(out=None):
    out = out or os.path.join(LIB, "libspecrel_scene.so")
    return ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", *FP_FLAGS, "-Wall", "-Wextra", "-I" + INC,
            "-o", out, os.path.join(CSRC, "scene", "scene.cpp"), "-lpthread"]


def hip_cmd(out=None):
    out = out or os.path.join(LIB, "libspecrel_hip.so")
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    return [hipcc, f"--offload-arch={HIP_ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared", *FP_FLAGS,
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-I" + INC,
            "-o", out, os.path.join(CSRC, "hip", "kernels.hip"), os.path.join(CSRC, "hip", "adjoint.hip"),
            os.path.join(CSRC, "hip", "extrude.hip"), os.path.join(CSRC, "hip", "bvh_build.hip"),
            os.path.join(CSRC, "hip", "capi.hip")]


def mathcheck_cmd(out=None):
    out = out or os.path.join(LIB, "libspecrel_mathcheck.so")
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    return [hipcc, f"--offload-arch={HIP_ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared", *FP_FLAGS,
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-I" + INC,
            "-o", out, os.path.join(CSRC, "diag", "mathcheck.hip")]


def build_all(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    cmds = [scene_cmd(), hip_cmd(), mathcheck_cmd()]
    with ThreadPoolExecutor(3) as ex:
        outs = list(ex.map(_run, cmds))
    if verbose:
        for c, o in zip(cmds, outs):
            print(" ".join(c))
            if o.strip():
                print(o)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
