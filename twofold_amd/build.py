"""Builds libtwofold_hip.so in place with hipcc for gfx950.

Flags: -ffp-contract=off keeps every multiply and add separate, as the reference builds with
(CMakeLists.txt:32-37), so results match bit for bit; packed fp32 instructions are switched off.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = [os.path.join(HERE, "csrc", "solver.hip")]
OUT = os.path.join(HERE, "libtwofold_hip.so")
SCENE_SOURCES = [os.path.join(HERE, "csrc", "scene.cpp")]
SCENE_OUT = os.path.join(HERE, "libtwofold_scene.so")
CXX = os.environ.get("CXX", "g++")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument",
         "-I" + os.path.join(ROOT, "include")]


def _run(cmd, out, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd + ["-o", out + ".tmp"], capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"{cmd[0]} failed")
    os.replace(out + ".tmp", out)


def build(verbose=False):
    """The HIP solver library, and the host scene library (narrowphase for driving it)."""
    _run([HIPCC] + FLAGS + SOURCES, OUT, verbose)
    _run([CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-I" + os.path.join(ROOT, "include")]
         + SCENE_SOURCES, SCENE_OUT, verbose)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
