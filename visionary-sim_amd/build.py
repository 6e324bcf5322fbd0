"""Builds the native parts of the package in-tree.

* ``lib/libvsim.so`` -- the gfx950 kernels and the C ABI of ``include/vsim.h``
  (hipcc, one object per source, compiled in parallel).
* ``lib/vsim.node`` -- the Node-API addon the JavaScript host loads (when the
  Node-API headers are present).

Floating point: ``-ffp-contract=off`` (no FMA contraction), IEEE divide and
square root, denormals kept, packed-FP32 instructions off. These make the
kernels' f32 arithmetic the oracle's bit for bit.

Usage: ``python visionary-sim_amd/build.py [--jobs N]``.
"""
import argparse
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("VSIM_ARCH", "gfx950")

FP_FLAGS = [
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
COMMON = ["-std=c++17", "-O3", "-fPIC", "-Wall", "-I" + os.path.join(ROOT, "include")]

SOURCES = ["kernels.hip", "lattice.hip", "grab.hip", "voxel.hip", "backend.cpp", "world.cpp", "server.cpp", "ply.cpp"]


def _run(cmds):
    procs = [(c, subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)) for c in cmds]
    failed = []
    for cmd, p in procs:
        out, _ = p.communicate()
        text = "\n".join(l for l in out.decode().splitlines() if "packed-fp32-ops" not in l and l.strip())
        if p.returncode != 0:
            failed.append((cmd, text))
        elif text:
            sys.stderr.write(text + "\n")
    if failed:
        for cmd, text in failed:
            sys.stderr.write(" ".join(cmd) + "\n" + text + "\n")
        raise RuntimeError("native build failed")


def build_lib():
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    objs, cmds = [], []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(BUILD, src + ".o")
        lang = ["-x", "hip", "--offload-arch=" + ARCH]
        cmds.append([HIPCC] + lang + COMMON + FP_FLAGS + ["-c", path, "-o", obj])
        objs.append(obj)
    _run(cmds)
    out = os.path.join(LIB, "libvsim.so")
    tmp = out + ".tmp"
    _run([[HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs])
    os.replace(tmp, out)
    return out


def build_addon():
    """Node-API addon over libvsim.so; skipped when the headers are missing."""
    src = os.path.join(HERE, "node", "addon.cpp")
    inc = "/usr/include/node"
    if not os.path.exists(src) or not os.path.exists(os.path.join(inc, "node_api.h")):
        return None
    out = os.path.join(LIB, "vsim.node")
    tmp = out + ".tmp"
    cmd = [shutil.which("g++") or "g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-I" + inc,
           "-I" + os.path.join(ROOT, "include"), src, "-o", tmp, "-L" + LIB, "-lvsim",
           "-Wl,-rpath,$ORIGIN"]
    _run([cmd])
    os.replace(tmp, out)
    return out


def build_oracle():
    """The CPU oracle (test infrastructure), oracle/Makefile."""
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--no-oracle", action="store_true")
    args = ap.parse_args()
    print(build_lib())
    addon = build_addon()
    if addon:
        print(addon)
    if not args.no_oracle:
        build_oracle()


if __name__ == "__main__":
    main()
