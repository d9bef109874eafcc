"""Builds libp4a.so (gfx950 kernels + host runtime) and the Node.js addon.

Every build compiles from source; outputs stay in the source tree
(particles4all_amd/lib/libp4a.so, particles4all_amd/node/p4a.node).
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "lib", "libp4a.so")
NODE_SRC = os.path.join(PKG, "node", "p4a_napi.cc")
NODE_OUT = os.path.join(PKG, "node", "p4a.node")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# One IEEE operation per source operation (the exact-math kernels depend on
# it); packed FP32 instructions disabled.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", INCLUDE,
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-I", INCLUDE]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l
                    and "is not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out)


def build_lib(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    jobs = []
    objs = []
    for src in ("kernels.hip", "sim.hip", "batch.hip", "mesh.hip"):
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        cmd = [HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        jobs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                           stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    scene_obj = os.path.join(BUILD, "scene.o")
    _run(["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, "scene.cpp"), "-o", scene_obj], verbose)
    objs.append(scene_obj)
    for cmd, p in jobs:
        out, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs, verbose)
    os.replace(tmp, LIB)
    return LIB


def build_node(verbose=False):
    """N-API addon over libp4a.so (Node.js 12, N-API headers in /usr/include/node)."""
    inc = "/usr/include/node"
    if not os.path.exists(NODE_SRC):
        return None
    if not os.path.exists(os.path.join(inc, "node_api.h")):
        if verbose:
            print("node_api.h not found; skipping the Node.js addon")
        return None
    libdir = os.path.dirname(LIB)
    tmp = NODE_OUT + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", inc, "-I", INCLUDE,
          NODE_SRC, "-o", tmp, "-L", libdir, "-lp4a", "-Wl,-rpath,$ORIGIN/../lib"], verbose)
    os.replace(tmp, NODE_OUT)
    return NODE_OUT


def build_oracle(verbose=False):
    """Checker only: compiles oracle/liboracle.so (see oracle/Makefile)."""
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-B", "-C", odir], verbose)
    return os.path.join(odir, "liboracle.so")


def build_all(verbose=False):
    build_lib(verbose)
    build_node(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", LIB)
