"""Build recipe for the native parts, in-tree (no JIT cache):

  libfluidphysics.so   HIP kernels + C ABI, hipcc --offload-arch=gfx950
  fp_napi.node         N-API addon for the Node.js host (g++, links the .so)

    python -m fluidphysics_amd.build
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # exact IEEE double for the cell coordinates, the hash's JS path and the knife
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-value", "-Wno-unused-result",
]

NODE_INCLUDE = "/usr/include/node"

SOURCES = ["fp_api.hip", "fp_pool.hip", "fp_grid.hip", "fp_fluid.hip", "fp_solid.hip", "fp_step.hip"]


def build_library(verbose=False):
    srcs = [os.path.join(HERE, "csrc", f) for f in SOURCES]
    out = os.path.join(HERE, "libfluidphysics.so")
    tmp = out + ".tmp"
    cmd = [HIPCC] + HIP_FLAGS + ["-I", os.path.join(ROOT, "include"), "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for %s" % " ".join(SOURCES))
    os.replace(tmp, out)
    return out


def build_addon(verbose=False):
    """The N-API addon; skipped (returns None) where the Node headers are absent."""
    src = os.path.join(HERE, "addon", "fp_napi.cc")
    if not os.path.exists(src) or not os.path.isdir(NODE_INCLUDE):
        return None
    out = os.path.join(HERE, "addon", "fp_napi.node")
    cxx = shutil.which("g++") or "g++"
    cmd = [cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-DNAPI_VERSION=6", "-DNODE_GYP_MODULE_NAME=fp_napi",
           "-I", NODE_INCLUDE, "-I", os.path.join(ROOT, "include"),
           "-o", out + ".tmp", src, "-L", HERE, "-lfluidphysics", "-Wl,-rpath,$ORIGIN/.."]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed for %s" % src)
    os.replace(out + ".tmp", out)
    return out


def build_all(verbose=False):
    lib = build_library(verbose)
    addon = build_addon(verbose)
    return lib, addon


if __name__ == "__main__":
    print(build_all(verbose=True))
