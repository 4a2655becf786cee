"""Build the gfx950 shared library (and, for tests, the CPU oracle).

    python -m fastkpc_amd.build            # library only
    python -m fastkpc_amd.build --oracle   # library + oracle

The library is compiled in-tree with hipcc for gfx950 only; each translation
unit is compiled in parallel and linked into fastkpc_amd/libkpcmi.so.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libkpcmi.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["kpc_dcov.hip", "kpc_dcov_lanczos.hip", "kpc_smooth.hip", "kpc_magic.hip", "kpc_hsic.hip",
           "kpc_perm.hip", "kpc_noalloc.hip", "kpc_orient.cpp",
           "kpc_host.cpp", "kpc_tune.cpp"]
COMMON = [
    "-std=c++17", "-O3", "--offload-arch=gfx950", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=off",
    "-I", os.path.join(ROOT, "include"),
]


def _compile(src, obj_dir):
    obj = os.path.join(obj_dir, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, r.stderr[-4000:]))
    return obj


def build_library(verbose=True):
    obj_dir = os.path.join(HERE, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(lambda s: _compile(s, obj_dir), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % r.stderr[-4000:])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_oracle(verbose=True):
    d = os.path.join(ROOT, "oracle")
    r = subprocess.run(["make", "-s", "-C", d, "-B", "liborc.so"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n%s" % r.stderr[-4000:])
    if verbose:
        print("built", os.path.join(d, "liborc.so"))


if __name__ == "__main__":
    build_library()
    if "--oracle" in sys.argv:
        build_oracle()
