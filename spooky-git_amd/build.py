"""Builds libspk.so (the HIP kernels + C ABI) in-tree for gfx950, and the host
side of a run: libspkio.so (HDF5 snapshots, include/spk_io.h) and the
`spooky` driver executable.

    python spooky-git_amd/build.py [-j N] [--lib-only]

Each csrc/*.hip file is compiled by hipcc into _build/, then linked into
spooky-git_amd/lib/libspk.so.  Packed-FP32 instructions are disabled for all
device code (-Xclang -target-feature -Xclang -packed-fp32-ops).  The host
files (host/*.cpp) are plain C++17 built with g++; the snapshot library links
the HDF5 C library found under HDF5_DIR (default /opt/conda).
"""
import argparse
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libspk.so")
ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
HOST = os.path.join(PKG, "host")
BINDIR = os.path.join(PKG, "bin")
INCLUDE = os.path.join(os.path.dirname(PKG), "include")
HDF5_DIR = os.environ.get("HDF5_DIR", "/opt/conda")

CFLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def compile_one(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + CFLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, r.stdout[-6000:], r.stderr[-6000:]))
    return obj


def build(jobs=None, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (r.stdout[-4000:], r.stderr[-4000:]))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def _run(cmd, what):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("%s failed:\n%s\n%s\n%s" % (what, " ".join(cmd), r.stdout[-4000:], r.stderr[-4000:]))


def build_host(verbose=False):
    """libspkio.so and bin/spooky; skipped (with a message) when the HDF5 C
    library is not present."""
    h5inc, h5lib = os.path.join(HDF5_DIR, "include"), os.path.join(HDF5_DIR, "lib")
    if not os.path.exists(os.path.join(h5inc, "hdf5.h")):
        if verbose:
            print("HDF5 not found under %s: libspkio.so and spooky not built" % HDF5_DIR)
        return None
    os.makedirs(BINDIR, exist_ok=True)
    flags = ["-std=c++17", "-O2", "-Wall", "-Wextra", "-fPIC", "-I" + INCLUDE, "-I" + HOST]
    io = os.path.join(LIBDIR, "libspkio.so")
    _run([CXX] + flags + ["-I" + h5inc, "-shared", "-o", io + ".tmp", os.path.join(HOST, "spk_io.cpp"),
                          "-L" + h5lib, "-lhdf5", "-Wl,-rpath," + h5lib], "libspkio link")
    os.replace(io + ".tmp", io)
    exe = os.path.join(BINDIR, "spooky")
    srcs = [os.path.join(HOST, f) for f in ("spk_driver.cpp", "spk_run.cpp", "spk_config.cpp")]
    _run([CXX] + flags + ["-o", exe + ".tmp"] + srcs +
         ["-L" + LIBDIR, "-lspk", "-lspkio", "-Wl,-rpath,$ORIGIN/../lib"], "spooky link")
    os.replace(exe + ".tmp", exe)
    if verbose:
        print("built", io, "and", exe)
    return exe


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("--lib-only", action="store_true")
    a = ap.parse_args()
    try:
        build(a.j, verbose=True)
        if not a.lib_only:
            build_host(verbose=True)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
