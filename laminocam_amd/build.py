"""Build recipe for the gfx950 HIP library (liblaminocam.so), in tree.

Every .hip source under csrc/ is compiled by hipcc for gfx950 and linked into
laminocam_amd/lib/liblaminocam.so (with zlib, for the file formats' deflate
streams).  Packed FP32 instructions are disabled
(`-target-feature -packed-fp32-ops`).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "liblaminocam.so")

ARCH = os.environ.get("LC_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def flags():
    return ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-result", "-Wno-unused-function",
            "-I", os.path.join(HERE, "..", "include")]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
    os.makedirs(OBJ, exist_ok=True)
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [hipcc()] + flags() + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def build(verbose=False, jobs=None):
    srcs = sources()
    jobs = jobs or min(8, len(srcs))
    with cf.ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(_compile, srcs))
    if verbose:
        for (o, err) in results:
            if err.strip():
                print(err, file=sys.stderr)
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + [o for o, _ in results] + ["-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


CXX_DIR = os.path.join(HERE, "..", "tests", "cxx")
CXX_TESTS = ("test_shim", "test_host_api", "test_toml", "test_config")
CXX_TEST_BIN = os.path.join(CXX_DIR, "_bin", "test_shim")


def build_cxx_test():
    """g++ build of the C++ API test programs against the in-tree library."""
    os.makedirs(os.path.join(CXX_DIR, "_bin"), exist_ok=True)
    cxx = os.environ.get("CXX") or shutil.which("g++") or "g++"
    out = []
    for name in CXX_TESTS:
        dst = os.path.join(CXX_DIR, "_bin", name)
        cmd = [cxx, "-std=c++17", "-O2", "-Wall", "-I", os.path.join(HERE, "..", "include"),
               os.path.join(CXX_DIR, name + ".cpp"), "-o", dst, "-L", LIBDIR, "-llaminocam",
               "-Wl,-rpath,$ORIGIN/../../../laminocam_amd/lib"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"C++ test build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        out.append(dst)
    return out


CLI_DIR = os.path.join(HERE, "cli")
BIN_DIR = os.path.join(HERE, "bin")
CLI_PROGRAMS = ("recon_tomo", "recon_lamino", "forward_scalar")


def build_cli():
    """g++ build of the reconstruction programs (recon_tomo, recon_lamino,
    forward_scalar) into laminocam_amd/bin/, linked to the in-tree library."""
    os.makedirs(BIN_DIR, exist_ok=True)
    cxx = os.environ.get("CXX") or shutil.which("g++") or "g++"
    out = []
    for name in CLI_PROGRAMS:
        dst = os.path.join(BIN_DIR, name)
        cmd = [cxx, "-std=c++17", "-O2", "-Wall", "-I", os.path.join(HERE, "..", "include"),
               os.path.join(CLI_DIR, name + ".cpp"), "-o", dst + ".tmp", "-L", LIBDIR, "-llaminocam",
               "-Wl,-rpath,$ORIGIN/../lib"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"program build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(dst + ".tmp", dst)
        out.append(dst)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_cxx_test())
    print(build_cli())
