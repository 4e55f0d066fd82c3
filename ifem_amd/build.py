"""Build recipe of the native pieces (run on the build host, never at import).

* ``build_library()`` compiles ``ifem_amd/csrc/*.hip`` for gfx950 into the
  in-tree ``ifem_amd/libifem_gpu.so`` (C ABI declared in include/ifem_gpu.h).
* ``build_facade()`` compiles the C++ reference-interface facade
  (``ifem_amd/facade``) against that library into ``libifem_facade.so`` and
  its test driver ``facade_test``.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ifem_amd")
LIB = os.path.join(PKG, "libifem_gpu.so")
FACADE_LIB = os.path.join(PKG, "libifem_facade.so")
HDF5_LIB = os.path.join(PKG, "libifem_hdf5.so")
# the image's HDF5 (C library and headers)
HDF5_ROOT = os.environ.get("IFEM_HDF5_ROOT", "/opt/conda")
FACADE_TEST = os.path.join(PKG, "facade_test")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-value", "-Wno-unused-result",
]


def _run(cmd, cwd=ROOT):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd[:3]) + " ...")
    return r


def build_library() -> str:
    """One object per source, compiled in parallel (no device symbol crosses
    a file), then linked into the shared library."""
    from concurrent.futures import ThreadPoolExecutor
    srcs = sorted(glob.glob(os.path.join(PKG, "csrc", "*.hip")))
    odir = os.path.join(ROOT, "_build", "obj")
    os.makedirs(odir, exist_ok=True)
    inc = "-I" + os.path.join(ROOT, "include")

    def compile_one(src):
        obj = os.path.join(odir, os.path.basename(src)[:-4] + ".o")
        _run([HIPCC, *HIP_FLAGS, inc, "-c", "-o", obj, src])
        return obj

    jobs = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_hdf5() -> str:
    """libifem_hdf5.so (include/ifem_hdf5.h): plain C over the HDF5 C
    library, found through its own run path so that nothing else is taken
    from that prefix."""
    inc = os.path.join(HDF5_ROOT, "include")
    lib = os.path.join(HDF5_ROOT, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        sys.stderr.write(f"HDF5 headers not found under {inc}: export library not built\n")
        return ""
    tmp = HDF5_LIB + ".tmp"
    _run(["gcc", "-O2", "-std=c11", "-Wall", "-fPIC", "-shared",
          "-I" + os.path.join(ROOT, "include"), "-I" + inc,
          os.path.join(PKG, "facade", "hdf5_writer.c"), "-o", tmp,
          "-L" + lib, "-l:libhdf5.so.103", "-Wl,-rpath," + lib, "-Wl,--enable-new-dtags"])
    os.replace(tmp, HDF5_LIB)
    return HDF5_LIB


def build_facade() -> str:
    fdir = os.path.join(PKG, "facade")
    srcs = sorted(glob.glob(os.path.join(fdir, "*.C")))
    if not srcs:
        return ""
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + fdir]
    tmp = FACADE_LIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", *inc, "-o", tmp, *srcs,
          "-L" + PKG, "-lifem_gpu", "-lifem_hdf5", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, FACADE_LIB)
    test_src = os.path.join(fdir, "test", "facade_test.cpp")
    if os.path.exists(test_src):
        tmp = FACADE_TEST + ".tmp"
        _run(["g++", "-O2", "-std=c++17", *inc, "-o", tmp, test_src,
              "-L" + PKG, "-lifem_facade", "-lifem_gpu", "-lifem_hdf5", "-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, FACADE_TEST)
    return FACADE_LIB


def build_oracle() -> None:
    _run(["make", "-C", os.path.join(ROOT, "oracle"), "-s"])


if __name__ == "__main__":
    build_library()
    build_hdf5()
    build_facade()
    build_oracle()
    print("built", LIB)
