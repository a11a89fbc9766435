"""Build recipe of the magpy_amd native code (run from anywhere).

  python magpy_amd/build.py        # run by path: the package itself imports
                                   # the binding this script builds

1. libmagpy_amd.so   hipcc, gfx950, from csrc/ensemble_run.hip and
                     csrc/adjoint.hip and csrc/dom.hip (kernels in
                     csrc/*_kernels.hpp);
                     exports the C ABI of include/magpy_amd.h.
2. core.*.so         Cython binding core.pyx, linked against libmagpy_amd.so
                     with an $ORIGIN rpath.

Outputs are written next to this file, so they travel with the source tree.
"""
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BUILD_DIR = os.path.join(ROOT, "_build")
LIB = os.path.join(HERE, "libmagpy_amd.so")
ARCH = os.environ.get("MAGPY_AMD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=" + ARCH,
    # packed FP32 (v_pk_*_f32) stays off in every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd, **kw):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, **kw)


def core_ext_path():
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(HERE, "core" + suffix)


SOURCES = ["ensemble_run.hip", "adjoint.hip", "dom.hip"]


def build_lib():
    srcs = [os.path.join(HERE, "csrc", s) for s in SOURCES]
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB + ".tmp"
    _run([hipcc] + HIPCC_FLAGS + ["-I" + os.path.join(ROOT, "include")] + srcs + ["-o", tmp])
    os.replace(tmp, LIB)


def build_core():
    os.makedirs(BUILD_DIR, exist_ok=True)
    pyx = os.path.join(HERE, "core.pyx")
    cpp = os.path.join(BUILD_DIR, "magpy_amd_core.cpp")
    _run([sys.executable, "-m", "cython", "-3", "--cplus", "-I", os.path.join(ROOT, "include"),
          pyx, "-o", cpp])
    inc = sysconfig.get_paths()["include"]
    out = core_ext_path()
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-I" + inc,
          "-I" + os.path.join(ROOT, "include"), cpp, "-o", tmp, "-L" + HERE, "-lmagpy_amd",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)


def build():
    build_lib()
    build_core()


if __name__ == "__main__":
    build()
