"""Build recipe of the MI355X integrator (no cmake/ninja needed).

  lib/libamb.so                                  HIP kernels + C ABI (hipcc, gfx950)
  activemodelbplus/integrator/integrator*.so     pybind11 module over the C ABI (g++)

Run ``python build.py`` from anywhere; ``build()`` is what ``__graft_entry__.build``
calls.  Objects go to ``_build/`` (git-ignored).
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(REPO, "include")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIBAMB = os.path.join(LIBDIR, "libamb.so")
PKG = os.path.join(HERE, "activemodelbplus", "integrator")

HIP_SOURCES = ["amb_step.hip", "amb_capi.hip", "amb_batch.hip", "amb_unet.hip", "amb_sk.hip", "amb_noalloc.hip"]
ARCH = os.environ.get("AMB_OFFLOAD_ARCH", "gfx950")

# Packed FP32 VALU ops are disabled for all device code.
HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=on",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
    f"-I{INCLUDE}", f"-I{CSRC}",
]


def _run(cmd, quiet=True):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    if not quiet:
        out = "\n".join(l for l in (res.stdout + res.stderr).splitlines() if "packed-fp32-ops" not in l)
        if out.strip():
            print(out)
    return res


def _hipcc():
    return os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def module_path() -> str:
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(PKG, "integrator" + suffix)


def build_lib(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)

    def compile_one(src):
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        _run([_hipcc(), *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj], quiet=not verbose)
        return obj

    with ThreadPoolExecutor(max_workers=min(6, len(HIP_SOURCES))) as ex:
        objs = list(ex.map(compile_one, HIP_SOURCES))
    tmp = LIBAMB + ".tmp"
    _run([_hipcc(), "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, *objs,
          "-L/opt/rocm/lib", "-lhipfft", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIBAMB)
    return LIBAMB


def build_module(verbose: bool = False) -> str:
    import pybind11

    os.makedirs(PKG, exist_ok=True)
    out = module_path()
    tmp = out + ".tmp"
    py_inc = sysconfig.get_paths()["include"]
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden",
           f"-I{pybind11.get_include()}", f"-I{py_inc}", f"-I{INCLUDE}",
           os.path.join(CSRC, "pyintegrator.cpp"), "-o", tmp,
           f"-L{LIBDIR}", "-lamb", "-Wl,-rpath,$ORIGIN/../../lib", "-Wl,-rpath,/opt/rocm/lib"]
    _run(cmd, quiet=not verbose)
    os.replace(tmp, out)
    return out


def build(verbose: bool = False) -> None:
    build_lib(verbose)
    build_module(verbose)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIBAMB)
    print(module_path())
