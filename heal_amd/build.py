"""Builds the HEAL MI355X backend in-tree.

``build_all()`` compiles every HIP source under ``heal_amd/csrc`` for gfx950
with hipcc and links ``heal_amd/libheal_amd.so`` (the C ABI declared in
``include/heal_amd.h``).  Object files go to ``heal_amd/_build``.  Run as
``python -m heal_amd.build``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libheal_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("HEAL_AMD_ARCH", "gfx950")

SOURCES = ["hx_runtime.hip", "hx_elementwise.hip", "hx_ntt.hip", "hx_keyswitch.hip", "hx_ntt_mfma.hip"]

HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    # no packed-FP32 instructions anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


API_DIR = os.path.join(HERE, "api")
API_LIB = os.path.join(HERE, "libheal_lattica.so")


def _torch_flags():
    import sysconfig

    import torch
    from torch.utils import cpp_extension as ce

    inc = [f"-I{p}" for p in ce.include_paths()] + [f"-I{sysconfig.get_paths()['include']}"]
    libdirs = ce.library_paths()
    abi = int(torch.compiled_with_cxx11_abi())
    cflags = ["-O2", "-std=c++17", "-fPIC", f"-D_GLIBCXX_USE_CXX11_ABI={abi}", "-w", *inc]
    ldflags = [f"-L{d}" for d in libdirs] + [f"-Wl,-rpath,{d}" for d in libdirs]
    return cflags, ldflags


def _module_path() -> str:
    import sysconfig

    return os.path.join(HERE, "lattica_hw" + sysconfig.get_config_var("EXT_SUFFIX"))


def build_api() -> list[str]:
    """The C++ HEAL API templates (libheal_lattica.so) and the Python module lattica_hw."""
    os.makedirs(BUILD, exist_ok=True)
    cflags, ldflags = _torch_flags()
    api_obj = os.path.join(BUILD, "lattica_hw_api.o")
    bind_obj = os.path.join(BUILD, "py_bindings.o")
    with cf.ThreadPoolExecutor(max_workers=2) as ex:
        jobs = [
            ex.submit(_run, ["g++", *cflags, "-c", os.path.join(API_DIR, "lattica_hw_api.cpp"), "-o", api_obj]),
            ex.submit(_run, ["g++", *cflags, "-DTORCH_API_INCLUDE_EXTENSION_H", "-DTORCH_EXTENSION_NAME=lattica_hw",
                             "-c", os.path.join(API_DIR, "py_bindings.cpp"), "-o", bind_obj]),
        ]
        for j in jobs:
            j.result()
    origin = "-Wl,-rpath,$ORIGIN"
    _run(["g++", "-shared", "-o", API_LIB + ".tmp", api_obj, f"-L{HERE}", "-lheal_amd", origin, *ldflags,
          "-ltorch", "-ltorch_cpu", "-lc10"])
    os.replace(API_LIB + ".tmp", API_LIB)
    mod = _module_path()
    _run(["g++", "-shared", "-o", mod + ".tmp", bind_obj, f"-L{HERE}", "-lheal_lattica", "-lheal_amd", origin,
          *ldflags, "-ltorch", "-ltorch_cpu", "-lc10", "-ltorch_python"])
    os.replace(mod + ".tmp", mod)
    return [API_LIB, mod]


def build_all(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, len(srcs)))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    build_api()
    return LIB


if __name__ == "__main__":
    print(build_all())
