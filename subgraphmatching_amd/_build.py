"""Builds the HIP library libsgm_hip.so in place (gfx950), then the C++ row driver
sgm_overall, which uses only the drop-in headers include/sgm/*.h and the library."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsgm_hip.so")
DRIVER = os.path.join(HERE, "sgm_overall")
ORDERS_DRIVER = os.path.join(HERE, "sgm_orders")
ROOT = os.path.dirname(HERE)
SOURCES = ["sgm_kernels.hip", "sgm_enum.hip", "sgm_costmodel.hip", "sgm_host.cpp", "sgm_capi.cpp",
           "sgm_batch.cpp", "sgm_costmodel_host.cpp", "sgm_noalloc.hip", "sgm_tree.cpp", "sgm_orders.cpp",
           "sgm_dynamic.cpp", "sgm_gup.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SGM_OFFLOAD_ARCH", "gfx950")
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         # no packed-fp32 instructions anywhere (none are needed: all arithmetic is integer)
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-fno-strict-aliasing", "-Wall", "-Wno-unused-function"]


def build(verbose: bool = False) -> str:
    """Compile every source and link libsgm_hip.so; returns the library path."""
    objs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = [HIPCC, "-x", "hip", *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
        if verbose and r.stderr.strip():
            print(r.stderr)
        objs.append(obj)
    tmp = LIB + ".tmp"
    r = subprocess.run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-pthread", "-o", tmp],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    build_driver()
    return LIB


def build_driver() -> str:
    """g++ builds of the C++ drivers against the drop-in headers (no HIP headers needed): the
    algorithm rows (sgm_overall) and the order comparison (sgm_orders)."""
    for src, exe in (("overall_performance.cpp", DRIVER), ("order_ablations.cpp", ORDERS_DRIVER)):
        tmp = exe + ".tmp"
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
               os.path.join(HERE, "cpp", src), "-o", tmp,
               "-L", HERE, "-lsgm_hip", "-Wl,-rpath,$ORIGIN", "-pthread"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"g++ failed on {src}:\n{r.stderr}")
        os.replace(tmp, exe)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose=True))
