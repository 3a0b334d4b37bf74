"""Build of the gfx950 shared library libefem_gpu.so (in-tree).

hipcc compiles the HIP kernels for gfx950 and the host C++ (the host side
links against the HIP runtime).  Objects go to efem_amd/_build/, the library
to efem_amd/libefem_gpu.so.  Packed-FP32 instructions are disabled for all
device code.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libefem_gpu.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("EFEM_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          "-Wno-unused-variable", "-Wno-unused-result", f"-I{ROOT / 'include'}"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-munsafe-fp-atomics"]

SOURCES = [
    ("device/kernels.hip", True),
    ("device/factor.hip", True),
    # the persistent batched solver, once per workgroup width
    ("device/batch.hip", True, "256"),
    ("device/batch.hip", True, "512"),
    ("device/batch.hip", True, "1024"),
    ("api.cpp", True),
    ("host/model.cpp", False),
    ("host/element.cpp", False),
    ("host/mesh.cpp", False),
    ("host/sparse.cpp", False),
    ("host/dense.cpp", False),
    ("host/input.cpp", False),
    ("host/vtk.cpp", False),
    ("host/devplan.cpp", False),
    ("host/fe.cpp", False),
]

# the C++ interfaces' test driver (tests/cpp), linked against the library
FE_DRIVER_SRC = ROOT / "tests" / "cpp" / "fe_driver.cpp"
FE_DRIVER = ROOT / "tests" / "cpp" / "fe_driver"


def _compile(src: str, device: bool, batch_width: str | None = None) -> Path:
    s = CSRC / src
    obj = BUILD / (src.replace("/", "_") + (f".{batch_width}" if batch_width else "") + ".o")
    cmd = [HIPCC, *COMMON]
    if batch_width:
        cmd += [f"-DEFEM_BATCH_BT={batch_width}"]
    if device:
        cmd += ["-x", "hip", *DEVICE]
    else:
        cmd += ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-fopenmp"]
    cmd += ["-c", str(s), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    jobs = min(len(SOURCES), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda a: _compile(*a), SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-fopenmp", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_fe_driver(verbose: bool = False) -> Path:
    """tests/cpp/fe_driver: a host program written against include/efem/fe.hpp."""
    tmp = FE_DRIVER.with_name(FE_DRIVER.name + ".tmp")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", str(FE_DRIVER_SRC), "-o", str(tmp), f"-L{PKG}", "-lefem_gpu",
           "-Wl,-rpath,$ORIGIN/../../efem_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"fe_driver build failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, FE_DRIVER)
    if verbose:
        print(f"built {FE_DRIVER}")
    return FE_DRIVER


def build_oracle(verbose: bool = False) -> None:
    """The CPU oracle (test infrastructure) and, when the reference tree is
    present, oracle/_ref (the reference's material sources)."""
    oracle = ROOT / "oracle"
    r = subprocess.run(["make", "-s", "-C", str(oracle), "liboracle.so"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{r.stdout}\n{r.stderr}")
    ref = Path(os.environ.get("EFEM_REFERENCE", "/root/reference"))
    if (ref / "src" / "material" / "NeoHookean.cpp").exists() and shutil.which("g++"):
        r = subprocess.run(["make", "-s", "-C", str(oracle), "ref", f"REF={ref}"],
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"oracle/_ref build failed\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built oracle")


if __name__ == "__main__":
    build(verbose=True)
    build_fe_driver(verbose=True)
    if "--oracle" in sys.argv:
        build_oracle(verbose=True)
