"""Build librfd.so (HIP, gfx950) in-tree.

Every .hip file under rockfit-dem_amd/csrc is compiled for gfx950 and linked
into rockfit-dem_amd/rfd/librfd.so; the C++ host driver under
rockfit-dem_amd/host is linked against it as rockfit-dem_amd/host/rfd_run.
Flags that matter for results:

* -ffp-contract=off: no FMA contraction, every FP32 product and sum rounds on
  its own, in the order written (the reference-layout kernels are bit-exact
  against a sequential CPU restatement because of this);
* -fhip-fp32-correctly-rounded-divide-sqrt: IEEE division and sqrt;
* -target-feature -packed-fp32-ops: no packed FP32 instructions.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
CSRC = PKG.parent / "csrc"
INCLUDE = PKG.parent.parent / "include"
LIB = PKG / "librfd.so"
OBJ = PKG.parent / "_build"
HOST = PKG.parent / "host"
RUN = HOST / "rfd_run"
# host code: the same no-contraction rule, so that its FP64 parameter
# arithmetic rounds as written
HOST_FLAGS = ["-O2", "-std=c++17", "-ffp-contract=off", "-Wall", f"-I{INCLUDE}"]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]


def sources() -> list[pathlib.Path]:
    return sorted(CSRC.glob("*.hip"))


def _compile(src: pathlib.Path) -> pathlib.Path:
    obj = OBJ / (src.stem + ".o")
    cmd = [HIPCC, *FLAGS, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = False, jobs: int | None = None) -> pathlib.Path:
    """Compile all HIP sources and link librfd.so; returns the library path."""
    OBJ.mkdir(parents=True, exist_ok=True)
    srcs = sources()
    if not srcs:
        raise RuntimeError(f"no HIP sources under {CSRC}")
    jobs = jobs or min(8, len(srcs))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    build_host()
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources, and {RUN}", file=sys.stderr)
    return LIB


def build_host() -> pathlib.Path:
    """Link the C++ host driver (host/*.cpp) against librfd.so."""
    srcs = sorted(HOST.glob("*.cpp"))
    tmp = RUN.with_suffix(".tmp")
    cmd = [HIPCC, *HOST_FLAGS, *map(str, srcs), "-o", str(tmp), f"-L{PKG}", "-lrfd",
           "-Wl,-rpath,$ORIGIN/../rfd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host driver build failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, RUN)
    return RUN


if __name__ == "__main__":
    build(verbose=True)
