"""Build recipe for the hgx HIP library (gfx950) and the oracle shared library.

``python -m hipgisaxs_amd.build`` compiles every ``csrc/*.hip`` with hipcc for gfx950 into
``hipgisaxs_amd/libhgx.so`` (in-tree, so it travels with the source tree), then the FP64 oracle
(``oracle/hgx_oracle.cpp``) with g++/OpenMP into ``oracle/liboracle.so`` (test infrastructure).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libhgx.so"
ORACLE_SRC = ROOT / "oracle" / "hgx_oracle.cpp"
ORACLE_LIB = ROOT / "oracle" / "liboracle.so"

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 instructions are switched off for every HIP translation unit.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=fast",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unknown-pragmas",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]
ORACLE_FLAGS = ["-O3", "-fopenmp", "-march=x86-64-v3", "-fPIC", "-shared", "-std=c++17"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def _compile(src: Path) -> Path:
    out = OBJ / (src.stem + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(out)])
    return out


def build_hip(jobs: int = 8) -> Path:
    """Compiles all csrc/*.hip for gfx950 and links libhgx.so."""
    OBJ.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    if not sources:
        raise RuntimeError("no HIP sources found")
    with ThreadPoolExecutor(max_workers=min(jobs, len(sources))) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    return LIB


PEAKS_SRC = PKG / "peaks.hip"
PEAKS_BIN = PKG / "hgx_peaks"


def build_peaks() -> Path:
    """The throughput microbenchmark (FP32/FP64 FMA, v_sin, v_exp, FP32 MFMA): a standalone executable."""
    tmp = PEAKS_BIN.with_suffix(".tmp")
    _run([HIPCC, "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", str(PEAKS_SRC), "-o", str(tmp)])
    os.replace(tmp, PEAKS_BIN)
    return PEAKS_BIN


ADAPTER_SRCS = [PKG / "host" / "hgx_adapter.cpp"]
ADAPTER_LIB = PKG / "libhgx_adapter.so"
HIG_SRCS = [PKG / "host" / n for n in ("hig_input.cpp", "hig_sim.cpp", "hig_fit.cpp", "hull.cpp")]
HIG_LIB = PKG / "libhgx_hig.so"
SHIM_SRC = ROOT / "tests" / "cpp" / "hig_adapter_shim.cpp"
SHIM_LIB = ROOT / "tests" / "cpp" / "libhig_shim.so"


def build_host_adapters() -> Path:
    """The plain-type library behind the reference-interface C++ classes (libhgx_adapter.so), the .hig
    driver (libhgx_hig.so; host code on the C ABI), and the test shim that compiles the classes of
    host/hig_gpu.hpp against stand-ins of the reference's type headers (tests/cpp/refapi) and calls
    them as the reference does."""
    cxx = shutil.which("g++") or "g++"
    common = ["-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", f"-I{ROOT / 'include'}", f"-I{PKG / 'host'}"]
    tmp = ADAPTER_LIB.with_suffix(".so.tmp")
    _run([cxx, *common, *map(str, ADAPTER_SRCS), "-o", str(tmp), f"-L{PKG}", "-lhgx", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, ADAPTER_LIB)
    tmp = HIG_LIB.with_suffix(".so.tmp")
    _run([cxx, *common, *map(str, HIG_SRCS), "-o", str(tmp), f"-L{PKG}", "-lhgx_adapter", "-lhgx",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, HIG_LIB)
    tmp = SHIM_LIB.with_suffix(".so.tmp")
    _run([cxx, *common, f"-I{ROOT / 'tests' / 'cpp' / 'refapi'}", str(SHIM_SRC), "-o", str(tmp), f"-L{PKG}",
          "-lhgx_adapter", "-lhgx", "-Wl,-rpath,$ORIGIN/../../hipgisaxs_amd"])
    os.replace(tmp, SHIM_LIB)
    return HIG_LIB


def build_oracle() -> Path:
    """Compiles the FP64 OpenMP oracle (test infrastructure) into oracle/liboracle.so."""
    cxx = shutil.which("g++") or "g++"
    tmp = ORACLE_LIB.with_suffix(".so.tmp")
    _run([cxx, *ORACLE_FLAGS, str(ORACLE_SRC), "-o", str(tmp)])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def build_all() -> None:
    build_hip()
    build_host_adapters()
    build_peaks()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "all" in what or "hip" in what:
        print("built", build_hip())
        print("built", build_host_adapters())
        print("built", build_peaks())
    if "all" in what or "oracle" in what:
        print("built", build_oracle())
