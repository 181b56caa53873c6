"""Build the nlmech_amd HIP library for gfx950 and its C++ host layer
(in-tree).

    python -m nlmech_amd.build          # or __graft_entry__.build()

Every source under nlmech_amd/csrc is compiled with hipcc for gfx950 into
nlmech_amd/_build/*.o and linked into nlmech_amd/lib/libnlmech_amd.so.  The
host layer (nlmech_amd/host: input deck, .vtu I/O, FDModel /
QuasiStaticModel) is compiled with g++ into nlmech_amd/lib/libnlmech_host.so,
and the programs under nlmech_amd/host/tools into nlmech_amd/bin/.  The build
always recompiles (it is also the "does it build" check).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
SRC = PKG / "csrc"
BUILD = PKG / "_build"
LIBDIR = PKG / "lib"
LIB = LIBDIR / "libnlmech_amd.so"
ARCH = os.environ.get("NLM_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

# -ffp-contract=off: no implicit multiply-add fusion anywhere, so the exact
# precision mode keeps the reference's rounding; fast paths call fma
# explicitly.  Packed FP32 instructions are disabled.
HIPFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-command-line-argument",
]


HOST = PKG / "host"
BINDIR = PKG / "bin"
HOST_LIB = LIBDIR / "libnlmech_host.so"
INCLUDE = PKG.parent / "include"
CXX = shutil.which("g++") or "g++"
CXXFLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
            "-Wextra", f"-I{INCLUDE}", f"-I{HOST}"]
ROCM_LIB = "/opt/rocm/lib"


def _run(cmd: list[str], what: str) -> str:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{what} failed:\n{' '.join(cmd)}\n{r.stdout}\n"
                           f"{r.stderr}")
    return r.stderr


def _host_compile(src: Path) -> Path:
    obj = BUILD / ("host_" + src.stem + ".o")
    _run([CXX, *CXXFLAGS, "-c", str(src), "-o", str(obj)], f"g++ {src.name}")
    return obj


def build_host() -> Path:
    """libnlmech_host.so and the nlm_fd / nlm_mesh / nlm_dc programs."""
    BUILD.mkdir(exist_ok=True)
    BINDIR.mkdir(exist_ok=True)
    srcs = sorted(HOST.glob("*.cpp"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_host_compile, srcs))
    tmp = HOST_LIB.with_suffix(".so.tmp")
    _run([CXX, "-shared", "-o", str(tmp), *[str(o) for o in objs],
          f"-L{LIBDIR}", "-lnlmech_amd", "-lz", "-Wl,-rpath,$ORIGIN",
          f"-Wl,-rpath-link,{ROCM_LIB}"], "link libnlmech_host.so")
    os.replace(tmp, HOST_LIB)
    for tool in sorted((HOST / "tools").glob("*.cpp")):
        exe = BINDIR / tool.stem
        tmp = exe.with_name(exe.name + ".tmp")
        _run([CXX, *CXXFLAGS, str(tool), "-o", str(tmp), f"-L{LIBDIR}",
              "-lnlmech_host", "-lnlmech_amd", "-lz",
              "-Wl,-rpath,$ORIGIN/../lib", f"-Wl,-rpath-link,{ROCM_LIB}"],
             f"g++ {tool.name}")
        os.replace(tmp, exe)
    return HOST_LIB


def sources() -> list[Path]:
    return sorted(list(SRC.glob("*.hip")) + list(SRC.glob("*.cpp")))


def _compile(src: Path) -> tuple[Path, str]:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, "-x", "hip", *HIPFLAGS, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{' '.join(cmd)}\n"
                           f"{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(_compile, srcs))
    if verbose:
        for obj, err in results:
            if err.strip():
                print(err, file=sys.stderr)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *[str(o) for o, _ in results]]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_host()
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
