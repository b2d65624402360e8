"""Builds the native libraries in-tree.

  ltpmg_amd/lib/libltpmg_hip.so   HIP engine for gfx950 (C ABI, include/ltpmg.h)
  ltpmg_amd/lib/libltpmg_gpu.so   the reference's eight extern entry points over the C ABI
                                  (the counterpart of its libgpu; exports only those)
  ltpmg_amd/lib/libltpmg_host.so  C++ host API (drop-in Scheduler/Database/Query/...)
  ltpmg_amd/lib/ltpmg             command-line driver (the reference's ltpmg binary)
  oracle/_build/liboracle.so      the CPU checker (test infrastructure)

Run ``python -m ltpmg_amd.build`` or ``__graft_entry__.build()``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "ltpmg_amd"
LIB = PKG / "lib"
OBJ = PKG / "_build"
CSRC = PKG / "csrc"
HOST = PKG / "host"
INCLUDE = ROOT / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed-FP32 instructions are disabled for all HIP code built here.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", f"-I{INCLUDE}",
]
HIP_SOURCES = ["engine.hip", "kernels_table.hip", "kernels_cc.hip", "kernels_order.hip", "kernels_append.hip"]
CXX = os.environ.get("CXX", "g++")
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-pthread", f"-I{INCLUDE}",
             f"-I{INCLUDE / 'ltpmg_host'}"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}")
    # keep warnings visible but short
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        sys.stderr.write(out + "\n")


def build_hip(jobs: int = 8) -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    LIB.mkdir(parents=True, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = OBJ / (Path(src).stem + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    out = LIB / "libltpmg_hip.so"
    tmp = LIB / "libltpmg_hip.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, out)
    return out


def build_oracle() -> Path:
    _run(["make", "-s", "-C", str(ROOT / "oracle")])
    return ROOT / "oracle" / "_build" / "liboracle.so"


# libltpmg_gpu.so: the eight entry points and what they call, linked only to
# libltpmg_hip.so; everything but the eight is kept local by the version script
GPU_SOURCES = ["gpu_shim.cpp", "pack.cpp", "oe_decompose.cpp"]


def _link(out: Path, objs: list[Path], *libs: str) -> None:
    tmp = out.with_name(out.name + ".tmp")
    _run([CXX, "-shared", "-pthread", "-o", str(tmp), *map(str, objs), *libs])
    os.replace(tmp, out)


def build_host(jobs: int = 8) -> Path | None:
    srcs = sorted(HOST.glob("*.cpp"))
    if not srcs:
        return None
    OBJ.mkdir(parents=True, exist_ok=True)
    LIB.mkdir(parents=True, exist_ok=True)
    cmds, obj = [], {}
    for src in srcs:
        if src.name == "main.cpp":
            continue
        obj[src.name] = OBJ / ("host_" + src.stem + ".o")
        cmds.append([CXX, *CXX_FLAGS, "-c", str(src), "-o", str(obj[src.name])])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    rpath = "-Wl,-rpath,$ORIGIN"
    _link(LIB / "libltpmg_gpu.so", [obj[n] for n in GPU_SOURCES], f"-Wl,--version-script={HOST / 'libltpmg_gpu.map'}",
          f"-L{LIB}", "-lltpmg_hip", rpath)
    host_objs = [o for n, o in obj.items() if n != "gpu_shim.cpp"]
    out = LIB / "libltpmg_host.so"
    _link(out, host_objs, f"-L{LIB}", "-lltpmg_gpu", "-lltpmg_hip", rpath)
    main = HOST / "main.cpp"
    if main.exists():
        exe = LIB / "ltpmg"
        _run([CXX, *CXX_FLAGS, str(main), "-o", str(exe) + ".tmp", f"-L{LIB}", "-lltpmg_host",
              "-lltpmg_gpu", "-lltpmg_hip", rpath])
        os.replace(str(exe) + ".tmp", exe)
    return out


def build_all() -> None:
    build_hip()
    build_oracle()
    build_host()


if __name__ == "__main__":
    build_all()
    print("built:", *(p.name for p in sorted(LIB.iterdir())))
