"""Build recipe for the native libraries (no cmake: direct g++/hipcc invocations).

Outputs (all in-tree, git-ignored, shipped with the source tree):
  isar_imaging_cuda_amd/lib/libisar_imaging.so   HIP pipeline + C-ABI (gfx950)
  isar_imaging_cuda_amd/lib/libisar_synth.so     synthetic-echo generator / .wbd writer (host)
  isar_imaging_cuda_amd/lib/isar_calling_app     drop-in caller reproducing calling_dll.cpp
  oracle/build/libisar_oracle.so                 FP64 oracle (test infrastructure, built here
                                                 so tests and the CPU baseline can load it)
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
LIB = PKG / "lib"
INC = ROOT / "include"
ORACLE = ROOT / "oracle"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
GXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"

# Packed FP32 instructions are disabled for every device compile: some of their forms have given
# wrong results on MI355X, intermittently and without an error.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++20", "-fPIC", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-ffp-contract=fast",
]
# Per-source additions. pa.hip: multiply-adds are fused only where one source expression has
# them, so the one-workgroup and the split autofocus kernels, which inline the same transform
# code into different surroundings, round identically. fft_debug.hip: the same, so that its
# one-transform-per-workgroup and its wave-grouped test kernels can be compared bit for bit.
HIP_FILE_FLAGS = {"pa.hip": ["-ffp-contract=on"], "fft_debug.hip": ["-ffp-contract=on"]}
HOST_FLAGS = ["-O3", "-std=c++20", "-fPIC", "-Wall", "-fopenmp"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")


def _newer(target: Path, deps: list[Path]) -> bool:
    if not target.exists():
        return False
    t = target.stat().st_mtime
    return all(d.stat().st_mtime <= t for d in deps if d.exists())


def build_oracle(force: bool = True) -> Path:
    out = ORACLE / "build" / "libisar_oracle.so"
    out.parent.mkdir(parents=True, exist_ok=True)
    srcs = [ORACLE / "isar_oracle.cpp"]
    if force or not _newer(out, srcs + [ORACLE / "isar_oracle.h"]):
        _run([GXX, *HOST_FLAGS, "-march=native", "-shared", "-I", str(ORACLE), *map(str, srcs),
              "-o", str(out)])
    return out


def build_synth(force: bool = True) -> Path:
    out = LIB / "libisar_synth.so"
    LIB.mkdir(parents=True, exist_ok=True)
    srcs = [CSRC / "host" / "synth.cpp"]
    if force or not _newer(out, srcs + [INC / "isar_synth.h"]):
        _run([GXX, *HOST_FLAGS, "-shared", "-I", str(INC), *map(str, srcs), "-o", str(out)])
    return out


def _hip_sources() -> list[Path]:
    return sorted((CSRC / "kernels").glob("*.hip")) + sorted((CSRC / "host").glob("*.hip"))


def _host_sources() -> list[Path]:
    return sorted(p for p in (CSRC / "host").glob("*.cpp") if p.name not in ("synth.cpp", "pipeline.cpp")) + \
        [CSRC / "host" / "pipeline.cpp"]


def build_imaging(force: bool = True, jobs: int = 8) -> Path | None:
    hip = _hip_sources()
    if not hip:
        return None
    out = LIB / "libisar_imaging.so"
    objdir = PKG / "_build"
    objdir.mkdir(parents=True, exist_ok=True)
    LIB.mkdir(parents=True, exist_ok=True)
    headers = list(INC.glob("*.h")) + list((CSRC).rglob("*.hpp"))
    incs = ["-I", str(INC), "-I", str(CSRC)]

    def compile_one(src: Path) -> Path:
        obj = objdir / (src.stem + (".hip.o" if src.suffix == ".hip" else ".o"))
        if force or not _newer(obj, [src] + headers):
            if src.suffix == ".hip":
                _run([HIPCC, *HIP_FLAGS, *HIP_FILE_FLAGS.get(src.name, []), *incs, "-c", str(src), "-o", str(obj)])
            else:
                _run([HIPCC, "-O3", "-std=c++20", "-fPIC", "-Wall", "-D__HIP_PLATFORM_AMD__", *incs,
                      "-I", "/opt/rocm/include", "-x", "c++", "-c", str(src), "-o", str(obj)])
        return obj

    srcs = hip + _host_sources() + [CSRC / "host" / "synth.cpp"]

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = out.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-fopenmp", *map(str, objs), "-o", str(tmp),
          "-Wl,-soname,libisar_imaging.so"])
    os.replace(tmp, out)
    app_src = CSRC / "app" / "isar_calling_app.cpp"
    if app_src.exists():
        app = LIB / "isar_calling_app"
        _run([GXX, "-O2", "-std=c++20", "-I", str(INC), str(app_src), "-o", str(app),
              "-L", str(LIB), "-lisar_imaging", f"-Wl,-rpath,{LIB}", "-Wl,-rpath,$ORIGIN"])
    return out


def build_emu(force: bool = True, sanitize: bool = True, jobs: int = 8) -> Path:
    """CPU build of the kernel sources through csrc/emu/hip_emu.hpp (test infrastructure).

    Host-only g++ build: the sanitizers apply to host code only (there is no device code)."""
    objdir = PKG / "_build" / "emu"
    objdir.mkdir(parents=True, exist_ok=True)
    out = objdir / "isar_emu_driver"
    flags = ["-std=c++20", "-O1", "-g", "-pthread", "-DISAR_EMU", "-fno-omit-frame-pointer",
             "-I", str(INC), "-I", str(CSRC)]
    if sanitize:
        flags += ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"]
    # the imaging kernels only: the matrix-core kernels of nn.hip and the GPU frame generator
    # have no host emulation
    srcs = [p for p in _hip_sources() if p.name not in ("nn.hip", "synth_gpu.hip")] + [CSRC / "host" / "pipeline.cpp",
                                                                 CSRC / "emu" / "emu_driver.cpp"]
    headers = list(INC.glob("*.h")) + list(CSRC.rglob("*.hpp"))

    def compile_one(src: Path) -> Path:
        obj = objdir / (src.stem + ".o")
        if force or not _newer(obj, [src] + headers):
            _run([GXX, *flags, "-x", "c++", "-c", str(src), "-o", str(obj)])
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, srcs))
    _run([GXX, *flags, *map(str, objs), "-o", str(out)])
    return out


def build_all(force: bool = True) -> dict[str, Path | None]:
    res = {"oracle": build_oracle(force), "synth": build_synth(force)}
    res["imaging"] = build_imaging(force)
    return res


if __name__ == "__main__":
    if "--emu" in sys.argv:
        print("emu:", build_emu(force="--lazy" not in sys.argv))
    else:
        for k, v in build_all(force="--lazy" not in sys.argv).items():
            print(f"{k}: {v}")
