"""Build the native parts of kite_amd in-tree.

* ``kite_amd/lib/libkite_gpu.so`` -- HIP kernels for gfx950 + the C ABI
  declared in ``include/kite_gpu.h``.
* ``kite_amd/lib/libkite_synth.so`` -- seeded synthetic FASTQ generator.
* ``kite_amd/lib/kite`` -- the C++ command-line host (count / dump / merge /
  histo / analyze / info) linked against libkite_gpu.so.
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
HOST = PKG / "host"
LIB = PKG / "lib"
OBJ = PKG / "lib" / "obj"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-g0",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
GPU_SOURCES = ["scan.hip", "count.hip", "merge.hip", "corr.hip", "encode.hip", "dump.hip", "inflate.hip", "hostlink.hip",
               "devmem.hip", "abi.hip"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        sys.stderr.write(out + "\n")


def _hip_object(src: str) -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    obj = OBJ / (Path(src).stem + ".o")
    _run([HIPCC, *HIP_FLAGS, "-I", str(ROOT / "include"), "-c", str(CSRC / src), "-o", str(obj)])
    return obj


def build_gpu() -> Path:
    srcs = [s for s in GPU_SOURCES if (CSRC / s).exists()]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_hip_object, srcs))
    out = LIB / "libkite_gpu.so"
    tmp = LIB / "libkite_gpu.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
          "-lhsa-runtime64", "-lpthread"])
    os.replace(tmp, out)
    return out


def build_synth() -> Path:
    out = LIB / "libkite_synth.so"
    _run(["gcc", "-O2", "-fPIC", "-shared", "-std=gnu11", "-o", str(out), str(CSRC / "synth.c"),
          "-lpthread"])
    return out


def build_host() -> Path | None:
    srcs = sorted(HOST.glob("*.cpp")) if HOST.exists() else []
    if not srcs:
        return None
    out = LIB / "kite"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", str(ROOT / "include"), "-I", str(HOST),
          "-o", str(out), *map(str, srcs), f"-L{LIB}", "-lkite_gpu", "-Wl,-rpath,$ORIGIN",
          "-lpthread", "-lz"])
    return out


def build_oracle() -> None:
    """The CPU oracle is test infrastructure; it is compiled here so the
    parity tests, smoke() and the CPU baseline leg of bench.py find it."""
    _run(["make", "-s", "-C", str(ROOT / "oracle")])


def build_all() -> None:
    LIB.mkdir(parents=True, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=3) as ex:
        futs = [ex.submit(build_gpu), ex.submit(build_synth), ex.submit(build_oracle)]
        for f in futs:
            f.result()
    build_host()
    shutil.rmtree(OBJ, ignore_errors=True)


if __name__ == "__main__":
    build_all()
    print("built:", ", ".join(sorted(p.name for p in LIB.iterdir())))
