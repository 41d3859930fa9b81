"""Builds the native libraries in-tree.

* ``thindb_amd/libthindb_gpu.so`` — the HIP engine (C-ABI of
  ``include/thindb_gpu.h``), compiled for gfx950 with hipcc.
* ``thindb_amd/libthindb_gen.so`` — the seeded synthetic-table generator
  (host C).

Nothing here is run at import time; ``__graft_entry__.build()`` and the
tests call :func:`build_all` explicitly.
"""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
GPU_LIB = PKG / "libthindb_gpu.so"
GEN_LIB = PKG / "libthindb_gen.so"

HIP_SOURCES = ["scan_agg.hip", "order.hip", "radix.hip", "batch.hip", "decode.hip", "distinct.hip", "runtime.cpp", "batch.cpp",
               "server.cpp", "jit.cpp", "segment.cpp", "noalloc.cpp"]
# Device headers the run-time compiler needs (plan-specialised scan kernels
# and batch kernels), embedded as string literals.
JIT_HEADERS = ["engine.h", "scan_common.h", "strip_generic.h", "order_device.h", "radix_device.h", "batch_device.h"]
ARCH = os.environ.get("THINDB_GPU_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-munsafe-fp-atomics",
    # Packed FP32 instructions stay off in all device code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _hipcc() -> str:
    for cand in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}")


def write_jit_headers(objdir: Path) -> Path:
    """Embeds the device headers as raw string literals for hipRTC."""
    out = ["// generated by thindb_amd/_build.py: device headers for run-time compilation"]
    for i, name in enumerate(JIT_HEADERS):
        text = (CSRC / name).read_text()
        if ")TDBGJIT\"" in text:
            raise RuntimeError(f"{name} contains the raw-string delimiter")
        out.append(f'static const char kJitHeader{i}[] = R"TDBGJIT({text})TDBGJIT";')
    out.append("static const char* const kJitHeaderSrc[] = {" + ", ".join(f"kJitHeader{i}" for i in range(len(JIT_HEADERS))) + "};")
    out.append("static const char* const kJitHeaderName[] = {" + ", ".join(f'"{n}"' for n in JIT_HEADERS) + "};")
    out.append(f"static const int kJitHeaders = {len(JIT_HEADERS)};")
    path = objdir / "jit_headers.inc"
    path.write_text("\n".join(out) + "\n")
    return path


def build_gpu(jobs: int = 4) -> Path:
    """Compiles every HIP source for gfx950 and links libthindb_gpu.so."""
    objdir = ROOT / "_build"
    objdir.mkdir(exist_ok=True)
    write_jit_headers(objdir)
    hipcc = _hipcc()
    objs = []
    procs = []
    for src in HIP_SOURCES:
        obj = objdir / (src + ".o")
        cmd = [hipcc, *HIPCC_FLAGS, "-I", str(objdir), "-x", "hip", "-c", str(CSRC / src), "-o", str(obj)]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        objs.append(obj)
    errors = []
    for cmd, p in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            errors.append(f"{' '.join(cmd)}\n{out}\n{err}")
    if errors:
        raise RuntimeError("HIP build failed:\n" + "\n".join(errors))
    tmp = GPU_LIB.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp), "-lhiprtc"])
    os.replace(tmp, GPU_LIB)
    return GPU_LIB


def build_gen() -> Path:
    tmp = GEN_LIB.with_suffix(".so.tmp")
    _run(["gcc", "-O2", "-fPIC", "-shared", "-pthread", str(CSRC / "gen.c"), "-o", str(tmp), "-lm"])
    os.replace(tmp, GEN_LIB)
    return GEN_LIB


def build_all() -> None:
    build_gen()
    build_gpu()


if __name__ == "__main__":
    build_all()
    print("built", GPU_LIB, GEN_LIB)
