"""Builds libafhip.so (the af_* C ABI + gfx950 kernels) in place.

    python -m arrayfire_amd.build          # or arrayfire_amd.build.build()

Every source under csrc/ is compiled with hipcc for gfx950 only; objects go to
arrayfire_amd/_build/, the library to arrayfire_amd/lib/libafhip.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libafhip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed-FP32 VALU instructions are disabled for every kernel (the compiler
# would otherwise pair independent FP32 FMAs into v_pk_fma_f32).
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def sources() -> list[str]:
    out = []
    for d, _, files in os.walk(CSRC):
        for f in sorted(files):
            if f.endswith((".hip", ".cpp")):
                out.append(os.path.join(d, f))
    return sorted(out)


def _obj(src: str) -> str:
    rel = os.path.relpath(src, CSRC).replace(os.sep, "_")
    return os.path.join(BUILD, rel + ".o")


def _compile(src: str) -> tuple[str, int, str]:
    obj = _obj(src)
    cmd = [HIPCC, *COMMON, "-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd.insert(1, "-x")
        cmd.insert(2, "hip")
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the GPU-only feature name: drop that note
    err = "\n".join(l for l in r.stderr.splitlines()
                    if "packed-fp32-ops" not in l)
    return src, r.returncode, err


def build(jobs: int | None = None, verbose: bool = False) -> str:
    """Compile every source (always) and link libafhip.so. Returns its path."""
    if not os.path.exists(HIPCC) and shutil.which("hipcc") is None:
        raise RuntimeError("hipcc not found")
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 4, len(srcs))
    failed = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, rc, err in ex.map(_compile, srcs):
            if rc != 0:
                failed.append((src, err))
            elif verbose and err.strip():
                print(err, file=sys.stderr)
    if failed:
        msg = "\n".join(f"--- {s}\n{e}" for s, e in failed)
        raise RuntimeError(f"hipcc failed:\n{msg}")
    tmp = LIB + ".tmp"
    link = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp,
            *[_obj(s) for s in srcs]]
    r = subprocess.run(link, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
