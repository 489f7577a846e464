"""Build recipe for the native parts of multigpubfs_amd (gfx950 only).

Produces, in-tree:
  multigpubfs_amd/lib/libmgbfs_hip.so   HIP kernels + engine + C ABI
  multigpubfs_amd/bin/mgbfs             command-line host (C++)
Usage: python -m multigpubfs_amd.build   (or __graft_entry__.build()).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "lib" / "libmgbfs_hip.so"
CLI = PKG / "bin" / "mgbfs"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instruction forms are disabled for every device compile.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-result",
          f"-I{ROOT / 'include'}", f"-I{CSRC}"] + DEVICE_FLAGS


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + proc.stdout)
    return proc.stdout


def build(verbose: bool = False) -> dict:
    BUILD.mkdir(exist_ok=True)
    LIB.parent.mkdir(exist_ok=True)
    CLI.parent.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    objects = [BUILD / (s.stem + ".o") for s in sources]
    headers = list(CSRC.glob("*.hpp")) + list((ROOT / "include").glob("*.h"))

    def compile_one(pair):
        src, obj = pair
        return _run([HIPCC, *COMMON, "-c", str(src), "-o", str(obj)])

    workers = min(len(sources), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=max(1, workers)) as pool:
        logs = list(pool.map(compile_one, zip(sources, objects)))
    _run([HIPCC, *COMMON, "-shared", "-o", str(LIB), *map(str, objects), "-lpthread"])
    cli_src = CSRC / "mgbfs_cli.cpp"
    if cli_src.exists():
        _run([HIPCC, *COMMON, str(cli_src), "-o", str(CLI), f"-L{LIB.parent}", "-lmgbfs_hip",
              "-Wl,-rpath,$ORIGIN/../lib", "-lpthread"])
    if verbose:
        for log in logs:
            if log.strip():
                print(log)
    return {"lib": str(LIB), "cli": str(CLI) if cli_src.exists() else None,
            "sources": [s.name for s in sources], "headers": [h.name for h in headers]}


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
