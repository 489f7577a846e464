"""The C-ABI library loads on a CPU-only host and exports every entry point
that include/*.h declares (no compute calls: there may be no GPU here)."""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "multigpubfs_amd" / "lib" / "libmgbfs_hip.so"


def declared_functions():
    names = set()
    for header in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
        for m in re.finditer(r"\b(?:int|void|uint32_t|const char\s*\*)\s*\**\s*(mgbfs_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_the_engine_and_kernel_entry_points():
    names = declared_functions()
    for required in ("mgbfs_engine_create", "mgbfs_engine_advance", "mgbfs_engine_destroy",
                     "mgbfs_packed_bitmap_expand", "mgbfs_hip_abi_version"):
        assert required in names


def test_library_exports_every_declared_symbol():
    if not LIB.exists():
        pytest.fail(f"{LIB} missing: run python -m multigpubfs_amd.build")
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB)], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = sorted(declared_functions() - exported)
    assert not missing, missing


def test_library_loads_and_reports_abi_version_without_a_gpu():
    lib = ctypes.CDLL(str(LIB))
    lib.mgbfs_hip_abi_version.restype = ctypes.c_uint32
    assert lib.mgbfs_hip_abi_version() == 1


def test_native_loader_fails_loudly_when_library_is_missing(tmp_path, monkeypatch):
    from multigpubfs_amd import _native
    monkeypatch.setattr(_native, "LIB_PATH", tmp_path / "missing.so")
    monkeypatch.setattr(_native, "_LIB", None)
    with pytest.raises(_native.NativeUnavailable):
        _native.load()
