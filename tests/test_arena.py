"""Host-only check of the batch engine's device-arena range bookkeeping
(multigpubfs_amd/csrc/arena.hpp), compiled with g++ and run on the CPU."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
def test_range_set_matches_occupancy_model(tmp_path):
    exe = tmp_path / "arena_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", str(ROOT / "multigpubfs_amd" / "csrc"),
                    str(ROOT / "tests" / "native" / "arena_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "arena ok" in out.stdout


def test_key_mixer_is_bijective_and_spreads_buckets(tmp_path):
    """The sorted backend's key mixer (multigpubfs_amd/csrc/common.hpp)."""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(hipcc).exists():
        pytest.skip("hipcc not installed")
    exe = tmp_path / "keymix_check"
    subprocess.run([hipcc, "-std=c++17", "-O1", "-I", str(ROOT / "multigpubfs_amd" / "csrc"),
                    str(ROOT / "tests" / "native" / "keymix_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "keymix ok" in out.stdout


def test_back_edge_masks_cover_repeated_generators(tmp_path):
    """Back-edge masks of the sorted backend (fill_back_edge_masks in
    multigpubfs_amd/csrc/common.hpp): mask values for LRX with a repeated
    generator, and a CPU model of the pruned layer loop against plain BFS."""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(hipcc).exists():
        pytest.skip("hipcc not installed")
    exe = tmp_path / "backedge_check"
    subprocess.run([hipcc, "-std=c++17", "-O1", "-I", str(ROOT / "multigpubfs_amd" / "csrc"),
                    str(ROOT / "tests" / "native" / "backedge_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "backedge ok" in out.stdout
