"""Resident graph workers and the CayleyPy catalog adapter on the CPU: the
job-file protocol of `mgbfs graph-session` (schema 1, sequence numbers, env
filtering, errors end the worker, shutdown), the Python Worker/Job/Pool
wrappers, and from_catalog's validation (multigpubfs/catalog.py:3-13)."""
import json
import subprocess
import sys
import time
import types
from pathlib import Path
from types import SimpleNamespace

import pytest

from multigpubfs_amd import from_catalog, lrx, persistent_graph_workers
from multigpubfs_amd.session import Worker

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "multigpubfs_amd" / "bin" / "mgbfs"
needs_exe = pytest.mark.skipif(not EXE.exists(), reason="mgbfs not built")


def start(root: Path):
    root.mkdir()
    return subprocess.Popen([str(EXE), "graph-session", str(root), "0", "0"], stdout=subprocess.PIPE,
                            stderr=subprocess.PIPE, text=True)


def write_job(root: Path, sequence: int, body: dict):
    tmp = root / "job.tmp"
    tmp.write_text(json.dumps(body))
    tmp.rename(root / f"job-{sequence:08}.json")


def response(root: Path, sequence: int, timeout=20):
    path = root / f"response-{sequence:08}.json"
    end = time.monotonic() + timeout
    while not path.exists():
        assert time.monotonic() < end, "no response"
        time.sleep(0.01)
    return json.loads(path.read_text())


@needs_exe
@pytest.mark.parametrize("body,code", [
    ({"schema": 1, "sequence": 5, "command": "inventory", "args": [], "env": {}}, "GENERIC_SESSION_SEQUENCE"),
    ({"schema": 2, "sequence": 0, "command": "inventory", "args": [], "env": {}}, "GENERIC_SESSION_SEQUENCE"),
    ({"schema": 1, "sequence": 0, "command": "graph", "args": [], "env": {"PATH": "/x"}}, "GENERIC_SESSION_ENV"),
    ({"schema": 1, "sequence": 0, "command": "reboot", "args": [], "env": {}}, "GENERIC_SESSION_COMMAND"),
    ({"schema": 1, "sequence": 0, "command": "graph", "args": ["only-one"], "env": {}}, "CLI_USAGE"),
])
def test_rejected_job_answers_and_ends_the_worker(tmp_path, body, code):
    root = tmp_path / "w"
    proc = start(root)
    write_job(root, 0, body)
    r = response(root, 0)
    assert r["schema"] == 1 and r["sequence"] == 0 and r["ok"] is False and r["error"].startswith(code)
    assert proc.wait(timeout=20) == 1  # never reused after an error
    assert json.loads(proc.stderr.read().strip().splitlines()[-1])["status"] == "ERROR"


@needs_exe
def test_shutdown_file_ends_an_idle_worker(tmp_path):
    root = tmp_path / "w"
    proc = start(root)
    time.sleep(0.2)
    (root / "shutdown").touch()
    assert proc.wait(timeout=20) == 0


@needs_exe
def test_worker_wrapper_phase_and_no_restart(tmp_path):
    w = Worker(EXE, {"PATH": "/usr/bin:/bin"}, 0, 0, tmp_path / "w")
    try:
        job = w.submit("reboot", [])
        assert job.wait(20) == 1 and job.error == "GENERIC_SESSION_COMMAND"
        phase = job.phase()
        assert phase["phase"] == "failed" and phase["pid"] == job.pid and phase["sequence"] == "0"
        with pytest.raises(RuntimeError, match="GENERIC_SESSION_FAILED_NO_RESTART"):
            w.submit("inventory", [])
    finally:
        w.close()


def test_sessions_do_not_nest():
    with persistent_graph_workers():
        with pytest.raises(RuntimeError, match="GENERIC_SESSION_NESTED"):
            with persistent_graph_workers():
                pass


@pytest.mark.parametrize("name,kwargs,code", [
    ("lrx", {}, "INVALID_CATALOG_CONSTRUCTOR"),
    ("Other.lrx", {}, "INVALID_CATALOG_CONSTRUCTOR"),
    ("PermutationGroups._private", {}, "INVALID_CATALOG_CONSTRUCTOR"),
    ("PermutationGroups.lrx", {"_x": 1}, "INVALID_CATALOG_ARGUMENTS"),
])
def test_catalog_validation(name, kwargs, code):
    with pytest.raises(ValueError, match=code):
        from_catalog(name, kwargs=kwargs)


def test_catalog_requires_cayleypy(monkeypatch):
    monkeypatch.setitem(sys.modules, "cayleypy", None)
    with pytest.raises(RuntimeError, match="CAYLEYPY_OPTIONAL_DEPENDENCY_REQUIRED"):
        from_catalog("PermutationGroups.lrx", args=[5])


def test_catalog_builds_through_from_cayleypy(monkeypatch):
    # Stand-in module with CayleyPy's attribute names (graph_definition.py:130-143).
    def lrx_def(n):
        return SimpleNamespace(generators_type=SimpleNamespace(name="PERMUTATION"),
                               generators_permutations=lrx(n).action["generators"], central_state=list(range(n)),
                               name=f"lrx-{n}", generator_names=["L", "R", "X"])
    fake = types.ModuleType("cayleypy")
    fake.PermutationGroups = SimpleNamespace(lrx=lrx_def)
    monkeypatch.setitem(sys.modules, "cayleypy", fake)
    g = from_catalog("PermutationGroups.lrx", args=[6])
    assert g.action == lrx(6).action and g.start == list(range(6))
    with pytest.raises(ValueError, match="UNKNOWN_CATALOG_CONSTRUCTOR"):
        from_catalog("PermutationGroups.nothing")
