"""Resident graph workers: one long-lived `mgbfs graph-session` process per
device, fed job files, so a sequence of BFS runs pays process start-up,
HIP initialisation and (for repeated graphs) engine allocation once.

Mirrors the reference's resident-rank protocol (multigpubfs/generic_session.py:
24-144, crates/mgbfs-runtime/src/generic_session.rs:6-71): job files
`job-NNNNNNNN.json` = {schema: 1, sequence, command, args, env}, answers
`response-NNNNNNNN.json` = {schema: 1, sequence, ok, result | error, cache},
`phase-NNNNNNNN.json` progress, `cancel-NNNNNNNN` to stop a job at the next
layer, `shutdown` to end the worker.  Only MGBFS_* / NCCL_* variables travel
with a job.  A worker whose job failed is never reused.

    with persistent_graph_workers():
        for g in graphs:
            run_graph(g, out / g.name, device=0)   # served by the resident worker
"""
from __future__ import annotations

import contextlib
import json
import os
import subprocess
import tempfile
import time
from pathlib import Path

_active = None


def _read_json(path: Path):
    try:
        return json.loads(path.read_text())
    except (FileNotFoundError, ValueError):
        return None


class Job:
    def __init__(self, worker, sequence, command):
        self.worker = worker
        self.sequence = sequence
        self.command = command
        self.pid = worker.process.pid
        self.returncode = None
        self.response = None

    def poll(self):
        if self.returncode is None:
            response = _read_json(self.worker.root / f"response-{self.sequence:08}.json")
            if response is not None:
                if response.get("schema") != 1 or response.get("sequence") != self.sequence:
                    raise RuntimeError("GENERIC_SESSION_RESPONSE_IDENTITY")
                self.response = response
                self.returncode = 0 if response.get("ok") is True else 1
            elif self.worker.process.poll() is not None:
                self.returncode = self.worker.process.returncode or 1
        return self.returncode

    def wait(self, timeout=None):
        start = time.monotonic()
        nap = 0.0005
        while self.poll() is None:
            if timeout is not None and time.monotonic() - start > timeout:
                raise subprocess.TimeoutExpired("resident graph job", timeout)
            time.sleep(nap)
            nap = min(nap * 2, 0.02)
        return self.returncode

    def phase(self):
        value = _read_json(self.worker.root / f"phase-{self.sequence:08}.json")
        if value is None:
            return None
        if value.get("schema") != 1 or value.get("pid") != self.pid or value.get("sequence") != str(self.sequence):
            raise RuntimeError("GENERIC_SESSION_PHASE_IDENTITY")
        return value

    def request_cancel(self):
        (self.worker.root / f"cancel-{self.sequence:08}").touch()

    @property
    def error(self):
        if self.response is not None and self.response.get("ok") is False:
            return self.response.get("error")
        return None if self.returncode in (None, 0) else "GENERIC_SESSION_WORKER_EXITED"


class Worker:
    def __init__(self, native, env, rank, device, root: Path):
        self.rank, self.device, self.root = rank, device, root
        root.mkdir(parents=True)
        self.sequence = 0
        self.job = None
        self.log = (root / "worker.log").open("w")
        startup = {k: v for k, v in env.items() if not k.startswith("MGBFS_") or k == "MGBFS_EXECUTABLE"}
        startup.update(RANK=str(rank), LOCAL_RANK=str(device))
        self.process = subprocess.Popen([str(native), "graph-session", str(root), str(rank), str(device)],
                                        env=startup, stdout=self.log, stderr=subprocess.STDOUT)

    def submit(self, command, args, env=None):
        if self.job is not None and self.job.poll() is None:
            raise RuntimeError("GENERIC_SESSION_CONCURRENT_JOB")
        if self.process.poll() is not None or (self.job is not None and self.job.returncode):
            raise RuntimeError("GENERIC_SESSION_FAILED_NO_RESTART")
        job = Job(self, self.sequence, command)
        env = dict(env or {}, MGBFS_JOB_PHASE_PATH=str(self.root / f"phase-{self.sequence:08}.json"),
                   MGBFS_JOB_SEQUENCE=str(self.sequence))
        path = self.root / f"job-{self.sequence:08}.json"
        tmp = path.with_suffix(".tmp")
        tmp.write_text(json.dumps({"schema": 1, "sequence": self.sequence, "command": command,
                                   "args": [str(a) for a in args],
                                   "env": {k: str(v) for k, v in env.items() if k.startswith(("MGBFS_", "NCCL_"))}}))
        os.replace(tmp, path)
        self.sequence += 1
        self.job = job
        return job

    def close(self):
        (self.root / "shutdown").touch()
        try:
            self.process.wait(timeout=5)
        except subprocess.TimeoutExpired:
            self.process.terminate()
            try:
                self.process.wait(timeout=10)
            except subprocess.TimeoutExpired:
                self.process.kill()
                self.process.wait()
        self.log.close()


class Pool:
    """One worker per selected device; reselecting other devices or another
    executable replaces the workers."""

    def __init__(self):
        self.temporary = tempfile.TemporaryDirectory(prefix="mgbfs-graph-session-",
                                                     dir=os.environ.get("MGBFS_GENERIC_SESSION_LOG_ROOT"))
        self.root = Path(self.temporary.name)
        self.workers = []
        self.identity = None
        self.generation = 0

    def select(self, native, devices, env=None):
        key = (str(native), tuple(devices))
        if key == self.identity and any(w.process.poll() is not None or (w.job and w.job.returncode)
                                        for w in self.workers):
            raise RuntimeError("GENERIC_SESSION_FAILED_GROUP_NO_RESTART")
        if key != self.identity:
            self.close_workers()
            root = self.root / str(self.generation)
            self.generation += 1
            env = dict(os.environ if env is None else env)
            try:
                for rank, device in enumerate(devices):
                    self.workers.append(Worker(native, env, rank, device, root / str(rank)))
                self.identity = key
            except BaseException:
                self.close_workers()
                raise
        return self.workers

    def worker_for(self, native, device, env=None):
        return self.select(native, [device], env)[0]

    def inventory(self, native, devices, env=None, seconds=30.0):
        jobs = [w.submit("inventory", [], env) for w in self.select(native, devices, env)]
        for job in jobs:
            if job.wait(seconds):
                raise RuntimeError("GENERIC_SESSION_INVENTORY_FAILED " + repr(job.response))
        return [job.response["result"] for job in jobs]

    def release_memory(self, seconds=30.0):
        jobs = [w.submit("release-memory", []) for w in self.workers]
        for job in jobs:
            if job.wait(seconds):
                raise RuntimeError("RESIDENT_MEMORY_HANDOFF_FAILED")

    def close_workers(self):
        for worker in self.workers:
            worker.close()
        self.workers = []
        self.identity = None

    def close(self):
        self.close_workers()
        self.temporary.cleanup()


@contextlib.contextmanager
def persistent_graph_workers():
    """Route run_graph (and so the CayleyPy backend) through resident workers."""
    global _active
    if _active is not None:
        raise RuntimeError("GENERIC_SESSION_NESTED")
    pool = Pool()
    _active = pool
    try:
        yield pool
    finally:
        _active = None
        pool.close()


def active_pool():
    return _active
