"""Layer archive (MGBFSAS2): the Python readers against frame sequences from
the reference's own archive tests, against corruptions of each field, and
against the bytes the native writer (multigpubfs_amd/csrc/archive.hpp)
produces for oracle BFS layers.

Frames here are built by `wire`, a writer written from the format alone
(the reference test's helper shape, tests/test_selected_tail.py:13-20):
header, then per frame head + payload + sha256(head + payload) chained.
"""
import hashlib
import io
import shutil
import struct
import subprocess
from pathlib import Path

import pytest

from multigpubfs_amd import archive as A
from oracle import oracle as O

ROOT = Path(__file__).resolve().parents[1]


def wire(frames, n=3, config=bytes(32), sequences=None, magic=b"MGBFSAS2"):
    header = magic + struct.pack("<Q", n) + config
    chain = hashlib.sha256(header).digest()
    out = bytearray(header)
    for i, (kind, depth, count, payload) in enumerate(frames):
        seq = i if sequences is None else sequences[i]
        head = b"MGBFSFR1" + struct.pack("<QQQQQ", kind, depth, count, len(payload), seq) + chain
        chain = hashlib.sha256(head + payload).digest()
        out.extend(head + payload + chain)
    return bytes(out)


def read(data, n=3):
    return A.read_prefix(io.BytesIO(data), n)


# ------------------------------------------- reference test vectors

def test_samples_and_empty_layer_commit():
    """tests/test_selected_tail.py:57-67 frame list."""
    frames = [(1, 0, 1000, bytes([0, 1, 2]) * 1000), (2, 0, 1000, b""), (2, 1, 0, b""), (3, 2, 1000, b"")]
    r = read(wire(frames))
    assert r["depths"] == 2 and not r["terminal_replaced"]
    assert r["layers"] == {0: bytes([0, 1, 2]) * 1000, 1: b""}
    assert r["config_digest"] == "00" * 32
    assert A.verify(wire(frames)) == {"depths": 2, "records": 1000}


def test_terminal_replacement_stream():
    """tests/test_selected_tail.py:69-77: the replacement carries 3000
    records, which the whole-archive check accepts and the bounded prefix
    reader (at most 1000 states of n bytes per frame) refuses."""
    frames = [(1, 0, 1, bytes([0, 1, 2])), (2, 0, 1, b""),
              (1, 1, 1000, bytes([2, 1, 0]) * 1000), (2, 1, 1000, b""),
              (4, 1, 0, b""), (1, 1, 1500, bytes([2, 1, 0]) * 1500),
              (1, 1, 1500, bytes([0, 2, 1]) * 1500), (2, 1, 3000, b""), (3, 2, 3001, b"")]
    data = wire(frames)
    assert A.verify(data) == {"depths": 2, "records": 3001}
    assert list(A.layers(data)) == [(0, bytes([0, 1, 2])),
                                    (1, bytes([2, 1, 0]) * 1500 + bytes([0, 2, 1]) * 1500)]
    with pytest.raises(RuntimeError, match="SPECIALIZED_ARCHIVE_ORDER_BOUND"):
        read(data)


def test_small_replacement_is_read_as_terminal_replaced():
    frames = [(1, 0, 1, bytes([0, 1, 2])), (2, 0, 1, b""),
              (1, 1, 2, bytes([2, 1, 0]) * 2), (2, 1, 2, b""),
              (4, 1, 0, b""), (1, 1, 1, bytes([1, 0, 2])), (2, 1, 1, b""), (3, 2, 2, b"")]
    r = read(wire(frames))
    assert r["terminal_replaced"] and r["depths"] == 2
    assert r["layers"] == {0: bytes([0, 1, 2]), 1: bytes([1, 0, 2])}
    assert A.verify(wire(frames)) == {"depths": 2, "records": 2}


def test_rust_selected_stream_vectors():
    """crates/mgbfs-runtime/tests/archive.rs:4-24 and :27-37."""
    frames = [(1, 0, 1, bytes([0, 1, 2])), (2, 0, 1, b""),
              (1, 1, 1000, bytes([1]) * 3000), (2, 1, 1000, b""),
              (4, 1, 0, b""), (1, 1, 3000, bytes([2]) * 9000), (2, 1, 3000, b""), (3, 2, 3001, b"")]
    data = wire(frames)
    A.verify(data)
    with pytest.raises(RuntimeError, match="ARCHIVE_TRUNCATED"):
        A.verify(data[:-1])
    damaged = bytearray(data)
    damaged[150] ^= 1
    with pytest.raises(RuntimeError, match="ARCHIVE_"):
        A.verify(bytes(damaged))
    empty = wire([(2, 0, 0, b""), (2, 1, 0, b""), (3, 2, 0, b"")])
    assert A.verify(empty) == {"depths": 2, "records": 0}
    assert read(empty)["layers"] == {0: b"", 1: b""}


def test_only_last_two_layers_are_kept():
    frames = []
    for d in range(5):
        frames += [(1, d, 1, bytes([d, d, d])), (2, d, 1, b"")]
    frames.append((3, 5, 5, b""))
    r = read(wire(frames))
    assert r["depths"] == 5 and sorted(r["layers"]) == [3, 4]
    assert A.decode_layer(r, 4, 3) == [[4, 4, 4]]
    assert A.decode_layer(r, 4, 3, labels=[10, 11, 12, 13, 14]) == [[14, 14, 14]]
    with pytest.raises(RuntimeError, match="SPECIALIZED_STATE_ALPHABET"):
        A.decode_layer(r, 4, 3, labels=[10, 11])


# ------------------------------------------------- corruptions

GOOD = [(1, 0, 2, bytes([0, 1, 2, 2, 1, 0])), (2, 0, 2, b""), (1, 1, 1, bytes([1, 2, 0])), (2, 1, 1, b""),
        (3, 2, 3, b"")]


def test_good_archive_reads():
    r = read(wire(GOOD))
    assert r["depths"] == 2 and r["layers"][0] == bytes([0, 1, 2, 2, 1, 0])


@pytest.mark.parametrize("frames,kwargs,code", [
    (GOOD, {"n": 4}, "SPECIALIZED_ARCHIVE_HEADER"),
    (GOOD, {"magic": b"MGBFSAR1"}, "SPECIALIZED_ARCHIVE_HEADER"),
    (GOOD, {"sequences": [0, 1, 3, 2, 4]}, "SPECIALIZED_ARCHIVE_ORDER_BOUND"),
    ([GOOD[0], GOOD[1], (1, 2, 1, bytes(3))] + GOOD[3:], {}, "SPECIALIZED_ARCHIVE_ORDER_BOUND"),
    ([(1, 0, 2, bytes(5))] + GOOD[1:], {}, "SPECIALIZED_ARCHIVE_RECORD_BOUND"),
    ([(1, 0, 0, b"")] + GOOD[1:], {}, "SPECIALIZED_ARCHIVE_RECORD_BOUND"),
    ([(1, 0, 600, bytes(1800)), (1, 0, 401, bytes(1203)), (2, 0, 1001, b""), (3, 1, 1001, b"")], {},
     "SPECIALIZED_ARCHIVE_RECORD_BOUND"),
    ([GOOD[0], (2, 0, 3, b"")] + GOOD[2:], {}, "SPECIALIZED_ARCHIVE_LAYER"),
    ([GOOD[0], (2, 0, 2, b"x")] + GOOD[2:], {}, "SPECIALIZED_ARCHIVE_LAYER"),
    (GOOD[:4] + [(3, 2, 4, b"")], {}, "SPECIALIZED_ARCHIVE_COMMIT"),
    (GOOD[:3] + [(3, 1, 2, b"")], {}, "SPECIALIZED_ARCHIVE_COMMIT"),
    ([(3, 0, 0, b"")], {}, "SPECIALIZED_ARCHIVE_COMMIT"),
    (GOOD[:4] + [(4, 1, 0, b""), (4, 0, 0, b"")], {}, "SPECIALIZED_ARCHIVE_REPLACEMENT"),
    ([(4, 0, 0, b"")], {}, "SPECIALIZED_ARCHIVE_ORDER_BOUND"),
    (GOOD[:4] + [(5, 2, 0, b"")], {}, "SPECIALIZED_ARCHIVE_FRAME"),
])
def test_reader_rejects(frames, kwargs, code):
    with pytest.raises(RuntimeError, match=code):
        read(wire(frames, **{k: v for k, v in kwargs.items() if k != "n"}), kwargs.get("n", 3))


def test_payload_flip_and_chain_break_and_truncation():
    data = bytearray(wire(GOOD))
    data[48 + 80] ^= 1  # first payload byte
    with pytest.raises(RuntimeError, match="SPECIALIZED_ARCHIVE_CHECKSUM"):
        read(bytes(data))
    with pytest.raises(RuntimeError, match="ARCHIVE_CHECKSUM"):
        A.verify(bytes(data))
    data = bytearray(wire(GOOD))
    data[48 + 60] ^= 1  # previous-chain field of the first frame
    with pytest.raises(RuntimeError, match="SPECIALIZED_ARCHIVE_ORDER_BOUND"):
        read(bytes(data))
    with pytest.raises(RuntimeError, match="ARCHIVE_CHAIN"):
        A.verify(bytes(data))
    with pytest.raises(EOFError, match="SPECIALIZED_ARCHIVE_TRUNCATED"):
        read(wire(GOOD[:4]))
    assert A.verify(wire(GOOD) + bytes(64)) == {"depths": 2, "records": 3}


def test_verify_hashed_form_carries_a_hash_plane():
    frames = [(1, 0, 1, bytes([0, 1, 2]) + bytes(16)), (2, 0, 1, b""), (3, 1, 1, b"")]
    assert A.verify(wire(frames, magic=b"MGBFSAR1")) == {"depths": 1, "records": 1}
    assert list(A.layers(wire(frames, magic=b"MGBFSAR1"))) == [(0, bytes([0, 1, 2]))]
    with pytest.raises(RuntimeError, match="ARCHIVE_RECORD_SHAPE"):
        A.verify(wire(frames[:1] + frames[1:], magic=b"MGBFSAS3"))
    with pytest.raises(RuntimeError, match="ARCHIVE_HEADER"):
        A.verify(wire(frames, magic=b"MGBFSXXX"))


# ------------------------------------------- native writer

@pytest.fixture(scope="module")
def writer(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not installed")
    exe = tmp_path_factory.mktemp("archive") / "archive_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", str(ROOT / "multigpubfs_amd" / "csrc"),
                    str(ROOT / "tests" / "native" / "archive_check.cpp"), "-o", str(exe)], check=True)
    return exe


def run_writer(exe, path, n, config, layers, commit=True):
    lines = [f"{path} {n} {config}"]
    for rows in layers:
        lines.append(f"L {len(rows)}")
        lines += [" ".join(map(str, row)) for row in rows]
    if commit:
        lines.append("C")
    return subprocess.run([str(exe)], input="\n".join(lines) + "\n", capture_output=True, text=True, timeout=60)


@pytest.mark.parametrize("n", [5, 8])
def test_native_writer_round_trips_oracle_layers(writer, tmp_path, n):
    graph = {"action": {"kind": "permutation", "generators": O.lrx_generators(n)}, "start": list(range(n))}
    layers = [rows[:1000] for rows in O.exact_layers(graph, 50000)]
    config = hashlib.sha256(b"lrx%d" % n).hexdigest()
    path = tmp_path / "a.mgbfsar1"
    out = run_writer(writer, path, n, config, layers)
    assert out.returncode == 0 and "archive ok" in out.stdout, out.stdout + out.stderr
    data = path.read_bytes()
    assert A.verify(data) == {"depths": len(layers), "records": sum(len(rows) for rows in layers)}
    r = A.read_archive(path, n)
    assert r["depths"] == len(layers) and r["config_digest"] == config
    for depth in (len(layers) - 2, len(layers) - 1):
        assert A.decode_layer(r, depth, n) == layers[depth]
    walked = list(A.layers(data))
    assert [d for d, _ in walked] == list(range(len(layers)))
    assert all(raw == bytes(v for row in layers[d] for v in row) for d, raw in walked)


@pytest.mark.parametrize("case,code", [
    ("too_many", "ARCHIVE_RECORD_BOUND"),
    ("wide_value", "ARCHIVE_STATE_WIDTH"),
    ("bad_config", "ARCHIVE_CONFIG"),
    ("no_layers", "ARCHIVE_UNCOMMITTED_LAYER"),
])
def test_native_writer_errors(writer, tmp_path, case, code):
    config = "ab" * 32
    layers = {"too_many": [[[0, 1, 2]] * 1001], "wide_value": [[[0, 1, 256]]], "bad_config": [[[0, 1, 2]]],
              "no_layers": []}[case]
    out = run_writer(writer, tmp_path / "a", 3, "zz" * 32 if case == "bad_config" else config, layers)
    assert out.returncode == 2 and f"ERROR {code}" in out.stdout, out.stdout + out.stderr


def test_native_writer_uncommitted_prefix_is_rejected(writer, tmp_path):
    path = tmp_path / "a"
    out = run_writer(writer, path, 3, "ab" * 32, [[[0, 1, 2]], [[1, 2, 0]]], commit=False)
    assert out.returncode == 0
    with pytest.raises(EOFError, match="SPECIALIZED_ARCHIVE_TRUNCATED"):
        A.read_archive(path, 3)
    with pytest.raises(RuntimeError, match="ARCHIVE_TRUNCATED"):
        A.verify(path.read_bytes())
    with pytest.raises(RuntimeError, match="ARCHIVE_TRUNCATED"):
        list(A.layers(path.read_bytes()))
