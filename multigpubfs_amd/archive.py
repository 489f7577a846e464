"""Readers of the layer archive that `mgbfs graph --archive PATH` writes.

Wire format (selected-samples form, magic `MGBFSAS2`):

    header  48 B   b"MGBFSAS2", state width n (u64 LE), 32-byte config digest
    frame   80 B   b"MGBFSFR1", kind, depth, count, payload bytes, sequence
                   (u64 LE each), chain hash of everything before
            payload
            32 B   sha256(frame head + payload) = the next chain hash
    chain0 = sha256(header)

Kinds: 1 records (count states, n bytes each, at most 1000 per layer),
2 layer end (count = the layer's records), 4 replacement of the last
layer's records (at most once, terminal layer only), 3 commit (depth = the
number of layers, count = all records).

`read_prefix` follows multigpubfs/specialized.py:54-79 (same checks, same
error codes, keeps the last two layers' records); `verify` follows
crates/mgbfs-runtime/src/archive.rs:716-800 (streaming check of all three
magics, codes ARCHIVE_*).
"""
from __future__ import annotations

import hashlib
import io
import struct

MAGIC_SELECTED = b"MGBFSAS2"
MAGIC_STATES = b"MGBFSAS3"
MAGIC_HASHED = b"MGBFSAR1"
FRAME_MAGIC = b"MGBFSFR1"
SAMPLE_LIMIT = 1000


def _exact(stream, size):
    data = bytearray()
    while len(data) < size:
        part = stream.read(size - len(data))
        if not part:
            raise EOFError("SPECIALIZED_ARCHIVE_TRUNCATED")
        data.extend(part)
    return bytes(data)


def read_prefix(stream, n):
    """Reads one committed archive of n-byte states from a binary stream.

    Returns {depths, layers: {depth: bytes} (last two layers), config_digest
    (hex), terminal_replaced}; raises RuntimeError with the reference's codes
    (specialized.py:54-79) and EOFError('SPECIALIZED_ARCHIVE_TRUNCATED')."""
    header = _exact(stream, 48)
    if header[:8] != MAGIC_SELECTED or struct.unpack_from("<Q", header, 8)[0] != n:
        raise RuntimeError("SPECIALIZED_ARCHIVE_HEADER")
    chain = hashlib.sha256(header).digest()
    seq = depth = total = rows = previous_rows = 0
    buffer = bytearray()
    layers = {}
    replaced = False
    while True:
        frame = _exact(stream, 80)
        kind, at, count, size, sequence = struct.unpack_from("<5Q", frame, 8)
        expected = depth - 1 if kind == 4 else depth
        if (frame[:8] != FRAME_MAGIC or at != expected or sequence != seq or frame[48:] != chain
                or size > SAMPLE_LIMIT * n):
            raise RuntimeError("SPECIALIZED_ARCHIVE_ORDER_BOUND")
        payload = _exact(stream, size)
        chain = hashlib.sha256(frame + payload).digest()
        if _exact(stream, 32) != chain:
            raise RuntimeError("SPECIALIZED_ARCHIVE_CHECKSUM")
        if kind == 1:
            if not count or size != count * n or rows + count > SAMPLE_LIMIT:
                raise RuntimeError("SPECIALIZED_ARCHIVE_RECORD_BOUND")
            buffer.extend(payload)
            rows += count
        elif kind == 2:
            if size or count != rows:
                raise RuntimeError("SPECIALIZED_ARCHIVE_LAYER")
            layers[depth] = bytes(buffer)
            for old in sorted(layers)[:-2]:
                del layers[old]
            total += rows
            previous_rows = rows
            rows = 0
            depth += 1
            buffer.clear()
        elif kind == 4:
            if size or count or rows or not depth or replaced:
                raise RuntimeError("SPECIALIZED_ARCHIVE_REPLACEMENT")
            replaced = True
            depth -= 1
            total -= previous_rows
        elif kind == 3:
            if size or rows or not depth or count != total:
                raise RuntimeError("SPECIALIZED_ARCHIVE_COMMIT")
            return dict(depths=depth, layers=layers, config_digest=header[16:].hex(), terminal_replaced=replaced)
        else:
            raise RuntimeError("SPECIALIZED_ARCHIVE_FRAME")
        seq += 1


def read_archive(path, n):
    with open(path, "rb") as stream:
        return read_prefix(stream, n)


def decode_layer(receipt, depth, n, labels=None):
    """States of one retained layer as lists of ints (label-mapped when given)."""
    raw = receipt["layers"].get(depth, b"")
    if len(raw) % n:
        raise RuntimeError("SPECIALIZED_STATE_SHAPE")
    rows = [list(raw[i:i + n]) for i in range(0, len(raw), n)]
    if labels is not None:
        if any(v >= len(labels) for v in raw):
            raise RuntimeError("SPECIALIZED_STATE_ALPHABET")
        rows = [[labels[v] for v in row] for row in rows]
    return rows


def _walk(data):
    """Checks frames one by one (archive.rs:716-800) and yields each as
    (kind, depth, count, payload, width) once it has passed; returns
    {depths, records} at the commit frame."""
    reader = io.BytesIO(data) if isinstance(data, (bytes, bytearray, memoryview)) else data

    def read(size, eof):
        chunk = reader.read(size)
        if len(chunk) != size:
            raise RuntimeError(eof)
        return chunk

    header = read(48, "ARCHIVE_HEADER")
    selected = header[:8] == MAGIC_SELECTED
    state_only = selected or header[:8] == MAGIC_STATES
    if not state_only and header[:8] != MAGIC_HASHED:
        raise RuntimeError("ARCHIVE_HEADER")
    width = struct.unpack_from("<Q", header, 8)[0]
    if width == 0 or width > 33025:
        raise RuntimeError("ARCHIVE_WIDTH")
    record_bytes = width + (0 if state_only else 16)
    chain = hashlib.sha256(header).digest()
    sequence = depth = count = total = previous = 0
    replaced = closed = False
    while True:
        head = read(80, "ARCHIVE_TRUNCATED")
        kind, at, records, size, seq = struct.unpack_from("<5Q", head, 8)
        expected = (depth - 1 if depth else None) if kind == 4 else depth
        if head[:8] != FRAME_MAGIC or head[48:] != chain or seq != sequence or at != expected:
            raise RuntimeError("ARCHIVE_CHAIN")
        if closed and kind != 3:
            raise RuntimeError("ARCHIVE_REPLACEMENT_PHASE")
        sha = hashlib.sha256(head)
        parts = []
        remaining = size
        while remaining:
            step = min(remaining, 1 << 16)
            parts.append(read(step, "ARCHIVE_TRUNCATED"))
            sha.update(parts[-1])
            remaining -= step
        stored = read(32, "ARCHIVE_TRUNCATED")
        chain = sha.digest()
        if stored != chain:
            raise RuntimeError("ARCHIVE_CHECKSUM")
        if kind == 1:
            if records == 0 or records * record_bytes != size:
                raise RuntimeError("ARCHIVE_RECORD_SHAPE")
            count += records
            if selected and not replaced and count > SAMPLE_LIMIT:
                raise RuntimeError("ARCHIVE_SELECTED_RECORD_PHASE")
        elif kind == 2:
            if size or records != count:
                raise RuntimeError("ARCHIVE_LAYER_COUNT")
            total += count
            previous = count
            count = 0
            depth += 1
            closed = closed or replaced
        elif kind == 3:
            if size or depth == 0 or count or records != total:
                raise RuntimeError("ARCHIVE_RUN_COUNT")
            return dict(depths=depth, records=total)
        elif kind == 4:
            if not selected or replaced or size or records or count or depth == 0:
                raise RuntimeError("ARCHIVE_REPLACEMENT_PHASE")
            depth -= 1
            total -= previous
            replaced = True
        else:
            raise RuntimeError("ARCHIVE_FRAME_KIND")
        yield kind, at, records, b"".join(parts), width
        sequence += 1


def verify(data):
    """Checks a whole archive held in memory or a binary stream; returns
    {depths, records} of the committed prefix (archive.rs:716-800).  Bytes
    after the commit frame (preallocated capacity) are ignored."""
    walk = _walk(data)
    while True:
        try:
            next(walk)
        except StopIteration as done:
            return done.value


def layers(data):
    """Yields (depth, states) for every layer of a committed archive, states
    being the layer's records concatenated (n bytes each; the hashed form's
    16-byte hash plane removed).  A replaced terminal layer is yielded once,
    with its replacement records.  The whole archive is checked as by
    `verify`; nothing is yielded past a frame that fails, and a layer is
    yielded only once the next frame shows it was not replaced."""
    pending = None
    records = bytearray()
    for kind, depth, count, payload, width in _walk(data):
        if kind == 1:
            if len(payload) != count * width:  # hashed form: state plane first
                payload = payload[:count * width]
            records.extend(payload)
            continue
        if kind == 4:
            pending = None
            continue
        if pending is not None:
            yield pending
            pending = None
        if kind == 2:
            pending = (depth, bytes(records))
            records.clear()
    if pending is not None:
        yield pending
