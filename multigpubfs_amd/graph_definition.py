"""Lossless schema-2 graph description (host-side contract, no BFS here).

Mirrors the reference's ``GraphDefinition`` contract
(multigpubfs/graph_definition.py:17-69, crates/mgbfs-core/src/graph_definition.rs:22-49):
same JSON keys, same validation error codes, same semantic digest, and the
same inverse-closure rule that selects three-bank retention (:84-108).  The
CPU successor/exact-layer oracle of the reference is *not* part of the
product: it lives in oracle/ as test infrastructure.
"""
from __future__ import annotations

import hashlib
import json

I64_MIN = -(1 << 63)
I64_MAX = (1 << 63) - 1
_KEYS = {"schema", "name", "generator_names", "action", "start", "expected_max_unique_states"}


def wrap_i64(x: int) -> int:
    """Two's-complement wrap of an integer to int64."""
    return ((x - I64_MIN) % (1 << 64)) + I64_MIN


def _int64(value):
    if type(value) is not int or value < I64_MIN or value > I64_MAX:
        raise ValueError("GRAPH_INT64_RANGE")
    return value


def _int64_list(values):
    if not isinstance(values, list):
        raise ValueError("GRAPH_VECTOR")
    return [_int64(v) for v in values]


def _plain_copy(value):
    """Copy of a JSON-shaped value (dicts, lists, tuples of scalars); other
    objects are shared, as they are immutable scalars in a valid graph."""
    if isinstance(value, dict):
        return {k: _plain_copy(v) for k, v in value.items()}
    if isinstance(value, (list, tuple)):
        out = [v if isinstance(v, (int, str, float)) or v is None else _plain_copy(v) for v in value]
        return out if isinstance(value, list) else tuple(out)
    return value


class GraphDefinition:
    """Schema-2 graph: ``{schema, name, generator_names, action, start,
    expected_max_unique_states}`` with a permutation or matrix action."""

    def __init__(self, value: dict):
        self._data = _plain_copy(value)
        self.validate()

    # -- constructors ----------------------------------------------------
    @classmethod
    def permutation(cls, generators, start, *, name="", generator_names=None, expected_max_unique_states=None):
        names = generator_names or [str(i) for i in range(len(generators))]
        return cls({"schema": 2, "name": name, "generator_names": names,
                    "action": {"kind": "permutation", "degree": len(start), "generators": generators},
                    "start": start, "expected_max_unique_states": expected_max_unique_states})

    @classmethod
    def matrix(cls, rows, cols, generators, start, *, name="", generator_names=None, expected_max_unique_states=None):
        names = generator_names or [str(i) for i in range(len(generators))]
        gens = [{"matrix": list(m), "modulo": mod} for m, mod in generators]
        return cls({"schema": 2, "name": name, "generator_names": names,
                    "action": {"kind": "matrix", "rows": rows, "cols": cols, "generators": gens},
                    "start": start, "expected_max_unique_states": expected_max_unique_states})

    @classmethod
    def from_dict(cls, value):
        return cls(value)

    # -- accessors -------------------------------------------------------
    @property
    def start(self):
        return _plain_copy(self._data["start"])

    @property
    def action(self):
        return _plain_copy(self._data["action"])

    @property
    def name(self):
        return self._data["name"]

    @property
    def generator_names(self):
        return list(self._data["generator_names"])

    @property
    def state_elements(self):
        return len(self._data["start"])

    @property
    def generator_count(self):
        return len(self._data["action"]["generators"])

    def to_dict(self):
        return _plain_copy(self._data)

    def to_json(self):
        return json.dumps(self._data, sort_keys=True, separators=(",", ":"), ensure_ascii=True)

    def digest(self):
        """sha256 of canonical JSON of {schema, action, start} (name-independent)."""
        core = {"schema": self._data["schema"], "action": self._data["action"], "start": self._data["start"]}
        return hashlib.sha256(json.dumps(core, sort_keys=True, separators=(",", ":")).encode()).hexdigest()

    # -- validation --------------------------------------------------------
    def validate(self):
        d = self._data
        if not isinstance(d, dict) or set(d) != _KEYS or d["schema"] != 2:
            raise ValueError("GRAPH_SCHEMA")
        if not isinstance(d["name"], str):
            raise ValueError("GRAPH_NAME")
        start = _int64_list(d["start"])
        action = d["action"]
        if not isinstance(action, dict) or not isinstance(action.get("generators"), list) or not action["generators"]:
            raise ValueError("GRAPH_GENERATORS")
        names = d["generator_names"]
        if (not isinstance(names, list) or len(names) != len(action["generators"])
                or any(not isinstance(x, str) for x in names)):
            raise ValueError("GRAPH_GENERATOR_NAMES")
        bound = d["expected_max_unique_states"]
        if bound is not None and (type(bound) is not int or bound < 1 or bound >= 1 << 64):
            raise ValueError("GRAPH_ORBIT_BOUND")
        kind = action.get("kind")
        if kind == "permutation":
            n = action.get("degree")
            if (set(action) != {"kind", "degree", "generators"} or type(n) is not int
                    or not 1 <= n <= 0xFFFFFFFF or len(start) != n):
                raise ValueError("GRAPH_PERMUTATION_SHAPE")
            identity = list(range(n))
            for g in action["generators"]:
                if _int64_list(g) != g or sorted(g) != identity:
                    raise ValueError("GRAPH_PERMUTATION_GENERATOR")
        elif kind == "matrix":
            rows, cols = action.get("rows"), action.get("cols")
            if (set(action) != {"kind", "rows", "cols", "generators"} or type(rows) is not int
                    or type(cols) is not int or min(rows, cols) < 1 or rows * cols > 0xFFFFFFFF
                    or len(start) != rows * cols):
                raise ValueError("GRAPH_MATRIX_SHAPE")
            for g in action["generators"]:
                if not isinstance(g, dict) or set(g) != {"matrix", "modulo"} or len(_int64_list(g["matrix"])) != rows * rows:
                    raise ValueError("GRAPH_MATRIX_GENERATOR")
                mod = g["modulo"]
                if type(mod) is not int or (mod != 0 and not 2 <= mod <= 1 << 31):
                    raise ValueError("GRAPH_MATRIX_MODULUS")
        else:
            raise ValueError("GRAPH_ACTION_KIND")
        return self

    # -- retention rule ------------------------------------------------------
    @staticmethod
    def _compose(g, h, n, mod):
        """Square n x n product g . h with the action's wrap-then-reduce rule."""
        right = [v % mod for v in h] if mod else list(h)
        out = []
        for i in range(n):
            for j in range(n):
                acc = wrap_i64(sum(g[i * n + k] * right[k * n + j] for k in range(n)))
                out.append(acc % mod if mod else acc)
        return out

    def inverse_closed(self) -> bool:
        """True when every generator's inverse acts as a generator, so every
        edge out of layer d lands in layers d-1, d or d+1 (three-bank
        retention is exact).  Rule of graph_definition.py:84-108."""
        action = self._data["action"]
        gens = action["generators"]
        if action["kind"] == "permutation":
            present = {tuple(g) for g in gens}
            for g in gens:
                inverse = [0] * len(g)
                for i, v in enumerate(g):
                    inverse[v] = i
                if tuple(inverse) not in present:
                    return False
            return True
        n = action["rows"]
        moduli = {g["modulo"] for g in gens}
        if len(moduli) != 1:
            return False
        mod = gens[0]["modulo"]
        if mod and mod & (mod - 1):
            # Wrapping could fake a modular inverse unless each row sum is safe.
            for g in gens:
                for i in range(n):
                    if sum(abs(v) * (mod - 1) for v in g["matrix"][i * n:(i + 1) * n]) > I64_MAX:
                        return False
        eye = [int(i == j) for i in range(n) for j in range(n)]
        for g in gens:
            if not any(self._compose(g["matrix"], h["matrix"], n, mod) == eye
                       and self._compose(h["matrix"], g["matrix"], n, mod) == eye for h in gens):
                return False
        # A modular inverse only inverts canonical values.
        if mod and any(not 0 <= x < mod for x in self._data["start"]):
            return False
        return True


def _flatten(value):
    if hasattr(value, "tolist"):
        value = value.tolist()
    if isinstance(value, (list, tuple)):
        out = []
        for item in value:
            out.extend(_flatten(item))
        return out
    if isinstance(value, bool):
        raise ValueError("GRAPH_NONINTEGER_SCALAR")
    try:
        converted = int(value)
    except (TypeError, ValueError, OverflowError):
        raise ValueError("GRAPH_NONINTEGER_SCALAR") from None
    if converted != value:
        raise ValueError("GRAPH_NONINTEGER_SCALAR")
    return [_int64(converted)]


def from_cayleypy(definition) -> GraphDefinition:
    """CayleyPy CayleyGraphDef (or an object holding one as .definition /
    .graph_def) -> GraphDefinition, preserving generators and start
    (graph_definition.py:130-143)."""
    if not hasattr(definition, "generators_type"):
        if hasattr(definition, "definition"):
            definition = definition.definition
        elif hasattr(definition, "graph_def"):
            definition = definition.graph_def
    kind = definition.generators_type.name
    start = _flatten(definition.central_state)
    options = {"name": definition.name, "generator_names": list(definition.generator_names)}
    if kind == "PERMUTATION":
        return GraphDefinition.permutation([_flatten(g) for g in definition.generators_permutations], start, **options)
    if kind == "MATRIX":
        gens = definition.generators_matrices
        first = gens[0].matrix
        n = int(first.shape[0]) if hasattr(first, "shape") else len(first)
        if len(start) % n:
            raise ValueError("GRAPH_MATRIX_SHAPE")
        return GraphDefinition.matrix(n, len(start) // n, [(_flatten(g.matrix), int(g.modulo)) for g in gens],
                                      start, **options)
    raise ValueError("GRAPH_ACTION_KIND")


def lrx(n: int, name: str | None = None) -> GraphDefinition:
    """LRX on S_n from the identity: L = cyclic left shift (child[i] =
    parent[i+1]), R = its inverse, X = swap of positions 0 and 1
    (multigpubfs/specialized.py:24, crates/mgbfs-core/src/matrix.rs:188-193)."""
    left = [(i + 1) % n for i in range(n)]
    right = [(i - 1) % n for i in range(n)]
    swap = [1, 0] + list(range(2, n))
    return GraphDefinition.permutation([left, right, swap], list(range(n)), name=name or f"lrx{n}",
                                       generator_names=["L", "R", "X"])
