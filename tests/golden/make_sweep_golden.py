"""Writes tests/golden/sweep256.json: the 256 sweep instances (inputs, from
multigpubfs_amd.sweep.make_sweep) and their exact layer tables computed by
the CPU oracle (oracle/mgbfs_oracle.c orc_packed_bfs with explicit
generators).  Run from the repository root."""
import json
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from multigpubfs_amd.sweep import make_sweep  # noqa: E402
from oracle import oracle as O  # noqa: E402

out = []
for i, g in enumerate(make_sweep(256)):
    gens = g["action"]["generators"]
    n = g["action"]["degree"]
    layers = O.packed_bfs(n, gens)["layers"]
    out.append({"graph": g, "layer_sizes": layers})
path = Path(__file__).resolve().parent / "sweep256.json"
path.write_text(json.dumps({"schema": "sweep-golden-v1", "seed": "0x4D47424600000000 + i",
                            "instances": out}, separators=(",", ":")) + "\n")
sizes = [sum(x["layer_sizes"]) for x in out]
print(path, len(out), "states min/max/total", min(sizes), max(sizes), sum(sizes))
