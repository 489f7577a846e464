"""Extract the reference's golden layer tables (data files, read as JSON).

Writes tests/golden/lrx_layers.json from
  data/reference/lrx13-layers.json                       (S13 LRX, 79 layers)
  docs/validation/data/2026-09-06-s10-profile-panel.json (S10 LRX, [0].layer_sizes)
Run once in a checkout that has /root/reference; the output is committed.
"""
import json
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
OUT = Path(__file__).resolve().parent / "lrx_layers.json"

s13 = json.loads((REF / "data/reference/lrx13-layers.json").read_text())
panel = json.loads((REF / "docs/validation/data/2026-09-06-s10-profile-panel.json").read_text())
s10 = panel[0]["layer_sizes"]
assert all(row["layer_sizes"] == s10 for row in panel)
doc = {
    "schema": "lrx-golden-layers-v1",
    "generators": ["L", "R", "X"],
    "start": "identity",
    "sources": {
        "13": "data/reference/lrx13-layers.json (layers)",
        "10": "docs/validation/data/2026-09-06-s10-profile-panel.json ([*].layer_sizes)",
    },
    "layers": {"13": s13["layers"], "10": s10},
}
OUT.write_text(json.dumps(doc, indent=1) + "\n")
print(OUT, {k: (len(v), sum(v)) for k, v in doc["layers"].items()})
