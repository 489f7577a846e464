"""Writes tests/golden/serve2000_small.json: the exact layer tables of the
permutation jobs of degree 8..10 in the default 2,000-job serving stream
(multigpubfs_amd.serving.make_jobs: seed 0x5345525645000000, 40 jobs/s),
computed by the CPU oracle (oracle/mgbfs_oracle.c orc_packed_bfs with the
job's generators).  Jobs are keyed by id; the stream itself is regenerated
from its seed by whoever reads the file.  Run from the repository root."""
import hashlib
import json
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from multigpubfs_amd.serving import SERVE_SEED, make_jobs  # noqa: E402
from oracle import oracle as O  # noqa: E402

jobs = make_jobs(2000)
tables = {}
for job in jobs:
    a = job["graph"]["action"]
    if a["kind"] == "permutation" and a["degree"] <= 10:
        tables[str(job["id"])] = O.packed_bfs(a["degree"], a["generators"])["layers"]
stream = hashlib.sha256(json.dumps([j["graph"] for j in jobs], sort_keys=True).encode()).hexdigest()
path = Path(__file__).resolve().parent / "serve2000_small.json"
path.write_text(json.dumps({"schema": "serve-golden-v1", "seed": hex(SERVE_SEED), "count": 2000,
                            "stream_sha256": stream, "tables": tables}, separators=(",", ":")) + "\n")
print(path, len(tables), "jobs; states", sum(sum(t) for t in tables.values()))
