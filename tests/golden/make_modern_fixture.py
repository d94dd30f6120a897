"""Compact the reference's frozen single-epoch fixture into tests/golden/.

Source (read as data, never executed):
  crates/umap-core/tests/fixtures/modern_layout_upstream.json (54 MB),
  generated by comparison/umap06/generate_modern_oracle.py:43-185 from
  umap-learn 0.6.0's numba kernels.

The fixture's inputs are closed-form (generate_modern_oracle.py:43-62), so the
compact file keeps every scalar parameter, the four-edge CSR and clocks, a
SHA-256 of each regenerated input array (so the tests can prove they rebuild
exactly the fixture's inputs) and, for every expected output matrix, only the
rows that differ from the initial state. All numbers are copied verbatim.

Usage: python tests/golden/make_modern_fixture.py [path-to-reference-root]
"""
import hashlib
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
SRC = REF / "crates/umap-core/tests/fixtures/modern_layout_upstream.json"


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main() -> None:
    fx = json.loads(SRC.read_text())
    out = {"source": str(SRC.relative_to(REF)), "provenance": fx["provenance"],
           "schedules": fx["schedules"], "cases": []}
    for c in fx["cases"]:
        rows, dim, refs = c["rows"], c["dim"], c["refs"]
        head = np.array(c["head"], np.float32)
        tail = np.array(c["tail"], np.float32)
        init_upd = np.array(c["initial_updates"], np.float32)
        entry = {k: c[k] for k in ("rows", "refs", "dim", "operation", "optimizer", "gamma",
                                   "epoch", "alpha", "momentum", "scale", "block", "offsets",
                                   "indices", "periods", "negative_periods", "next_positive",
                                   "next_negative", "expected_positive", "expected_negative")}
        entry["offsets"] = c["offsets"]
        entry["sha256"] = {
            "head": digest(head), "tail": digest(tail),
            "from_order": digest(np.array(c["from_order"], np.int32)),
            "to_order": digest(np.array(c["to_order"], np.int32)),
            "ranks": digest(np.array(c["ranks"], np.int32)),
            "initial_updates": digest(init_upd),
        }
        zeros = np.zeros_like(head)
        for key, base in (("expected_head", head), ("expected_updates", init_upd),
                          ("expected_final_updates", init_upd), ("expected_first", zeros),
                          ("expected_second", zeros)):
            arr = np.array(c[key], np.float64)
            changed = np.where((arr.astype(np.float32) != base).any(axis=1))[0]
            entry[key] = {str(int(r)): c[key][int(r)] for r in changed}
        out["cases"].append(entry)
    dst = HERE / "modern_layout_upstream_compact.json"
    dst.write_text(json.dumps(out, separators=(",", ":")))
    print(f"wrote {dst} ({dst.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
