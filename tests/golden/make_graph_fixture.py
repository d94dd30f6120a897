"""Extract the fuzzy-graph golden cases from the reference's fixture
crates/umap/tests/fixtures/standard_umap_reference.json (generated by
umap-learn 0.6.0, see its "reference" block) into tests/golden/.

Each case is data: kNN ids and distances [n, k], the smoothing parameters,
and umap-learn's rho, sigma, directed memberships (dense n x n) and fuzzy
union (dense n x n). The reference checks its fuzzy_graph against them at
1e-5 + 1e-5 |expected| (crates/umap/tests/test_standard_graph.rs:88-97,
325-400). Run from the repository root where the reference checkout is
available:

    python tests/golden/make_graph_fixture.py /root/reference
"""
import json
import sys
from pathlib import Path

ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
src = ref / "crates/umap/tests/fixtures/standard_umap_reference.json"
data = json.loads(src.read_text())
out = {
    "source": "crates/umap/tests/fixtures/standard_umap_reference.json (graph_cases)",
    "generator": data["reference"],
    "graph_cases": data["graph_cases"],
}
dst = Path(__file__).resolve().parent / "fuzzy_graph_cases.json"
dst.write_text(json.dumps(out, separators=(",", ":")))
print(f"wrote {dst} ({dst.stat().st_size} bytes)")
