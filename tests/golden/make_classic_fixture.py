"""Extract the compatibility-optimizer probes from the reference's golden
fixture crates/umap/tests/fixtures/standard_umap_reference.json (generated by
umap-learn 0.6.0, see its "reference" block) into tests/golden/.

The probes are data: per-epoch head/tail embeddings, clocks and per-head
Tausworthe states of umap-learn's serial optimizer. Run from the repository
root where the reference checkout is available:

    python tests/golden/make_classic_fixture.py /root/reference
"""
import json
import sys
from pathlib import Path

ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
src = ref / "crates/umap/tests/fixtures/standard_umap_reference.json"
data = json.loads(src.read_text())
out = {
    "source": "crates/umap/tests/fixtures/standard_umap_reference.json (optimizer_probes)",
    "generator": data["reference"],
    "optimizer_probes": data["optimizer_probes"],
}
dst = Path(__file__).resolve().parent / "classic_optimizer_probes.json"
dst.write_text(json.dumps(out, separators=(",", ":")))
print(f"wrote {dst} ({dst.stat().st_size} bytes)")
