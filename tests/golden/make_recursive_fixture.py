"""Copy the reference's recursive-initialisation fixture into tests/golden/.

Source (read as data, never executed):
  crates/umap-core/tests/fixtures/recursive_oracle.json, the umap-learn
  0.6.0 vectors the reference's own tests check base_case, prolongate and
  procrustes_align against (crates/umap-core/src/recursive_init/
  oracle_tests.rs:82-155, tolerance 1e-6 + 1e-5 |expected|).

Every case is copied verbatim (inputs and expected outputs).

Usage: python tests/golden/make_recursive_fixture.py [path-to-reference-root]
"""
import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
SRC = REF / "crates/umap-core/tests/fixtures/recursive_oracle.json"


def main() -> None:
    fx = json.loads(SRC.read_text())
    out = {"source": str(SRC.relative_to(REF)), "provenance": fx["provenance"]}
    for key in ("base_cases", "prolongation_cases", "procrustes_cases"):
        out[key] = fx[key]
    (HERE / "recursive_oracle_cases.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
