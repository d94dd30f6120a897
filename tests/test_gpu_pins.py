"""The reference's golden data and known answers, through the device paths.

- The fit graph on the device (umaphip_fit_graph, no pruning) against
  umap-learn 0.6.0's graph_cases fuzzy union at the reference's
  1e-5 + 1e-5 |expected| (crates/umap/tests/test_standard_graph.rs:88-97,
  325-400).
- The device kNN against the reference's tie and self-exclusion known
  answers (crates/umap/tests/test_standard_dispatch.rs:79-106).
- The resident plan's refresh cadence, epochs + [H > 1024]
  (crates/umap/src/standard_umap/modern_runtime.rs:1894-1899).
"""
import json
from pathlib import Path

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


@pytest.mark.parametrize("i", range(3))
def test_device_fit_graph_golden_cases(i):
    from cuscan_amd import graph as G

    case = json.loads((GOLDEN / "fuzzy_graph_cases.json").read_text())["graph_cases"][i]
    n, k = case["n"], case["k"]
    print(f"[gpu] device fit graph, golden case {case['name']}", flush=True)
    g = G.fit_graph_device(np.asarray(case["indices"], np.int32).reshape(n, k),
                           np.asarray(case["distances"], np.float32).reshape(n, k),
                           local_connectivity=case["local_connectivity"],
                           bandwidth=case["bandwidth"], mix=case["set_op_mix_ratio"],
                           epochs=200, rate=5.0, prune_divisor=0.0)
    dense = np.zeros((n, n))
    off = g["offsets"].astype(np.int64)
    for r in range(n):
        cols = g["indices"][off[r]:off[r + 1]].astype(np.int64)
        assert (np.diff(cols) > 0).all() and (cols != r).all()
        dense[r, cols] = g["weights"][off[r]:off[r + 1]]
    want = np.asarray(case["symmetric"], np.float64).reshape(n, n)
    assert ((dense > 0) == (want > 0)).all(), case["name"]
    assert (np.abs(dense - want) <= 1e-5 + 1e-5 * np.abs(want)).all(), case["name"]


def test_device_knn_reference_known_answers():
    from cuscan_amd.knn import KnnIndex

    print("[gpu] device kNN known answers", flush=True)
    ref = np.array([[0, 0], [1, 0], [-1, 0], [0, 0]], np.float32)
    index = KnnIndex(ref, max_queries=8)
    try:
        idx, dist = index.query(np.array([[-1, 0], [0, 0]], np.float32), 3, exclude_self=True,
                                query_offset=2)
    finally:
        index.close()
    assert idx.ravel().tolist() == [0, 3, 1, 0, 1, 2]
    assert dist.ravel().tolist() == [1.0, 1.0, 2.0, 0.0, 1.0, 1.0]
    index = KnnIndex(np.array([[0, 0], [0, 0], [2, 0]], np.float32), max_queries=8)
    try:
        idx, dist = index.query(np.zeros((1, 2), np.float32), 3)
    finally:
        index.close()
    assert idx.ravel().tolist() == [0, 1, 2]
    assert dist.ravel().tolist() == [0.0, 0.0, 2.0]


@pytest.mark.parametrize("h", [1000, 1100])
def test_plan_refresh_cadence(h):
    from cuscan_amd import graph as G
    from cuscan_amd import layout as L
    from tests.gpu_helpers import random_graph

    print(f"[gpu] plan refresh cadence h={h}", flush=True)
    epochs = 22
    g = random_graph(h, h, 4, 3, epochs)
    params = L.LayoutParams(epochs=epochs, a=1.58, b=0.9, negative_selection_range=300)
    _, _, _, stats = L.optimize_compact(G.random_init(h, 2, 1), None, g, params, shared=True)
    assert stats["refreshes"] == epochs + (1 if h > 1024 else 0)
