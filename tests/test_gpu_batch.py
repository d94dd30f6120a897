"""Batched layouts (many graphs per launch: one persistent cooperative launch
for every epoch, or a hipGraph-captured loop of per-epoch launches).

Every graph must match its own single-graph plan: clocks bit for bit;
coordinates bit for bit for small graphs (the plan runs the same tile kernel
with the same per-row order) and within 2e-3 for large graphs (the plan's
edge-balanced sweep sums in another order; global negatives do not depend
on positions, so runs do not diverge). Graph replays are checked against each
other and against plain launches.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.gpu_helpers import random_graph  # noqa: E402

pytestmark = pytest.mark.gpu
A_CURVE, B_CURVE = 1.5769434, 0.8950608


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def make_batch(sizes, epochs, exclude=False):
    from cuscan_amd.layout import LayoutParams

    graphs, heads = [], []
    for g, n in enumerate(sizes):
        graphs.append(random_graph(n, n, 5 + g % 4, 100 + g, epochs))
        heads.append(np.random.default_rng(200 + g).uniform(-10, 10, (n, 2)).astype(np.float32))
    params = LayoutParams(epochs=epochs, a=A_CURVE, b=B_CURVE, exclude_graph_neighbors=exclude)
    seeds = [42 + g for g in range(len(sizes))]
    return graphs, heads, params, seeds


@pytest.mark.parametrize("exclude", [False, True])
def test_batch_matches_single_graph_plans_and_replays(exclude):
    from cuscan_amd import layout as L
    from cuscan_amd.batch import LayoutBatch

    sizes = [300, 1500, 800, 2600, 60, 1025, 4000, 1024]
    epochs = 30
    graphs, heads, params, seeds = make_batch(sizes, epochs, exclude)
    print(f"[gpu] batch of {len(sizes)} graphs exclude={exclude}", flush=True)
    batch = LayoutBatch(graphs, heads, params, seeds)
    shape, ms = batch.tune(probe_epochs=5)
    assert shape in (0, 1, 2) and all(t > 0 for t in ms)
    batch.set_shape(0)
    batch.run(use_graph=True)
    r1, _, p1, n1 = batch.download(clocks=True)
    batch.run(use_graph=True)
    r2, _, p2, n2 = batch.download(clocks=True)
    batch.run(use_graph=False)
    r3, _, p3, n3 = batch.download(clocks=True)
    for a, b, c in zip(r1, r2, r3):
        assert (a == b).all() and (a == c).all()
    assert (p1 == p2).all() and (p1 == p3).all() and (n1 == n3).all()
    # one persistent cooperative launch for every epoch equals one launch per
    # epoch; the autotuner also runs in that form
    assert not batch.persistent
    batch.set_persistent(True)
    pshape, pms = batch.tune(probe_epochs=5)
    assert pshape in (0, 1, 2) and all(t > 0 for t in pms)
    batch.set_shape(0)
    batch.run(use_graph=True)
    r4, _, p4, n4 = batch.download(clocks=True)
    batch.set_persistent(False)
    for a, b in zip(r1, r4):
        assert (a == b).all()
    assert (p1 == p4).all() and (n1 == n4).all()
    # the clock schedule (the default) and clocks advanced inside the sweeps
    # give the same bits, in both launch forms
    assert batch.clock_schedule == epochs
    batch.set_clock_schedule(False)
    assert batch.clock_schedule == 0
    batch.run(use_graph=True)
    r5, _, p5, n5 = batch.download(clocks=True)
    batch.set_persistent(True)
    batch.run()
    r6, _, p6, n6 = batch.download(clocks=True)
    batch.set_persistent(False)
    batch.set_clock_schedule(True)
    assert batch.clock_schedule == epochs
    for a, b, c in zip(r1, r5, r6):
        assert (a == b).all() and (a == c).all()
    assert (p1 == p5).all() and (p1 == p6).all() and (n1 == n5).all() and (n1 == n6).all()
    e0 = 0
    for g, (graph, head) in enumerate(zip(graphs, heads)):
        p = L.LayoutParams(epochs=epochs, a=A_CURVE, b=B_CURVE, exclude_graph_neighbors=exclude,
                           seed=seeds[g])
        emb, npos, nneg, _ = L.optimize_compact(head, None, graph, p, shared=True)
        E = len(graph["indices"])
        assert (p1[e0:e0 + E] == npos).all() and (n1[e0:e0 + E] == nneg).all()
        e0 += E
        if sizes[g] <= 1024:
            assert (r1[g] == emb).all(), f"graph {g}"
        else:
            np.testing.assert_allclose(r1[g], emb, atol=2e-3, rtol=0)
    # every tile shape gives the same clocks and matching coordinates
    for s in (1, 2):
        batch.set_shape(s)
        batch.run(use_graph=True)
        rs, _, ps, _ = batch.download(clocks=True)
        assert (ps == p1).all()
        for a, b in zip(rs, r1):
            np.testing.assert_allclose(a, b, atol=2e-3, rtol=0)
    batch.close()


def test_batch_rejects_projection_mode_graphs():
    from cuscan_amd.batch import LayoutBatch
    from cuscan_amd.layout import LayoutParams

    graphs, heads, _, seeds = make_batch([1200], 5)
    with pytest.raises(ValueError, match="negative_selection_range"):
        LayoutBatch(graphs, heads, LayoutParams(epochs=5, a=A_CURVE, b=B_CURVE,
                                                negative_selection_range=500), seeds)
