"""Parity of the edge-balanced unit sweep (the benchmarked epoch kernel)
inside the resident plan, one traced epoch at a time.

The plan records, at the chosen epoch, the state the sweep starts from
(positions, negative order, window ranks, repulsion scale), each row's
assembled gradient (the value the lazy Adam step consumes) and every
negative each firing edge drew. The oracle then restates
gradient_kernel (crates/umap/src/kernels/standard_umap/modern_layout.rs:
128-257) on that state with the clocks as they stand before the epoch:

- negatives: per edge, the count, every candidate id and its exclusion flag
  bit-exact against orc_trace_negatives;
- gradient: within 1e-5 normwise of the FP64 oracle; and per row, within
  1e-5 of the row's FP64 value plus TERM_EPS of the row's sum of term
  norms. A row whose terms nearly cancel cannot be resolved to 1e-5 of its
  (small) sum by any f32 arithmetic, the reference's included: its error
  scales with the terms, so the per-row bar is stated in their units. A
  wrong, missing or extra term moves a row by a whole term and fails it.
  The reference's own f32 arithmetic and summation order is held to the
  same per-row bar on the same inputs.

Paths covered: rows in CSR order (global negatives, exclusion; sparse rows
and hubs split over units) and in projection order (windows; 24- and
32-byte records; term counts from the clock schedule, whole-run and in
7-epoch passes, or clocks advanced by the sweep itself), the term-parallel
path and the per-edge path (rate 24, more terms than a unit holds).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
from tests.gpu_helpers import (advance_graph, hub_graph, row_relative_error,  # noqa: E402
                               term_relative_error)

pytestmark = pytest.mark.gpu
GRAD_TOL = 1e-5
TERM_EPS = 1e-6   # per-row error in units of the row's sum of term norms


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _graph_from_data(x, k=15, rate=5.0, epochs=200):
    from cuscan_amd import graph as G
    from cuscan_amd.knn import KnnIndex

    n = len(x)
    index = KnnIndex(torch.from_numpy(x).cuda(), max_queries=n)
    idx, dist = index.query(x, k - 1, exclude_self=True)
    index.close()
    fg = G.fuzzy_graph(np.concatenate([np.arange(n, dtype=np.int32)[:, None], idx], 1),
                       np.concatenate([np.zeros((n, 1), np.float32), dist], 1))
    return G.compact_graph(n, n, fg["heads"], fg["tails"], fg["weights"], epochs=epochs,
                           rate=rate, prune_divisor=float(epochs))


@pytest.fixture(scope="module")
def cfg2_graph():
    """100,000 x 50, 50 clusters (centres in [-10, 10]^50, sigma in [0.5, 2]), seed 42."""
    rng = np.random.default_rng(42)
    centers = rng.uniform(-10, 10, (50, 50))
    sig = rng.uniform(0.5, 2.0, 50)
    lab = rng.integers(0, 50, 100_000)
    x = (centers[lab] + rng.normal(0, 1, (100_000, 50)) * sig[lab, None]).astype(np.float32)
    print("[gpu] unit sweep: cfg2 graph", flush=True)
    return _graph_from_data(x)


@pytest.fixture(scope="module")
def mixture_rows():
    """100,000 x 784 rows of the 10-cluster 0..255 mixture (centres in
    [32, 224]^784, sigma 24, rounded and clipped), seed 42."""
    rng = np.random.default_rng(42)
    centers = rng.uniform(32, 224, (10, 784))
    lab = rng.integers(0, 10, 100_000)
    return np.clip(np.round(centers[lab] + rng.normal(0, 24, (100_000, 784))), 0,
                   255).astype(np.float32)


@pytest.fixture(scope="module")
def mixture_graph(mixture_rows):
    print("[gpu] unit sweep: mixture graph", flush=True)
    return _graph_from_data(mixture_rows)


@pytest.fixture(scope="module")
def mixture_graph_rate24(mixture_rows):
    print("[gpu] unit sweep: mixture graph, rate 24", flush=True)
    return _graph_from_data(mixture_rows, rate=24.0)


def traced_epoch(g, params, epoch, *, cap=16, compact=None, schedule=None):
    from cuscan_amd import graph as G
    from cuscan_amd import layout as L

    n = len(g["offsets"]) - 1
    head = G.random_init(n, 2, 42)
    plan = L.LayoutPlan(head, None, g, params, shared=True)
    try:
        if compact is not None:
            plan.set_record_format(compact)
        if schedule is not None:
            plan.set_clock_schedule(schedule)
        plan.set_trace(epoch, cap)
        plan.run()
        _, npos, nneg, stats = plan.download(clocks=True)
        tr = plan.trace()
    finally:
        plan.close()
    assert stats["edge_balanced"] == 1
    tr["clock_schedule"] = stats["clock_schedule"]
    return tr, npos, nneg


def check_trace(g, params, epoch, tr):
    """Negatives bit-exact and the gradient against the FP64 oracle."""
    n = len(tr["head"])
    gt = advance_graph(g, epoch)
    R = min(n, max(params.negative_selection_range, 200))
    src = np.arange(n, dtype=np.int32)
    if R < n:
        inv = np.empty(n, np.int64)
        inv[tr["order"]] = np.arange(n)
        assert (tr["ranks"] == inv).all(), "window ranks are not the inverse of the order"
        # The order in force at epoch t is the stable argsort of the epoch's
        # positions projected on direction t (refresh 0 precedes epoch 0).
        rng = O.StdRng(params.seed)
        for _ in range(epoch):
            rng.direction(2)
        want, _ = O.stable_argsort(O.project(tr["head"], rng.direction(2)))
        assert (tr["order"] == want).all(), "projection order differs from the stable argsort"

    def epoch_args():
        return O.Epoch(head=tr["head"], tail=tr["head"], offsets=g["offsets"],
                       indices=g["indices"], source_order=src, negative_order=tr["order"],
                       source_ranks=tr["ranks"], periods=g["periods"],
                       next_positive=gt["next_positive"],
                       negative_periods=g["negative_periods"],
                       next_negative=gt["next_negative"],
                       updates=np.zeros((n, 2), np.float32), m=np.zeros((n, 2)),
                       v=np.zeros((n, 2)), epoch=epoch, a=params.a, b=params.b,
                       gamma=params.repulsion, scale=tr["scale"], negative_range=R,
                       shared=True, negatives=params.negative_sample_rate != 0,
                       exclude=params.exclude_graph_neighbors)

    # negatives: per edge, count, candidates and flags in sample order
    edges, cands, kept, fires = epoch_args().trace_negatives()
    count = tr["count"]
    assert int((count >= 0).sum()) == fires
    exp_count = np.bincount(edges, minlength=len(count)) if len(edges) else np.zeros_like(count)
    fired = count >= 0
    assert (count[fired] == exp_count[fired]).all()
    assert (exp_count[~fired] == 0).all()
    order = np.argsort(edges, kind="stable")
    se, sc, sk = edges[order], cands[order], kept[order]
    slot = np.arange(len(se)) - np.searchsorted(se, se, side="left")
    assert (tr["candidates"][se, slot] == sc).all(), "negative candidate mismatch"
    assert (tr["kept"][se, slot] == sk).all(), "exclusion flag mismatch"
    # gradient
    g64, mag = epoch_args().gradient_block64_mag(0, n, 1.0)
    grad = tr["gradient"].astype(np.float64)
    ref_norm = float(np.linalg.norm(g64))
    normwise = float(np.linalg.norm(grad - g64)) / ref_norm if ref_norm > 0 else \
        float(np.abs(grad).max())
    err = row_relative_error(grad, g64)
    term_err = term_relative_error(grad, g64, mag)
    ref32 = epoch_args()
    ref32.gradient_block(0, n, 1.0, threads=8)
    err32 = row_relative_error(ref32.updates, g64)
    term_err32 = term_relative_error(ref32.updates, g64, mag)
    print(f"epoch {epoch}: fires {fires}, negatives {len(edges)}, gradient normwise "
          f"{normwise:.3e}; per row: relative {err:.3e}, in term units {term_err:.3e} "
          f"(reference-order f32: {err32:.3e}, {term_err32:.3e}); "
          f"ranked {tr['rank_ordered']}, 24-byte {tr['compact_records']}, "
          f"units term-parallel/per-edge {tr['term_parallel_units']}/{tr['per_edge_units']}",
          flush=True)
    assert normwise <= GRAD_TOL
    assert term_err <= TERM_EPS
    assert term_err32 <= TERM_EPS
    return fires, len(edges)


def _params(**kw):
    from cuscan_amd import graph as G
    from cuscan_amd import layout as L

    a, b = G.fit_curve(0.1, 1.0)
    return L.LayoutParams(epochs=200, a=a, b=b, **kw)


@pytest.mark.parametrize("epoch", [1, 50, 199])
def test_cfg2_csr_order_exclusion(cfg2_graph, epoch):
    """Global negatives with exclusion: rows swept in CSR order, term counts
    from the clock schedule (epochs 1 and 199) or clocks advanced by the
    sweep from the records and firing shadow (epoch 50)."""
    params = _params(exclude_graph_neighbors=True)
    print(f"[gpu] unit sweep cfg2 epoch {epoch}", flush=True)
    span = 0 if epoch == 50 else None
    tr, npos, nneg = traced_epoch(cfg2_graph, params, epoch, schedule=span)
    assert not tr["rank_ordered"] and tr["clock_schedule"] == (0 if span == 0 else 200)
    want = advance_graph(cfg2_graph, 200)
    assert (npos == want["next_positive"]).all() and (nneg == want["next_negative"]).all()
    fires, _ = check_trace(cfg2_graph, params, epoch, tr)
    assert fires > 0


@pytest.mark.parametrize("epoch", [1, 50, 199])
def test_mixture_windows_rank_order(mixture_graph, epoch):
    """Projection windows (Q = 20,000 of 100,000 rows, adapted scale):
    rows swept in projection order with 24-byte records, term counts from
    the whole-run clock schedule (the default)."""
    params = _params(negative_selection_range=20_000)
    print(f"[gpu] unit sweep mixture epoch {epoch}", flush=True)
    tr, npos, nneg = traced_epoch(mixture_graph, params, epoch)
    assert tr["rank_ordered"] and tr["compact_records"] and tr["clock_schedule"] == 200
    check_trace(mixture_graph, params, epoch, tr)
    want = advance_graph(mixture_graph, 200)
    assert (npos == want["next_positive"]).all() and (nneg == want["next_negative"]).all()


@pytest.mark.parametrize("epoch", [1, 199])
def test_mixture_windows_clocks_in_sweep(mixture_graph, epoch):
    """Schedule off: the sweep reads the 24-byte records and advances the
    clocks itself."""
    params = _params(negative_selection_range=20_000)
    print(f"[gpu] unit sweep mixture, clocks in the sweep, epoch {epoch}", flush=True)
    tr, npos, nneg = traced_epoch(mixture_graph, params, epoch, schedule=0)
    assert tr["rank_ordered"] and tr["compact_records"] and tr["clock_schedule"] == 0
    check_trace(mixture_graph, params, epoch, tr)
    want = advance_graph(mixture_graph, 200)
    assert (npos == want["next_positive"]).all() and (nneg == want["next_negative"]).all()


@pytest.mark.parametrize("epoch", [50, 199])
def test_mixture_windows_32_byte_records(mixture_graph, epoch):
    params = _params(negative_selection_range=20_000, exclude_graph_neighbors=True)
    print(f"[gpu] unit sweep mixture 32-byte epoch {epoch}", flush=True)
    # 7-epoch schedule passes: epochs 50 and 199 are slabs 1 and 3 of a pass
    tr, npos, nneg = traced_epoch(mixture_graph, params, epoch, compact=False, schedule=7)
    assert tr["rank_ordered"] and not tr["compact_records"] and tr["clock_schedule"] == 7
    check_trace(mixture_graph, params, epoch, tr)
    want = advance_graph(mixture_graph, 200)
    assert (npos == want["next_positive"]).all() and (nneg == want["next_negative"]).all()


def test_schedule_forms_identical(mixture_graph):
    """Whole runs with the schedule in one pass, in 7- and 1-epoch passes and
    off, with 24- and 32-byte records: the same embedding and clocks, bit for
    bit."""
    from cuscan_amd import graph as G
    from cuscan_amd import layout as L

    params = _params(negative_selection_range=20_000, exclude_graph_neighbors=True)
    head = G.random_init(len(mixture_graph["offsets"]) - 1, 2, 42)
    outs = []
    for compact, span in [(True, -1), (True, 0), (True, 7), (False, 1), (False, 0)]:
        print(f"[gpu] schedule forms: 24-byte {compact}, span {span}", flush=True)
        plan = L.LayoutPlan(head, None, mixture_graph, params, shared=True)
        try:
            plan.set_record_format(compact)
            plan.set_clock_schedule(span)
            plan.run()
            y, npos, nneg, stats = plan.download(clocks=True)
        finally:
            plan.close()
        assert stats["clock_schedule"] == {-1: 200, 0: 0}.get(span, span)
        assert stats["rank_ordered_epochs"] == 200
        outs.append((y, npos, nneg))
    for y, npos, nneg in outs[1:]:
        assert np.array_equal(y.view(np.uint32), outs[0][0].view(np.uint32))
        assert (npos == outs[0][1]).all() and (nneg == outs[0][2]).all()


def test_schedule_needs_small_rate():
    """Rates above 64 keep the clocks in the sweep; asking for a schedule
    is an error."""
    from cuscan_amd import graph as G
    from cuscan_amd import layout as L
    from tests.gpu_helpers import hub_graph

    g = hub_graph(5000, 8, 3, 200, rate=65.0, isolated=())
    params = _params(negative_selection_range=1_000, negative_sample_rate=65)
    plan = L.LayoutPlan(G.random_init(5000, 2, 1), None, g, params, shared=True)
    try:
        with pytest.raises(Exception, match="clock schedule"):
            plan.set_clock_schedule(5)
        plan.set_clock_schedule(-1)
        plan.run()
        _, _, _, stats = plan.download(clocks=True)
        assert stats["clock_schedule"] == 0 and stats["rank_ordered_epochs"] == 200
    finally:
        plan.close()


@pytest.mark.parametrize("q", [20_000, 200_000])
def test_mixture_rate24_per_edge_path(mixture_graph_rate24, q):
    """Rate 24: about 16 negatives per firing edge, so units hold more terms
    than LDS and take the per-edge path (windows: rank order; global: CSR
    order)."""
    params = _params(negative_selection_range=q, negative_sample_rate=24,
                     exclude_graph_neighbors=q < 100_000)
    print(f"[gpu] unit sweep rate 24, Q={q}", flush=True)
    tr, _, _ = traced_epoch(mixture_graph_rate24, params, 120, cap=64)
    assert tr["per_edge_units"] > 0 and tr["term_parallel_units"] > 0
    assert tr["rank_ordered"] == (q < 100_000)
    check_trace(mixture_graph_rate24, params, 120, tr)


@pytest.mark.parametrize("q,isolated", [(200_000, (3, 4, 2999)), (1_000, (3, 4, 2999)),
                                        (1_000, ())])
def test_hub_graph(q, isolated):
    """Hub rows spread over many units (fragments), with and without rows
    that have no edge (CSR order through the nonzero-row table), global and
    windowed negatives, exclusion."""
    g = hub_graph(5000, 8, 3, 200, isolated=isolated)
    params = _params(negative_selection_range=q, exclude_graph_neighbors=True)
    print(f"[gpu] unit sweep hub graph Q={q} isolated={isolated}", flush=True)
    for epoch in (7, 120):
        tr, _, _ = traced_epoch(g, params, epoch)
        assert tr["rank_ordered"] == (q < 5000 and not isolated)
        check_trace(g, params, epoch, tr)


def test_trace_needs_unit_path():
    from cuscan_amd import graph as G
    from cuscan_amd import layout as L
    from tests.gpu_helpers import random_graph

    g = random_graph(600, 600, 6, 1, 50)
    params = _params()
    params.epochs = 50
    plan = L.LayoutPlan(G.random_init(600, 2, 1), None, g, params, shared=True)
    try:
        with pytest.raises(ValueError, match="edge-balanced sweep"):
            plan.set_trace(3, 8)
    finally:
        plan.close()
