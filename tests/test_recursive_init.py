"""Recursive initialisation's host steps (no device work).

Pinned against:
- the reference's umap-learn 0.6.0 vectors for base_case, prolongate and
  procrustes_align (crates/umap-core/tests/fixtures/recursive_oracle.json,
  copied by tests/golden/make_recursive_fixture.py; tolerance 1e-6 +
  1e-5 |expected| as in crates/umap-core/src/recursive_init/oracle_tests.rs:
  66-80);
- the known answers of the reference's graph-step unit tests
  (crates/umap-core/src/recursive_init/graph.rs:540-714) and level tests
  (mod.rs:672-695);
- the pinned rescale vectors and closed forms of
  crates/umap-core/src/init_rescale.rs:223-358;
- the sampled PCA anchor against numpy's exact SVD on data whose rank is
  below the randomised range width (where the randomised PCA is exact), in
  both orientations.
A run whose graph is a base case does no device work, so the whole
recursion is exercised here for graphs up to base_init_threshold rows.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from cuscan_amd import recursive as R

FIXTURE = json.loads((Path(__file__).parent / "golden" / "recursive_oracle_cases.json").read_text())


def close_oracle(actual, expected):
    expected = np.asarray(expected, np.float32)
    assert actual.shape == expected.shape
    assert np.isfinite(actual).all()
    tol = 1e-6 + 1e-5 * np.abs(expected)
    assert (np.abs(actual - expected) <= tol).all(), (actual, expected)


def csr(case_graph):
    return (np.array(case_graph["offsets"], np.int64), np.array(case_graph["indices"], np.int32),
            np.array(case_graph["weights"], np.float32))


# ------------------------------------------------------------- fixture vectors

@pytest.mark.parametrize("case", FIXTURE["base_cases"], ids=lambda c: c["name"])
def test_base_case_matches_upstream(case):
    o, i, w = csr(case["graph"])
    n = len(o) - 1
    mask = np.array(case["sample_mask"], bool)
    d = len(case["anchor"][0])
    values = np.zeros((n, d), np.float32)
    values[mask] = np.array(case["anchor"], np.float32)
    out = R.base_case(o, i, w, values, mask, case["depth"])
    close_oracle(out, case["result"])


@pytest.mark.parametrize("case", FIXTURE["prolongation_cases"], ids=lambda c: c["name"])
def test_prolongation_matches_upstream(case):
    o, i, w = csr(case["graph"])
    out = R.prolongate(o, i, w, np.array(case["labels"], np.int32),
                       np.array(case["coarse"], np.float32))
    close_oracle(out, case["centered_f32"])


@pytest.mark.parametrize("case", FIXTURE["procrustes_cases"], ids=lambda c: c["name"])
def test_procrustes_matches_upstream(case):
    anchor = np.array(case["anchor"], np.float32)
    out = R.procrustes(np.array(case["embedding"], np.float32), anchor,
                       np.ones(len(anchor), bool))
    close_oracle(out, case["aligned"])


def test_procrustes_is_a_proper_rotation_of_random_layouts():
    """Closed form: for a layout that is a rotated, shifted, scaled anchor the
    alignment recovers the anchor's orientation (rotation only, about the
    origin), and a reflected layout is never reflected back."""
    rng = np.random.default_rng(5)
    for d in (2, 3, 5):
        anchor = rng.normal(size=(40, d)).astype(np.float32)
        q, _ = np.linalg.qr(rng.normal(size=(d, d)))
        if np.linalg.det(q) < 0:
            q[:, 0] = -q[:, 0]
        layout = (anchor.astype(np.float64) @ q.T).astype(np.float32)
        out = R.procrustes(layout, anchor, np.ones(40, bool))
        np.testing.assert_allclose(out, anchor, atol=2e-5)
        # the transform is a rotation: pairwise distances preserved
        refl = layout.copy()
        refl[:, 0] = -refl[:, 0]
        out2 = R.procrustes(refl, anchor, np.ones(40, bool))
        dist = lambda a: np.linalg.norm(a[:, None] - a[None], axis=-1)
        np.testing.assert_allclose(dist(out2), dist(refl), atol=5e-5)


# ------------------------------------------------------------- graph steps

def test_votes_ignore_unlabeled_and_break_ties_in_first_occurrence_order():
    o, i, w = [0, 3, 3, 3, 3], [1, 2, 3], [1.0, 1.0, 9.0]
    labels, _ = R.propagate(o, i, w, [-1, 7, 4, -1], share=False)
    assert labels[0] == 4
    labels, _ = R.propagate(o, i, w, [-1, 7, 4, -1], share=True)
    assert labels[0] == -1


def test_outliers_are_sequential_and_remap_is_ordered():
    labels = R.outliers([0, 1, 2, 2, 2], [2, 0], [-1, -1, 9, 2], seed=123)
    assert labels.tolist() == [9, 9, 9, 2]
    labels, parts = R.remap(labels)
    assert parts == 2 and labels.tolist() == [1, 1, 1, 0]


def test_remap_assigns_unlabeled_rows_after_existing_parts():
    assert R.remap([])[1] == 0
    labels, parts = R.remap([-1, 8, -1, 3, 8])
    assert parts == 4 and labels.tolist() == [2, 1, 3, 0, 1]
    labels, parts = R.remap([-1, -1, -1])
    assert parts == 3 and labels.tolist() == [0, 1, 2]


def test_aggregation_preserves_internal_edges_and_transition_isolates():
    o, i, w = [0, 2, 4, 5, 5], [1, 2, 0, 2, 0], [0.5, 0.25, 0.5, 0.75, 0.25]
    s = R.aggregate(o, i, np.array(w, np.float32).astype(np.float64), [0, 0, 1, 2], 3)
    assert s["offsets"].tolist() == [0, 2, 3, 3]
    assert s["indices"].tolist() == [0, 1, 0]
    assert s["weights"].tolist() == [1.0, 1.0, 0.25]
    t = R.transition(s["offsets"], s["indices"], s["weights"])
    assert t["offsets"].tolist() == [0, 1, 2, 2]
    assert t["indices"].tolist() == [1, 0]
    assert t["weights"].tolist() == [1.0, 1.0]


def test_hub_order_and_diverse_fill_are_deterministic():
    o, i, w = [0, 2, 3, 4], [1, 2, 0, 0], [1.0, 1.0, 1.0, 0.5]
    assert R.hubs(o, i, w, 2, "degree", 0).tolist() == [1, 0, -1]
    assert R.hubs(o, i, w, 2, "diverse", 0).tolist() == [0, 1, -1]


def test_random_hubs_use_the_splitmix_hash():
    """Random hubs: label l sits at row splitmix64(seed + l) mod n."""
    def mix(x):
        m = (1 << 64) - 1
        x = (x + 0x9E3779B97F4A7C15) & m
        x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & m
        x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & m
        return x ^ (x >> 31)
    n = 50
    o = np.zeros(n + 1, np.int64)
    labels = R.hubs(o, [], [], 7, "random", 99)
    expect = np.full(n, -1)
    for lab in range(7):
        expect[mix(99 + lab) % n] = lab
    assert labels.tolist() == expect.tolist()


def test_asymmetric_transition_and_graph_steps():
    o, i, w = [0, 2, 3, 5, 5], [1, 2, 2, 0, 1], [0.25, 0.75, 0.5, 0.2, 0.8]
    labels, _ = R.propagate(o, i, w, [0, -1, 1, -1], share=False)
    again, _ = R.propagate(o, i, w, [0, -1, 1, -1], share=False)
    assert labels.tolist() == again.tolist()
    s = R.aggregate(o, i, np.array(w, np.float32).astype(np.float64), [0, 0, 1, 2], 3)
    assert np.diff(s["offsets"]).tolist() == [2, 1, 0]     # count_coarse_edges
    t = R.transition(s["offsets"], s["indices"], s["weights"])
    assert t["offsets"].tolist() == [0, 1, 2, 2]
    assert t["indices"].tolist() == [1, 0]
    assert t["weights"].tolist() == [1.0, 1.0]
    t = R.transition([0, 2, 3, 3], [1, 2, 0], [1.0, 3.0, 2.0])
    assert t["offsets"].tolist() == [0, 2, 3, 4]
    assert t["indices"].tolist() == [1, 2, 0, 0]
    assert t["weights"].tolist() == [1.0, 0.75, 1.0, 0.75]


def test_transition_matches_a_dense_restatement():
    """Random summed graphs (with self loops and one-way edges) against the
    definition: row-normalise without the diagonal, fuzzy union, clamp."""
    rng = np.random.default_rng(3)
    for n in (7, 30):
        dense = np.where(rng.random((n, n)) < 0.3, rng.random((n, n)) * 3, 0.0)
        dense[rng.random(n) < 0.2] = 0.0          # rows without edges
        o = np.concatenate([[0], np.cumsum((dense != 0).sum(1))]).astype(np.int64)
        idx = np.nonzero(dense)[1].astype(np.int32)
        wts = dense[dense != 0]
        t = R.transition(o, idx, wts)
        off = dense.copy()
        np.fill_diagonal(off, 0.0)
        tot = off.sum(1, keepdims=True)
        p = np.where(tot > 0, off / np.where(tot > 0, tot, 1), 0.0)
        u = np.clip(p + p.T - p * p.T, 0, 1).astype(np.float32)
        got = np.zeros((n, n), np.float32)
        for r in range(n):
            got[r, t["indices"][t["offsets"][r]:t["offsets"][r + 1]]] = \
                t["weights"][t["offsets"][r]:t["offsets"][r + 1]]
        np.testing.assert_allclose(got, u, rtol=1e-6, atol=0)
        for r in range(n):                                 # sorted columns
            cols = t["indices"][t["offsets"][r]:t["offsets"][r + 1]]
            assert (np.diff(cols) > 0).all()


def test_aggregate_matches_a_dense_restatement():
    rng = np.random.default_rng(4)
    n, parts = 40, 9
    dense = np.where(rng.random((n, n)) < 0.2, rng.random((n, n)), 0.0)
    np.fill_diagonal(dense, 0.0)
    labels = rng.integers(0, parts, n).astype(np.int32)
    o = np.concatenate([[0], np.cumsum((dense != 0).sum(1))]).astype(np.int64)
    s = R.aggregate(o, np.nonzero(dense)[1].astype(np.int32), dense[dense != 0], labels, parts)
    onehot = np.eye(parts)[labels]
    expect = onehot.T @ dense @ onehot
    got = np.zeros((parts, parts))
    for p in range(parts):
        sl = slice(s["offsets"][p], s["offsets"][p + 1])
        got[p, s["indices"][sl]] = s["weights"][sl]
        assert (np.diff(s["indices"][sl]) > 0).all()
    np.testing.assert_allclose(got, expect, rtol=1e-12)
    assert ((got != 0) == (onehot.T @ (dense != 0) @ onehot != 0)).all()


def test_propagation_labels_every_reachable_row():
    """A path graph labelled at one end: with weight-1 edges every row is
    labelled within n sweeps; deeper levels need 10% of the row total."""
    n = 12
    rows = [[j for j in (r - 1, r + 1) if 0 <= j < n] for r in range(n)]
    o = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int64)
    i = np.array([j for r in rows for j in r], np.int32)
    labels = np.full(n, -1, np.int32)
    labels[0] = 5
    out, sweeps = R.propagate(o, i, np.ones(len(i), np.float32), labels, share=False)
    assert (out == 5).all() and sweeps <= 100
    out, _ = R.propagate(o, i, np.full(len(i), 0.5, np.float32), labels, share=False)
    assert (out[1:] == -1).all()          # a single 0.5 vote never reaches 1.0
    out, _ = R.propagate(o, i, np.full(len(i), 0.5, np.float32), labels, share=True)
    assert (out == 5).all()


# ------------------------------------------------------------- level params

def test_level_targets_use_requested_parts_and_realized_clamps():
    cfg = R.params(base_init_threshold=2)
    assert R.level_params(2, 16.0, 1.0, config=cfg) is None
    assert R.level_params(3, 16.0, 1.0, config=cfg).parts == 2
    first = R.level_params(400, 16.0, 1.0, config=cfg)
    assert (first.parts, first.epochs, first.a) == (100, 50, 2.0)
    second = R.level_params(60, previous=first, config=cfg)
    assert (second.parts, second.epochs) == (25, 70)
    # the curve is raised to 1/4 per level; good while coarsening continues
    assert second.a == np.float32(16.0 ** (1 / 16))
    assert first.good_initialization == (100 // 4 > 2)


def test_recursive_params_are_validated():
    o = np.array([0, 1, 2], np.int64)
    i = np.array([1, 0], np.int32)
    w = np.ones(2, np.float32)
    x = np.zeros((2, 3), np.float32)
    for bad in (dict(coarsening_ratio=5), dict(base_init_threshold=1),
                dict(n_embedding_epochs=0), dict(anchor_sample_size=0)):
        with pytest.raises(ValueError):
            R.recursive_init(o, i, w, x, 2, 1.0, 1.0, 0, R.params(**bad))


def test_malformed_and_asymmetric_graphs_are_rejected():
    x = np.random.default_rng(0).normal(size=(3, 2)).astype(np.float32)
    good = ([0, 1, 2, 2], [1, 0], [1.0, 1.0])
    R.recursive_init(*good, x, 2, 1.0, 1.0, 0)              # base case: host only
    for o, i, w in (([0, 1, 1, 1], [1], [1.0]),             # asymmetric
                    ([0, 1, 2, 2], [0, 1], [1.0, 1.0]),     # self loops
                    ([0, 1, 2, 2], [1, 0], [np.inf, 1.0]),  # non-finite
                    ([0, 1, 2, 2], [1, 0], [0.0, 1.0]),     # non-positive
                    ([0, 2, 2, 2], [2, 1], [1.0, 1.0])):    # unsorted
        with pytest.raises(ValueError, match="invalid recursive CSR graph"):
            R.recursive_init(o, i, w, x, 2, 1.0, 1.0, 0)


# ------------------------------------------------------------- PCA anchor

def pca_anchor(x, d):
    """Exact PCA scores with sklearn's svd_flip on the left vectors, each
    axis mapped to [0, 10]."""
    xc = x.astype(np.float64) - x.astype(np.float64).mean(0)
    u, s, vt = np.linalg.svd(xc, full_matrices=False)
    out = np.zeros((x.shape[0], d), np.float32)
    for a in range(d):
        col = u[:, a]
        sign = -1.0 if col[np.argmax(np.abs(col))] < 0 else 1.0
        v = (sign * s[a] * col).astype(np.float32)
        v = v - v.min()
        v = v / v.max() * np.float32(10)
        out[:, a] = v
    return out


@pytest.mark.parametrize("n,features", [(300, 40), (25, 60)])   # both orientations
def test_anchor_equals_exact_pca_on_low_rank_data(n, features):
    rng = np.random.default_rng(11)
    x = (rng.normal(size=(n, 4)) * [9.0, 5.0, 2.0, 1.0]) @ rng.normal(size=(4, features))
    x = x.astype(np.float32)
    values, mask = R.anchor(x, 2, 16384, seed=7)
    assert mask.all()
    np.testing.assert_allclose(values, pca_anchor(x, 2), atol=2e-4)
    assert values.min() == 0.0 and np.isclose(values.max(0), 10.0).all()


def test_anchor_samples_a_reservoir_and_fills_extra_axes():
    rng = np.random.default_rng(12)
    x = rng.normal(size=(500, 3)).astype(np.float32)
    values, mask = R.anchor(x, 2, 64, seed=3)
    assert mask.sum() == 64
    assert (values[~mask] == 0).all()
    again, mask2 = R.anchor(x, 2, 64, seed=3)
    assert (mask == mask2).all() and (values == again).all()
    other, mask3 = R.anchor(x, 2, 64, seed=4)
    assert (mask != mask3).any()
    # one feature, two components: the second axis is uniform on [0, 10)
    values, mask = R.anchor(x[:, :1], 2, 16384, seed=3)
    assert values[:, 1].min() >= 0.0 and values[:, 1].max() < 10.0
    assert np.isclose(values[:, 0].max(), 10.0) and values[:, 0].min() == 0.0


def test_base_case_only_run_is_the_scaled_anchor():
    """n <= base_init_threshold: run() is base_case(graph, anchor, 1); with
    every row sampled that is the centred PCA anchor scaled so each axis'
    5-95% span is 3 log10(n) log2(2)."""
    rng = np.random.default_rng(13)
    n = 200
    x = rng.normal(size=(n, 5)).astype(np.float32)
    o = np.concatenate([[0], np.arange(1, n + 1)]).astype(np.int64)
    o[-1] = o[-2]
    # a symmetric ring graph
    rows = [sorted({(r - 1) % n, (r + 1) % n}) for r in range(n)]
    o = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int64)
    i = np.array([j for r in rows for j in r], np.int32)
    w = np.ones(len(i), np.float32)
    out = R.recursive_init(o, i, w, x, 2, 1.577, 0.895, 42)
    values, mask = R.anchor(x, 2, 16384, seed=42)
    expect = R.base_case(o, i, w, values, mask, 1)
    assert (out == expect).all()
    for a in range(2):
        q = np.quantile(out[:, a].astype(np.float64), [0.05, 0.95])
        assert np.isclose(q[1] - q[0], 3 * np.log10(n), rtol=1e-5)


# ------------------------------------------------------------- init rescale

def rescale(values, n, d, heads, tails, weights, mode, factor, repulsion=1.0):
    o = np.zeros(n + 1, np.int64)
    for h in heads:
        o[h + 1] += 1
    o = np.cumsum(o)
    v = np.array(values, np.float32).reshape(n, d)
    assert list(heads) == sorted(heads)
    return R.init_rescale(v, mode, repulsion=repulsion, scale_factor=factor, offsets=o,
                          indices=np.array(tails, np.int32),
                          weights=np.array(weights, np.float32)).reshape(-1)


def close_rescale(a, b):
    assert abs(a - b) <= 1e-5 + 1e-5 * abs(b), (a, b)


def test_pinned_upstream_rescale_oracle():
    cases = {
        "isotropic": [-0.7508402, -1.0726289, -0.32178867, -1.0726289, -0.10726289,
                      -0.21452577, 1.1798917, 2.3597834],
        "per_axis": [-1.0130817, -0.7949736, -0.43417785, -0.7949736, -0.14472595,
                     -0.15899472, 1.5919855, 1.7489419],
        "graph_edge": [-0.875, -1.25, -0.375, -1.25, -0.125, -0.25, 1.375, 2.75],
    }
    for mode, expected in cases.items():
        out = rescale([0.0, 4.0, 2.0, 4.0, 3.0, 8.0, 9.0, 20.0], 4, 2, [0, 0, 1, 2, 3],
                      [1, 2, 0, 3, 2], [2.0, 3.0, 1.0, 1.0, 4.0], mode, 1.25)
        for a, e in zip(out, expected):
            close_rescale(a, e)


def test_directed_weighted_median_boundaries_and_zero():
    out = rescale([0.0, 1.0, 3.0, 10.0], 4, 1, [0, 0, 0, 1], [1, 2, 3, 0],
                  [1.0, 5.0, 1.0, 1.0], "graph_edge", 2.0, repulsion=0.0)
    close_rescale(out[0], -3.5 * 2.0 / 3.0)
    out = rescale([0.0, 2.0, 5.0], 3, 1, [0, 0], [1, 2], [1.0, 1.0], "graph_edge", 2.0)
    close_rescale(out[0], -7.0 / 3.0)
    out = rescale([2.0, 2.0, 5.0], 3, 1, [0, 1], [1, 0], [1.0, 1.0], "graph_edge", 2.0)
    close_rescale(out[0], -2.0)
    close_rescale(out[2], 4.0)
    out = rescale([1.0, 3.0], 2, 1, [], [], [], "graph_edge", 2.0)
    assert out.tolist() == [-1.0, 1.0]


def test_linear_quantiles_anisotropy_and_translation():
    t = 3.0 * np.log10(3.0)
    out = rescale([10.0, 8.0, 0.0, 8.0, 5.0, 8.0], 3, 2, [], [], [], "per_axis", 1.0)
    close_rescale(out[0], t / 1.8)
    close_rescale(out[2], -t / 1.8)
    for k in (1, 3, 4, 5):
        close_rescale(out[k], 0.0)
    iso = rescale([0.0, 0.0, 2.0, 20.0, 4.0, 40.0], 3, 2, [], [], [], "isotropic", 1.0)
    close_rescale(iso[1], iso[0] * 10.0)
    close_rescale(iso[0], -2.0 * t / np.sqrt(3.6 * 36.0))
    moved = rescale([100.0, -3.0, 102.0, 17.0, 104.0, 37.0], 3, 2, [], [], [], "isotropic", 2.0)
    for a, b in zip(moved, iso):
        close_rescale(a, 2.0 * b)
    zero = rescale([0.0, 0.0, 2.0, 20.0, 4.0, 40.0], 3, 2, [], [], [], "isotropic", 1.0,
                   repulsion=0.0)
    assert (zero == 0).all()
    assert (rescale([4.0] * 6, 3, 2, [], [], [], "per_axis", 1.0) == 0).all()
