"""Pins of the oracle (the CPU restatement) from the reference's own golden
data and known answers, and from invariants a plausible slip would break.

- Fuzzy graph: umap-learn 0.6.0's graph_cases (rho, sigma, directed and
  symmetric memberships) at the reference's 1e-5 + 1e-5 |expected|
  (crates/umap/tests/test_standard_graph.rs:88-97, 325-400), for the oracle
  and for the product's host fuzzy graph.
- Exact kNN: the reference's tie and self-exclusion known answers
  (crates/umap/tests/test_standard_dispatch.rs:79-106) and brute force on
  integer data, where every squared distance is an exact integer and the
  order is fixed by (distance, index).
- Small-layout retry walk: small_retry_index(2, 1, a, 5) = [4, 0, 1, 2, 3]
  and the limit max(count, 1) * N (crates/umap-core/src/modern_layout.rs:
  481-502, 1305-1313), seen through the oracle's negative trace.
- Refresh cadence: epochs + [H > 1024] projection refreshes
  (crates/umap/src/standard_umap/modern_runtime.rs:1894-1899).
- Adaptation resultants: zero without repulsion, zero when every candidate
  is excluded, zero for mirrored candidates, and five copies of one
  candidate equal five times the fixture-pinned force without its 16 cap.
- FP64 gradient magnitudes: at least the row's gradient norm, equal to it
  for one term; the per-row error in term units flags a dropped term.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_helpers import advance_graph, random_graph, term_relative_error

GOLDEN = Path(__file__).resolve().parent / "golden"


def graph_cases():
    return json.loads((GOLDEN / "fuzzy_graph_cases.json").read_text())["graph_cases"]


def close(actual, expected, label):
    actual = np.asarray(actual, np.float64)
    expected = np.asarray(expected, np.float64)
    assert actual.shape == expected.shape, label
    tol = 1e-5 + 1e-5 * np.abs(expected)
    bad = np.abs(actual - expected) > tol
    assert not bad.any(), f"{label}: {actual[bad]} vs {expected[bad]}"


def check_graph_case(case, result):
    n, k = case["n"], case["k"]
    idx = np.asarray(case["indices"], np.int64).reshape(n, k)
    name = case["name"]
    close(result["rho"], case["rho"], f"{name} rho")
    close(result["sigma"], case["sigma"], f"{name} sigma")
    dense_dir = np.asarray(case["directed"], np.float64).reshape(n, n)
    want_dir = np.where((idx >= 0) & (idx < n),
                        dense_dir[np.arange(n)[:, None], np.clip(idx, 0, n - 1)], 0.0)
    close(np.asarray(result["directed"]).reshape(n, k), want_dir, f"{name} directed")
    dense = np.zeros((n, n))
    pairs = []
    for h, t, w in zip(result["heads"], result["tails"], result["weights"]):
        if w == 0.0:
            continue
        assert 0 <= h < n and 0 <= t < n and h != t
        pairs.append((int(h), int(t)))
        dense[h, t] = w
    assert pairs == sorted(set(pairs)), f"{name}: edges sorted and unique"
    close(dense.ravel(), case["symmetric"], f"{name} symmetric")


@pytest.mark.parametrize("i", range(3))
def test_oracle_fuzzy_graph_golden_cases(i):
    case = graph_cases()[i]
    n, k = case["n"], case["k"]
    g = O.fuzzy_graph(np.asarray(case["indices"], np.int32).reshape(n, k),
                      np.asarray(case["distances"], np.float32).reshape(n, k),
                      local_connectivity=case["local_connectivity"],
                      bandwidth=case["bandwidth"], mix=case["set_op_mix_ratio"])
    u = g["used"]
    check_graph_case(case, dict(g, heads=g["heads"][:u], tails=g["tails"][:u],
                                weights=g["weights"][:u]))


@pytest.mark.parametrize("i", range(3))
def test_host_fuzzy_graph_golden_cases(i):
    from cuscan_amd import graph as G

    case = graph_cases()[i]
    n, k = case["n"], case["k"]
    g = G.fuzzy_graph(np.asarray(case["indices"], np.int32).reshape(n, k),
                      np.asarray(case["distances"], np.float32).reshape(n, k),
                      local_connectivity=case["local_connectivity"],
                      bandwidth=case["bandwidth"], mix=case["set_op_mix_ratio"])
    check_graph_case(case, g)


def test_exact_knn_reference_known_answers():
    ref = np.array([[0, 0], [1, 0], [-1, 0], [0, 0]], np.float32)
    qry = np.array([[-1, 0], [0, 0]], np.float32)
    idx, dist = O.exact_knn(qry, ref, 3, exclude_self=True, query_offset=2, threads=1)
    assert idx.ravel().tolist() == [0, 3, 1, 0, 1, 2]
    assert dist.ravel().tolist() == [1.0, 1.0, 2.0, 0.0, 1.0, 1.0]
    ref = np.array([[0, 0], [0, 0], [2, 0]], np.float32)
    idx, dist = O.exact_knn(np.zeros((1, 2), np.float32), ref, 3, threads=1)
    assert idx.ravel().tolist() == [0, 1, 2]
    assert dist.ravel().tolist() == [0.0, 0.0, 2.0]


@pytest.mark.parametrize("exclude,offset", [(False, 0), (True, 0), (True, 37)])
def test_exact_knn_brute_force_integer_data(exclude, offset):
    rng = np.random.default_rng(5)
    ref = rng.integers(0, 6, (300, 7)).astype(np.float32)
    ref[100:140] = ref[:40]                     # duplicate rows: exact distance ties
    qry = ref[offset:offset + 60] if exclude else rng.integers(0, 6, (60, 7)).astype(np.float32)
    k = 12
    idx, dist = O.exact_knn(qry, ref, k, exclude_self=exclude, query_offset=offset, threads=2)
    d2 = ((qry.astype(np.float64)[:, None, :] - ref.astype(np.float64)[None]) ** 2).sum(-1)
    for q in range(len(qry)):
        ids = np.arange(len(ref))
        keep = ids != offset + q if exclude else np.ones(len(ref), bool)
        order = np.lexsort((ids[keep], d2[q, keep]))[:k]
        assert idx[q].tolist() == ids[keep][order].tolist()
        assert np.array_equal(dist[q], np.sqrt(d2[q, keep][order]).astype(np.float32))


def _retry_epoch(cols_of_row0):
    """Five rows, row 0 holding three edges (edge 2 fires at epoch 1 with one
    negative), identity negative order, exclusion on: a small layout."""
    n = 5
    cols = list(cols_of_row0)
    offsets = np.array([0, len(cols)] + [len(cols)] * (n - 1), np.uint32)
    E = len(cols)
    npos = np.full(E, 100.0)
    npos[2] = 1.0
    nneg = np.full(E, 100.0)
    nneg[2] = 0.5
    head = np.arange(2 * n, dtype=np.float32).reshape(n, 2)
    return O.Epoch(head=head, tail=head, offsets=offsets, indices=np.array(cols, np.uint32),
                   source_order=np.arange(n, dtype=np.int32),
                   negative_order=np.arange(n, dtype=np.int32),
                   source_ranks=np.arange(n, dtype=np.int32), periods=np.ones(E),
                   next_positive=npos, negative_periods=np.full(E, 0.5),
                   next_negative=nneg, updates=np.zeros((n, 2), np.float32),
                   m=np.zeros((n, 2)), v=np.zeros((n, 2)), epoch=1, a=1.0, b=1.0, gamma=1.0,
                   scale=1.0, negative_range=n, shared=True, negatives=True, exclude=True)


def test_small_retry_walk_known_answer():
    # edge 2 of row 0 -> column 4; base = 2 * (1 + 1) mod 5 = 4: candidates
    # 4 (neighbour), 0 (self), 1, 2 (neighbours) are dropped, 3 repels
    edges, cands, kept, fires = _retry_epoch([1, 2, 4]).trace_negatives()
    assert fires == 1
    assert edges.tolist() == [2] * 5
    assert cands.tolist() == [4, 0, 1, 2, 3]
    assert kept.tolist() == [0, 0, 0, 0, 1]


def test_small_retry_limit_fully_excluded_row():
    # row 0 holds 1..4 and is row 0 itself: every candidate is dropped, and
    # the walk stops after max(count, 1) * N = 5 attempts
    edges, cands, kept, fires = _retry_epoch_full().trace_negatives()
    assert fires == 1 and edges.tolist() == [2] * 5
    assert cands.tolist() == [4, 0, 1, 2, 3] and not kept.any()


def _retry_epoch_full():
    n = 5
    cols = [1, 2, 3, 4]
    offsets = np.array([0, 4, 4, 4, 4, 4], np.uint32)
    npos = np.array([100.0, 100.0, 1.0, 100.0])
    nneg = np.array([100.0, 100.0, 0.5, 100.0])
    head = np.arange(2 * n, dtype=np.float32).reshape(n, 2)
    return O.Epoch(head=head, tail=head, offsets=offsets, indices=np.array(cols, np.uint32),
                   source_order=np.arange(n, dtype=np.int32),
                   negative_order=np.arange(n, dtype=np.int32),
                   source_ranks=np.arange(n, dtype=np.int32), periods=np.ones(4),
                   next_positive=npos, negative_periods=np.full(4, 0.5),
                   next_negative=nneg, updates=np.zeros((n, 2), np.float32),
                   m=np.zeros((n, 2)), v=np.zeros((n, 2)), epoch=1, a=1.0, b=1.0, gamma=1.0,
                   scale=1.0, negative_range=n, shared=True, negatives=True, exclude=True)


@pytest.mark.parametrize("h", [1000, 1100])
def test_refresh_cadence(h):
    epochs = 22
    g = random_graph(h, h, 4, 3, epochs)
    head = np.random.default_rng(0).uniform(-5, 5, (h, 2)).astype(np.float32)
    _, _, _, stats = O.optimize_compact(head, None, g, epochs=epochs, a=1.58, b=0.9,
                                        negative_selection_range=300, threads=2)
    assert stats["refreshes"] == epochs + (1 if h > 1024 else 0)


def _one_sample(emb, offsets, indices, cand_rank, gamma, exclude, n):
    order = np.arange(n, dtype=np.int32)
    return O.resultants(emb, order, offsets, indices, sample_ranks=np.array([0], np.int32),
                        local_offsets=np.full(5, cand_rank, np.int32),
                        global_ranks=np.full(5, cand_rank, np.int32), raw_range=n,
                        a=1.58, b=0.9, gamma=gamma, exclude=exclude)


def test_resultant_invariants():
    n = 6
    emb = np.array([[0, 0], [1, 0.5], [-1, -0.5], [2, 1], [-2, -1], [3, -4]], np.float32)
    offsets = np.array([0, 2, 2, 2, 2, 2, 2], np.uint32)
    indices = np.array([1, 5], np.uint32)
    # gamma = 0: no repulsion at all
    assert not _one_sample(emb, offsets, indices, 3, 0.0, False, n).any()
    # every candidate excluded (a neighbour of source row 0)
    assert not _one_sample(emb, offsets, indices, 5, 1.5, True, n).any()
    assert _one_sample(emb, offsets, indices, 5, 1.5, False, n).all()
    # mirrored candidates (rows 1/2 and 3/4 around row 0) and the source
    # itself (q = 0) cancel
    order = np.arange(n, dtype=np.int32)
    res = O.resultants(emb, order, offsets, indices, np.array([0], np.int32),
                       np.array([1, 2, 3, 4, 0], np.int32), np.array([1, 2, 3, 4, 0], np.int32),
                       raw_range=n, a=1.58, b=0.9, gamma=1.0, exclude=False)
    assert np.abs(res).max() < 1e-12
    # five copies of one candidate: five times the (fixture-pinned) f32 force
    # with scale 1; without the min(gamma, 16) cap above 16
    for gamma in (1.5, 24.0):
        r = _one_sample(emb, offsets, indices, 5, gamma, False, n)
        grad = np.zeros(2, np.float32)
        O.lib().orc_accumulate_force(grad, np.ascontiguousarray(emb[0]),
                                     np.ascontiguousarray(emb[5]), 2, 1.58, 0.9, gamma, 1.0, 0, 1)
        want = 5.0 * float(np.linalg.norm(grad.astype(np.float64))) * max(gamma / 16.0, 1.0)
        assert abs(r[0] - want) <= 1e-5 * want and abs(r[1] - want) <= 1e-5 * want


def test_gradient_magnitude_and_term_units():
    n = 3000
    g = random_graph(n, n, 6, 11, 200)
    epoch = 60
    gt = advance_graph(g, epoch)
    rng = np.random.default_rng(1)
    head = rng.uniform(-10, 10, (n, 2)).astype(np.float32)
    order = rng.permutation(n).astype(np.int32)

    def epoch_args(negatives=True):
        return O.Epoch(head=head, tail=head, offsets=g["offsets"], indices=g["indices"],
                       source_order=np.arange(n, dtype=np.int32), negative_order=order,
                       source_ranks=np.arange(n, dtype=np.int32), periods=g["periods"],
                       next_positive=gt["next_positive"], negative_periods=g["negative_periods"],
                       next_negative=gt["next_negative"], updates=np.zeros((n, 2), np.float32),
                       m=np.zeros((n, 2)), v=np.zeros((n, 2)), epoch=epoch, a=1.58, b=0.9,
                       gamma=1.0, scale=1.0, negative_range=n, shared=True,
                       negatives=negatives, exclude=True)

    g64, mag = epoch_args().gradient_block64_mag(0, n, 1.0)
    assert np.array_equal(g64, epoch_args().gradient_block64(0, n, 1.0))
    assert (mag >= np.linalg.norm(g64, axis=1) * (1 - 1e-12)).all()
    assert (mag > 0).sum() > n // 2
    # attraction only: a row with one firing edge has magnitude = its norm
    a64, amag = epoch_args(False).gradient_block64_mag(0, n, 1.0)
    fires = gt["next_positive"] <= epoch
    per_row = np.add.reduceat(fires.astype(np.int64), g["offsets"][:-1].astype(np.int64))
    one = per_row == 1
    assert one.sum() > 10
    assert np.allclose(amag[one], np.linalg.norm(a64[one], axis=1), rtol=1e-14, atol=0)
    # the f32 restatement sits far inside the per-row bar ...
    ref32 = epoch_args(False)
    ref32.gradient_block(0, n, 1.0, threads=2)
    assert term_relative_error(ref32.updates, a64, amag) < 1e-6
    # ... and one attraction term dropped from a row with several (planted
    # here, on the CPU) fails it
    row = int(np.flatnonzero(per_row >= 3)[0])
    s0 = int(g["offsets"][row])
    e = s0 + int(np.flatnonzero(fires[s0:int(g["offsets"][row + 1])])[0])
    term = np.zeros(2, np.float32)
    O.lib().orc_accumulate_force(term, np.ascontiguousarray(head[row]),
                                 np.ascontiguousarray(head[g["indices"][e]]), 2, 1.58, 0.9, 1.0,
                                 1.0, 0, 0)
    bad = ref32.updates.astype(np.float64).copy()
    bad[row] -= term
    assert term_relative_error(bad, a64, amag) > 1e-2
