"""Host graph/initialisation steps of the product against the oracle, and the
oracle's versions against independent numpy restatements / invariants."""
import math

import numpy as np
import pytest

from cuscan_amd import graph as G
from oracle import oracle as O


def blobs(n, d, centers, seed, scale=10.0, sigma=1.0):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-scale, scale, (centers, d))
    return (c[rng.integers(0, centers, n)] + rng.normal(0, sigma, (n, d))).astype(np.float32)


def fit_knn(x, k):
    idx, dist = O.exact_knn(x, x, k - 1, exclude_self=True, threads=4)
    n = x.shape[0]
    return (np.concatenate([np.arange(n, dtype=np.int32)[:, None], idx], 1),
            np.concatenate([np.zeros((n, 1), np.float32), dist], 1))


def numpy_memberships(idx, dist, bipartite):
    """Independent restatement: rho = smallest positive distance, sigma by
    the same bisection, memberships exp(-max(d - rho, 0) / sigma)."""
    n, k = idx.shape
    target = math.log2(k)
    mean = dist.astype(np.float64).mean()
    out = np.zeros_like(dist)
    rhos, sigmas = [], []
    for r in range(n):
        d = dist[r].astype(np.float32)
        pos = d[d > 0]
        rho = np.float32(pos.min()) if pos.size else np.float32(0)
        lo, hi, mid = np.float32(0), np.float32(3.4028235e38), np.float32(1)
        for _ in range(64):
            delta = d[1:] - rho
            val = np.where(delta <= 0, np.float32(1),
                           np.exp(-np.maximum(delta, 0) / (mid if mid != 0 else np.float32(1))))
            s = np.float32(0)
            for v in val.astype(np.float32):
                s = np.float32(s + v)
            if abs(s - np.float32(target)) < 1e-5:
                break
            if s > target:
                hi = mid
                mid = np.float32((lo + hi) * np.float32(0.5))
            elif hi == np.float32(3.4028235e38):
                lo = mid
                mid = np.float32(mid * 2)
            else:
                lo = mid
                mid = np.float32((lo + hi) * np.float32(0.5))
        floor_mean = np.float32(d.sum(dtype=np.float32) / np.float32(k)) if rho > 0 else np.float32(mean)
        mid = max(mid, np.float32(floor_mean * np.float32(1e-3)))
        m = np.where((d <= rho) | (mid == 0), np.float32(1),
                     np.exp(-np.maximum(d - rho, 0) / mid).astype(np.float32))
        bad = (idx[r] < 0) | ((idx[r] == r) & (not bipartite))
        m[bad] = 0
        out[r] = m
        rhos.append(rho)
        sigmas.append(mid)
    return out, np.array(rhos, np.float32), np.array(sigmas, np.float32)


@pytest.mark.parametrize("bipartite", [False, True])
def test_fuzzy_graph_product_equals_oracle_and_oracle_matches_numpy(bipartite):
    x = blobs(400, 8, 5, 1)
    if bipartite:
        q = blobs(60, 8, 5, 2)
        idx, dist = O.exact_knn(q, x, 15)
        kw = dict(local_connectivity=0.0, bipartite=True, n_reference=400)
    else:
        idx, dist = fit_knn(x, 15)
        kw = dict(local_connectivity=1.0)
    o = O.fuzzy_graph(idx, dist, **kw)
    p = G.fuzzy_graph(idx, dist, **kw)
    u = o["used"]
    assert u == len(p["heads"])
    for key in ("heads", "tails", "weights"):
        assert (p[key] == o[key][:u]).all(), key
    for key in ("rho", "sigma", "directed"):
        assert (p[key].view(np.uint32) == o[key].view(np.uint32)).all(), key
    if not bipartite:
        m, rho, sigma = numpy_memberships(idx, dist, bipartite)
        assert (rho == o["rho"]).all()
        # numpy's vectorised float32 exp differs from libm by an ulp, which can
        # move the bisection's stopping step: sigma agrees to the 1e-5 stop rule.
        np.testing.assert_allclose(o["sigma"], sigma, rtol=2e-5)
        np.testing.assert_allclose(o["directed"], m, rtol=2e-5, atol=1e-6)
        # symmetrisation: w = f + r - f*r over each unordered pair, both directions
        W = {}
        for r in range(idx.shape[0]):
            for s in range(idx.shape[1]):
                if idx[r, s] != r:
                    W[(r, int(idx[r, s]))] = W.get((r, int(idx[r, s])), 0.0) + float(o["directed"][r, s])
        for hd, tl, w in zip(o["heads"][:u], o["tails"][:u], o["weights"][:u]):
            f, rr = W.get((hd, tl), 0.0), W.get((tl, hd), 0.0)
            assert abs(w - (f + rr - f * rr)) < 1e-6
        # per-head segments sorted by tail, both directions present
        assert (np.diff(o["heads"][:u]) >= 0).all()
        pairs = set(zip(o["heads"][:u].tolist(), o["tails"][:u].tolist()))
        assert all((t, h) in pairs for h, t in pairs)
        # memberships over slots 1..k sum to log2(k) unless sigma is floored
        s = o["directed"][:, 1:].sum(1)
        floored = o["sigma"] <= o["sigma"].min() * 1.0001
        assert (np.abs(s[~floored] - math.log2(15)) < 1e-3).all()


def test_fuzzy_graph_rejects_invalid_ids():
    idx = np.array([[0, 1], [1, 7]], np.int32)
    dist = np.array([[0, 1], [0, 1]], np.float32)
    with pytest.raises(Exception, match="invalid neighbor"):
        G.fuzzy_graph(idx, dist)


@pytest.mark.parametrize("transform", [False, True])
def test_compact_graph_product_equals_oracle(transform):
    rng = np.random.default_rng(4)
    rows, cols = (300, 300) if not transform else (50, 300)
    n = 3000
    h = rng.integers(0, rows, n)
    t = rng.integers(0, cols, n)
    w = rng.uniform(0, 1, n).astype(np.float32)
    w[rng.integers(0, n, 300)] = 0
    divisor = 200.0
    thr = np.float32(w.max()) / np.float32(divisor)
    w_pruned = np.where(w < thr, np.float32(0), w)
    o = O.compact_graph(rows, cols, h, t, w_pruned, 200, 5.0)
    p = G.compact_graph(rows, cols, h, t, w, epochs=200, rate=5.0, prune_divisor=divisor,
                        transform=transform)
    for key in o:
        assert (p[key] == o[key]).all(), key


def test_fit_curve_product_equals_oracle():
    for md, sp in [(0.1, 1.0), (0.0, 1.0), (0.5, 2.0), (0.05, 0.5)]:
        assert G.fit_curve(md, sp) == O.fit_curve(md, sp)


@pytest.mark.parametrize("n,d,mode", [(2000, 2, 0), (1000, 3, 1), (2, 2, 0), (1, 2, 0)])
def test_random_init_product_equals_oracle_and_closed_form(n, d, mode):
    p = G.random_init(n, d, 42, mode=["isotropic", "per_axis"][mode])
    o = O.random_init(n, d, 42, mode=mode)
    assert (p.view(np.uint32) == o.view(np.uint32)).all()
    if n < 3:
        return
    target = 3.0 * math.log10(n)
    spans = np.array([np.quantile(o[:, c].astype(np.float64), 0.95) -
                      np.quantile(o[:, c].astype(np.float64), 0.05) for c in range(d)])
    if mode == 0:
        assert abs(math.exp(np.log(spans).mean()) - target) < 1e-4 * target
    else:
        np.testing.assert_allclose(spans, target, rtol=1e-4)
    assert np.abs(o.mean(0)).max() < 1e-4 * target
    raw = O.StdRng(42)
    first = raw.range_f32(-10.0, 10.0)
    assert -10.0 <= first < 10.0


def test_transform_init_product_equals_oracle_and_closed_form():
    rng = np.random.default_rng(9)
    emb = rng.normal(size=(100, 2)).astype(np.float32)
    idx = rng.integers(0, 100, (20, 6)).astype(np.int32)
    w = rng.uniform(0.1, 0.9, (20, 6)).astype(np.float32)
    w[3, 2] = 1.0                      # copy rule
    idx[5, :] = [7, 7, 8, 9, 10, 11]   # duplicates are coalesced
    w[5, :2] = [0.5, 0.5]              # coalesced weight exactly 1 -> copy id 7
    o = O.transform_init(idx, w, emb)
    p = G.transform_init(idx, w, emb)
    assert (p.view(np.uint32) == o.view(np.uint32)).all()
    assert (o[3] == emb[idx[3, 2]]).all() and (o[5] == emb[7]).all()
    r = 0
    expect = (emb[idx[r]].astype(np.float64) * (w[r] / w[r].sum())[:, None]).sum(0)
    np.testing.assert_allclose(o[r], expect, rtol=1e-5)
    with pytest.raises(Exception, match="no positive"):
        G.transform_init(idx[:1], np.zeros((1, 6), np.float32), emb)
