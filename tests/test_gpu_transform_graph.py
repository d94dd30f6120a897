"""Transform front-end on the device (bipartite memberships, initial
placement, pruning at max / epochs, compact CSR with clocks) against the host
chain fuzzy_graph(bipartite) -> transform_init -> compact_graph, which the
host tests pin bit for bit to the oracle. The device uses its own expf, so
memberships agree to float rounding: the CSR structure must be identical
and the periods and placements within 1e-5 relative."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def host_chain(idx, dist, emb, lc, epochs, rate):
    from cuscan_amd import graph as G

    n = emb.shape[0]
    g = G.fuzzy_graph(idx, dist, local_connectivity=lc, bandwidth=1.0, bipartite=True,
                      n_reference=n)
    init = G.transform_init(idx, g["directed"], emb)
    cg = G.compact_graph(idx.shape[0], n, g["heads"], g["tails"], g["weights"], epochs=epochs,
                         rate=rate, prune_divisor=float(epochs), transform=True)
    cg["init"] = init
    return cg


def knn_case(q, n, k, seed, dup=False):
    rng = np.random.default_rng(seed)
    emb = rng.uniform(-10, 10, (n, 2)).astype(np.float32)
    idx = np.stack([rng.choice(n, k, replace=False) for _ in range(q)]).astype(np.int32)
    dist = np.sort(rng.gamma(2.0, 1.5, (q, k)).astype(np.float32), axis=1)
    if dup:
        idx[::5, 3] = idx[::5, 1]          # a reference listed twice in a row
        dist[::7, 0] = 0.0                  # a zero distance (membership exactly 1)
    return idx, dist, emb


@pytest.mark.parametrize("q,n,k,lc,epochs,dup", [
    (1, 50, 15, 0.0, 100, False),
    (37, 3000, 15, 0.0, 100, True),
    (1300, 4000, 15, 0.0, 100, True),
    (200, 500, 64, 0.5, 30, True),
    (64, 100, 1, 0.0, 100, False),
    (500, 2000, 10, 1.5, 7, True),
])
def test_device_transform_graph_matches_host_chain(q, n, k, lc, epochs, dup):
    from cuscan_amd import graph as G

    print(f"[gpu] transform graph q={q} n={n} k={k}", flush=True)
    idx, dist, emb = knn_case(q, n, k, 100 + q, dup)
    want = host_chain(idx, dist, emb, lc, epochs, 5.0)
    got = G.transform_graph_device(idx, dist, emb, local_connectivity=lc, epochs=epochs, rate=5.0)
    assert (got["offsets"] == want["offsets"]).all()
    assert (got["indices"] == want["indices"]).all()
    for key in ("periods", "negative_periods", "next_positive", "next_negative"):
        np.testing.assert_allclose(got[key], want[key], rtol=1e-5, err_msg=key)
    np.testing.assert_allclose(got["init"], want["init"], rtol=1e-5, atol=1e-5)


def test_device_transform_graph_errors():
    from cuscan_amd import graph as G

    idx, dist, emb = knn_case(8, 50, 5, 3)
    bad = idx.copy()
    bad[2, 1] = 50                          # outside the references
    with pytest.raises(RuntimeError, match="invalid neighbor indices"):
        G.transform_graph_device(bad, dist, emb, epochs=10, rate=5.0)
    none = np.full_like(idx, -1)            # no neighbours at all
    with pytest.raises(ValueError, match="empty after positive-edge pruning"):
        G.transform_graph_device(none, dist, emb, epochs=10, rate=5.0)


def fit_case(n, k, seed):
    from cuscan_amd.knn import KnnIndex

    rng = np.random.default_rng(seed)
    c = rng.uniform(-10, 10, (8, 6))
    x = (c[rng.integers(0, 8, n)] + rng.normal(0, 1, (n, 6))).astype(np.float32)
    idx, dist = KnnIndex(x, max_queries=n).query(x, k - 1, exclude_self=True)
    indices = np.concatenate([np.arange(n, dtype=np.int32)[:, None], idx], axis=1)
    distances = np.concatenate([np.zeros((n, 1), np.float32), dist], axis=1)
    return indices, distances


@pytest.mark.parametrize("n,k,lc,mix,epochs", [
    (500, 15, 1.0, 1.0, 200),
    (3000, 15, 1.0, 1.0, 200),
    (2000, 10, 1.5, 0.7, 50),
    (300, 64, 0.0, 1.0, 500),
])
def test_device_fit_graph_matches_host_chain(n, k, lc, mix, epochs):
    from cuscan_amd import graph as G

    print(f"[gpu] fit graph n={n} k={k}", flush=True)
    indices, distances = fit_case(n, k, 7 + n)
    g = G.fuzzy_graph(indices, distances, local_connectivity=lc, bandwidth=1.0, mix=mix)
    want = G.compact_graph(n, n, g["heads"], g["tails"], g["weights"], epochs=epochs, rate=5.0,
                           prune_divisor=float(epochs))
    got = G.fit_graph_device(indices, distances, local_connectivity=lc, mix=mix, epochs=epochs,
                             rate=5.0, prune_divisor=float(epochs))
    assert (got["offsets"] == want["offsets"]).all()
    assert (got["indices"] == want["indices"]).all()
    # a membership's last-ulp difference moves the bisection's sigma, and a
    # union weight carries both rows' memberships: 1e-4 relative
    np.testing.assert_allclose(got["weights"], want["weights"], rtol=1e-4)
    for key in ("periods", "negative_periods", "next_positive", "next_negative"):
        np.testing.assert_allclose(got[key], want[key], rtol=1e-4, err_msg=key)


def test_device_fit_graph_rejects_invalid_ids():
    from cuscan_amd import graph as G

    indices, distances = fit_case(200, 8, 3)
    indices[5, 3] = 200
    with pytest.raises(RuntimeError, match="invalid neighbor indices"):
        G.fit_graph_device(indices, distances, epochs=100, rate=5.0, prune_divisor=100.0)
