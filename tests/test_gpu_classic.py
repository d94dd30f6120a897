"""Compatibility (classic) SGD on the device.

The contract is embedding quality (crates/umap/src/kernels/standard_umap/
layout.rs:42-47): coordinates come from concurrent atomic updates. What is
deterministic is pinned exactly: the f64 clocks and the per-head Tausworthe
states after every epoch (umap-learn's recorded frames, and the oracle's
whole runs). Whole runs are checked by trustworthiness@15 within 0.01 of the
oracle's serial run on the same graph and initialisation.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from cuscan_amd.quality import trustworthiness  # noqa: E402
from oracle import oracle as O  # noqa: E402
from tests.classic_probes import A_PROBE, B_PROBE, cases, replay  # noqa: E402

pytestmark = pytest.mark.gpu
CASES = cases()


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _dev(a, dtype):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda:0")


@pytest.mark.parametrize("i", range(len(CASES)))
def test_classic_epoch_abi_matches_umap_learn_frames(i):
    from cuscan_amd import device as D

    case, edge_count, frames = CASES[i]
    print(f"[gpu] classic probe {case['name']} edges={edge_count}", flush=True)

    def run(st, epoch, alpha):
        head = _dev(st["head"], torch.float32)
        tail = head if st["shared"] else _dev(st["tail"], torch.float32)
        npos, nneg = _dev(st["npos"], torch.float64), _dev(st["nneg"], torch.float64)
        rng = _dev(st["rng"], torch.int64)
        D.classic_epoch(head=head, tail=tail, tails=_dev(st["tails"], torch.int32),
                        weights=_dev(st["weights"], torch.float32),
                        epochs_per_sample=_dev(st["per"], torch.float64),
                        epoch_of_next_sample=npos,
                        epochs_per_negative_sample=_dev(st["nper"], torch.float64),
                        epoch_of_next_negative_sample=nneg, rng_state=rng,
                        head_offsets=_dev(st["offsets"], torch.int32),
                        head_edges=_dev(st["ids"], torch.int32),
                        mask=torch.zeros(max(len(st["tails"]), 1), dtype=torch.int32,
                                         device="cuda:0"),
                        epoch=epoch, a=A_PROBE, b=B_PROBE, gamma=st["gamma"], alpha=alpha,
                        negative_sample_rate=st["rate"], move_other=st["shared"],
                        shared=st["shared"])
        torch.cuda.synchronize()
        st["head"] = head.cpu().numpy()
        st["npos"], st["nneg"] = npos.cpu().numpy(), nneg.cpu().numpy()
        st["rng"] = rng.cpu().numpy()

    def coords(st, want_head, want_tail):
        # every probe has a single head vertex with edges, so the atomic
        # updates are applied in serial order
        np.testing.assert_allclose(st["head"], want_head, rtol=2e-5, atol=2e-5)

    replay(case, edge_count, frames, run, coords)


def blobs(n_per, d, centers, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-10, 10, (centers, d))
    return np.vstack([c[i] + rng.normal(0, 1, (n_per, d)) for i in range(centers)]).astype(np.float32)


def _graph(x, k=15):
    from cuscan_amd import graph as G

    n = len(x)
    idx, dist = O.exact_knn(x, x, k - 1, exclude_self=True, threads=8)
    indices = np.concatenate([np.arange(n, dtype=np.int32)[:, None], idx], 1)
    distances = np.concatenate([np.zeros((n, 1), np.float32), dist], 1)
    return G.fuzzy_graph(indices, distances)


@pytest.mark.parametrize("epochs", [60, 200])
def test_classic_fit_clocks_and_generators_exact_quality_within_001(epochs):
    from cuscan_amd import graph as G
    from cuscan_amd.classic import optimize_classic

    x = blobs(200, 10, 10, 42)
    g = _graph(x)
    a, b = G.fit_curve(0.1, 1.0)
    init = G.random_init(len(x), 2, 42)
    print(f"[gpu] classic fit 2000x10 epochs={epochs}", flush=True)
    emb, gp, gn, grng = optimize_classic(init, None, g["heads"], g["tails"], g["weights"],
                                         n_epochs=epochs, default_epochs=500, a=a, b=b,
                                         learning_rate=1.0, gamma=1.0, negative_sample_rate=5,
                                         seed=42, shared=True, return_state=True)
    ref, rp, rn, rrng = O.classic_optimize(init, None, g["heads"], g["tails"], g["weights"],
                                           n_epochs=epochs, default_epochs=500, a=a, b=b,
                                           learning_rate=1.0, gamma=1.0, rate=5, seed=42,
                                           shared=True)
    assert (gp == rp).all() and (gn == rn).all()
    assert (grng == rrng).all()
    t_gpu, _ = trustworthiness(x, emb, k=15)
    t_cpu, _ = trustworthiness(x, ref, k=15)
    print(f"classic trustworthiness gpu={t_gpu:.4f} cpu={t_cpu:.4f}")
    assert np.isfinite(emb).all()
    assert abs(t_gpu - t_cpu) <= 0.01


def test_classic_transform_keeps_tail_and_matches_generators():
    from cuscan_amd import graph as G
    from cuscan_amd.classic import optimize_classic

    x = blobs(150, 8, 6, 7)
    q = blobs(20, 8, 6, 8)[:60]
    n = len(x)
    idx, dist = O.exact_knn(q, x, 15)
    g = G.fuzzy_graph(idx, dist, local_connectivity=0.0, bipartite=True, n_reference=n)
    fitted = G.random_init(n, 2, 3)
    init = G.transform_init(idx, g["directed"], fitted)
    a, b = G.fit_curve(0.1, 1.0)
    kw = dict(n_epochs=30, default_epochs=30, a=a, b=b, learning_rate=0.25, gamma=1.0, seed=42,
              shared=False)
    emb, gp, gn, grng = optimize_classic(init, fitted, g["heads"], g["tails"], g["weights"],
                                         negative_sample_rate=5, return_state=True, **kw)
    ref, rp, rn, rrng = O.classic_optimize(init, fitted, g["heads"], g["tails"], g["weights"],
                                           rate=5, **kw)
    assert (gp == rp).all() and (gn == rn).all() and (grng == rrng).all()
    # transform rows only see the fixed tail: each head row is updated by
    # its own lane alone, so coordinates follow the serial run closely
    np.testing.assert_allclose(emb, ref, rtol=0, atol=1e-3)


def test_umap_compatibility_fit_and_transform():
    from cuscan_amd import UMAP

    x = blobs(150, 10, 8, 5)
    model = UMAP(n_neighbors=15, optimizer="compatibility", init="random", n_epochs=120,
                 random_state=3)
    print("[gpu] umap compatibility fit 1200x10", flush=True)
    emb = model.fit_transform(x)
    ref = O.fit(x, n_epochs=120, seed=3, threads=8, optimizer="compatibility")
    t_gpu, _ = trustworthiness(x, emb, k=15)
    t_cpu, _ = trustworthiness(x, ref, k=15)
    print(f"umap compatibility trustworthiness gpu={t_gpu:.4f} cpu={t_cpu:.4f}")
    assert abs(t_gpu - t_cpu) <= 0.01
    new = model.transform(x[:40] + 0.01)
    assert new.shape == (40, 2) and np.isfinite(new).all()
