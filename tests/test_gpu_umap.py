"""End-to-end UMAP fit/transform on the device against the CPU reference
pipeline (oracle): full stochastic runs are checked by trustworthiness@15
within 0.01 (the reference's acceptance margin for whole runs)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from cuscan_amd.quality import trustworthiness  # noqa: E402
from oracle import oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def blobs(n_per, d, centers, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-10, 10, (centers, d))
    return np.vstack([c[i] + rng.normal(0, 1, (n_per, d)) for i in range(centers)]).astype(np.float32)


def test_fit_config1_quality_matches_cpu_reference():
    """2,000 x 10 blobs (10 clusters, seed 42), Adam, init=random, 200 epochs."""
    from cuscan_amd import UMAP

    x = blobs(200, 10, 10, 42)
    print("[gpu] umap fit 2000x10", flush=True)
    emb = UMAP(n_neighbors=15, optimizer="adam", init="random", n_epochs=200,
               negative_sample_rate=5, random_state=42).fit_transform(x)
    ref = O.fit(x, n_epochs=200, seed=42, threads=8)
    t_gpu, p_gpu = trustworthiness(x, emb, k=15)
    t_cpu, p_cpu = trustworthiness(x, ref, k=15)
    print(f"trustworthiness gpu={t_gpu:.4f} cpu={t_cpu:.4f} preservation {p_gpu:.4f}/{p_cpu:.4f}")
    assert emb.shape == (2000, 2) and emb.dtype == np.float32
    assert abs(t_gpu - t_cpu) <= 0.01
    assert abs(p_gpu - p_cpu) <= 0.02
    assert t_gpu > 0.9


@pytest.mark.parametrize("optimizer,q", [("adam", 200_000), ("momentum", 200_000), ("adam", 600)])
def test_fit_then_transform(optimizer, q):
    from cuscan_amd import UMAP

    x = blobs(300, 12, 8, 3)
    model = UMAP(optimizer=optimizer, init="random", n_epochs=150, random_state=3,
                 negative_selection_range=q)
    emb = model.fit_transform(x)
    t_fit, _ = trustworthiness(x, emb, k=15)
    assert t_fit > 0.9
    xq = blobs(40, 12, 8, 3)[::7] + np.float32(0.01)
    out = model.transform(xq)
    again = model.transform(xq)
    assert out.shape == (len(xq), 2)
    assert (out == again).all(), "transform must be repeatable"
    assert (model.embedding_ == emb).all(), "transform must not move the fitted embedding"
    # each transformed row lands near the fitted points it neighbours in input space
    d_in = ((xq[:, None, :] - x[None]) ** 2).sum(-1)
    nn = np.argsort(d_in, axis=1)[:, :15]
    d_out = np.linalg.norm(out[:, None, :] - emb[nn], axis=-1).mean(1)
    spread = np.linalg.norm(emb - emb.mean(0), axis=1).mean()
    assert (d_out < 0.5 * spread).all()
    assert out.dtype == np.float32
    assert model.transform(xq.astype(np.float64)).dtype == np.float64


def test_seeded_fit_is_reproducible():
    from cuscan_amd import UMAP

    x = blobs(150, 6, 4, 9)
    a = UMAP(init="random", n_epochs=60, random_state=5).fit_transform(x)
    b = UMAP(init="random", n_epochs=60, random_state=5).fit_transform(x)
    assert (a == b).all()


@pytest.mark.parametrize("optimizer", ["adam", "momentum"])
def test_unique_fits_distinct_rows(optimizer):
    """modern_unique_windows_use_distinct_rows (estimator.rs:3647-3672):
    duplicates share their distinct row's coordinates, which are those of a
    fit of the distinct rows alone; transform runs against the distinct
    rows."""
    from cuscan_amd import UMAP

    base = blobs(60, 6, 4, 11)
    rng = np.random.default_rng(5)
    dup = base[rng.integers(0, len(base), 120)]
    x = np.concatenate([base, dup])[rng.permutation(len(base) + len(dup))]
    kw = dict(optimizer=optimizer, init="random", n_epochs=60, random_state=9,
              negative_selection_range=50)
    model = UMAP(unique=True, **kw)
    emb = model.fit_transform(x)
    assert emb.shape == (len(x), 2)
    from cuscan_amd.umap import _distinct_rows

    rows, inverse = _distinct_rows(x)
    alone = UMAP(**kw).fit_transform(rows)
    assert (emb == alone[inverse]).all()
    xq = base[::9] + np.float32(0.05)
    assert (model.transform(xq) == UMAP(**kw).fit(rows).transform(xq)).all()
    tiny = np.array([[0.0], [1.0], [2.0], [0.0], [1.0]], np.float32)
    e = UMAP(unique=True, optimizer=optimizer, n_neighbors=2, n_epochs=4, init="random",
             negative_selection_range=1).fit_transform(tiny)
    assert (e[0] == e[3]).all() and (e[1] == e[4]).all()
