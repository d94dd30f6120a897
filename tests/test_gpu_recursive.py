"""Recursive initialisation with its coarse layouts on the device.

- The native recursion equals, bit for bit, a Python composition of the same
  host steps and device layouts written from the reference's descend
  (crates/umap-core/src/recursive_init/mod.rs:125-320) and the coarse-layout
  callback (crates/umap/src/standard_umap/estimator.rs:2529-2585 with
  ModernLayoutParams::recursive_level, crates/umap-core/src/
  modern_layout.rs:55-71, and compact_graph_from_csr, :210-300). This pins
  the composition: the per-depth seed streams, the level schedule, the
  negative range, the clocks and the order of prolongation and alignment.
- init="auto" equals init="recursive" (python/tests/test_standard_umap.py:
  791-807), and recursive fits reach the random-init fit's quality.
The individual host steps are pinned on the CPU (tests/test_recursive_init.py).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from cuscan_amd import graph as G  # noqa: E402
from cuscan_amd import layout as L  # noqa: E402
from cuscan_amd import recursive as R  # noqa: E402
from cuscan_amd.quality import trustworthiness  # noqa: E402
from oracle import oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu
M64 = (1 << 64) - 1


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def blobs(n_per, d, centers, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-10, 10, (centers, d))
    return np.vstack([c[i] + rng.normal(0, 1, (n_per, d)) for i in range(centers)]).astype(np.float32)


def fit_graph(x, k=15, epochs=200):
    n = len(x)
    idx, dist = O.exact_knn(x, x, k - 1, exclude_self=True, threads=8)
    idx = np.concatenate([np.arange(n, dtype=np.int32)[:, None], idx], 1)
    dist = np.concatenate([np.zeros((n, 1), np.float32), dist], 1)
    return G.fit_graph_device(idx, dist, epochs=epochs, rate=5, prune_divisor=float(epochs))


# ------------------------------------------------- restatement of descend

def stream_seed(seed, domain, depth):
    """mod.rs:486-491."""
    v = (seed ^ ((depth * 0x9E3779B97F4A7C15) & M64) ^ domain) & M64
    v = ((v ^ (v >> 30)) * 0xBF58476D1CE4E5B9) & M64
    v = ((v ^ (v >> 27)) * 0x94D049BB133111EB) & M64
    return v ^ (v >> 31)


def coarse_negative_range(rows):
    """mod.rs:563-567."""
    half = rows // 2
    return min(max(half + (rows % 2) * (half % 2), 1), rows)


def compact_from_csr(g, epochs, rate):
    """compact_graph_from_csr with edge_periods (modern_layout.rs:188-300)."""
    w = g["weights"].astype(np.float32)
    samples = np.float32(epochs) * (w / w.max())
    period = np.float64(epochs) / samples.astype(np.float64)
    neg = (period / np.float64(rate)) * 1.5
    return dict(offsets=g["offsets"].astype(np.uint32), indices=g["indices"].astype(np.uint32),
                weights=w, periods=period, negative_periods=neg, next_positive=period.copy(),
                next_negative=neg.copy())


def python_run(g, x, d, a, b, cfg, seed, method):
    values, mask = R.anchor(x, d, cfg.anchor_sample_size, seed)
    level = R.level_params(len(x), a, b, config=cfg)
    return descend(g, None, (values, mask), level, 1, cfg, seed, method)


def descend(g, summed, anchor, level, depth, cfg, seed, method):
    o, i, w = g["offsets"], g["indices"], g["weights"]
    if level is None:
        return R.base_case(o, i, w, anchor[0], anchor[1], depth)
    labels = R.hubs(o, i, w, level.parts, method, stream_seed(seed, 2, depth))
    labels, _ = R.propagate(o, i, w, labels, depth > 1)
    labels = R.outliers(o, i, labels, stream_seed(seed, 3, depth))
    labels, parts = R.remap(labels)
    src = summed if summed is not None else dict(offsets=o, indices=i,
                                                 weights=w.astype(np.float64))
    s = R.aggregate(src["offsets"], src["indices"], src["weights"], labels, parts)
    t = R.transition(s["offsets"], s["indices"], s["weights"])
    values, mask = anchor
    d = values.shape[1]
    sums = np.zeros((parts, d))
    counts = np.zeros(parts, np.int64)
    for r in np.nonzero(mask)[0]:
        sums[labels[r]] += values[r].astype(np.float64)
        counts[labels[r]] += 1
    cmask = counts > 0
    cvals = np.zeros((parts, d), np.float32)
    cvals[cmask] = (sums[cmask] / counts[cmask, None]).astype(np.float32)
    nxt = R.level_params(parts, previous=level, config=cfg)
    coarse = descend(t, s, (cvals, cmask), nxt, depth + 1, cfg, seed, method)
    if len(t["weights"]):
        params = L.LayoutParams(epochs=level.epochs, a=level.a, b=level.b, optimizer="adam",
                                repulsion=4.0, initial_alpha=0.5, negative_sample_rate=1,
                                negative_selection_range=coarse_negative_range(parts),
                                negative_sample_scale=None, adaptation_samples=128,
                                exclude_graph_neighbors=False,
                                good_initialization=bool(level.good_initialization),
                                seed=stream_seed(seed, 4, depth))
        coarse, _, _, _ = L.optimize_compact(coarse, None, compact_from_csr(t, level.epochs, 1.0),
                                             params, shared=True)
    result = R.prolongate(o, i, w, labels, coarse)
    return R.procrustes(result, values, mask)


@pytest.mark.parametrize("n_per,ratio,base,method", [
    (600, 4, 256, "degree"),       # 6,000 rows: two coarse layouts
    (300, 2, 200, "diverse"),      # 3,000 rows, ratio 2: three coarse layouts
    (200, 3, 300, "random"),
])
def test_native_recursion_equals_the_composed_steps(n_per, ratio, base, method):
    x = blobs(n_per, 12, 10, 21)
    g = fit_graph(x)
    cfg = R.params(ratio, base, 20, method, 4096)
    a, b = G.fit_curve(0.1, 1.0)
    print(f"[gpu] recursive n={len(x)} ratio={ratio} base={base} {method}", flush=True)
    native, stats = R.recursive_init(g["offsets"], g["indices"], g["weights"], x, 2, a, b, 77,
                                     cfg, return_stats=True)
    composed = python_run(g, x, 2, a, b, cfg, 77, method)
    assert stats["depth"] >= 2 and stats["layout_ms"] > 0
    assert stats["rows"][0] == len(x)
    assert np.isfinite(native).all()
    assert (native == composed).all()
    again = R.recursive_init(g["offsets"], g["indices"], g["weights"], x, 2, a, b, 77, cfg)
    assert (again == native).all(), "the recursion must be deterministic"


def test_recursive_auto_matches_multilevel_adam_fit():
    from cuscan_amd import UMAP

    x = np.random.default_rng(11).normal(size=(64, 3)).astype(np.float32)
    options = dict(optimizer="adam", n_epochs=5, random_state=11,
                   recursive_init_kwds={"base_init_threshold": 8, "n_embedding_epochs": 2})
    recursive = UMAP(init="recursive", **options).fit_transform(x)
    auto = UMAP(init="auto", **options).fit_transform(x)
    assert recursive.shape == (64, 2)
    assert np.isfinite(recursive).all()
    np.testing.assert_array_equal(recursive, auto)


@pytest.mark.parametrize("optimizer", ["adam", "momentum", "compatibility"])
def test_recursive_fit_quality(optimizer):
    """Default recursive settings on 8,000 rows (three levels): the fit is as
    trustworthy as the random-init fit, and the multilevel initialisation is
    already a good embedding by itself."""
    from cuscan_amd import UMAP

    x = blobs(800, 16, 10, 5)
    print(f"[gpu] recursive fit quality {optimizer}", flush=True)
    rec = UMAP(optimizer=optimizer, n_epochs=100, random_state=4)
    emb = rec.fit_transform(x)
    rnd = UMAP(optimizer=optimizer, init="random", n_epochs=100, random_state=4).fit_transform(x)
    t_rec, _ = trustworthiness(x, emb, k=15)
    t_rnd, _ = trustworthiness(x, rnd, k=15)
    print(f"trustworthiness recursive={t_rec:.4f} random={t_rnd:.4f} "
          f"levels={rec.recursive_stats_['rows']}")
    assert rec.recursive_stats_["depth"] >= 2
    assert t_rec >= t_rnd - 0.01
    assert t_rec > 0.9


def test_graph_edge_rescale_sets_the_median_edge_length():
    """init_rescale='graph_edge': the weight-median edge length of the initial
    layout over the pruned fit graph equals init_scale_factor."""
    from cuscan_amd import UMAP

    x = blobs(100, 6, 5, 8)
    model = UMAP(init="random", init_rescale="graph_edge", init_scale_factor=1.25, n_epochs=1,
                 learning_rate=0.0, random_state=2)
    model._validate()
    g = fit_graph(x, epochs=500)
    a, b = G.fit_curve(0.1, 1.0)
    init = model._initialize(x, g, len(x), 2, a, b, 2)
    heads = np.repeat(np.arange(len(x)), np.diff(g["offsets"].astype(np.int64)))
    lengths = np.linalg.norm(init[heads] - init[g["indices"]], axis=1)
    order = np.argsort(lengths, kind="stable")
    cum = np.cumsum(g["weights"][order].astype(np.float64))
    median = lengths[order][np.searchsorted(cum, cum[-1] / 2.0)]
    assert abs(median - 1.25) < 1e-5
    assert np.allclose(init.mean(0), 0.0, atol=1e-4)
