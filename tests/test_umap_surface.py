"""Validation surface of the UMAP estimator (python/tests/test_standard_umap.py
:45-62 and python/cuscan/standard_umap.py:130-249). No device needed: every
case is rejected before any device work."""
import numpy as np
import pytest

from cuscan_amd.umap import UMAP


def blobs(n=24):
    rng = np.random.default_rng(31)
    centers = np.array([[-7.0, 0, 0], [0, 7.0, 0], [7.0, 0, 0]], np.float32)
    return np.vstack([c + rng.normal(0, 0.22, (n, 3)) for c in centers]).astype(np.float32)


@pytest.mark.parametrize(("name", "value"), [
    ("optimizer", "unknown"), ("optimizer", True),
    ("negative_selection_range", True), ("negative_selection_range", 0),
    ("negative_selection_range", 2**31), ("negative_selection_range", 1.5),
    ("negative_sample_scale", True), ("negative_sample_scale", -1),
    ("negative_sample_scale", float("nan")), ("negative_sample_scale", float("inf")),
    ("negative_sample_scale", 1e100), ("negative_sample_scale", 1e-100),
    ("exclude_graph_neighbors", 1), ("negative_sample_scale_adaptation_samples", -1),
    ("negative_sample_scale_adaptation_samples", 2**31), ("init_scale_factor", 0.0),
    ("random_state", -1), ("n_epochs", 0), ("init", "pca"), ("init_rescale", "nope"),
    ("recursive_coarsening_ratio", 5), ("memory_limit_bytes", 0),
])
def test_invalid_settings_raise_value_error(name, value):
    kwargs = {"init": "random", name: value}
    with pytest.raises(ValueError, match=name):
        UMAP(**kwargs).fit(blobs(4))


@pytest.mark.parametrize("bad", [np.zeros((3,)), [[1.0, np.nan]], [["a", "b"]]])
def test_invalid_input_raises(bad):
    with pytest.raises(ValueError, match="X must"):
        UMAP(init="random").fit(bad)


def test_transform_before_fit_raises():
    with pytest.raises(ValueError, match="not fitted"):
        UMAP().transform(blobs(2))


def test_unsupported_inits_are_explicit():
    with pytest.raises(NotImplementedError, match="init='spectral'"):
        UMAP(init="spectral").fit(blobs(4))


def test_defaults_match_reference_config():
    u = UMAP()
    assert (u.n_components, u.n_neighbors, u.min_dist, u.spread) == (2, 15, 0.1, 1.0)
    assert (u.optimizer, u.negative_selection_range, u.negative_sample_rate) == ("adam", 200_000, 5)
    assert (u.negative_sample_scale_adaptation_samples, u.transform_seed) == (128, 42)
    assert u.exclude_graph_neighbors is False and u.negative_sample_scale is None


@pytest.mark.parametrize(("kwds", "match"), [
    ([("base_init_threshold", 8)], "recursive_init_kwds must be a dict"),
    ({"levels": 3}, "recursive_init_kwds supports only"),
    ({"base_init_threshold": 1}, "base_init_threshold"),
    ({"base_init_threshold": True}, "base_init_threshold"),
    ({"n_embedding_epochs": 0}, "n_embedding_epochs"),
    ({"anchor_sample_size": 2**31}, "anchor_sample_size"),
    ({"hub_selection": "pagerank"}, "hub_selection"),
])
def test_recursive_init_kwds_are_validated(kwds, match):
    """recursive_init_kwds checks of python/cuscan/standard_umap.py:146-182."""
    with pytest.raises(ValueError, match=match):
        UMAP(recursive_init_kwds=kwds).fit(blobs(4))


def test_auto_init_resolves_to_recursive():
    assert UMAP().init == "auto"
    u = UMAP(recursive_init_kwds={"base_init_threshold": 8, "hub_selection": "diverse"})
    u._validate()
    assert u._recursive == {"base_init_threshold": 8, "n_embedding_epochs": 50,
                            "hub_selection": "diverse", "anchor_sample_size": 16384}


def test_distinct_rows_first_occurrence_and_signed_zero():
    """distinct_training_rows (crates/umap/src/standard_umap/estimator.rs:
    38-77): first-occurrence order, +0 and -0 one key (the first occurrence
    kept), every row mapped to its distinct row."""
    from cuscan_amd.umap import _distinct_rows

    x = np.array([[1.0, -0.0], [2.0, 0.0], [1.0, 0.0], [3.0, 5.0], [2.0, -0.0], [3.0, 5.0]],
                 np.float32)
    rows, inverse = _distinct_rows(x)
    assert rows.tolist() == [[1.0, 0.0], [2.0, 0.0], [3.0, 5.0]]
    assert np.signbit(rows[0, 1]) and not np.signbit(rows[1, 1])
    assert inverse.tolist() == [0, 1, 0, 2, 1, 2]
    assert (rows[inverse] == x).all()
    # a plain brute-force restatement on random rows with many repeats
    rng = np.random.default_rng(3)
    y = rng.integers(0, 3, (500, 3)).astype(np.float32)
    y[rng.random(y.shape) < 0.2] *= -1.0
    rows, inverse = _distinct_rows(y)
    seen = {}
    for i, r in enumerate(y):
        key = tuple(0 if v == 0 else float(v) for v in r)
        seen.setdefault(key, len(seen))
        assert inverse[i] == seen[key]
    assert len(rows) == len(seen)


def test_unique_needs_two_distinct_rows():
    with pytest.raises(ValueError, match="two distinct rows"):
        UMAP(unique=True, init="random").fit(np.ones((5, 3), np.float32))
