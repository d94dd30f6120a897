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
    with pytest.raises(NotImplementedError, match="init='random'"):
        UMAP(init="spectral").fit(blobs(4))


def test_defaults_match_reference_config():
    u = UMAP()
    assert (u.n_components, u.n_neighbors, u.min_dist, u.spread) == (2, 15, 0.1, 1.0)
    assert (u.optimizer, u.negative_selection_range, u.negative_sample_rate) == ("adam", 200_000, 5)
    assert (u.negative_sample_scale_adaptation_samples, u.transform_seed) == (128, 42)
    assert u.exclude_graph_neighbors is False and u.negative_sample_scale is None
