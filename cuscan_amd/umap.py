"""Drop-in ``UMAP`` estimator (python/cuscan/standard_umap.py:19-356).

Keyword arguments, validation messages, dtype rules and defaults follow the
reference's Python surface and ``UMAPConfig`` (crates/umap-core/src/
config.rs:37-66). Fit and transform run the HIP pipeline: exact kNN on the
device (BF16 MFMA prefilter + FP32 rerank), the fuzzy graph and compaction
on the device, the recursive initialisation (host coarsening, coarse
layouts on the device), and the modern layout (Adam or momentum) on the
device.

Scope of this build: ``init`` in {"recursive", "auto", "random"} (the
spectral solver is not provided); ``optimizer`` in {"adam",
"momentum", "compatibility"} (the last runs the classic SGD epoch on the
device); neighbour search is always exact, which equals the
reference with ``neighbor_search="exact"`` (its "approximate" NN-descent is
not part of this path); model serialisation is not provided.
"""
from __future__ import annotations

from typing import Literal

import numpy as np

DEFAULT_RANDOM_SEED = 42        # crates/umap/src/standard_umap/estimator.rs:28
AUTO_EXACT_MAX_ROWS = 4096      # estimator.rs:26 (informational: search is exact here)


def _output_dtype(X) -> np.dtype:
    dtype = getattr(X, "dtype", None)
    if dtype is not None and np.issubdtype(dtype, np.floating):
        return np.dtype(dtype)
    return np.dtype(np.float32)


def _is_int(v) -> bool:
    return isinstance(v, (int, np.integer)) and not isinstance(v, (bool, np.bool_))


def _distinct_rows(x):
    """distinct_training_rows (crates/umap/src/standard_umap/estimator.rs:
    38-77): the distinct rows in order of first occurrence, keyed on their
    f32 bit patterns with +0 and -0 equal, each kept as its first occurrence,
    and for every input row the index of its distinct row."""
    bits = np.ascontiguousarray(x, np.float32).view(np.uint32).copy()
    bits[x == 0.0] = 0
    _, first, inverse = np.unique(bits, axis=0, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty(len(first), np.int64)
    rank[order] = np.arange(len(first))
    rows = np.ascontiguousarray(x[first[order]])
    if rows.shape[0] < 2:
        raise ValueError("unique fitting requires at least two distinct rows")
    return rows, rank[inverse.reshape(-1)]


class UMAP:
    """Standard nonparametric UMAP estimator on a HIP device."""

    def __init__(
        self,
        n_components: int = 2,
        n_neighbors: int = 15,
        min_dist: float = 0.1,
        spread: float = 1.0,
        local_connectivity: float = 1.0,
        set_op_mix_ratio: float = 1.0,
        n_epochs: int | None = None,
        learning_rate: float = 1.0,
        negative_sample_rate: int = 5,
        repulsion_strength: float = 1.0,
        init: Literal["spectral", "random", "recursive", "auto"] = "auto",
        random_state: int | None = None,
        transform_seed: int = 42,
        transform_queue_size: float = 4.0,
        backend: Literal["auto", "cpu", "cuda", "hip"] = "auto",
        device_index: int = 0,
        neighbor_search: Literal["auto", "exact", "approximate"] = "auto",
        unique: bool = False,
        memory_limit_bytes: int | None = None,
        init_rescale: Literal["isotropic", "per_axis", "graph_edge"] = "isotropic",
        init_scale_factor: float = 1.0,
        optimizer: Literal["compatibility", "adam", "momentum"] = "adam",
        negative_selection_range: int = 200_000,
        negative_sample_scale: float | None = None,
        exclude_graph_neighbors: bool = False,
        negative_sample_scale_adaptation_samples: int = 128,
        recursive_coarsening_ratio: int = 4,
        recursive_init_kwds: dict | None = None,
    ) -> None:
        self.n_components = n_components
        self.n_neighbors = n_neighbors
        self.min_dist = min_dist
        self.spread = spread
        self.local_connectivity = local_connectivity
        self.set_op_mix_ratio = set_op_mix_ratio
        self.n_epochs = n_epochs
        self.learning_rate = learning_rate
        self.negative_sample_rate = negative_sample_rate
        self.repulsion_strength = repulsion_strength
        self.init = init
        self.random_state = random_state
        self.transform_seed = transform_seed
        self.transform_queue_size = transform_queue_size
        self.backend = backend
        self.device_index = device_index
        self.neighbor_search = neighbor_search
        self.unique = unique
        self.memory_limit_bytes = memory_limit_bytes
        self.init_rescale = init_rescale
        self.init_scale_factor = init_scale_factor
        self.optimizer = optimizer
        self.negative_selection_range = negative_selection_range
        self.negative_sample_scale = negative_sample_scale
        self.exclude_graph_neighbors = exclude_graph_neighbors
        self.negative_sample_scale_adaptation_samples = negative_sample_scale_adaptation_samples
        self.recursive_coarsening_ratio = recursive_coarsening_ratio
        self.recursive_init_kwds = recursive_init_kwds

    # ------------------------------------------------------------ validation
    @staticmethod
    def _input_array(X):
        try:
            array = np.asarray(X, dtype=np.float32)
        except (TypeError, ValueError) as error:
            raise ValueError("X must be a finite numeric 2-D array") from error
        if array.ndim != 2:
            raise ValueError("X must be a 2-D array")
        if not np.isfinite(array).all():
            raise ValueError("X must contain only finite values")
        return np.ascontiguousarray(array)

    def _validate(self):
        """Checks of python/cuscan/standard_umap.py:130-249 and UMAPConfig::validate
        (crates/umap-core/src/config.rs:70-148)."""
        if not isinstance(self.init, str) or self.init not in ("spectral", "random", "recursive", "auto"):
            raise ValueError("init must be 'spectral', 'random', 'recursive', or 'auto'")
        ratio = self.recursive_coarsening_ratio
        if not _is_int(ratio) or int(ratio) not in (2, 3, 4):
            raise ValueError("recursive_coarsening_ratio must be 2, 3, or 4")
        self._recursive = self._recursive_config()
        if not isinstance(self.optimizer, str) or self.optimizer not in ("compatibility", "adam", "momentum"):
            raise ValueError("optimizer must be 'compatibility', 'adam', or 'momentum'")
        q = self.negative_selection_range
        if not _is_int(q) or not 0 < int(q) <= 2**31 - 1:
            raise ValueError("negative_selection_range must be a positive int32 integer")
        scale = self.negative_sample_scale
        if scale is not None:
            if isinstance(scale, (bool, np.bool_)) or not isinstance(
                    scale, (int, float, np.integer, np.floating)):
                raise ValueError("negative_sample_scale must be a finite non-negative float or None")
            try:
                scale_float = float(scale)
            except (OverflowError, ValueError):
                raise ValueError("negative_sample_scale must be representable as float32") from None
            if (not np.isfinite(scale_float) or scale_float < 0
                    or scale_float > float(np.finfo(np.float32).max)
                    or (scale_float != 0 and np.float32(scale_float) == 0)):
                raise ValueError("negative_sample_scale must be a finite non-negative float32 value")
        if not isinstance(self.exclude_graph_neighbors, (bool, np.bool_)):
            raise ValueError("exclude_graph_neighbors must be a boolean")
        s = self.negative_sample_scale_adaptation_samples
        if not _is_int(s) or not 0 <= int(s) <= 2**31 - 1:
            raise ValueError("negative_sample_scale_adaptation_samples must be a non-negative int32 integer")
        if self.init_rescale not in ("isotropic", "per_axis", "graph_edge"):
            raise ValueError("init_rescale must be 'isotropic', 'per_axis', or 'graph_edge'")
        f = self.init_scale_factor
        if isinstance(f, (bool, np.bool_)) or not isinstance(f, (int, float, np.integer, np.floating)) \
                or not np.isfinite(f) or f <= 0.0:
            raise ValueError("init_scale_factor must be finite and greater than zero")
        if self.random_state is not None and (not _is_int(self.random_state) or self.random_state < 0):
            raise ValueError("random_state must be a non-negative integer or None")
        if self.n_epochs is not None and (not _is_int(self.n_epochs) or self.n_epochs <= 0):
            raise ValueError("n_epochs must be a positive integer or None")
        if self.memory_limit_bytes is not None and (
                not _is_int(self.memory_limit_bytes) or not 0 < int(self.memory_limit_bytes) <= 2**64 - 1):
            raise ValueError("memory_limit_bytes must be a positive uint64 integer or None")
        if not _is_int(self.n_components) or self.n_components <= 0:
            raise ValueError("n_components must be greater than zero")
        if not _is_int(self.n_neighbors) or self.n_neighbors < 2:
            raise ValueError("n_neighbors must be at least two")
        if not np.isfinite(self.min_dist) or self.min_dist < 0:
            raise ValueError("min_dist must be finite and nonnegative")
        if not np.isfinite(self.spread) or self.spread <= 0:
            raise ValueError("spread must be finite and greater than zero")
        if self.min_dist > self.spread:
            raise ValueError("min_dist must not exceed spread")
        if not np.isfinite(self.local_connectivity) or self.local_connectivity < 0:
            raise ValueError("local_connectivity must be finite and nonnegative")
        if not np.isfinite(self.set_op_mix_ratio) or not 0 <= self.set_op_mix_ratio <= 1:
            raise ValueError("set_op_mix_ratio must be between zero and one")
        if not np.isfinite(self.learning_rate) or self.learning_rate < 0:
            raise ValueError("learning_rate must be finite and nonnegative")
        if not np.isfinite(self.repulsion_strength) or self.repulsion_strength < 0:
            raise ValueError("repulsion_strength must be finite and nonnegative")
        if not _is_int(self.negative_sample_rate) or self.negative_sample_rate < 0:
            raise ValueError("negative_sample_rate must be a non-negative integer")
        if self.backend not in ("auto", "cpu", "cuda", "hip"):
            raise ValueError("backend must be 'auto', 'cpu', or 'cuda'")
        if self.backend == "cpu":
            raise ValueError("backend='cpu' is not available: this build runs on HIP devices")
        if self.neighbor_search not in ("auto", "exact", "approximate"):
            raise ValueError("neighbor_search must be 'auto', 'exact', or 'approximate'")
        if self.init == "spectral":
            raise NotImplementedError(
                "init='spectral' (the block Krylov spectral solver) is not part of this build; "
                "pass init='recursive', 'auto' or 'random'")

    def _recursive_config(self) -> dict:
        """recursive_init_kwds checks of python/cuscan/standard_umap.py:146-182."""
        supported = ("base_init_threshold", "n_embedding_epochs", "hub_selection",
                     "anchor_sample_size")
        kwds = self.recursive_init_kwds
        if kwds is not None and not isinstance(kwds, dict):
            raise ValueError("recursive_init_kwds must be a dict or None")
        kwds = {} if kwds is None else kwds
        if any(key not in supported for key in kwds):
            raise ValueError(f"recursive_init_kwds supports only: {', '.join(supported)}")
        config = {
            "base_init_threshold": kwds.get("base_init_threshold", 1024),
            "n_embedding_epochs": kwds.get("n_embedding_epochs", 50),
            "hub_selection": kwds.get("hub_selection", "degree"),
            "anchor_sample_size": kwds.get("anchor_sample_size", 16384),
        }
        for key, minimum, maximum in (("base_init_threshold", 2, 2**31 - 1),
                                      ("n_embedding_epochs", 1, np.iinfo(np.uintp).max),
                                      ("anchor_sample_size", 1, 2**31 - 1)):
            value = config[key]
            if not _is_int(value) or not minimum <= int(value) <= maximum:
                raise ValueError(f"{key} must be an integer in [{minimum}, {maximum}]")
            config[key] = int(value)
        if not isinstance(config["hub_selection"], str) or config["hub_selection"] not in (
                "degree", "diverse", "random"):
            raise ValueError("hub_selection must be 'degree', 'diverse', or 'random'")
        return config

    def _initialize(self, x, graph, n, d, a, b, seed):
        """Initial coordinates (estimator.rs:577-870): the recursive
        multilevel layout (init 'recursive'/'auto') or U(-10, 10) draws, then
        the configured rescale. graph is the pruned positive fit graph CSR."""
        from . import recursive as R

        kw = dict(repulsion=self.repulsion_strength, scale_factor=self.init_scale_factor,
                  offsets=graph["offsets"], indices=graph["indices"], weights=graph["weights"])
        if self.init in ("recursive", "auto"):
            cfg = R.params(int(self.recursive_coarsening_ratio), **self._recursive)
            values, self.recursive_stats_ = R.recursive_init(
                graph["offsets"], graph["indices"], graph["weights"], x, d, a, b, seed, cfg,
                return_stats=True)
            return R.init_rescale(values, self.init_rescale, **kw)
        from . import graph as G

        if self.init_rescale != "graph_edge":
            return G.random_init(n, d, seed, repulsion=self.repulsion_strength,
                                 scale_factor=self.init_scale_factor, mode=self.init_rescale)
        values = G.random_init(n, d, seed, mode="raw")
        return R.init_rescale(values, "graph_edge", **kw)

    def _device(self):
        import torch

        from . import _lib

        if _lib.lib().umaphip_device_count() < 1 or not torch.cuda.is_available():
            raise RuntimeError("no HIP device is available")
        torch.cuda.set_device(int(self.device_index))

    # ------------------------------------------------------------------- fit
    def fit(self, X) -> "UMAP":
        from . import graph as G
        from . import layout as L
        from .knn import KnnIndex

        x = self._input_array(X)
        self._validate()
        features = x.shape[1]
        # unique: fit the distinct rows; every input row takes its distinct
        # row's coordinates (FittedUMAP::embedding, estimator.rs:1320-1330)
        inverse = None
        if self.unique:
            x, inverse = _distinct_rows(x)
        self._device()
        n = x.shape[0]
        d = int(self.n_components)
        seed = DEFAULT_RANDOM_SEED if self.random_state is None else int(self.random_state)
        self.n_features_in_ = features
        self._train = x
        self._inverse = inverse
        if n == 1:
            self._set_embedding(np.zeros((1, d), np.float32))
            self._index, self._a, self._b = None, 0.0, 0.0
            return self
        self_inclusive = min(int(self.n_neighbors), n - 1)
        nonself = self_inclusive - 1
        if nonself == 0:
            raise ValueError("fit neighbor graph is empty for two samples")
        index = KnnIndex(x, max_queries=min(n, 65536))
        idx, dist = index.query(x, nonself, exclude_self=True)
        indices = np.concatenate([np.arange(n, dtype=np.int32)[:, None], idx], axis=1)
        distances = np.concatenate([np.zeros((n, 1), np.float32), dist], axis=1)
        default_epochs = 500 if n <= 10_000 else 200
        n_epochs = int(self.n_epochs) if self.n_epochs is not None else default_epochs
        divisor = float(n_epochs if n_epochs > 10 else default_epochs)
        a, b = G.fit_curve(self.min_dist, self.spread)
        # fuzzy union, pruning and the compact CSR on the device; the pruned
        # positive graph also feeds the recursive and graph_edge inits
        compact = G.fit_graph_device(indices, distances,
                                     local_connectivity=self.local_connectivity,
                                     bandwidth=1.0, mix=self.set_op_mix_ratio, epochs=n_epochs,
                                     rate=self.negative_sample_rate, prune_divisor=divisor)
        init = self._initialize(x, compact, n, d, a, b, seed)
        good = self.init in ("recursive", "auto")
        if self.optimizer == "compatibility":
            from .classic import optimize_classic

            g = G.fuzzy_graph(indices, distances, local_connectivity=self.local_connectivity,
                              bandwidth=1.0, mix=self.set_op_mix_ratio)
            emb = optimize_classic(init, None, g["heads"], g["tails"], g["weights"],
                                   n_epochs=n_epochs, default_epochs=default_epochs, a=a, b=b,
                                   learning_rate=self.learning_rate,
                                   gamma=self.repulsion_strength,
                                   negative_sample_rate=int(self.negative_sample_rate), seed=seed,
                                   shared=True)
        else:
            params = self._layout_params(n_epochs, a, b, self.learning_rate, seed, good=good)
            emb, _, _, self.layout_stats_ = L.optimize_compact(init, None, compact, params,
                                                               shared=True)
        self._set_embedding(emb)
        self._index, self._a, self._b = index, a, b
        return self

    def _set_embedding(self, emb):
        """The fitted rows' coordinates (the transform's reference
        embedding) and embedding_, expanded to every input row under
        unique=True."""
        self._fit_embedding = emb
        self._embedding_dev = None
        self.embedding_ = emb if self._inverse is None else emb[self._inverse]

    def _layout_params(self, epochs, a, b, alpha, seed, good=False):
        from .layout import LayoutParams

        return LayoutParams(
            epochs=epochs, a=a, b=b, optimizer=self.optimizer, repulsion=self.repulsion_strength,
            initial_alpha=alpha, negative_sample_rate=int(self.negative_sample_rate),
            negative_selection_range=int(self.negative_selection_range),
            negative_sample_scale=None if self.negative_sample_scale is None else float(self.negative_sample_scale),
            adaptation_samples=int(self.negative_sample_scale_adaptation_samples),
            exclude_graph_neighbors=bool(self.exclude_graph_neighbors), good_initialization=good,
            seed=seed)

    def fit_transform(self, X):
        dtype = _output_dtype(X)
        return self.fit(X).embedding_.astype(dtype, copy=False)

    # ------------------------------------------------------------- transform
    def transform(self, X):
        """FittedUMAP::transform (crates/umap/src/standard_umap/estimator.rs:
        1367-2264): query kNN, bipartite graph, weighted placement, then a
        transform layout against the immutable fitted embedding."""
        if not hasattr(self, "embedding_"):
            raise ValueError("This UMAP instance is not fitted yet")
        dtype = _output_dtype(X)
        x = self._input_array(X)
        if x.shape[1] != self.n_features_in_:
            raise ValueError(f"expected {self.n_features_in_} features, got {x.shape[1]}")
        return self._transform(x).astype(dtype, copy=False)

    def _transform(self, x):
        from . import graph as G
        from . import layout as L

        q = x.shape[0]
        n = self._fit_embedding.shape[0]
        if q == 0:
            return np.zeros((0, self._fit_embedding.shape[1]), np.float32)
        if self._index is None:
            return np.zeros((q, self._fit_embedding.shape[1]), np.float32)
        k = min(int(self.n_neighbors), n)
        idx, dist = self._index.query(x, k)
        epochs = int(self.n_epochs) // 3 if self.n_epochs is not None else (100 if q <= 10_000 else 30)
        tseed = int(self.transform_seed)
        lc = max(self.local_connectivity - 1.0, 0.0)
        if self.optimizer != "compatibility":
            # the device front-end (the transform server builds its graphs
            # with the same kernels): memberships, placement, pruned CSR
            compact = G.transform_graph_device(idx, dist, self._embedding_device(),
                                               local_connectivity=lc, epochs=max(epochs, 1),
                                               rate=self.negative_sample_rate)
            if epochs == 0:
                return compact["init"]
            params = self._layout_params(epochs, self._a, self._b, self.learning_rate / 4.0, tseed)
            emb, _, _, _ = L.optimize_compact(compact["init"], self._fit_embedding, compact, params,
                                              shared=False)
            return emb
        g = G.fuzzy_graph(idx, dist, local_connectivity=lc, bandwidth=1.0,
                          mix=self.set_op_mix_ratio, bipartite=True, n_reference=n)
        if len(g["weights"]) == 0 or not (g["weights"] > 0).any():
            raise ValueError("transform fuzzy graph is empty after positive-edge pruning")
        init = G.transform_init(idx, g["directed"], self._fit_embedding)
        if epochs == 0:
            return init
        from .classic import optimize_classic

        # compatibility transform: prepare_schedule(.., n_epochs, n_epochs, ..)
        # (estimator.rs:2158-2166), learning rate / 4, transform_seed
        return optimize_classic(init, self._fit_embedding, g["heads"], g["tails"], g["weights"],
                                n_epochs=epochs, default_epochs=epochs, a=self._a, b=self._b,
                                learning_rate=self.learning_rate / 4.0,
                                gamma=self.repulsion_strength,
                                negative_sample_rate=int(self.negative_sample_rate),
                                seed=tseed, shared=False)

    def _embedding_device(self):
        """The fitted embedding as a resident device tensor."""
        import torch

        cached = getattr(self, "_embedding_dev", None)
        if cached is None:
            cached = torch.as_tensor(np.ascontiguousarray(self._fit_embedding, np.float32),
                                     device=torch.device("cuda:0"))
            self._embedding_dev = cached
        return cached

    def save(self, path, overwrite: bool = False) -> None:
        raise NotImplementedError("model serialisation (HDF5 schema v5) is not part of this build")

    @classmethod
    def load(cls, path, backend="auto", device_index=0):
        raise NotImplementedError("model serialisation (HDF5 schema v5) is not part of this build")
