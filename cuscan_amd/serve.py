"""Continuous-batching transform server over a resident fitted model.

``TransformServer.from_model(model)`` keeps the fitted references, the
embedding and the per-epoch projection tables of the embedding in HBM and
serves ``UMAP.transform`` (crates/umap/src/standard_umap/estimator.rs:
1367-2264) for many concurrent requests: requests are admitted while a
fixed device arena has room, their kNN runs in batches on its own stream,
and each batch's layouts (every epoch of every request) run in one launch.
Every request gets exactly the coordinates ``model.transform`` computes for
it alone.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class TransformServer:
    def __init__(self, reference, embedding, *, a, b, n_neighbors=15, epochs=100,
                 learning_rate=1.0, repulsion=1.0, local_connectivity=1.0, set_op_mix_ratio=1.0,
                 negative_sample_rate=5, negative_selection_range=200_000,
                 negative_sample_scale=None, adaptation_samples=128,
                 exclude_graph_neighbors=False, transform_seed=42, max_batch_rows=8192,
                 memory_bytes=0):
        ref = np.ascontiguousarray(reference, np.float32)
        emb = np.ascontiguousarray(embedding, np.float32)
        if ref.ndim != 2 or emb.ndim != 2 or len(ref) != len(emb):
            raise ValueError("reference and embedding must be 2-D with the same rows")
        self.features = ref.shape[1]
        self.d = emb.shape[1]
        p = _lib.ServerParams()
        p.n_neighbors, p.epochs = int(n_neighbors), int(epochs)
        p.a, p.b = float(a), float(b)
        p.learning_rate, p.repulsion = float(learning_rate), float(repulsion)
        p.local_connectivity, p.set_op_mix_ratio = float(local_connectivity), float(set_op_mix_ratio)
        p.negative_sample_rate = int(negative_sample_rate)
        p.negative_selection_range = int(negative_selection_range)
        p.has_scale = 0 if negative_sample_scale is None else 1
        p.scale = 0.0 if negative_sample_scale is None else float(negative_sample_scale)
        p.adaptation_samples = int(adaptation_samples)
        p.exclude_graph_neighbors = int(bool(exclude_graph_neighbors))
        p.transform_seed = int(transform_seed) & ((1 << 64) - 1)
        p.max_batch_rows, p.memory_bytes = int(max_batch_rows), int(memory_bytes)
        handle = C.c_void_p()
        _lib.check(_lib.lib().umaphip_server_create(ref.ctypes.data, ref.shape[0], ref.shape[1],
                                                    emb.ctypes.data, emb.shape[1], C.byref(p),
                                                    C.byref(handle)))
        self._h = handle

    @classmethod
    def from_model(cls, model, **kw):
        """Serve a fitted ``cuscan_amd.UMAP`` with its own transform settings."""
        if not hasattr(model, "embedding_"):
            raise ValueError("This UMAP instance is not fitted yet")
        epochs = int(model.n_epochs) // 3 if model.n_epochs is not None else 100
        return cls(model._train, model._fit_embedding, a=model._a, b=model._b,
                   n_neighbors=model.n_neighbors, epochs=epochs, learning_rate=model.learning_rate,
                   repulsion=model.repulsion_strength, local_connectivity=model.local_connectivity,
                   set_op_mix_ratio=model.set_op_mix_ratio,
                   negative_sample_rate=model.negative_sample_rate,
                   negative_selection_range=model.negative_selection_range,
                   negative_sample_scale=model.negative_sample_scale,
                   adaptation_samples=model.negative_sample_scale_adaptation_samples,
                   exclude_graph_neighbors=model.exclude_graph_neighbors,
                   transform_seed=model.transform_seed, **kw)

    def submit(self, x) -> tuple:
        x = np.ascontiguousarray(x, np.float32)
        if x.ndim != 2 or x.shape[1] != self.features:
            raise ValueError(f"expected rows with {self.features} features")
        t = C.c_int64()
        _lib.check(_lib.lib().umaphip_server_submit(self._h, x.ctypes.data, x.shape[0],
                                                    C.byref(t)))
        return (t.value, x.shape[0])

    def wait(self, ticket, timeout: float = 60.0):
        """(embedding, latency_ms) once done; None if the timeout passes."""
        tid, q = ticket
        out = np.empty((q, self.d), np.float32)
        lat = C.c_double()
        done = C.c_int32()
        _lib.check(_lib.lib().umaphip_server_wait(self._h, tid, float(timeout), out.ctypes.data,
                                                  C.byref(lat), C.byref(done)))
        return (out, lat.value) if done.value else None

    def poll(self, ticket):
        return self.wait(ticket, 0.0)

    def stats(self) -> dict:
        s = _lib.ServerCounters()
        _lib.check(_lib.lib().umaphip_server_stats(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in _lib.ServerCounters._fields_}

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.lib().umaphip_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        # at interpreter shutdown the binding may already be torn down
        try:
            self.close()
        except (TypeError, AttributeError):
            pass
