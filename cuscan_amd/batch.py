"""Batched layouts: many independent modern Adam fits in one persistent
cooperative launch that runs every epoch (grid barrier between epochs), or in
one launch per epoch with the epoch loop captured once as a hipGraph, with a
tile-shape autotuner.

Each graph follows optimize_compact (crates/umap/src/standard_umap/
modern_runtime.rs:1011-1378) in global negative mode with its own seed;
small graphs (<= 1024 rows) keep the small-layout semantics.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .layout import LayoutParams

SHAPES = ("64 rows x 256 lanes", "32 rows x 128 lanes", "16 rows x 64 lanes")


class LayoutBatch:
    def __init__(self, graphs, heads, params: LayoutParams, seeds, persistent: bool = False):
        """graphs: compact graphs (dicts as from graph.compact_graph);
        heads: initial [rows, 2] embeddings; seeds: one layout seed each;
        persistent: one cooperative launch for all epochs (else one launch
        per epoch)."""
        if not (len(graphs) == len(heads) == len(seeds)) or not graphs:
            raise ValueError("graphs, heads and seeds must be non-empty and equally long")
        rows = np.array([len(g["offsets"]) - 1 for g in graphs], np.int64)
        edges = np.array([len(g["indices"]) for g in graphs], np.int64)
        for h, r in zip(heads, rows):
            if np.shape(h) != (r, 2):
                raise ValueError("each head must be [rows, 2]")
        cat = lambda key, dt: np.ascontiguousarray(np.concatenate([np.asarray(g[key], dt) for g in graphs]))
        self._arrays = dict(
            rows=rows, edges=edges, offsets=cat("offsets", np.uint32), indices=cat("indices", np.uint32),
            periods=cat("periods", np.float64), npos=cat("next_positive", np.float64),
            nper=cat("negative_periods", np.float64), nneg=cat("next_negative", np.float64),
            head=np.ascontiguousarray(np.concatenate([np.asarray(h, np.float32) for h in heads])),
            seeds=np.array([int(s) & ((1 << 64) - 1) for s in seeds], np.uint64))
        a = self._arrays
        self.rows, self.edges = rows, edges
        p = params.to_c()
        handle = C.c_void_p()
        _lib.check(_lib.lib().umaphip_batch_create(
            len(graphs), a["rows"].ctypes.data, a["edges"].ctypes.data, a["offsets"].ctypes.data,
            a["indices"].ctypes.data, a["periods"].ctypes.data, a["npos"].ctypes.data,
            a["nper"].ctypes.data, a["nneg"].ctypes.data, a["head"].ctypes.data, 2, C.byref(p),
            a["seeds"].ctypes.data, C.byref(handle)))
        self._h = handle
        self.shape = 0
        self.set_persistent(persistent)

    def set_persistent(self, persistent: bool) -> None:
        """Launch form of later runs and tuning (same results either way)."""
        _lib.check(_lib.lib().umaphip_batch_set_persistent(self._h, int(bool(persistent))))
        self.persistent = bool(persistent)

    def set_clock_schedule(self, enabled: bool) -> None:
        """Clock schedule on (the default when the batch allows it) or off
        (same results either way)."""
        _lib.check(_lib.lib().umaphip_batch_set_clock_schedule(self._h, int(bool(enabled))))

    @property
    def clock_schedule(self) -> int:
        """Epochs the clock schedule covers (0: off)."""
        return int(_lib.lib().umaphip_batch_clock_schedule(self._h))

    def tune(self, probe_epochs: int = 20):
        """Times each tile shape on the first epochs; keeps the fastest.
        Returns (shape index, [ms per shape])."""
        shape = C.c_int32()
        ms = (C.c_double * 3)()
        _lib.check(_lib.lib().umaphip_batch_tune(self._h, int(probe_epochs), C.byref(shape), ms))
        self.shape = shape.value
        return self.shape, list(ms)

    def set_shape(self, shape: int) -> None:
        _lib.check(_lib.lib().umaphip_batch_set_shape(self._h, int(shape)))
        self.shape = int(shape)

    def run(self, use_graph: bool = True) -> None:
        _lib.check(_lib.lib().umaphip_batch_run(self._h, int(bool(use_graph))))

    def download(self, clocks: bool = False):
        """(list of per-graph embeddings, device ms[, next_positive, next_negative])."""
        head = np.empty((int(self.rows.sum()), 2), np.float32)
        E = int(self.edges.sum())
        npos = np.empty(max(E, 1), np.float64) if clocks else None
        nneg = np.empty(max(E, 1), np.float64) if clocks else None
        ms = C.c_double()
        _lib.check(_lib.lib().umaphip_batch_download(
            self._h, head.ctypes.data, None if npos is None else npos.ctypes.data,
            None if nneg is None else nneg.ctypes.data, C.byref(ms)))
        parts = np.split(head, np.cumsum(self.rows)[:-1])
        if clocks:
            return parts, ms.value, npos[:E], nneg[:E]
        return parts, ms.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.lib().umaphip_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        # at interpreter shutdown the binding may already be torn down
        try:
            self.close()
        except (TypeError, AttributeError):
            pass
