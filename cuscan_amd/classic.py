"""Compatibility (classic) SGD layout: cuscan's ``optimizer="compatibility"``.

``optimize_classic`` mirrors crates/umap/src/standard_umap/estimator.rs
``optimize`` (:2868-2982) with its schedule (``prepare_schedule``
:2753-2798), head CSR (:3080-3131) and generator seeding (:2800-2829); every
epoch runs the HIP clock and vertex kernels of the library (the atomic
vertex pass of crates/umap/src/kernels/standard_umap/layout.rs:68-336).
Coordinates are produced by concurrent atomic updates, so they are not
unique; the clocks and the per-row generator states are.
"""
from __future__ import annotations

import numpy as np

from . import _lib


def _p(a: np.ndarray | None):
    return None if a is None else a.ctypes.data


def optimize_classic(head, tail, heads, tails, weights, *, n_epochs, default_epochs, a, b,
                     learning_rate=1.0, gamma=1.0, negative_sample_rate=5, seed=42,
                     shared=True, return_state=False):
    """Run every epoch on the device from host arrays (the fuzzy graph's COO).

    Returns the embedding, or (embedding, next_positive, next_negative,
    rng_state) when ``return_state``.
    """
    head = np.array(head, dtype=np.float32, order="C", copy=True)
    h, d = head.shape
    tail_arr = None
    n = h
    if not shared:
        tail_arr = np.ascontiguousarray(tail, np.float32)
        n = tail_arr.shape[0]
        if tail_arr.shape[1] != d:
            raise ValueError("tail embedding dimension differs from head")
    heads = np.ascontiguousarray(heads, np.int32)
    tails = np.ascontiguousarray(tails, np.int32)
    weights = np.ascontiguousarray(weights, np.float32)
    E = len(heads)
    if len(tails) != E or len(weights) != E:
        raise ValueError("layout edge arrays have different lengths")
    npos = np.zeros(max(E, 1), np.float64)
    nneg = np.zeros(max(E, 1), np.float64)
    rng = np.zeros(3 * h, np.int64)
    _lib.check(_lib.lib().umaphip_classic_optimize(
        h, n, d, _p(head), _p(tail_arr), _p(heads), _p(tails), _p(weights), E, int(n_epochs),
        int(default_epochs), float(a), float(b), float(learning_rate), float(gamma),
        int(negative_sample_rate), int(seed) & ((1 << 64) - 1), int(bool(shared)), _p(npos),
        _p(nneg), _p(rng)))
    if return_state:
        return head, npos[:E], nneg[:E], rng.reshape(h, 3)
    return head
