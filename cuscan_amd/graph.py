"""Host-array front of the graph and initialisation steps (C ABI, host C++)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def fuzzy_graph(indices, distances, *, local_connectivity=1.0, bandwidth=1.0, mix=1.0,
                bipartite=False, n_reference=None):
    """StandardUmapOps::fuzzy_graph (crates/umap/src/backend.rs:89-111).

    Returns dict(heads, tails, weights) trimmed to the used COO entries, plus
    rho, sigma and the directed memberships.
    """
    indices = np.ascontiguousarray(indices, np.int32)
    distances = np.ascontiguousarray(distances, np.float32)
    n, k = indices.shape
    nref = n if n_reference is None else int(n_reference)
    cap = 2 * n * k
    heads = np.empty(cap, np.int32)
    tails = np.empty(cap, np.int32)
    weights = np.empty(cap, np.float32)
    rho = np.empty(n, np.float32)
    sigma = np.empty(n, np.float32)
    directed = np.empty((n, k), np.float32)
    used = C.c_int64()
    _lib.check(_lib.lib().umaphip_fuzzy_graph(
        _p(indices), _p(distances), n, k, float(local_connectivity), float(bandwidth), float(mix),
        int(bool(bipartite)), nref, _p(heads), _p(tails), _p(weights), _p(rho), _p(sigma),
        _p(directed), C.byref(used)))
    u = used.value
    return dict(heads=heads[:u], tails=tails[:u], weights=weights[:u], rho=rho, sigma=sigma,
                directed=directed)


def compact_graph(rows, columns, heads, tails, weights, *, epochs, rate, prune_divisor=0.0,
                  transform=False):
    """Pruning + compact_graph: the layout CSR with f64 periods and clocks."""
    heads = np.ascontiguousarray(heads, np.int32)
    tails = np.ascontiguousarray(tails, np.int32)
    weights = np.ascontiguousarray(weights, np.float32)
    ne = len(weights)
    cap = max(ne, 1)
    offsets = np.empty(rows + 1, np.uint32)
    indices = np.empty(cap, np.uint32)
    w = np.empty(cap, np.float32)
    arrays = [np.empty(cap, np.float64) for _ in range(4)]
    out = C.c_int64()
    _lib.check(_lib.lib().umaphip_compact_graph(
        rows, columns, ne, _p(heads), _p(tails), _p(weights), float(prune_divisor), int(epochs),
        float(rate), int(bool(transform)), _p(offsets), _p(indices), _p(w), *map(_p, arrays),
        C.byref(out)))
    E = out.value
    return dict(offsets=offsets, indices=indices[:E].copy(), weights=w[:E].copy(),
                periods=arrays[0][:E].copy(), negative_periods=arrays[1][:E].copy(),
                next_positive=arrays[2][:E].copy(), next_negative=arrays[3][:E].copy())


def fit_curve(min_dist=0.1, spread=1.0):
    a, b = C.c_float(), C.c_float()
    _lib.check(_lib.lib().umaphip_fit_curve(float(min_dist), float(spread), C.byref(a),
                                            C.byref(b)))
    return a.value, b.value


def random_init(n, d, seed, *, repulsion=1.0, scale_factor=1.0, mode="isotropic"):
    out = np.empty((n, d), np.float32)
    m = {"isotropic": 0, "per_axis": 1}[mode]
    _lib.check(_lib.lib().umaphip_random_init(n, d, int(seed) & ((1 << 64) - 1), float(repulsion),
                                              float(scale_factor), m, _p(out)))
    return out


def transform_init(indices, directed, embedding):
    indices = np.ascontiguousarray(indices, np.int32)
    directed = np.ascontiguousarray(directed, np.float32)
    emb = np.ascontiguousarray(embedding, np.float32)
    q, k = indices.shape
    out = np.empty((q, emb.shape[1]), np.float32)
    _lib.check(_lib.lib().umaphip_transform_init(_p(indices), _p(directed), q, k, _p(emb),
                                                 emb.shape[0], emb.shape[1], _p(out)))
    return out
