"""Recursive (multilevel) initialisation, ``init="recursive"`` / ``"auto"``.

Thin wrappers over the native recursion (crates/umap-core/src/recursive_init/
mod.rs:48-320, graph.rs, geometry.rs; driver crates/umap/src/standard_umap/
estimator.rs:2506-2589). Graphs are square CSR dicts with ``offsets``
(int64 [n + 1]), ``indices`` (int32) and ``weights`` (float32); the summed
graphs of the aggregation carry float64 weights.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

HUB_SELECTION = {"degree": 0, "diverse": 1, "random": 2}
RESCALE = {"isotropic": 0, "per_axis": 1, "graph_edge": 2}
STAGES = ("anchor", "hubs", "propagation", "outliers", "aggregation", "transition",
          "prolongation", "base_case")


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def _csr(offsets, indices, weights, wdtype=np.float32):
    o = np.ascontiguousarray(offsets, np.int64)
    i = np.ascontiguousarray(indices, np.int32)
    w = np.ascontiguousarray(weights, wdtype)
    if len(i) == 0:                      # valid pointers for empty graphs
        i, w = np.zeros(1, np.int32), np.zeros(1, wdtype)
    return o, i, w


def params(coarsening_ratio=4, base_init_threshold=1024, n_embedding_epochs=50,
           hub_selection="degree", anchor_sample_size=16384) -> _lib.RecursiveParams:
    """RecursiveInitConfig (crates/umap-core/src/config.rs:382-402) defaults."""
    return _lib.RecursiveParams(int(coarsening_ratio), int(base_init_threshold),
                                int(n_embedding_epochs), HUB_SELECTION[hub_selection],
                                int(anchor_sample_size))


def recursive_init(offsets, indices, weights, data, d, a, b, seed, config=None,
                   return_stats=False):
    """recursive_init::run: coordinates [n, d] before the final rescale."""
    o, i, w = _csr(offsets, indices, weights)
    x = np.ascontiguousarray(data, np.float32)
    n, features = x.shape
    out = np.empty((n, d), np.float32)
    stats = _lib.RecursiveStats()
    cfg = config if config is not None else params()
    _lib.check(_lib.lib().umaphip_recursive_init(n, _p(o), _p(i), _p(w), _p(x), features, d,
                                                 float(a), float(b), C.byref(cfg),
                                                 int(seed) & ((1 << 64) - 1), _p(out),
                                                 C.byref(stats)))
    if not return_stats:
        return out
    depth = stats.depth
    return out, dict(depth=depth, rows=list(stats.rows[:depth]), edges=list(stats.edges[:depth]),
                     sweeps=list(stats.sweeps[:depth]), host_ms=stats.host_ms,
                     layout_ms=stats.layout_ms,
                     stage_ms=dict(zip(STAGES, (round(v, 2) for v in stats.stage_ms))))


def init_rescale(values, mode, *, repulsion=1.0, scale_factor=1.0, offsets=None, indices=None,
                 weights=None):
    """init_rescale::rescale in place on a float32 [n, d] array."""
    v = values
    assert v.dtype == np.float32 and v.flags.c_contiguous
    n, d = v.shape
    if offsets is None:
        o = np.zeros(n + 1, np.int64)
        i, w = np.zeros(1, np.int32), np.zeros(1, np.float32)
    else:
        o, i, w = _csr(offsets, indices, weights)
    _lib.check(_lib.lib().umaphip_init_rescale(_p(v), n, d, _p(o), _p(i), _p(w), RESCALE[mode],
                                               float(repulsion), float(scale_factor)))
    return v


def level_params(n, a=1.0, b=1.0, previous=None, config=None):
    """LevelParams::first (previous None) or previous.next(n); None at a base case."""
    cfg = config if config is not None else params()
    out = _lib.RecursiveLevel()
    has = C.c_int32()
    prev = C.byref(previous) if previous is not None else None
    _lib.check(_lib.lib().umaphip_recursive_level_params(int(n), prev, float(a), float(b),
                                                         C.byref(cfg), C.byref(out),
                                                         C.byref(has)))
    return out if has.value else None


def hubs(offsets, indices, weights, parts, method="degree", seed=0):
    o, i, w = _csr(offsets, indices, weights)
    n = len(o) - 1
    labels = np.empty(n, np.int32)
    _lib.check(_lib.lib().umaphip_recursive_hubs(n, _p(o), _p(i), _p(w), int(parts),
                                                 HUB_SELECTION[method], int(seed), _p(labels)))
    return labels


def propagate(offsets, indices, weights, labels, share):
    o, i, w = _csr(offsets, indices, weights)
    lab = np.array(labels, np.int32)
    sweeps = C.c_int32()
    _lib.check(_lib.lib().umaphip_recursive_propagate(len(o) - 1, _p(o), _p(i), _p(w),
                                                      int(bool(share)), _p(lab),
                                                      C.byref(sweeps)))
    return lab, sweeps.value


def outliers(offsets, indices, labels, seed):
    o, i, _ = _csr(offsets, indices, np.zeros(len(indices), np.float32))
    lab = np.array(labels, np.int32)
    _lib.check(_lib.lib().umaphip_recursive_outliers(len(o) - 1, _p(o), _p(i), int(seed),
                                                     _p(lab)))
    return lab


def remap(labels):
    lab = np.array(labels, np.int32)
    parts = C.c_int64()
    buf = lab if len(lab) else np.zeros(1, np.int32)
    _lib.check(_lib.lib().umaphip_recursive_remap(_p(buf), len(lab), C.byref(parts)))
    return lab, parts.value


def aggregate(offsets, indices, weights, labels, parts):
    o, i, w = _csr(offsets, indices, weights, np.float64)
    n = len(o) - 1
    lab = np.ascontiguousarray(labels, np.int32)
    cap = max(int(o[-1]), 1)
    oo = np.empty(parts + 1, np.int64)
    oi = np.empty(cap, np.int32)
    ow = np.empty(cap, np.float64)
    _lib.check(_lib.lib().umaphip_recursive_aggregate(n, _p(o), _p(i), _p(w), _p(lab), int(parts),
                                                      _p(oo), _p(oi), _p(ow)))
    E = int(oo[-1])
    return dict(offsets=oo, indices=oi[:E].copy(), weights=ow[:E].copy())


def transition(offsets, indices, weights):
    o, i, w = _csr(offsets, indices, weights, np.float64)
    n = len(o) - 1
    cap = max(2 * int(o[-1]), 1)
    oo = np.empty(n + 1, np.int64)
    oi = np.empty(cap, np.int32)
    ow = np.empty(cap, np.float32)
    _lib.check(_lib.lib().umaphip_recursive_transition(n, _p(o), _p(i), _p(w), _p(oo), _p(oi),
                                                       _p(ow)))
    E = int(oo[-1])
    return dict(offsets=oo, indices=oi[:E].copy(), weights=ow[:E].copy())


def anchor(data, d, sample_size, seed):
    x = np.ascontiguousarray(data, np.float32)
    n, features = x.shape
    values = np.empty((n, d), np.float32)
    mask = np.empty(n, np.uint8)
    _lib.check(_lib.lib().umaphip_recursive_anchor(_p(x), n, features, d, int(sample_size),
                                                   int(seed) & ((1 << 64) - 1), _p(values),
                                                   _p(mask)))
    return values, mask.astype(bool)


def base_case(offsets, indices, weights, anchor_values, mask, depth):
    o, i, w = _csr(offsets, indices, weights)
    n = len(o) - 1
    a = np.ascontiguousarray(anchor_values, np.float32)
    m = np.ascontiguousarray(mask, np.uint8)
    d = a.shape[1]
    out = np.empty((n, d), np.float32)
    _lib.check(_lib.lib().umaphip_recursive_base_case(n, _p(o), _p(i), _p(w), _p(a), _p(m), d,
                                                      int(depth), _p(out)))
    return out


def prolongate(offsets, indices, weights, labels, coarse):
    o, i, w = _csr(offsets, indices, weights)
    n = len(o) - 1
    lab = np.ascontiguousarray(labels, np.int32)
    c = np.ascontiguousarray(coarse, np.float32)
    parts, d = c.shape
    out = np.empty((n, d), np.float32)
    _lib.check(_lib.lib().umaphip_recursive_prolongate(n, _p(o), _p(i), _p(w), _p(lab), _p(c),
                                                       parts, d, _p(out)))
    return out


def procrustes(layout, anchor_values, mask):
    out = np.array(layout, np.float32, order="C")
    a = np.ascontiguousarray(anchor_values, np.float32)
    m = np.ascontiguousarray(mask, np.uint8)
    n, d = out.shape
    _lib.check(_lib.lib().umaphip_recursive_procrustes(_p(out), n, _p(a), _p(m), d))
    return out
