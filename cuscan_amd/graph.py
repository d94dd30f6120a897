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
    m = {"isotropic": 0, "per_axis": 1, "raw": -1}[mode]
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


def transform_graph_device(indices, distances, embedding, *, local_connectivity=0.0, epochs,
                           rate):
    """The transform front-end on the device (bipartite memberships, initial
    placement, pruning at max / epochs, compact CSR with clocks) for one
    request. indices/distances: [q, k] from the kNN (host or device);
    embedding: the fitted [n, d] (host or device). Returns the compact graph
    dict (host arrays, as compact_graph) plus 'init' [q, d]."""
    import torch

    dev = torch.device("cuda:0")
    idx = torch.as_tensor(np.ascontiguousarray(indices, np.int32) if not isinstance(indices, torch.Tensor)
                          else indices, device=dev).to(torch.int32).contiguous()
    dist = torch.as_tensor(np.ascontiguousarray(distances, np.float32)
                           if not isinstance(distances, torch.Tensor) else distances,
                           device=dev).to(torch.float32).contiguous()
    emb = torch.as_tensor(np.ascontiguousarray(embedding, np.float32)
                          if not isinstance(embedding, torch.Tensor) else embedding,
                          device=dev).to(torch.float32).contiguous()
    q, k = idx.shape
    n, d = emb.shape
    cap = max(q * k, 1)
    offsets = torch.empty(q + 1, dtype=torch.int32, device=dev)
    out_idx = torch.empty(cap, dtype=torch.int32, device=dev)
    arrays = [torch.empty(cap, dtype=torch.float64, device=dev) for _ in range(4)]
    init = torch.empty((q, d), dtype=torch.float32, device=dev)
    n_edges = C.c_int64()
    stream = torch.cuda.current_stream(dev).cuda_stream
    _lib.check(_lib.lib().umaphip_transform_graph(
        C.c_void_p(idx.data_ptr()), C.c_void_p(dist.data_ptr()), q, k, C.c_void_p(emb.data_ptr()),
        n, d, float(local_connectivity), int(epochs), float(rate),
        C.c_void_p(offsets.data_ptr()), C.c_void_p(out_idx.data_ptr()),
        *[C.c_void_p(a.data_ptr()) for a in arrays], C.c_void_p(init.data_ptr()),
        C.byref(n_edges), C.c_void_p(stream)))
    E = n_edges.value
    host = lambda t: t.cpu().numpy()
    return dict(offsets=host(offsets).view(np.uint32), indices=host(out_idx[:E]).view(np.uint32),
                periods=host(arrays[0][:E]), negative_periods=host(arrays[1][:E]),
                next_positive=host(arrays[2][:E]), next_negative=host(arrays[3][:E]),
                init=host(init))


def fit_graph_device(indices, distances, *, local_connectivity=1.0, bandwidth=1.0, mix=1.0,
                     epochs, rate, prune_divisor):
    """The fit graph on the device: fuzzy union, pruning below max /
    prune_divisor and the compact CSR with clocks (as fuzzy_graph followed by
    compact_graph). indices/distances: [n, k] with self first (host or
    device). Returns the compact graph dict (host arrays, as compact_graph)."""
    import torch

    dev = torch.device("cuda:0")
    idx = (indices if isinstance(indices, torch.Tensor)
           else torch.as_tensor(np.ascontiguousarray(indices, np.int32))).to(dev, torch.int32).contiguous()
    dist = (distances if isinstance(distances, torch.Tensor)
            else torch.as_tensor(np.ascontiguousarray(distances, np.float32))).to(dev, torch.float32).contiguous()
    n, k = idx.shape
    cap = max(2 * n * k, 1)
    offsets = torch.empty(n + 1, dtype=torch.int32, device=dev)
    out_idx = torch.empty(cap, dtype=torch.int32, device=dev)
    w = torch.empty(cap, dtype=torch.float32, device=dev)
    arrays = [torch.empty(cap, dtype=torch.float64, device=dev) for _ in range(4)]
    n_edges = C.c_int64()
    stream = torch.cuda.current_stream(dev).cuda_stream
    _lib.check(_lib.lib().umaphip_fit_graph(
        C.c_void_p(idx.data_ptr()), C.c_void_p(dist.data_ptr()), n, k, float(local_connectivity),
        float(bandwidth), float(mix), int(epochs), float(rate), float(prune_divisor),
        C.c_void_p(offsets.data_ptr()), C.c_void_p(out_idx.data_ptr()), C.c_void_p(w.data_ptr()),
        *[C.c_void_p(a.data_ptr()) for a in arrays], C.byref(n_edges), C.c_void_p(stream)))
    E = n_edges.value
    host = lambda t: t.cpu().numpy()
    return dict(offsets=host(offsets).view(np.uint32), indices=host(out_idx[:E]).view(np.uint32),
                weights=host(w[:E]), periods=host(arrays[0][:E]),
                negative_periods=host(arrays[1][:E]), next_positive=host(arrays[2][:E]),
                next_negative=host(arrays[3][:E]))
