"""Device-pointer front of the C ABI, with torch tensors as HBM plumbing.

These wrappers mirror cuscan's operator API ``StandardUmapOps``
(crates/umap/src/backend.rs:11-262) and the projection / adaptation
launchers of crates/umap/src/kernels/standard_umap/modern_layout.rs. Tensors
must live on a HIP device; work is enqueued on torch's current stream.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _p(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("tensor must be on the HIP device")
    return t.data_ptr()


def _need(t: torch.Tensor, dtype, name: str, contiguous: bool = True):
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
    if contiguous and not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


def epoch_args(*, head, fixed_tail, offsets, indices, source_order, negative_order,
               source_ranks, periods, next_positive, negative_periods, next_negative, updates,
               m, v, epoch, a, b, gamma, scale, negative_range, alpha, mu, correction1,
               correction2, momentum, shared, negatives, exclude_graph_neighbors):
    for name, t in (("offsets", offsets), ("indices", indices), ("source_order", source_order),
                    ("negative_order", negative_order)):
        _need(t, torch.int32, name)
    for name, t in (("periods", periods), ("next_positive", next_positive),
                    ("negative_periods", negative_periods), ("next_negative", next_negative)):
        _need(t, torch.float64, name)
    for name, t in (("head", head), ("updates", updates)):
        _need(t, torch.float32, name, contiguous=False)
    A = _lib.EpochArgs()
    A.h, A.d = head.shape
    A.n = negative_order.shape[0]
    A.head, (A.head_s0, A.head_s1) = _p(head), head.stride()
    tail = head if fixed_tail is None else fixed_tail
    A.fixed_tail, (A.tail_s0, A.tail_s1) = _p(tail), tail.stride()
    A.offsets, A.indices = _p(offsets), _p(indices)
    A.source_order, A.negative_order = _p(source_order), _p(negative_order)
    A.source_ranks = _p(source_ranks)
    A.periods, A.next_positive = _p(periods), _p(next_positive)
    A.negative_periods, A.next_negative = _p(negative_periods), _p(next_negative)
    A.updates, (A.upd_s0, A.upd_s1) = _p(updates), updates.stride()
    if m is not None:
        A.m, (A.m_s0, A.m_s1) = _p(m), m.stride()
    if v is not None:
        A.v, (A.v_s0, A.v_s1) = _p(v), v.stride()
    A.epoch = int(epoch)
    A.a, A.b, A.gamma, A.scale = float(a), float(b), float(gamma), float(scale)
    A.negative_range = int(negative_range)
    A.alpha, A.mu = float(alpha), float(mu)
    A.correction1, A.correction2 = float(correction1), float(correction2)
    A.momentum, A.shared = int(bool(momentum)), int(bool(shared))
    A.negatives, A.exclude_graph_neighbors = int(bool(negatives)), int(bool(exclude_graph_neighbors))
    A.n_edges = indices.shape[0]
    return A


def modern_epoch(**kw) -> None:
    """One epoch in place (StandardUmapOps::modern_epoch, backend.rs:172-232)."""
    A = epoch_args(**kw)
    _lib.check(_lib.lib().umaphip_modern_epoch(C.byref(A), _stream()))


def debug_negatives(cap: int = 64, **kw):
    """Per-edge negative candidates visited at this epoch (test hook)."""
    A = epoch_args(**kw)
    E = A.n_edges
    dev = kw["head"].device
    count = torch.empty(E, dtype=torch.int32, device=dev)
    cand = torch.full((E, cap), -7, dtype=torch.int32, device=dev)
    kept = torch.zeros((E, cap), dtype=torch.uint8, device=dev)
    _lib.check(_lib.lib().umaphip_debug_negatives(C.byref(A), cap, _p(count), _p(cand), _p(kept),
                                                  _stream()))
    return count, cand, kept


def project(x: torch.Tensor, direction: torch.Tensor) -> torch.Tensor:
    _need(x, torch.float32, "x", contiguous=False)
    out = torch.empty(x.shape[0], dtype=torch.float32, device=x.device)
    _lib.check(_lib.lib().umaphip_project(_p(x), x.shape[0], x.shape[1], x.stride(0), x.stride(1),
                                          _p(direction.contiguous()), _p(out), _stream()))
    return out


def stable_argsort(values: torch.Tensor):
    """(order, sorted, lowest non-finite row or -1)."""
    _need(values, torch.float32, "values")
    n = values.shape[0]
    nbytes = C.c_size_t()
    _lib.check(_lib.lib().umaphip_argsort_workspace_bytes(n, C.byref(nbytes)))
    ws = torch.empty(max(nbytes.value, 1), dtype=torch.uint8, device=values.device)
    order = torch.empty(n, dtype=torch.int32, device=values.device)
    out = torch.empty(n, dtype=torch.float32, device=values.device)
    status = torch.empty(1, dtype=torch.int64, device=values.device)
    _lib.check(_lib.lib().umaphip_stable_argsort_f32(_p(values), n, _p(order), _p(out),
                                                     _p(status), _p(ws), nbytes.value, _stream()))
    return order, out, status


def projection_ranks(order, sorted_values, query=None, *, shared: bool):
    n = order.shape[0]
    h = n if shared else query.shape[0]
    ranks = torch.empty(h, dtype=torch.int32, device=order.device)
    _lib.check(_lib.lib().umaphip_projection_ranks(_p(order), _p(sorted_values), n, _p(query), h,
                                                   _p(ranks), int(shared), _stream()))
    return ranks


def sample_resultants(embedding, order, offsets, indices, sample_ranks, local_offsets,
                      global_ranks, raw_range, a, b, gamma, exclude):
    s = sample_ranks.shape[0]
    out = torch.empty(2 * s, dtype=torch.float64, device=embedding.device)
    _lib.check(_lib.lib().umaphip_sample_resultants(
        _p(embedding.contiguous()), embedding.shape[0], embedding.shape[1], _p(order),
        _p(offsets), _p(indices), _p(sample_ranks), _p(local_offsets), _p(global_ranks), s,
        int(raw_range), float(a), float(b), float(gamma), int(bool(exclude)), _p(out),
        _stream()))
    return out


def classic_epoch(*, head, tail, tails, weights, epochs_per_sample, epoch_of_next_sample,
                  epochs_per_negative_sample, epoch_of_next_negative_sample, rng_state,
                  head_offsets, head_edges, mask, epoch, a, b, gamma, alpha,
                  negative_sample_rate, move_other, shared):
    """StandardUmapOps::optimize_epoch (crates/umap/src/backend.rs:234-262) on
    device tensors; head, clocks and rng_state are updated in place."""
    _need(head, torch.float32, "head")
    for name, t in (("tails", tails), ("head_offsets", head_offsets),
                    ("head_edges", head_edges), ("mask", mask)):
        _need(t, torch.int32, name)
    for name, t in (("epochs_per_sample", epochs_per_sample),
                    ("epoch_of_next_sample", epoch_of_next_sample),
                    ("epochs_per_negative_sample", epochs_per_negative_sample),
                    ("epoch_of_next_negative_sample", epoch_of_next_negative_sample)):
        _need(t, torch.float64, name)
    _need(weights, torch.float32, "weights")
    _need(rng_state, torch.int64, "rng_state")
    A = _lib.ClassicArgs()
    A.h, A.d = head.shape
    A.n = A.h if shared else tail.shape[0]
    A.n_edges = tails.numel()
    A.head = _p(head)
    A.tail = None if shared else _p(tail)
    A.tails, A.weights = _p(tails), _p(weights)
    A.epochs_per_sample = _p(epochs_per_sample)
    A.epoch_of_next_sample = _p(epoch_of_next_sample)
    A.epochs_per_negative_sample = _p(epochs_per_negative_sample)
    A.epoch_of_next_negative_sample = _p(epoch_of_next_negative_sample)
    A.rng_state = _p(rng_state)
    A.head_offsets, A.head_edges, A.mask = _p(head_offsets), _p(head_edges), _p(mask)
    A.epoch = int(epoch)
    A.a, A.b, A.gamma, A.alpha = float(a), float(b), float(gamma), float(alpha)
    A.negative_sample_rate = float(negative_sample_rate)
    A.move_other, A.shared = int(bool(move_other)), int(bool(shared))
    _lib.check(_lib.lib().umaphip_classic_epoch(C.byref(A), _stream()))


def cross_entropy(*, head, tail, offsets, indices, weights, a, b) -> float:
    """FP64 fuzzy-set cross entropy of the embedding over the positive edges
    (tail None for a fit); returns a Python float."""
    _need(head, torch.float32, "head")
    _need(offsets, torch.int32, "offsets")
    _need(indices, torch.int32, "indices")
    _need(weights, torch.float32, "weights")
    h, d = head.shape
    nbytes = C.c_size_t()
    _lib.check(_lib.lib().umaphip_cross_entropy_workspace_bytes(h, C.byref(nbytes)))
    ws = torch.empty(max(nbytes.value // 8, 1), dtype=torch.float64, device=head.device)
    out = torch.zeros(1, dtype=torch.float64, device=head.device)
    n = h if tail is None else tail.shape[0]
    _lib.check(_lib.lib().umaphip_cross_entropy(
        _p(head), h, _p(tail), n, d, _p(offsets), _p(indices), _p(weights), float(a), float(b),
        _p(ws), nbytes.value, _p(out), _stream()))
    return float(out.item())
