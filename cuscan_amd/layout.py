"""Python front of the modern layout (host arrays in/out, HIP underneath).

``optimize_compact`` mirrors cuscan's ``modern_runtime::optimize_compact``
(crates/umap/src/standard_umap/modern_runtime.rs:1011-1035): a compact CSR
graph with its f64 clocks, an initial head embedding and (for transform) a
fixed tail embedding go in; the optimised embedding and the advanced clocks
come out. ``LayoutPlan`` keeps everything resident in HBM so the same layout
can be re-run without re-uploading (benchmarks, serving).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib


@dataclass
class LayoutParams:
    """ModernLayoutParams (crates/umap-core/src/modern_layout.rs:9-70)."""

    epochs: int
    a: float
    b: float
    optimizer: str = "adam"
    repulsion: float = 1.0
    initial_alpha: float = 1.0
    negative_sample_rate: int = 5
    negative_selection_range: int = 200_000
    negative_sample_scale: float | None = None
    adaptation_samples: int = 128
    exclude_graph_neighbors: bool = False
    good_initialization: bool = False
    seed: int = 42

    def to_c(self) -> _lib.LayoutParams:
        if self.optimizer not in ("adam", "momentum"):
            raise ValueError("optimizer must be 'adam' or 'momentum' for the modern layout")
        p = _lib.LayoutParams()
        p.optimizer = 1 if self.optimizer == "momentum" else 0
        p.epochs = int(self.epochs)
        p.a, p.b = float(self.a), float(self.b)
        p.repulsion = float(self.repulsion)
        p.initial_alpha = float(self.initial_alpha)
        p.negative_sample_rate = int(self.negative_sample_rate)
        p.negative_selection_range = int(self.negative_selection_range)
        p.has_scale = 0 if self.negative_sample_scale is None else 1
        p.scale = 0.0 if self.negative_sample_scale is None else float(self.negative_sample_scale)
        p.adaptation_samples = int(self.adaptation_samples)
        p.exclude_graph_neighbors = int(bool(self.exclude_graph_neighbors))
        p.good_initialization = int(bool(self.good_initialization))
        p.seed = int(self.seed) & ((1 << 64) - 1)
        return p


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _graph_arrays(graph):
    return (np.ascontiguousarray(graph["offsets"], np.uint32),
            np.ascontiguousarray(graph["indices"], np.uint32),
            np.ascontiguousarray(graph["periods"], np.float64),
            np.ascontiguousarray(graph["negative_periods"], np.float64))


def optimize_compact(head, tail, graph, params: LayoutParams, *, shared: bool):
    """Run every epoch on the device. Returns (embedding, next_positive,
    next_negative, stats)."""
    head = np.array(head, dtype=np.float32, order="C", copy=True)
    h, d = head.shape
    tail_arr = None
    n = h
    if not shared:
        tail_arr = np.ascontiguousarray(tail, np.float32)
        n = tail_arr.shape[0]
        if tail_arr.shape[1] != d:
            raise ValueError("tail embedding dimension differs from head")
    offsets, indices, periods, negperiods = _graph_arrays(graph)
    npos = np.array(graph["next_positive"], np.float64, copy=True)
    nneg = np.array(graph["next_negative"], np.float64, copy=True)
    if len(offsets) != h + 1:
        raise ValueError("CSR offsets must have rows + 1 entries")
    stats = _lib.LayoutStats()
    p = params.to_c()
    _lib.check(_lib.lib().umaphip_optimize_compact(
        h, n, d, _ptr(head), None if tail_arr is None else _ptr(tail_arr), _ptr(offsets),
        _ptr(indices), len(indices), _ptr(periods), _ptr(npos), _ptr(negperiods), _ptr(nneg),
        C.byref(p), int(shared), C.byref(stats)))
    return head, npos, nneg, {f: getattr(stats, f) for f, _ in _lib.LayoutStats._fields_}


class LayoutPlan:
    """A device-resident layout that can be re-run from its initial state."""

    def __init__(self, head, tail, graph, params: LayoutParams, *, shared: bool):
        self._keep = []
        head = np.ascontiguousarray(head, np.float32)
        self.h, self.d = head.shape
        tail_arr = None
        self.n = self.h
        if not shared:
            tail_arr = np.ascontiguousarray(tail, np.float32)
            self.n = tail_arr.shape[0]
        offsets, indices, periods, negperiods = _graph_arrays(graph)
        npos = np.ascontiguousarray(graph["next_positive"], np.float64)
        nneg = np.ascontiguousarray(graph["next_negative"], np.float64)
        self.n_edges = len(indices)
        p = params.to_c()
        handle = C.c_void_p()
        _lib.check(_lib.lib().umaphip_plan_create(
            self.h, self.n, self.d, _ptr(head), None if tail_arr is None else _ptr(tail_arr),
            _ptr(offsets), _ptr(indices), self.n_edges, _ptr(periods), _ptr(npos),
            _ptr(negperiods), _ptr(nneg), C.byref(p), int(shared), C.byref(handle)))
        self._handle = handle

    def set_objective(self, weights, every: int) -> None:
        """Evaluate the FP64 cross entropy every `every` epochs (0 disables)."""
        w = None if weights is None else np.ascontiguousarray(weights, np.float32)
        if w is not None and len(w) != self.n_edges:
            raise ValueError("objective weights need one value per edge")
        self._keep.append(w)
        _lib.check(_lib.lib().umaphip_plan_set_objective(self._handle,
                                                         None if w is None else _ptr(w),
                                                         int(every)))

    def objective(self):
        """The last run's objective trace (FP64)."""
        count = C.c_int64()
        _lib.check(_lib.lib().umaphip_plan_objective(self._handle, None, 0, C.byref(count)))
        out = np.zeros(count.value, np.float64)
        if count.value:
            _lib.check(_lib.lib().umaphip_plan_objective(self._handle, _ptr(out), count.value,
                                                         C.byref(count)))
        return out

    def set_record_format(self, compact: bool) -> None:
        """24-byte (compact) or 32-byte edge records for the edge-balanced sweep."""
        _lib.check(_lib.lib().umaphip_plan_set_record_format(self._handle, int(bool(compact))))

    def set_clock_schedule(self, span: int = -1) -> None:
        """Clock schedule of a plan on the edge-balanced sweep: span < 0 automatic (the
        default), 0 off (the sweep advances the clocks from the records),
        span > 0 at most that many epochs per schedule pass."""
        _lib.check(_lib.lib().umaphip_plan_set_clock_schedule(self._handle, int(span)))

    def set_trace(self, epoch: int, negative_cap: int = 16) -> None:
        """Record one epoch of the edge-balanced sweep on later runs (epoch < 0: off)."""
        self._trace_cap = int(negative_cap)
        _lib.check(_lib.lib().umaphip_plan_set_trace(self._handle, int(epoch), int(negative_cap)))

    def trace(self) -> dict:
        """The last run's trace of the epoch chosen by ``set_trace``: the
        state the sweep started from, each row's assembled gradient and the
        negatives every firing edge drew (count -1: the edge did not fire)."""
        cap = self._trace_cap
        out = dict(head=np.zeros((self.h, 2), np.float32), order=np.zeros(self.n, np.int32),
                   ranks=np.zeros(self.h, np.int32), scale=np.zeros(1, np.float32),
                   gradient=np.zeros((self.h, 2), np.float32),
                   count=np.zeros(self.n_edges, np.int32),
                   candidates=np.zeros((self.n_edges, cap), np.int32),
                   kept=np.zeros((self.n_edges, cap), np.uint8))
        info = np.zeros(5, np.int32)
        _lib.check(_lib.lib().umaphip_plan_trace(
            self._handle, _ptr(out["head"]), _ptr(out["order"]), _ptr(out["ranks"]),
            _ptr(out["scale"]), _ptr(out["gradient"]), _ptr(out["count"]),
            _ptr(out["candidates"]), _ptr(out["kept"]), _ptr(info)))
        out["scale"] = float(out["scale"][0])
        out.update(rank_ordered=bool(info[0]), compact_records=bool(info[1]),
                   row_cap=int(info[2]), term_parallel_units=int(info[3]),
                   per_edge_units=int(info[4]))
        return out

    def set_profiling(self, enabled: bool = True) -> None:
        _lib.check(_lib.lib().umaphip_plan_set_profiling(self._handle, int(bool(enabled))))

    @property
    def stream(self) -> int:
        return _lib.lib().umaphip_plan_stream(self._handle)

    def run(self) -> None:
        _lib.check(_lib.lib().umaphip_plan_run(self._handle))

    def download(self, clocks: bool = False):
        head = np.zeros((self.h, self.d), np.float32)
        npos = np.zeros(self.n_edges, np.float64) if clocks else None
        nneg = np.zeros(self.n_edges, np.float64) if clocks else None
        stats = _lib.LayoutStats()
        _lib.check(_lib.lib().umaphip_plan_download(
            self._handle, _ptr(head), None if npos is None else _ptr(npos),
            None if nneg is None else _ptr(nneg), C.byref(stats)))
        out = {f: getattr(stats, f) for f, _ in _lib.LayoutStats._fields_}
        return (head, npos, nneg, out) if clocks else (head, out)

    def close(self) -> None:
        if getattr(self, "_handle", None):
            _lib.lib().umaphip_plan_destroy(self._handle)
            self._handle = None

    def __del__(self):
        # at interpreter shutdown the binding may already be torn down
        try:
            self.close()
        except (TypeError, AttributeError):
            pass


def orders_at_epoch(h, n, params: LayoutParams, *, shared: bool, epoch: int):
    """The (source, negative) orders in force at ``epoch`` (host RNG replay)."""
    n = h if shared else n
    src = np.zeros(h, np.int32)
    neg = np.zeros(n, np.int32)
    p = params.to_c()
    _lib.check(_lib.lib().umaphip_orders_at_epoch(h, n, C.byref(p), int(shared), int(epoch),
                                                  _ptr(src), _ptr(neg)))
    return src, neg
