"""ctypes front-end of the CPU restatement (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, always as the checker or the CPU baseline, never as the product.
The library is built by ``make -C oracle`` into ``oracle/_build/liborc.so``.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "_build" / "liborc.so"
_lib = None

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


class Rng(C.Structure):
    _fields_ = [
        ("key", C.c_uint32 * 8),
        ("block", C.c_uint64),
        ("buf", C.c_uint32 * 64),
        ("idx", C.c_int),
    ]


class EpochArgs(C.Structure):
    _fields_ = [
        ("h", C.c_int64), ("n", C.c_int64), ("d", C.c_int64),
        ("head", C.c_void_p), ("head_s0", C.c_int64), ("head_s1", C.c_int64),
        ("tail", C.c_void_p), ("tail_s0", C.c_int64), ("tail_s1", C.c_int64),
        ("offsets", C.c_void_p), ("indices", C.c_void_p),
        ("source_order", C.c_void_p), ("negative_order", C.c_void_p),
        ("source_ranks", C.c_void_p),
        ("periods", C.c_void_p), ("next_positive", C.c_void_p),
        ("negative_periods", C.c_void_p), ("next_negative", C.c_void_p),
        ("updates", C.c_void_p), ("upd_s0", C.c_int64), ("upd_s1", C.c_int64),
        ("m", C.c_void_p), ("v", C.c_void_p),
        ("epoch", C.c_uint32),
        ("a", C.c_float), ("b", C.c_float), ("gamma", C.c_float), ("scale", C.c_float),
        ("negative_range", C.c_uint32),
        ("alpha", C.c_float), ("mu", C.c_float),
        ("correction1", C.c_float), ("correction2", C.c_float),
        ("momentum", C.c_int), ("shared", C.c_int), ("negatives", C.c_int),
        ("exclude", C.c_int),
    ]


class LayoutParams(C.Structure):
    _fields_ = [
        ("optimizer", C.c_int), ("epochs", C.c_int),
        ("a", C.c_float), ("b", C.c_float), ("repulsion", C.c_float),
        ("initial_alpha", C.c_float),
        ("negative_sample_rate", C.c_int64), ("negative_selection_range", C.c_int64),
        ("has_scale", C.c_int), ("scale", C.c_float),
        ("adaptation_samples", C.c_int64), ("exclude", C.c_int), ("good", C.c_int),
        ("seed", C.c_uint64), ("obj_every", C.c_int), ("obj_weights", C.c_void_p),
        ("obj_out", C.c_void_p),
    ]


class LayoutStats(C.Structure):
    _fields_ = [
        ("epochs_run", C.c_int64), ("firing_edges", C.c_int64),
        ("negative_samples", C.c_int64), ("adaptations", C.c_int64),
        ("refreshes", C.c_int64), ("final_scale", C.c_float),
    ]


class ScaleState(C.Structure):
    _fields_ = [("prior", C.c_double), ("current", C.c_double), ("anchor", C.c_double),
                ("has_anchor", C.c_int)]


def lib():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise RuntimeError(f"oracle library missing: run `make -C {_HERE}`")
        L = C.CDLL(str(_LIB_PATH))
        v = C.c_void_p
        L.orc_rng_seed.argtypes = [C.POINTER(Rng), C.c_uint64]
        L.orc_rng_u32.restype = C.c_uint32
        L.orc_rng_u32.argtypes = [C.POINTER(Rng)]
        L.orc_rng_u64.restype = C.c_uint64
        L.orc_rng_u64.argtypes = [C.POINTER(Rng)]
        L.orc_rng_f64.restype = C.c_double
        L.orc_rng_f64.argtypes = [C.POINTER(Rng)]
        L.orc_rng_range_u32.restype = C.c_uint32
        L.orc_rng_range_u32.argtypes = [C.POINTER(Rng), C.c_uint32]
        L.orc_rng_range_f32.restype = C.c_float
        L.orc_rng_range_f32.argtypes = [C.POINTER(Rng), C.c_float, C.c_float]
        L.orc_rng_shuffle_i32.argtypes = [C.POINTER(Rng), i32p, C.c_size_t]
        L.orc_gaussian_pair.argtypes = [C.POINTER(Rng), C.POINTER(C.c_double),
                                        C.POINTER(C.c_double)]
        L.orc_chacha_block.argtypes = [u32p, C.c_int, u32p]
        L.orc_adam_schedule.argtypes = [C.c_int, C.c_double, C.c_int, f64p, f64p]
        L.orc_momentum_schedule.argtypes = [C.c_int, C.c_double, C.c_int, f64p, f64p]
        L.orc_negative_index.restype = C.c_uint64
        L.orc_negative_index.argtypes = [C.c_uint64] * 4
        L.orc_effective_negative_range.restype = C.c_uint64
        L.orc_effective_negative_range.argtypes = [C.c_uint64] * 2
        L.orc_window_negative_index.restype = C.c_uint64
        L.orc_window_negative_index.argtypes = [C.c_uint64] * 6
        L.orc_adaptation_due.restype = C.c_int
        L.orc_adaptation_due.argtypes = [C.c_int, C.c_int] + [C.c_uint64] * 5
        L.orc_resolved_negative_scale.restype = C.c_float
        L.orc_resolved_negative_scale.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_float,
                                                  C.c_int]
        L.orc_compact_graph.restype = C.c_long
        L.orc_compact_graph.argtypes = [C.c_int64, C.c_int64, C.c_int64, i64p, i64p, f32p,
                                        C.c_int, C.c_float, u32p, u32p, f32p, f64p, f64p,
                                        f64p, f64p]
        L.orc_advance_clock.restype = C.c_long
        L.orc_advance_clock.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double),
                                        C.c_double, C.c_double, C.c_uint64]
        L.orc_accumulate_force.argtypes = [f32p, f32p, f32p, C.c_int, C.c_float, C.c_float,
                                           C.c_float, C.c_float, C.c_int, C.c_int]
        L.orc_gradient_block.argtypes = [C.POINTER(EpochArgs), C.c_int64, C.c_int64,
                                         C.c_float, C.c_int]
        L.orc_gradient_block64.argtypes = [C.POINTER(EpochArgs), C.c_int64, C.c_int64,
                                           C.c_double, f64p]
        L.orc_gradient_block64_mag.argtypes = [C.POINTER(EpochArgs), C.c_int64, C.c_int64,
                                               C.c_double, f64p, f64p]
        L.orc_adam_apply.argtypes = [C.POINTER(EpochArgs), C.c_int]
        L.orc_momentum_apply_block.argtypes = [C.POINTER(EpochArgs), C.c_int64, C.c_int64]
        L.orc_momentum_decay.argtypes = [C.POINTER(EpochArgs)]
        L.orc_modern_epoch.argtypes = [C.POINTER(EpochArgs), C.c_int]
        L.orc_trace_negatives.restype = C.c_int64
        L.orc_trace_negatives.argtypes = [C.POINTER(EpochArgs), C.c_int64, v, v, v,
                                          C.POINTER(C.c_int64)]
        L.orc_projection_direction.argtypes = [C.POINTER(Rng), C.c_int, f32p]
        L.orc_project.argtypes = [f32p, C.c_int64, C.c_int64, C.c_int64, C.c_int64, f32p,
                                  f32p]
        L.orc_radix_key_f32.restype = C.c_uint32
        L.orc_radix_key_f32.argtypes = [C.c_float]
        L.orc_stable_argsort_f32.restype = C.c_int64
        L.orc_stable_argsort_f32.argtypes = [f32p, C.c_int64, i32p, f32p]
        L.orc_inverse_ranks.argtypes = [i32p, C.c_int64, i32p]
        L.orc_lower_bound_ranks.argtypes = [f32p, C.c_int64, f32p, C.c_int64, i32p]
        L.orc_resultants.argtypes = [f32p, C.c_int64, C.c_int64, i32p, u32p, u32p, i32p, i32p,
                                     i32p, C.c_int64, C.c_uint32, C.c_float, C.c_float,
                                     C.c_float, C.c_int, f64p]
        L.orc_force_ratio.restype = C.c_double
        L.orc_force_ratio.argtypes = [f64p, f64p, C.c_int64]
        L.orc_scale_init.argtypes = [C.POINTER(ScaleState), C.c_float]
        L.orc_scale_observe.restype = C.c_float
        L.orc_scale_observe.argtypes = [C.POINTER(ScaleState), C.c_double]
        L.orc_optimize_compact.restype = C.c_int
        L.orc_optimize_compact.argtypes = [C.c_int64, C.c_int64, C.c_int64, f32p, v, u32p, u32p,
                                           f64p, f64p, f64p, f64p, C.POINTER(LayoutParams),
                                           C.c_int, C.c_int, C.c_int, C.POINTER(LayoutStats)]
        L.orc_exact_knn.argtypes = [f32p, C.c_int64, f32p, C.c_int64, C.c_int64, C.c_int64,
                                    C.c_int, C.c_int64, C.c_int, i32p, f32p]
        L.orc_fuzzy_graph.restype = C.c_long
        L.orc_fuzzy_graph.argtypes = [i32p, f32p, C.c_int64, C.c_int64, C.c_float, C.c_float,
                                      C.c_float, C.c_int, C.c_int64, i32p, i32p, f32p, f32p,
                                      f32p, f32p]
        L.orc_fit_curve.restype = C.c_int
        L.orc_fit_curve.argtypes = [C.c_float, C.c_float, C.POINTER(C.c_float),
                                    C.POINTER(C.c_float)]
        L.orc_random_init.restype = C.c_int
        L.orc_random_init.argtypes = [C.c_int64, C.c_int64, C.c_uint64, C.c_float, C.c_float,
                                      C.c_int, f32p]
        L.orc_transform_init.restype = C.c_int
        L.orc_transform_init.argtypes = [i32p, f32p, C.c_int64, C.c_int64, f32p, C.c_int64, f32p]
        L.orc_cross_entropy.restype = C.c_double
        L.orc_cross_entropy.argtypes = [f32p, C.c_int64, v, C.c_int64, u32p, u32p, f32p,
                                        C.c_double, C.c_double]
        L.orc_taus_next.restype = C.c_int64
        L.orc_taus_next.argtypes = [i64p]
        L.orc_classic_epoch.argtypes = [C.c_int64, C.c_int64, C.c_int64, f32p, v, i32p, f32p,
                                        C.c_int64, f64p, f64p, f64p, f64p, i64p, i32p, i32p,
                                        C.c_uint32, C.c_float, C.c_float, C.c_float, C.c_float,
                                        C.c_float, C.c_int]
        L.orc_classic_optimize.restype = C.c_int
        L.orc_classic_optimize.argtypes = [C.c_int64, C.c_int64, C.c_int64, f32p, v, i32p, i32p,
                                           f32p, C.c_int64, C.c_int32, C.c_int32, C.c_float,
                                           C.c_float, C.c_float, C.c_float, C.c_int32, C.c_uint64,
                                           C.c_int, f64p, f64p, i64p]
        _lib = L
    return _lib


def threads_default() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


# ------------------------------------------------------------------ RNG
class StdRng:
    """rand's StdRng restated (ChaCha12, seed_from_u64)."""

    def __init__(self, seed: int):
        self.state = Rng()
        lib().orc_rng_seed(C.byref(self.state), C.c_uint64(seed))

    def u32(self) -> int:
        return lib().orc_rng_u32(C.byref(self.state))

    def u64(self) -> int:
        return lib().orc_rng_u64(C.byref(self.state))

    def f64(self) -> float:
        return lib().orc_rng_f64(C.byref(self.state))

    def range_u32(self, bound: int) -> int:
        return lib().orc_rng_range_u32(C.byref(self.state), bound)

    def range_f32(self, lo: float, hi: float) -> float:
        return lib().orc_rng_range_f32(C.byref(self.state), lo, hi)

    def shuffle(self, a: np.ndarray) -> None:
        assert a.dtype == np.int32 and a.flags.c_contiguous
        lib().orc_rng_shuffle_i32(C.byref(self.state), a, a.size)

    def gaussian_pair(self) -> tuple[float, float]:
        x, y = C.c_double(), C.c_double()
        lib().orc_gaussian_pair(C.byref(self.state), C.byref(x), C.byref(y))
        return x.value, y.value

    def direction(self, d: int) -> np.ndarray:
        out = np.zeros(d, np.float32)
        lib().orc_projection_direction(C.byref(self.state), d, out)
        return out


def chacha_block(words: np.ndarray, rounds: int) -> np.ndarray:
    out = np.zeros(16, np.uint32)
    lib().orc_chacha_block(np.ascontiguousarray(words, np.uint32), rounds, out)
    return out


# ------------------------------------------------------- building blocks
def adam_schedule(epochs: int, alpha0: float, good: bool):
    a = np.zeros(epochs, np.float64)
    m = np.zeros(epochs, np.float64)
    lib().orc_adam_schedule(epochs, alpha0, int(good), a, m)
    return a, m


def momentum_schedule(epochs: int, alpha0: float, good: bool):
    a = np.zeros(epochs, np.float64)
    m = np.zeros(epochs, np.float64)
    lib().orc_momentum_schedule(epochs, alpha0, int(good), a, m)
    return a, m


def compact_graph(rows, columns, row, column, weight, epochs, rate):
    row = np.ascontiguousarray(row, np.int64)
    column = np.ascontiguousarray(column, np.int64)
    weight = np.ascontiguousarray(weight, np.float32)
    ne = len(weight)
    cap = max(ne, 1)
    offsets = np.zeros(rows + 1, np.uint32)
    indices = np.zeros(cap, np.uint32)
    weights = np.zeros(cap, np.float32)
    arrays = [np.zeros(cap, np.float64) for _ in range(4)]
    r = lib().orc_compact_graph(rows, columns, ne, row, column, weight, epochs, rate, offsets,
                                indices, weights, *arrays)
    if r < 0:
        raise ValueError(f"compact_graph failed ({r})")
    return dict(offsets=offsets, indices=indices[:r], weights=weights[:r],
                periods=arrays[0][:r].copy(), negative_periods=arrays[1][:r].copy(),
                next_positive=arrays[2][:r].copy(), next_negative=arrays[3][:r].copy())


def advance_clock(next_positive, next_negative, period, negative_period, epoch):
    p = C.c_double(next_positive)
    n = C.c_double(next_negative)
    r = lib().orc_advance_clock(C.byref(p), C.byref(n), period, negative_period, epoch)
    return (None if r < 0 else int(r)), p.value, n.value


class Epoch:
    """Holds numpy arrays for one orc_epoch_args call (keeps them alive)."""

    def __init__(self, *, head, tail, offsets, indices, source_order, negative_order,
                 source_ranks, periods, next_positive, negative_periods, next_negative,
                 updates, m, v, epoch, a, b, gamma, scale, negative_range, alpha=0.0, mu=0.0,
                 correction1=1.0, correction2=1.0, momentum=False, shared=True,
                 negatives=True, exclude=False):
        self.head = np.array(head, copy=True, order="C", dtype=np.float32)
        self.tail = np.ascontiguousarray(tail, np.float32)
        self.offsets = np.ascontiguousarray(offsets, np.uint32)
        self.indices = np.ascontiguousarray(indices, np.uint32)
        self.source_order = np.ascontiguousarray(source_order, np.int32)
        self.negative_order = np.ascontiguousarray(negative_order, np.int32)
        self.source_ranks = np.ascontiguousarray(source_ranks, np.int32)
        self.periods = np.ascontiguousarray(periods, np.float64)
        self.next_positive = np.array(next_positive, copy=True, order="C", dtype=np.float64)
        self.negative_periods = np.ascontiguousarray(negative_periods, np.float64)
        self.next_negative = np.array(next_negative, copy=True, order="C", dtype=np.float64)
        self.updates = np.array(updates, copy=True, order="C", dtype=np.float32)
        self.m = np.array(m, copy=True, order="C", dtype=np.float32)
        self.v = np.array(v, copy=True, order="C", dtype=np.float32)
        h, d = self.head.shape
        A = EpochArgs()
        A.h, A.n, A.d = h, len(self.negative_order), d
        A.head, A.head_s0, A.head_s1 = self.head.ctypes.data, d, 1
        A.tail, A.tail_s0, A.tail_s1 = self.tail.ctypes.data, self.tail.shape[1], 1
        A.offsets = self.offsets.ctypes.data
        A.indices = self.indices.ctypes.data
        A.source_order = self.source_order.ctypes.data
        A.negative_order = self.negative_order.ctypes.data
        A.source_ranks = self.source_ranks.ctypes.data
        A.periods = self.periods.ctypes.data
        A.next_positive = self.next_positive.ctypes.data
        A.negative_periods = self.negative_periods.ctypes.data
        A.next_negative = self.next_negative.ctypes.data
        A.updates, A.upd_s0, A.upd_s1 = self.updates.ctypes.data, d, 1
        A.m, A.v = self.m.ctypes.data, self.v.ctypes.data
        A.epoch = epoch
        A.a, A.b, A.gamma, A.scale = a, b, gamma, scale
        A.negative_range = negative_range
        A.alpha, A.mu, A.correction1, A.correction2 = alpha, mu, correction1, correction2
        A.momentum, A.shared = int(momentum), int(shared)
        A.negatives, A.exclude = int(negatives), int(exclude)
        self.args = A

    def gradient_block(self, start, end, multiplier, threads=1):
        lib().orc_gradient_block(C.byref(self.args), start, end, multiplier, threads)

    def gradient_block64(self, start, end, multiplier):
        g = np.zeros(self.head.shape, np.float64)
        lib().orc_gradient_block64(C.byref(self.args), start, end, multiplier, g)
        return g

    def gradient_block64_mag(self, start, end, multiplier):
        """(FP64 gradient, per-row sum of the terms' Euclidean norms)."""
        g = np.zeros(self.head.shape, np.float64)
        mag = np.zeros(self.head.shape[0], np.float64)
        lib().orc_gradient_block64_mag(C.byref(self.args), start, end, multiplier, g, mag)
        return g, mag

    def adam_apply(self, threads=1):
        lib().orc_adam_apply(C.byref(self.args), threads)

    def momentum_apply_block(self, start, end):
        lib().orc_momentum_apply_block(C.byref(self.args), start, end)

    def momentum_decay(self):
        lib().orc_momentum_decay(C.byref(self.args))

    def modern_epoch(self, threads=1):
        lib().orc_modern_epoch(C.byref(self.args), threads)

    def trace_negatives(self):
        fires = C.c_int64()
        total = lib().orc_trace_negatives(C.byref(self.args), 0, None, None, None,
                                          C.byref(fires))
        cap = max(int(total), 1)
        e = np.zeros(cap, np.int64)
        c = np.zeros(cap, np.int32)
        k = np.zeros(cap, np.uint8)
        lib().orc_trace_negatives(C.byref(self.args), cap, e.ctypes.data, c.ctypes.data,
                                  k.ctypes.data, C.byref(fires))
        return e[:total], c[:total], k[:total], int(fires.value)


# ----------------------------------------------------------- projection
def project(x, direction):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros(x.shape[0], np.float32)
    lib().orc_project(x, x.shape[0], x.shape[1], x.shape[1], 1,
                      np.ascontiguousarray(direction, np.float32), out)
    return out


def stable_argsort(values):
    values = np.ascontiguousarray(values, np.float32)
    order = np.zeros(len(values), np.int32)
    s = np.zeros(len(values), np.float32)
    bad = lib().orc_stable_argsort_f32(values, len(values), order, s)
    if bad >= 0:
        raise FloatingPointError(f"non-finite reference projection at row {bad}")
    return order, s


def radix_key(value: float) -> int:
    return lib().orc_radix_key_f32(value)


def inverse_ranks(order):
    ranks = np.zeros(len(order), np.int32)
    lib().orc_inverse_ranks(np.ascontiguousarray(order, np.int32), len(order), ranks)
    return ranks


def lower_bound_ranks(sorted_values, query):
    q = np.ascontiguousarray(query, np.float32)
    r = np.zeros(len(q), np.int32)
    lib().orc_lower_bound_ranks(np.ascontiguousarray(sorted_values, np.float32),
                                len(sorted_values), q, len(q), r)
    return r


def resultants(emb, order, offsets, indices, sample_ranks, local_offsets, global_ranks,
               raw_range, a, b, gamma, exclude):
    emb = np.ascontiguousarray(emb, np.float32)
    s = len(sample_ranks)
    out = np.zeros(2 * s, np.float64)
    lib().orc_resultants(emb, emb.shape[0], emb.shape[1], np.ascontiguousarray(order, np.int32),
                         np.ascontiguousarray(offsets, np.uint32),
                         np.ascontiguousarray(indices, np.uint32),
                         np.ascontiguousarray(sample_ranks, np.int32),
                         np.ascontiguousarray(local_offsets, np.int32),
                         np.ascontiguousarray(global_ranks, np.int32), s, raw_range, a, b,
                         gamma, int(exclude), out)
    return out


def force_ratio(global_, local):
    g = np.ascontiguousarray(global_, np.float64)
    l = np.ascontiguousarray(local, np.float64)
    return lib().orc_force_ratio(g, l, len(g))


class ScaleAdaptation:
    def __init__(self, scale: float):
        self.s = ScaleState()
        lib().orc_scale_init(C.byref(self.s), scale)

    def observe(self, ratio: float) -> float:
        return lib().orc_scale_observe(C.byref(self.s), ratio)

    @property
    def current(self) -> float:
        return self.s.current


# ---------------------------------------------------------------- driver
def optimize_compact(head, tail, graph, *, epochs, a, b, repulsion=1.0, initial_alpha=1.0,
                     optimizer="adam", negative_sample_rate=5, negative_selection_range=200_000,
                     negative_sample_scale=None, adaptation_samples=128, exclude=False,
                     good=False, seed=42, shared=True, threads=None, stop_epoch=0,
                     collect_stats=True, objective_weights=None, objective_every=0):
    """Run the restated epoch loop in place on copies; returns (head, clocks, stats).
    collect_stats replays each epoch's negatives once more to count them.
    With objective_every > 0, stats["objective"] holds the FP64 cross entropy
    after every objective_every-th epoch and after the last."""
    head = np.array(head, np.float32, order="C", copy=True)
    h, d = head.shape
    tail_arr = None if shared else np.ascontiguousarray(tail, np.float32)
    n = h if shared else tail_arr.shape[0]
    npos = np.array(graph["next_positive"], np.float64, copy=True)
    nneg = np.array(graph["next_negative"], np.float64, copy=True)
    P = LayoutParams()
    P.optimizer = 1 if optimizer == "momentum" else 0
    P.epochs = epochs
    P.a, P.b, P.repulsion, P.initial_alpha = a, b, repulsion, initial_alpha
    P.negative_sample_rate = negative_sample_rate
    P.negative_selection_range = negative_selection_range
    P.has_scale = 0 if negative_sample_scale is None else 1
    P.scale = 0.0 if negative_sample_scale is None else negative_sample_scale
    P.adaptation_samples = adaptation_samples
    P.exclude = int(exclude)
    P.good = int(good)
    P.seed = seed
    obj_w = obj_out = None
    if objective_every > 0:
        obj_w = np.ascontiguousarray(objective_weights, np.float32)
        obj_out = np.zeros(epochs // objective_every + 1, np.float64)
        P.obj_every = objective_every
        P.obj_weights = obj_w.ctypes.data
        P.obj_out = obj_out.ctypes.data
    S = LayoutStats()
    rc = lib().orc_optimize_compact(
        h, n, d, head, None if tail_arr is None else tail_arr.ctypes.data,
        np.ascontiguousarray(graph["offsets"], np.uint32),
        np.ascontiguousarray(graph["indices"], np.uint32),
        np.ascontiguousarray(graph["periods"], np.float64), npos,
        np.ascontiguousarray(graph["negative_periods"], np.float64), nneg, C.byref(P),
        int(shared), threads or threads_default(), stop_epoch,
        C.byref(S) if collect_stats else None)
    if rc != 0:
        raise RuntimeError(f"oracle layout failed ({rc})")
    stats = {f: getattr(S, f) for f, _ in LayoutStats._fields_}
    if obj_out is not None:
        last = min(stop_epoch, epochs) if stop_epoch > 0 else epochs
        stats["objective"] = obj_out[:last // objective_every + (last % objective_every != 0)]
    return head, npos, nneg, stats


# ---------------------------------------------------------------- graph
def exact_knn(query, reference, k, exclude_self=False, query_offset=0, threads=None):
    q = np.ascontiguousarray(query, np.float32)
    r = np.ascontiguousarray(reference, np.float32)
    idx = np.zeros((q.shape[0], k), np.int32)
    dist = np.zeros((q.shape[0], k), np.float32)
    lib().orc_exact_knn(q, q.shape[0], r, r.shape[0], q.shape[1], k, int(exclude_self),
                        query_offset, threads or threads_default(), idx, dist)
    return idx, dist


def fuzzy_graph(indices, distances, local_connectivity=1.0, bandwidth=1.0, mix=1.0,
                bipartite=False, n_reference=None):
    indices = np.ascontiguousarray(indices, np.int32)
    distances = np.ascontiguousarray(distances, np.float32)
    n, k = indices.shape
    nref = n if n_reference is None else n_reference
    cap = 2 * n * k
    heads = np.zeros(cap, np.int32)
    tails = np.zeros(cap, np.int32)
    weights = np.zeros(cap, np.float32)
    rho = np.zeros(n, np.float32)
    sigma = np.zeros(n, np.float32)
    directed = np.zeros((n, k), np.float32)
    used = lib().orc_fuzzy_graph(indices, distances, n, k, local_connectivity, bandwidth, mix,
                                 int(bipartite), nref, heads, tails, weights, rho, sigma,
                                 directed)
    return dict(heads=heads, tails=tails, weights=weights, rho=rho, sigma=sigma,
                directed=directed, used=int(used))


def fit_curve(min_dist=0.1, spread=1.0):
    a, b = C.c_float(), C.c_float()
    if lib().orc_fit_curve(min_dist, spread, C.byref(a), C.byref(b)) != 0:
        raise ValueError("curve fitting failed")
    return a.value, b.value


def random_init(n, d, seed, repulsion=1.0, factor=1.0, mode=0):
    out = np.zeros((n, d), np.float32)
    lib().orc_random_init(n, d, seed, repulsion, factor, mode, out)
    return out


def transform_init(indices, directed, embedding):
    indices = np.ascontiguousarray(indices, np.int32)
    directed = np.ascontiguousarray(directed, np.float32)
    emb = np.ascontiguousarray(embedding, np.float32)
    q, k = indices.shape
    out = np.zeros((q, emb.shape[1]), np.float32)
    if lib().orc_transform_init(indices, directed, q, k, emb, emb.shape[1], out) != 0:
        raise FloatingPointError("transform initialization row has no positive memberships")
    return out


def cross_entropy(head, tail, graph, weights, a, b):
    """FP64 fuzzy-set cross entropy over the graph's positive edges. a and b
    are the layout's f32 curve parameters, widened to f64 as the device does."""
    a, b = float(np.float32(a)), float(np.float32(b))
    head = np.ascontiguousarray(head, np.float32)
    tail_arr = None if tail is None else np.ascontiguousarray(tail, np.float32)
    return float(lib().orc_cross_entropy(
        head, head.shape[0], None if tail_arr is None else tail_arr.ctypes.data, head.shape[1],
        np.ascontiguousarray(graph["offsets"], np.uint32),
        np.ascontiguousarray(graph["indices"], np.uint32),
        np.ascontiguousarray(weights, np.float32), float(a), float(b)))


def taus_next(state):
    """Advance a 3-word Tausworthe state in place; returns the signed draw."""
    assert state.dtype == np.int64 and state.flags.c_contiguous and state.size == 3
    return int(lib().orc_taus_next(state))


def classic_epoch(head, tail, tails, weights, per, npos, nper, nneg, rng, head_offsets,
                  head_edges, *, epoch, a, b, gamma, alpha, rate, shared):
    """One serial compatibility epoch in place on the given arrays."""
    h, d = head.shape
    n = h if shared else tail.shape[0]
    tail_arr = None if shared else np.ascontiguousarray(tail, np.float32)
    lib().orc_classic_epoch(h, n, d, head, None if tail_arr is None else tail_arr.ctypes.data,
                            tails, weights, len(tails), per, npos, nper, nneg, rng, head_offsets,
                            head_edges, epoch, a, b, gamma, alpha, rate, int(shared))


def classic_optimize(head, tail, heads, tails, weights, *, n_epochs, default_epochs, a, b,
                     learning_rate=1.0, gamma=1.0, rate=5, seed=42, shared=True):
    """Whole serial compatibility layout. Returns (head, npos, nneg, rng)."""
    head = np.array(head, np.float32, order="C", copy=True)
    h, d = head.shape
    tail_arr = None if shared else np.ascontiguousarray(tail, np.float32)
    E = len(heads)
    npos = np.zeros(max(E, 1), np.float64)
    nneg = np.zeros(max(E, 1), np.float64)
    rng = np.zeros(3 * h, np.int64)
    lib().orc_classic_optimize(h, h if shared else tail_arr.shape[0], d, head,
                               None if tail_arr is None else tail_arr.ctypes.data,
                               np.ascontiguousarray(heads, np.int32),
                               np.ascontiguousarray(tails, np.int32),
                               np.ascontiguousarray(weights, np.float32), E, n_epochs,
                               default_epochs, a, b, learning_rate, gamma, rate, seed,
                               int(shared), npos, nneg, rng)
    return head, npos[:E], nneg[:E], rng.reshape(h, 3)


def fit(x, *, n_neighbors=15, n_components=2, min_dist=0.1, spread=1.0, n_epochs=None,
        learning_rate=1.0, negative_sample_rate=5, repulsion=1.0, optimizer="adam",
        negative_selection_range=200_000, exclude=False, seed=42, threads=None):
    """CPU reference fit (init="random"): exact kNN -> fuzzy graph -> prune ->
    compact -> random init + rescale -> modern layout. Restates the call
    chain of crates/umap/src/standard_umap/estimator.rs:117-1005."""
    x = np.ascontiguousarray(x, np.float32)
    n = x.shape[0]
    k = min(n_neighbors, n - 1)
    idx, dist = exact_knn(x, x, k - 1, exclude_self=True, threads=threads)
    indices = np.concatenate([np.arange(n, dtype=np.int32)[:, None], idx], 1)
    distances = np.concatenate([np.zeros((n, 1), np.float32), dist], 1)
    g = fuzzy_graph(indices, distances)
    u = g["used"]
    heads, tails, w = g["heads"][:u], g["tails"][:u], g["weights"][:u]
    default_epochs = 500 if n <= 10_000 else 200
    epochs = n_epochs or default_epochs
    divisor = epochs if epochs > 10 else default_epochs
    thr = np.float32(w.max()) / np.float32(divisor)
    w = np.where(w < thr, np.float32(0), w)
    a, b = fit_curve(min_dist, spread)
    init = random_init(n, n_components, seed, repulsion)
    if optimizer == "compatibility":
        emb, _, _, _ = classic_optimize(init, None, heads, tails, g["weights"][:u], n_epochs=epochs,
                                        default_epochs=default_epochs, a=a, b=b,
                                        learning_rate=learning_rate, gamma=repulsion,
                                        rate=negative_sample_rate, seed=seed, shared=True)
        return emb
    graph = compact_graph(n, n, heads, tails, w, epochs, float(negative_sample_rate))
    emb, _, _, _ = optimize_compact(init, None, graph, epochs=epochs, a=a, b=b,
                                    repulsion=repulsion, initial_alpha=learning_rate,
                                    optimizer=optimizer, negative_sample_rate=negative_sample_rate,
                                    negative_selection_range=negative_selection_range,
                                    exclude=exclude, seed=seed, threads=threads,
                                    collect_stats=False)
    return emb
