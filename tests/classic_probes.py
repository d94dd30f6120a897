"""Shared driver for the compatibility-optimizer probes (umap-learn 0.6.0
golden frames extracted by tests/golden/make_classic_fixture.py).

Each frame is replayed from the previous frame's recorded state, as the
reference's own test does (crates/umap/tests/test_standard_layout.rs:
163-323): clocks must agree to 1e-5 and the per-head Tausworthe states
exactly; the recorded constants are a = 1.5769435, b = 0.8950609. The
recorded clocks are f32 (the reference test runs the clock kernel on f32
tensors, with the negative count formed in f64), so the replay feeds the
f64 kernels f32 values and rounds their results to f32 after every epoch:
for f32 operands the f64 sum rounded once equals the f32 sum.
"""
import json
from pathlib import Path

import numpy as np

A_PROBE, B_PROBE = 1.5769435, 0.8950609
PROBES = json.loads((Path(__file__).resolve().parent / "golden" /
                     "classic_optimizer_probes.json").read_text())["optimizer_probes"]


def initial_rng(case):
    head = np.asarray(case["head_initial"], np.float32)
    base = np.asarray(case["rng_state"], np.int64)
    bits = head[:, 0].astype(np.float64).view(np.int64)
    return (base[None, :] + bits[:, None]).astype(np.int64)


def head_csr(heads, weights, rows):
    offsets = np.zeros(rows + 1, np.int32)
    ids = []
    for r in range(rows):
        ids += [e for e, (hh, w) in enumerate(zip(heads, weights)) if w > 0 and hh == r]
        offsets[r + 1] = len(ids)
    return offsets, np.asarray(ids or [0], np.int32)


def cases():
    """(case, edge_count, frames) for every probe and ordered prefix."""
    out = []
    for case in PROBES:
        out.append((case, len(case["edges_head_tail"]), case["epochs"]))
        for prefix in case["ordered_event_prefixes"]:
            out.append((case, prefix["through_edge_index"] + 1, prefix["epochs"]))
    return out


def replay(case, edge_count, frames, run_epoch, check_coordinates):
    """run_epoch(state dict, epoch, alpha) advances the state in place."""
    edges = np.asarray(case["edges_head_tail"], np.int32)[:edge_count]
    heads, tails = edges[:, 0].copy(), edges[:, 1].copy()
    weights = np.asarray(case["positive_weights"], np.float32)[:edge_count]
    per32 = np.asarray(case["epochs_per_sample"], np.float32)[:edge_count]
    rate = float(case["negative_sample_rate"])
    per = per32.astype(np.float64)
    nper = (per32 / np.float32(rate)).astype(np.float64)
    shared = bool(case["move_other"])
    st = dict(head=np.asarray(case["head_initial"], np.float32).copy(),
              tail=np.asarray(case["tail_initial"], np.float32).copy(),
              tails=tails, weights=weights, per=per, npos=per.copy(), nper=nper,
              nneg=nper.copy(), rng=initial_rng(case), shared=shared, rate=rate,
              gamma=float(case["gamma"]))
    st["offsets"], st["ids"] = head_csr(heads, weights, st["head"].shape[0])
    for frame in frames:
        run_epoch(st, int(frame["epoch"]), float(frame["alpha_used"]))
        st["npos"] = st["npos"].astype(np.float32).astype(np.float64)
        st["nneg"] = st["nneg"].astype(np.float32).astype(np.float64)
        np.testing.assert_allclose(st["npos"], frame["epoch_of_next_sample"], rtol=1e-5, atol=1e-5)
        np.testing.assert_allclose(st["nneg"], frame["epoch_of_next_negative_sample"], rtol=1e-5,
                                   atol=1e-5)
        assert (st["rng"] == np.asarray(frame["rng_state_per_sample"], np.int64)).all()
        want_head = np.asarray(frame["head_embedding"], np.float32)
        want_tail = np.asarray(frame["tail_embedding"], np.float32)
        assert np.isfinite(st["head"]).all()
        check_coordinates(st, want_head, want_tail)
        # continue from the recorded state
        st["head"] = want_head.copy()
        st["tail"] = want_tail.copy()
