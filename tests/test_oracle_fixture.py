"""Pin the oracle against the reference's frozen umap-learn single-epoch fixture.

Mirrors crates/umap-core/tests/test_modern_oracle.rs:35-200 (host harness)
and crates/umap/src/kernels/standard_umap/modern_layout.rs:1068-1251
(kernel-semantics fixture test). Tolerance 3e-5 absolute on coordinates and
moments, exact on clocks (as in the reference).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import fixture_cases as F

FIX = F.load()
TOL = 3e-5


def test_schedules_match_upstream():
    for s in FIX["schedules"]:
        fn = O.adam_schedule if s["optimizer"] == "adam" else O.momentum_schedule
        alpha, mom = fn(s["epochs"], 1.0, s["good"])
        assert np.abs(alpha - np.array(s["alpha"])).max() < 1e-12
        assert np.abs(mom - np.array(s["momentum"])).max() < 1e-12


def _corrections(epoch):
    c1 = np.float32(1.0 - float(np.float32(0.9)) ** (epoch + 1))
    c2 = np.float32(1.0 - float(np.float32(0.99)) ** (epoch + 1))
    return float(c1), float(c2)


def _host_harness(case, x):
    """test_modern_oracle.rs's blocked host harness over the oracle's
    building blocks (advance_clock, accumulate_force, negative_index)."""
    rows, dim, refs, epoch = case["rows"], case["dim"], case["refs"], case["epoch"]
    small = rows <= 1024
    adam = case["optimizer"] == "adam"
    fit = case["operation"] == "fit"
    alpha, gamma, scale = case["alpha"], case["gamma"], case["scale"]
    offsets, indices = case["offsets"], case["indices"]
    npos = list(case["next_positive"])
    nneg = list(case["next_negative"])
    head = x["head"].copy()
    tail = x["tail"]
    updates = x["updates"].copy()
    first = np.zeros_like(head)
    second = np.zeros_like(head)
    raw = np.zeros_like(head)
    c1, c2 = _corrections(epoch)
    block = case["block"]
    for start in range(0, rows, block):
        end = min(start + block, rows)
        order = [x["to_order"][o] if (adam and fit) else o for o in x["from_order"][start:end]]
        for row in order:
            source = head[row].copy()
            grad = np.zeros(dim, np.float32)
            for edge in range(offsets[row], offsets[row + 1]):
                count, npos[edge], nneg[edge] = O.advance_clock(
                    npos[edge], nneg[edge], case["periods"][edge],
                    case["negative_periods"][edge], epoch)
                if count is None:
                    continue
                ref = head if fit else tail
                O.lib().orc_accumulate_force(grad, source,
                                             np.ascontiguousarray(ref[indices[edge]]), dim,
                                             1.0, 1.0, gamma, scale, int(small), 0)
                for sample in range(count):
                    rank = O.lib().orc_negative_index(edge, epoch, sample, refs)
                    target = x["to_order"][rank]
                    O.lib().orc_accumulate_force(grad, source, np.ascontiguousarray(ref[target]),
                                                 dim, 1.0, 1.0, gamma, scale, int(small), 1)
            if adam:
                updates[row] += grad
            else:
                updates[row] += np.float32(alpha) * grad
        for row in order:
            raw[row] = updates[row]
            if adam:
                g = updates[row]
                nz = g != 0
                first[row, nz] = np.float32(0.9) * first[row, nz] + (np.float32(1) - np.float32(0.9)) * g[nz]
                second[row, nz] = np.float32(0.99) * second[row, nz] + (np.float32(1) - np.float32(0.99)) * g[nz] * g[nz]
                mh = first[row, nz] / np.float32(c1)
                vh = second[row, nz] / np.float32(c2)
                head[row, nz] += np.float32(alpha) * mh / (np.sqrt(vh) + np.float32(1e-4))
                updates[row] = 0
            else:
                head[row] += updates[row]
    if not adam:
        updates *= np.float32(case["momentum"])
    return head, raw, updates, first, second, npos, nneg


@pytest.mark.parametrize("idx", range(32))
def test_host_harness_reproduces_fixture(idx):
    case = FIX["cases"][idx]
    x = F.inputs(case)
    head, raw, final, first, second, npos, nneg = _host_harness(case, x)
    assert npos == case["expected_positive"]
    assert nneg == case["expected_negative"]
    zeros = np.zeros_like(x["head"])
    np.testing.assert_allclose(head, F.expected(case, "expected_head", x["head"]), atol=TOL, rtol=0)
    np.testing.assert_allclose(raw, F.expected(case, "expected_updates", x["updates"]), atol=TOL, rtol=0)
    np.testing.assert_allclose(final, F.expected(case, "expected_final_updates", x["updates"]), atol=TOL, rtol=0)
    np.testing.assert_allclose(first, F.expected(case, "expected_first", zeros), atol=TOL, rtol=0)
    np.testing.assert_allclose(second, F.expected(case, "expected_second", zeros), atol=TOL, rtol=0)


def kernel_epoch(case, x, threads=1):
    """Run the restated kernels (gradient_kernel + adam/momentum) once."""
    rows, epoch = case["rows"], case["epoch"]
    fit = case["operation"] == "fit"
    adam = case["optimizer"] == "adam"
    c1, c2 = _corrections(epoch)
    ep = O.Epoch(head=x["head"], tail=x["tail"], offsets=case["offsets"], indices=case["indices"],
                 source_order=x["from_order"], negative_order=x["to_order"], source_ranks=x["ranks"],
                 periods=case["periods"], next_positive=case["next_positive"],
                 negative_periods=case["negative_periods"], next_negative=case["next_negative"],
                 updates=x["updates"], m=np.zeros_like(x["head"]), v=np.zeros_like(x["head"]),
                 epoch=epoch, a=1.0, b=1.0, gamma=case["gamma"], scale=case["scale"],
                 negative_range=case["refs"], alpha=case["alpha"], mu=case["momentum"],
                 correction1=c1, correction2=c2, momentum=not adam, shared=fit)
    if adam:
        ep.gradient_block(0, rows, 1.0, threads)
        raw = ep.updates.copy()
        ep.adam_apply(threads)
    else:
        block = rows if rows <= 1024 else rows // 2
        for start in range(0, rows, block):
            end = min(start + block, rows)
            ep.gradient_block(start, end, case["alpha"], threads)
            ep.momentum_apply_block(start, end)
        raw = ep.updates.copy()
        ep.momentum_decay()
    return ep, raw


@pytest.mark.parametrize("idx", range(32))
def test_kernel_semantics_match_fixture(idx):
    case = FIX["cases"][idx]
    x = F.inputs(case)
    ep, raw = kernel_epoch(case, x)
    assert ep.next_positive.tolist() == case["expected_positive"]
    assert ep.next_negative.tolist() == case["expected_negative"]
    zeros = np.zeros_like(x["head"])
    exp_head = F.expected(case, "expected_head", x["head"])
    if idx in F.UPSTREAM_BLOCKED_ONLY:
        # Jacobi Adam reads row 1024's pre-step position for edge 0's
        # negatives; upstream's blocked Adam reads it after block one.
        assert np.abs(ep.head - exp_head).max() > TOL
        return
    np.testing.assert_allclose(ep.head, exp_head, atol=TOL, rtol=0)
    np.testing.assert_allclose(raw, F.expected(case, "expected_updates", x["updates"]), atol=TOL, rtol=0)
    np.testing.assert_allclose(ep.updates, F.expected(case, "expected_final_updates", x["updates"]), atol=TOL, rtol=0)
    if case["optimizer"] == "adam":
        np.testing.assert_allclose(ep.m, F.expected(case, "expected_first", zeros), atol=TOL, rtol=0)
        np.testing.assert_allclose(ep.v, F.expected(case, "expected_second", zeros), atol=TOL, rtol=0)


def test_gradient64_agrees_with_f32_kernel_on_fixture():
    case = FIX["cases"][1]
    x = F.inputs(case)
    ep, raw = kernel_epoch(case, x)
    ep64 = O.Epoch(head=x["head"], tail=x["tail"], offsets=case["offsets"], indices=case["indices"],
                   source_order=x["from_order"], negative_order=x["to_order"], source_ranks=x["ranks"],
                   periods=case["periods"], next_positive=case["next_positive"],
                   negative_periods=case["negative_periods"], next_negative=case["next_negative"],
                   updates=np.zeros_like(x["head"]), m=np.zeros_like(x["head"]), v=np.zeros_like(x["head"]),
                   epoch=case["epoch"], a=1.0, b=1.0, gamma=case["gamma"], scale=case["scale"],
                   negative_range=case["refs"], shared=True)
    g64 = ep64.gradient_block64(0, case["rows"], 1.0)
    np.testing.assert_allclose(raw, g64, rtol=1e-6, atol=1e-7)
    assert ep64.next_positive.tolist() == case["expected_positive"]
