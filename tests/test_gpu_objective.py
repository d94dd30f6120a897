"""FP64 cross-entropy objective on the device against the oracle's
compensated FP64 sum (1e-11 relative), for fits and transforms, through the
device entry point and through a plan's per-epoch trace."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
from tests.gpu_helpers import random_graph  # noqa: E402

pytestmark = pytest.mark.gpu
A_CURVE, B_CURVE = 1.5769434, 0.8950608


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _dev(a, dtype):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda:0")


@pytest.mark.parametrize("shared", [True, False])
def test_cross_entropy_matches_oracle(shared):
    from cuscan_amd import device as D

    h, n = 3000, (3000 if shared else 5000)
    g = random_graph(h, n, 9, 7, 50, symmetric=shared)
    rng = np.random.default_rng(3)
    head = rng.normal(0, 4, (h, 2)).astype(np.float32)
    tail = None if shared else rng.normal(0, 4, (n, 2)).astype(np.float32)
    w = rng.uniform(0.01, 1.0, len(g["indices"])).astype(np.float32)
    print(f"[gpu] objective shared={shared}", flush=True)
    got = D.cross_entropy(head=_dev(head, torch.float32),
                          tail=None if tail is None else _dev(tail, torch.float32),
                          offsets=_dev(g["offsets"].astype(np.int32), torch.int32),
                          indices=_dev(g["indices"].astype(np.int32), torch.int32),
                          weights=_dev(w, torch.float32), a=A_CURVE, b=B_CURVE)
    want = O.cross_entropy(head, tail, g, w, A_CURVE, B_CURVE)
    assert abs(got - want) <= 1e-11 * abs(want)


def test_plan_objective_trace_matches_oracle_and_decreases():
    from cuscan_amd import layout as L

    h, epochs = 3000, 40
    g = random_graph(h, h, 8, 11, epochs)
    w = np.asarray(1.0 / g["periods"], np.float32)   # weight / max weight of each edge
    head = np.random.default_rng(5).uniform(-10, 10, (h, 2)).astype(np.float32)
    plan = L.LayoutPlan(head, None, g, L.LayoutParams(epochs=epochs, a=A_CURVE, b=B_CURVE),
                        shared=True)
    plan.set_objective(w, 10)
    print("[gpu] objective trace", flush=True)
    plan.run()
    emb, _ = plan.download()
    trace = plan.objective()
    plan.close()
    assert len(trace) == 4
    want = O.cross_entropy(emb, None, g, w, A_CURVE, B_CURVE)
    assert abs(trace[-1] - want) <= 1e-11 * abs(want)
    start = O.cross_entropy(head, None, g, w, A_CURVE, B_CURVE)
    assert trace[-1] < trace[0] < start
