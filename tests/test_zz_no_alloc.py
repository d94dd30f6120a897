"""The allocation-free argsort (umaphip_stable_argsort_f32_noalloc) through
the C ABI against umaphip_stable_argsort_f32 with a caller workspace: order,
sorted values and the non-finite status bit-identical, over two calls in a
row (the second smaller, on the scratch the first left behind), a call
with a non-finite value, an empty call, and the size limit."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu
MAX_KEYS = 1 << 20
UMAPHIP_CAPACITY = 4


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _p(t):
    return C.c_void_p(t.data_ptr())


def noalloc(values):
    from cuscan_amd import _lib

    n = values.shape[0]
    order = torch.full((max(n, 1),), -7, dtype=torch.int32, device="cuda:0")
    out = torch.full((max(n, 1),), -7.0, dtype=torch.float32, device="cuda:0")
    status = torch.full((1,), -7, dtype=torch.int64, device="cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    rc = _lib.lib().umaphip_stable_argsort_f32_noalloc(_p(values), n, _p(order), _p(out),
                                                       _p(status), C.c_void_p(stream))
    torch.cuda.synchronize()
    return rc, order[:n], out[:n], status


def with_workspace(values):
    from cuscan_amd import device as D

    order, out, status = D.stable_argsort(values)
    torch.cuda.synchronize()
    return order, out, status


def keys(n, seed):
    rng = np.random.default_rng(seed)
    v = np.round(rng.normal(0, 50, n)).astype(np.float32)   # many ties
    v[rng.integers(0, n, n // 50)] = -0.0
    v[rng.integers(0, n, n // 50)] = 0.0
    return torch.from_numpy(v).cuda()


def test_noalloc_argsort_matches_workspace_argsort():
    for n, seed in ((1_000_000, 1), (5_000, 2)):
        print(f"[gpu] no-alloc argsort n={n}", flush=True)
        v = keys(n, seed)
        rc, order, out, status = noalloc(v)
        assert rc == 0
        eo, es, est = with_workspace(v)
        assert torch.equal(order, eo)
        assert torch.equal(out.view(torch.int32), es.view(torch.int32))
        assert int(status.item()) == int(est.item()) == -1
        # a stable argsort: sorted ascending, ties in row order
        o = order.cpu().numpy()
        assert np.array_equal(o, np.argsort(np.where(v.cpu().numpy() == 0, 0, v.cpu().numpy()),
                                            kind="stable"))


def test_noalloc_argsort_status_empty_and_limit():
    v = keys(3000, 3)
    v[12] = float("nan")
    v[7] = float("inf")
    rc, _, _, status = noalloc(v)
    _, _, est = with_workspace(v)
    assert rc == 0 and int(status.item()) == int(est.item()) == 7
    rc, _, _, _ = noalloc(torch.empty(0, dtype=torch.float32, device="cuda:0"))
    assert rc == 0
    big = torch.zeros(MAX_KEYS + 1, dtype=torch.float32, device="cuda:0")
    rc, _, _, status = noalloc(big)
    assert rc == UMAPHIP_CAPACITY and int(status.item()) == -7   # nothing ran


def test_noalloc_argsort_allocates_nothing():
    v = keys(200_000, 4)
    noalloc(v)
    free0, _ = torch.cuda.mem_get_info()
    for _ in range(3):
        rc, _, _, _ = noalloc(v)
        assert rc == 0
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0
