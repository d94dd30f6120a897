"""Continuous-batching transform server.

A request's result must not depend on what else is in flight: each result
is compared bit for bit alone and in a mix of concurrent requests. The
server runs the tile sweep with the same per-row summation order as
UMAP.transform (all epochs of a request in one launch), so the served
result also equals the plain per-request transform bit for bit, for small
(global negatives) and large (projection windows) requests alike.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def blobs(n, d, centers, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-10, 10, (centers, d))
    return (c[rng.integers(0, centers, n)] + rng.normal(0, 1, (n, d))).astype(np.float32)


@pytest.fixture(scope="module")
def model():
    from cuscan_amd import UMAP

    x = blobs(3000, 16, 12, 1)
    print("[gpu] serve: fit 3000x16", flush=True)
    return UMAP(n_neighbors=15, n_epochs=60, init="random", negative_selection_range=500,
                random_state=1).fit(x)


def test_server_results_equal_alone_in_mix_and_plain(model):
    from cuscan_amd.serve import TransformServer

    small = blobs(37, 16, 12, 2)
    large = blobs(1300, 16, 12, 3)
    with TransformServer.from_model(model, max_batch_rows=4096) as srv:
        print("[gpu] serve: alone", flush=True)
        alone_s, _ = srv.wait(srv.submit(small))
        alone_l, _ = srv.wait(srv.submit(large))
        print("[gpu] serve: mix", flush=True)
        rng = np.random.default_rng(9)
        tickets, order = [], []
        for i in range(24):
            if i == 7:
                tickets.append(srv.submit(small)); order.append("s")
            if i == 15:
                tickets.append(srv.submit(large)); order.append("l")
            q = int(np.clip(rng.lognormal(np.log(40), 1.0), 1, 600))
            tickets.append(srv.submit(blobs(q, 16, 12, 100 + i))); order.append("o")
        results = [srv.wait(t) for t in tickets]
        stats = srv.stats()
    assert all(r is not None for r in results)
    mix_s = results[order.index("s")][0]
    mix_l = results[order.index("l")][0]
    assert (mix_s == alone_s).all()
    assert (mix_l == alone_l).all()
    assert stats["peak_active"] > 1 and stats["requests"] == 2 + len(tickets)
    print("[gpu] serve: plain", flush=True)
    plain_s = model.transform(small)
    assert (plain_s == alone_s).all()
    plain_l = model.transform(large)
    assert (plain_l == alone_l).all() and np.isfinite(alone_l).all()


def test_server_admission_by_memory_completes_every_request(model):
    from cuscan_amd.serve import TransformServer

    # an arena that holds two 200-row requests (about 120 KB each) at a time
    with TransformServer.from_model(model, max_batch_rows=1024, memory_bytes=300_000) as srv:
        xs = [blobs(200, 16, 12, 40 + i) for i in range(12)]
        print("[gpu] serve: memory pressure", flush=True)
        tickets = [srv.submit(x) for x in xs]
        outs = [srv.wait(t) for t in tickets]
        stats = srv.stats()
        alone = srv.wait(srv.submit(xs[5]))[0]
    assert all(o is not None and np.isfinite(o[0]).all() for o in outs)
    assert stats["admission_waits"] > 0
    assert (outs[5][0] == alone).all()


def test_server_rejects_invalid_requests(model):
    from cuscan_amd.serve import TransformServer

    with TransformServer.from_model(model, max_batch_rows=256) as srv:
        with pytest.raises(ValueError):
            srv.submit(np.zeros((3, 5), np.float32))
        with pytest.raises(RuntimeError):
            srv.submit(np.zeros((300, 16), np.float32))
        bad = np.zeros((2, 16), np.float32)
        bad[0, 0] = np.nan
        with pytest.raises(ValueError):
            srv.submit(bad)
