"""GPU exact kNN (BF16 MFMA prefilter + FP32 rerank + certificate and
exact FP32 rescan) against the oracle.

Bar: indices and distances bit-identical to the reference's exact kernel
(oracle restatement) for every query and any data: integer-valued data
(exact in BF16 after the shift), continuous data, data far from the origin
(where an uncentred BF16 prefilter loses true neighbours), duplicates and
ties, self-exclusion with offsets, and sentinels. The exact rescan itself is
checked by forcing it for every query, over one and several rescan rounds.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def image_like(n, d, clusters, seed, sigma=24.0):
    rng = np.random.default_rng(seed)
    centers = rng.uniform(32, 224, (clusters, d))
    lab = rng.integers(0, clusters, n)
    x = np.rint(np.clip(centers[lab] + rng.normal(0, sigma, (n, d)), 0, 255))
    return x.astype(np.float32)


@pytest.mark.parametrize("n,d,q,k", [(20_000, 784, 300, 15), (3000, 784, 1, 15),
                                     (5000, 64, 4096, 15), (40, 32, 7, 15)])
def test_integer_data_bit_exact(n, d, q, k):
    from cuscan_amd.knn import KnnIndex

    ref = image_like(n, d, 10, n + d)
    qry = image_like(q, d, 10, 7 * n + q)
    print(f"[gpu] knn n={n} d={d} q={q}", flush=True)
    idx = KnnIndex(ref, max_queries=max(q, 1))
    gi, gd = idx.query(qry, k)
    ei, ed = O.exact_knn(qry, ref, k)
    assert (gi == ei).all()
    assert (gd.view(np.uint32) == ed.view(np.uint32)).all()


# Grids above 256 workgroups (query blocks x slices) run the DMA-staged scan
# (two workgroups per CU, one wave per SIMD above 800 features); 40,000
# references and 4,000 queries (32 blocks) take 16 slices, 512 workgroups.
# 39,990 references end the last slice in a ragged chunk, and 300,000 with
# 512 queries (4 blocks x 128 slices) add the sample bound.
@pytest.mark.parametrize("n,d,q", [(40_000, 1000, 4000), (40_000, 500, 4000),
                                   (39_990, 200, 4000), (40_000, 20, 4000),
                                   (300_000, 200, 512)])
def test_large_grid_scan_widths_bit_exact(n, d, q):
    from cuscan_amd.knn import KnnIndex

    ref = image_like(n, d, 10, n + d)
    qry = image_like(q, d, 10, 3 * n + q)
    print(f"[gpu] knn large grid n={n} d={d} q={q}", flush=True)
    idx = KnnIndex(ref, max_queries=q)
    gi, gd = idx.query(qry, 15)
    idx.close()
    ei, ed = O.exact_knn(qry, ref, 15, threads=8)
    assert (gi == ei).all()
    assert (gd.view(np.uint32) == ed.view(np.uint32)).all()


def test_fit_graph_exclude_self_and_ties():
    from cuscan_amd.knn import KnnIndex

    x = image_like(6000, 100, 5, 3)
    x[100:110] = x[5]          # duplicated rows: distance-0 ties broken by index
    idx = KnnIndex(x, max_queries=6000)
    gi, gd = idx.query(x, 14, exclude_self=True)
    ei, ed = O.exact_knn(x, x, 14, exclude_self=True)
    assert (gi == ei).all() and (gd.view(np.uint32) == ed.view(np.uint32)).all()
    assert not (gi == np.arange(6000)[:, None]).any()


def test_fewer_references_than_k_uses_sentinels():
    from cuscan_amd.knn import KnnIndex

    ref = image_like(5, 16, 2, 1)
    qry = image_like(3, 16, 2, 2)
    gi, gd = KnnIndex(ref, max_queries=8).query(qry, 15)
    ei, ed = O.exact_knn(qry, ref, 15)
    assert (gi == ei).all() and (gd.view(np.uint32) == ed.view(np.uint32)).all()
    assert (gi[:, 5:] == 5).all()


def test_continuous_data_recall():
    from cuscan_amd.knn import KnnIndex

    rng = np.random.default_rng(0)
    centers = rng.uniform(-10, 10, (50, 50))
    ref = (centers[rng.integers(0, 50, 20_000)] + rng.normal(0, 1.0, (20_000, 50))).astype(np.float32)
    qry = (centers[rng.integers(0, 50, 500)] + rng.normal(0, 1.0, (500, 50))).astype(np.float32)
    gi, gd = KnnIndex(ref, max_queries=500).query(qry, 15)
    ei, ed = O.exact_knn(qry, ref, 15)
    assert (gi == ei).all() and (gd.view(np.uint32) == ed.view(np.uint32)).all()


@pytest.mark.parametrize("q", [1, 37, 64, 100, 384, 700])
def test_most_slices_bit_exact(q):
    """1,048,600 references: batches of up to 128 rows take 512 slices (the
    most a batch takes) of 2,064 references, the last one ragged, so the
    scan runs 512 DMA-staged workgroups (64-query blocks up to 64 rows) and
    the merge lists more than 256 slices per query; 384 and 700 rows take
    170 and 85 slices (3 and 6 query blocks: 510 cells, the slice count
    rounded down so that every cell is resident at once). The thresholds
    start from the 32,768-row sample bound; the merged prefilter lists equal
    those of a 9,472-row batch (6 slices, no bound)."""
    from cuscan_amd.knn import KnnIndex

    ref = image_like(1_048_600, 20, 10, 21)
    qry = image_like(9472, 20, 10, 22)
    print(f"[gpu] knn most slices q={q}", flush=True)
    idx = KnnIndex(ref, max_queries=9472)
    gi, gd = idx.query(qry[:q], 15)
    ei, ed = O.exact_knn(qry[:q], ref, 15, threads=8)
    assert (gi == ei).all() and (gd.view(np.uint32) == ed.view(np.uint32)).all()
    qt = torch.from_numpy(qry).cuda()
    small_ids, small_n = idx.candidates_device(qt[:q])
    big_ids, big_n = idx.candidates_device(qt)
    assert torch.equal(small_n, big_n[:q]) and torch.equal(small_ids, big_ids[:q])
    idx.close()


def test_most_slices_exclude_self():
    """Fit-graph style queries (40 rows of the 1,048,600 references, starting
    at row 1,000, excluding self) through the 64-query-block scan with the
    self test."""
    from cuscan_amd.knn import KnnIndex

    x = image_like(1_048_600, 20, 10, 23)
    print("[gpu] knn most slices exclude self", flush=True)
    idx = KnnIndex(x, max_queries=40)
    gi, gd = idx.query(x[1000:1040], 14, exclude_self=True, query_offset=1000)
    idx.close()
    ei, ed = O.exact_knn(x[1000:1040], x, 14, exclude_self=True, query_offset=1000, threads=8)
    assert (gi == ei).all() and (gd.view(np.uint32) == ed.view(np.uint32)).all()
    assert not (gi == np.arange(1000, 1040)[:, None]).any()


@pytest.mark.parametrize("q", [1, 64, 700, 4096])
def test_sample_bound_bit_exact(q):
    """150,000 references: batches of up to 8,192 rows take at least 8
    slices, so their thresholds start from the 8,192-row sample bound.
    Results stay bit-identical to the exact kernel, and the merged prefilter
    lists equal those of a 9,472-row batch (7 slices, no bound): the bound
    drops no candidate of the top 64."""
    from cuscan_amd.knn import KnnIndex

    ref = image_like(150_000, 64, 10, 11)
    qry = image_like(9472, 64, 10, 12)
    print(f"[gpu] knn sample bound q={q}", flush=True)
    idx = KnnIndex(ref, max_queries=9472)
    # slices per batch (KnnIndex::slices_for): ceil(512 / query blocks),
    # at most references / 2048; the bound needs at least 8 slices
    slices = lambda rows: min(-(-512 // -(-rows // 128)), 150_000 // 2048)
    assert slices(q) >= 8 and slices(9472) == 7
    gi, gd = idx.query(qry[:q], 15)
    ei, ed = O.exact_knn(qry[:q], ref, 15, threads=8)
    assert (gi == ei).all() and (gd.view(np.uint32) == ed.view(np.uint32)).all()
    qt = torch.from_numpy(qry).cuda()
    small_ids, small_n = idx.candidates_device(qt[:q])
    big_ids, big_n = idx.candidates_device(qt)
    assert torch.equal(small_n, big_n[:q]) and torch.equal(small_ids, big_ids[:q])
    assert (small_n == 64).all()
    idx.close()


@pytest.mark.parametrize("q", [1, 100, 1024])
def test_large_sample_bound_bit_exact(q):
    """262,200 references (at least 8 x 32,768): batches of up to 1,024 rows
    seed their thresholds from a 32,768-row sample. The merged prefilter
    lists equal those of a 9,472-row batch (7 slices, no bound), so the
    tighter bound drops no candidate of the top 64, and the results stay
    bit-identical to the exact kernel."""
    from cuscan_amd.knn import KnnIndex

    ref = image_like(262_200, 32, 10, 31)
    qry = image_like(9472, 32, 10, 32)
    print(f"[gpu] knn large sample bound q={q}", flush=True)
    idx = KnnIndex(ref, max_queries=9472)
    gi, gd = idx.query(qry[:q], 15)
    ei, ed = O.exact_knn(qry[:q], ref, 15, threads=8)
    assert (gi == ei).all() and (gd.view(np.uint32) == ed.view(np.uint32)).all()
    qt = torch.from_numpy(qry).cuda()
    small_ids, small_n = idx.candidates_device(qt[:q])
    big_ids, big_n = idx.candidates_device(qt)
    assert torch.equal(small_n, big_n[:q]) and torch.equal(small_ids, big_ids[:q])
    assert (small_n == 64).all()
    idx.close()


def test_sample_bound_exclude_self_with_offset():
    """Fit-graph style queries (rows of the reference set itself, excluding
    self) through the sample bound, for a batch that starts at row 1,000."""
    from cuscan_amd.knn import KnnIndex

    x = image_like(150_000, 48, 8, 5)
    idx = KnnIndex(x, max_queries=256)
    gi, gd = idx.query(x[1000:1256], 14, exclude_self=True, query_offset=1000)
    ei, ed = O.exact_knn(x[1000:1256], x, 14, exclude_self=True, query_offset=1000, threads=8)
    assert (gi == ei).all() and (gd.view(np.uint32) == ed.view(np.uint32)).all()
    assert not (gi == np.arange(1000, 1256)[:, None]).any()
    idx.close()


@pytest.mark.parametrize("offset", [0.0, 1000.0])
def test_offset_continuous_data_exact(offset):
    """20,000 references, 50 features, unit-variance clusters shifted by
    `offset`: without centring, BF16 rounding of the offset swamps the
    neighbour gaps. Separate queries and fit-style self queries."""
    from cuscan_amd.knn import KnnIndex

    rng = np.random.default_rng(3)
    centers = rng.uniform(-3, 3, (20, 50))
    ref = (centers[rng.integers(0, 20, 20_000)] + rng.normal(0, 1.0, (20_000, 50)) +
           offset).astype(np.float32)
    qry = (centers[rng.integers(0, 20, 2000)] + rng.normal(0, 1.0, (2000, 50)) +
           offset).astype(np.float32)
    print(f"[gpu] knn offset {offset}", flush=True)
    idx = KnnIndex(ref, max_queries=20_000)
    gi, gd = idx.query(qry, 15)
    rescanned, eps = idx.certificate_stats()
    ei, ed = O.exact_knn(qry, ref, 15, threads=8)
    print(f"offset {offset}: {rescanned} of 2000 queries rescanned, rounding ratio {eps:.2e}")
    assert (gi == ei).all() and (gd.view(np.uint32) == ed.view(np.uint32)).all()
    assert 0.0 < eps < 2.0 ** -8
    gi, gd = idx.query(ref[:3000], 14, exclude_self=True)
    ei, ed = O.exact_knn(ref[:3000], ref, 14, exclude_self=True, threads=8)
    assert (gi == ei).all() and (gd.view(np.uint32) == ed.view(np.uint32)).all()
    idx.close()


@pytest.mark.parametrize("n,d,q,k,exclude", [(20_000, 50, 300, 15, False),
                                             (6000, 100, 1500, 14, True),
                                             (3000, 16, 9000, 15, False),
                                             (5, 16, 3, 15, False)])
def test_forced_exact_rescan_bit_exact(n, d, q, k, exclude):
    """Every query through the exact FP32 rescan: one round, two rounds
    (1,500 queries, device-side count) and a batch above 8,192 rows (host
    count), with duplicates, self-exclusion and sentinels."""
    from cuscan_amd.knn import KnnIndex

    rng = np.random.default_rng(n + q)
    ref = (rng.normal(0, 1.0, (n, d)) * 3 + 100).astype(np.float32)
    if n > 200:
        ref[100:140] = ref[:40]
    qry = ref[:q] if exclude else (rng.normal(0, 1.0, (q, d)) * 3 + 100).astype(np.float32)
    print(f"[gpu] knn forced rescan n={n} q={q}", flush=True)
    idx = KnnIndex(ref, max_queries=q)
    idx.set_rescan(True)
    gi, gd = idx.query(qry, k, exclude_self=exclude)
    assert idx.certificate_stats()[0] == q
    ei, ed = O.exact_knn(qry, ref, k, exclude_self=exclude, threads=8)
    assert (gi == ei).all() and (gd.view(np.uint32) == ed.view(np.uint32)).all()
    idx.close()


def test_certificate_tight_on_integer_mixture():
    """100,000 x 784 rows of the 10-cluster 0..255 mixture as a fit graph:
    the shifted references are exact in BF16 and no query needs the rescan."""
    from cuscan_amd.knn import KnnIndex

    x = image_like(100_000, 784, 10, 42)
    print("[gpu] knn certificate on the 100k mixture", flush=True)
    idx = KnnIndex(x, max_queries=8192)
    qt = torch.from_numpy(x[:8192]).cuda()
    gi, gd = idx.query_device(qt, 14, exclude_self=True)
    rescanned, eps = idx.certificate_stats()
    print(f"mixture: {rescanned} of 8192 rescanned, rounding ratio {eps}")
    assert eps == 0.0 and rescanned == 0
    ei, ed = O.exact_knn(x[:512], x, 14, exclude_self=True, threads=8)
    assert (gi[:512].cpu().numpy() == ei).all()
    assert (gd[:512].cpu().numpy().view(np.uint32) == ed.view(np.uint32)).all()
    idx.close()


def test_certificate_catches_prefilter_misses():
    """Feature 0 is bimodal (+-5000 plus a fine offset in [0, 30)), so its
    mean-centred BF16 rounding (up to 16) swamps the fine offsets that
    decide the neighbours: the prefilter's 64 miss true neighbours, the
    certificate flags those queries and the exact rescan restores the
    reference's result bit for bit."""
    from cuscan_amd.knn import KnnIndex

    rng = np.random.default_rng(9)

    def rows(m):
        x = rng.normal(0, 1.0, (m, 8))
        x[:, 0] = rng.choice([-5000.0, 5000.0], m) + rng.uniform(0, 30, m)
        return x.astype(np.float32)

    ref, qry = rows(5000), rows(400)
    idx = KnnIndex(ref, max_queries=400)
    qt = torch.from_numpy(qry).cuda()
    cand, counts = idx.candidates_device(qt)
    ei, ed = O.exact_knn(qry, ref, 15, threads=8)
    cand = cand.cpu().numpy()
    missed = sum(len(set(e) - set(c)) for e, c in zip(ei, cand))
    gi, gd = idx.query(qry, 15)
    rescanned, eps = idx.certificate_stats()
    print(f"bimodal: prefilter missed {missed} true neighbours; {rescanned} of 400 rescanned, "
          f"rounding ratio {eps:.2e}")
    assert missed > 0 and rescanned > 0
    assert (gi == ei).all() and (gd.view(np.uint32) == ed.view(np.uint32)).all()
    idx.close()
