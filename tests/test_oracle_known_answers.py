"""Closed-form and known-answer pins of the oracle's building blocks.

Each test restates an inline reference test (cited) or a mathematical
identity that a plausible slip (a dropped term, a swapped index, a wrong
sign) would break.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

U64 = (1 << 64) - 1


def f32(x):
    return float(np.float32(x))


# -------------------------------------------- umap-core/src/modern_layout.rs:1696-1730
def test_schedule_boundaries():
    a, _ = O.adam_schedule(1, 4.0, False)
    assert f32(a[0]) == f32(0.2)
    adam, _ = O.adam_schedule(400, 1.0, False)
    for n, value in [(0, 2.0), (99, 0.10019), (100, 0.2), (399, 0.0505)]:
        assert abs(adam[n] - value) < 1e-12
        assert f32(adam[n]) == f32(value)
    good, _ = O.adam_schedule(400, 2.0, True)
    for n, value in [(0, 1.0), (49, 0.10036), (50, 0.2)]:
        assert abs(good[n] - value) < 1e-12
    m_a, m_mu = O.momentum_schedule(400, 2.0, True)
    for n, alpha, mu in [(0, 0.505, 0.0), (199, 0.0050125, 0.4999875), (200, 0.005, 0.5),
                         (399, 0.005, 0.5)]:
        assert abs(m_a[n] - alpha) < 1e-12 and abs(m_mu[n] - mu) < 1e-12
    m_a, m_mu = O.momentum_schedule(4, 2.0, False)
    assert (m_a[0], m_mu[0], m_a[3], m_mu[3]) == (0.5, 0.375, 0.03125, 0.4921875)


# ---------------------------------------------------------------- :1732-1768
def test_compact_padding_duplicates_clocks_and_hashes():
    rows = [2, 0, 0, 0, 0, 2]
    cols = [3, 2**40, 2, 1, 2, 0]
    w = [0.5, 0.0, 0.25, 1.0, 0.25, 0.0]
    g = O.compact_graph(4, 4, rows, cols, w, 100, 3.0)
    assert g["offsets"].tolist() == [0, 2, 2, 3, 3]
    assert g["indices"].tolist() == [1, 2, 3]
    assert g["periods"].tolist() == [1.0, 2.0, 2.0]
    assert g["negative_periods"].tolist() == [0.5, 1.0, 1.0]
    npos, nneg = g["next_positive"].copy(), g["next_negative"].copy()
    for n in range(6):
        for e in range(3):
            cnt, npos[e], nneg[e] = O.advance_clock(npos[e], nneg[e], g["periods"][e],
                                                    g["negative_periods"][e], n)
            for p in range(cnt or 0):
                assert O.lib().orc_negative_index(e, n, p, 4) == (e * (n + p + 1)) % 4
    c = O.compact_graph(1, 1, [0], [0], [1.0], 100, 3.0)
    s = [c["next_positive"][0], c["next_negative"][0]]
    res = []
    for n in (0, 1, 5):
        cnt, s[0], s[1] = O.advance_clock(s[0], s[1], c["periods"][0], c["negative_periods"][0], n)
        res.append(cnt)
    assert res == [None, 1, 8]


def test_compact_rejections():
    with pytest.raises(ValueError):
        O.compact_graph(1, 1, [], [], [], 100, 5.0)
    with pytest.raises(ValueError):
        O.compact_graph(1, 1, [0], [1], [1.0], 100, 5.0)
    with pytest.raises(ValueError):
        O.compact_graph(1, 1, [0], [0], [np.nan], 100, 5.0)
    with pytest.raises(ValueError):
        O.compact_graph(1, 1, [0], [0], [1.0], 100, -1.0)


# ---------------------------------------------------------------- :1770-1828
def test_fractional_clocks_and_zero_rate():
    g = O.compact_graph(1, 2, [0, 0], [0, 1], [1.0, 0.3], 7, 5.0)
    samples = np.float32(7.0) * (np.float32(0.3) / np.float32(1.0))
    assert g["periods"][1] == 7.0 / float(samples)
    assert g["periods"][1] != float(np.float32(1.0) / np.float32(0.3))
    cnt, p, n = O.advance_clock(2.0 + 1e-10, g["next_negative"][1], g["periods"][1],
                                g["negative_periods"][1], 2)
    assert cnt is None
    cnt, p, n = O.advance_clock(2.0 + 1e-10, g["next_negative"][1], g["periods"][1],
                                g["negative_periods"][1], 3)
    assert cnt == 2
    z = O.compact_graph(1, 2, [0, 0], [0, 1], [1.0, 0.3], 7, 0.0)
    assert z["negative_periods"].tolist() == [math.inf, math.inf]
    s = [z["next_positive"][0], z["next_negative"][0]]
    out = []
    for t in (0, 1, 100):
        cnt, s[0], s[1] = O.advance_clock(s[0], s[1], z["periods"][0], z["negative_periods"][0], t)
        out.append(cnt)
    assert out == [None, 0, 0]
    assert s[1] == math.inf and s[0] == 3.0


def test_negative_index_u128_identity():
    rng = np.random.default_rng(5)
    vals = [0, 1, 2, 17, 199, 2**31 - 1, 2**32 + 3, U64]
    vals += [int(x) for x in rng.integers(0, 2**63, 20)]
    for m in (1, 2, 7, 17, 199, 200_000, 2**32 - 5, U64):
        for e in vals[:8]:
            for t in vals[:6]:
                for p in (0, 1, 5, U64):
                    got = O.lib().orc_negative_index(e, t, p, m)
                    assert got == ((e % m) * ((t % m + p % m + 1) % m)) % m
    x = U64
    assert O.lib().orc_negative_index(x, x, x, 17) == ((x % 17) * ((x % 17 + x % 17 + 1) % 17)) % 17


def test_negative_windows():
    assert O.lib().orc_effective_negative_range(7, 100) == 100
    assert O.lib().orc_effective_negative_range(7, 300) == 200
    assert O.lib().orc_effective_negative_range(299, 300) == 299
    assert O.lib().orc_effective_negative_range(300, 300) == 300
    W = O.lib().orc_window_negative_index
    for rank, start in [(0, 0), (150, 50), (299, 100)]:
        assert W(0, 0, 0, rank, 300, 200) == start
        assert W(199, 0, 0, rank, 300, 200) == start + 199
    assert W(0, 0, 0, 50, 100, 7) == 47
    for rank in (0, 50, 99):
        assert W(99, 0, 0, rank, 100, 100) == 99
    x = U64
    assert W(x, x, x, 0, 300, 199) == ((x % 199) * ((2 * (x % 199) + 1) % 199)) % 199


def test_adaptation_due_cadence():
    due = [t for t in range(200) if O.lib().orc_adaptation_due(1, 1, 128, 200_000, 10**6, t, 200)]
    assert due == list(range(10, 101, 10))
    assert not O.lib().orc_adaptation_due(0, 1, 128, 10, 100, 10, 200)
    assert not O.lib().orc_adaptation_due(1, 0, 128, 10, 100, 10, 200)
    assert not O.lib().orc_adaptation_due(1, 1, 0, 10, 100, 10, 200)
    assert not O.lib().orc_adaptation_due(1, 1, 128, 100, 100, 10, 200)


# ---------------------------------------------------------------- :1830-1882
def test_scale_resolution_and_forces():
    R = O.lib().orc_resolved_negative_scale
    assert R(200_000, 1025, 0, 0.0, 1) == 1.0
    assert R(100, 200, 0, 0.0, 1) == f32(0.5 ** 0.75)
    assert R(200_000, 1025, 1, 3.0, 1) == 3.0
    assert R(7, 100, 0, 0.0, 0) == 1.0
    assert R(7, 400, 0, 0.0, 0) == 0.5
    assert R(7, 400, 1, 0.0, 0) == 0.0
    src = np.array([1.0, 0.0], np.float32)
    tgt = np.zeros(2, np.float32)
    F = O.lib().orc_accumulate_force

    def force(source, target, gamma, scale, small, negative, start=None):
        g = np.zeros(2, np.float32) if start is None else start.copy()
        F(g, np.ascontiguousarray(source), np.ascontiguousarray(target), 2, 1.0, 1.0, gamma,
          scale, small, negative)
        return g

    assert abs(force(src, tgt, 1.0, 3.0, 1, 0)[0] + 0.5) < 1e-6
    assert abs(force(src, tgt, 1.0, 3.0, 0, 0)[0] + 1.0) < 1e-6
    small = force(src, tgt, 1.0, 3.0, 1, 1)
    assert abs(small[0] - math.tanh(f32(1.0) / 1.001)) < 1e-6
    assert abs(force(src, tgt, 1.0, 3.0, 0, 1)[0] - math.tanh(3.0 / 1.001)) < 1e-6
    before = small.copy()
    small = force(src, src, 1.0, 1.0, 1, 1, small)
    small = force(src, tgt, 0.0, 1.0, 1, 1, small)
    assert (small == before).all()


def test_force_clip_closed_form_large_gamma():
    # Repulsion for gamma > 16 saturates at 16 * tanh(norm / gamma) / norm.
    src = np.array([0.3, -0.4], np.float32)
    tgt = np.zeros(2, np.float32)
    g = np.zeros(2, np.float32)
    a, b, gamma = 1.5, 0.9, 20.0
    O.lib().orc_accumulate_force(g, src, tgt, 2, a, b, gamma, 1.0, 0, 1)
    q = 0.25
    c = 2 * gamma * b / ((0.001 + q) * (1 + a * q ** b))
    norm = math.sqrt(c * c * q)
    expect = c * 16 * math.tanh(norm / gamma) / norm
    np.testing.assert_allclose(g, expect * src, rtol=2e-6)


# ---------------------------------------------------------------- :1385-1405
def test_force_ratio_known_answers():
    assert O.force_ratio([], []) == 1.0
    assert O.force_ratio([0.0, 1.0], [1.0, 0.0]) == 1.0
    assert abs(O.force_ratio([2.0] * 3, [1.0] * 3) - 2.0) < 1e-12
    assert abs(O.force_ratio([1.0, 1.0, 1.0, math.e ** 4], [1.0] * 4) - math.exp(0.7)) < 1e-12
    assert O.force_ratio([math.nan], [1.0]) < 0
    assert O.force_ratio([1.0], [-1.0]) < 0
    # Against numpy's linear quantiles on random data.
    rng = np.random.default_rng(3)
    g = rng.lognormal(size=128)
    l = rng.lognormal(size=128)
    logs = np.log(g) - np.log(l)
    lo, hi = np.quantile(logs, [0.1, 0.9])
    assert abs(O.force_ratio(g, l) - math.exp(np.clip(logs, lo, hi).mean())) < 1e-12


def test_scale_adaptation_known_answers():
    s = O.ScaleAdaptation(0.5)
    assert s.observe(2.0) == 0.5
    assert abs(s.observe(100.0) - f32(0.5 * 1.1 ** 0.1)) < 1e-6
    before = f32(s.current)
    assert abs(s.observe(0.001) - before * f32(0.9) ** f32(0.1)) < 1e-6
    z = O.ScaleAdaptation(0.0)
    for r in (1.0, 2.0, 0.5, 2.2e-308, 1.7e308):
        assert z.observe(r) == 0.0


# ---------------------------------------------------------------- :1885-1919
def test_lazy_adam_step():
    alpha = f32(O.adam_schedule(10, 1.0, False)[0][0])
    h = np.zeros((1, 1), np.float32)
    ep = O.Epoch(head=h, tail=h, offsets=[0, 0], indices=[], source_order=[0], negative_order=[0],
                 source_ranks=[0], periods=[], next_positive=[], negative_periods=[],
                 next_negative=[], updates=np.array([[2.0]], np.float32), m=np.zeros((1, 1)),
                 v=np.zeros((1, 1)), epoch=0, a=1, b=1, gamma=1, scale=1, negative_range=1,
                 alpha=alpha, correction1=f32(1 - 0.9), correction2=f32(1 - f32(0.99)))
    ep.args.correction1 = f32(1.0 - float(np.float32(0.9)))
    ep.args.correction2 = f32(1.0 - float(np.float32(0.99)))
    ep.adam_apply()
    assert abs(ep.head[0, 0] - alpha * 2.0 / 2.0001) < 1e-6
    assert ep.updates[0, 0] == 0.0
    saved = (ep.head.copy(), ep.m.copy(), ep.v.copy())
    ep.adam_apply()  # zero gradient: lazy step leaves everything untouched
    assert (ep.head == saved[0]).all() and (ep.m == saved[1]).all() and (ep.v == saved[2]).all()


# ------------------------------------------------------------- projection sort
def test_radix_key_monotone_and_signed_zero():
    vals = np.array([-np.inf, -3.5, -1e-30, -0.0, 0.0, 1e-30, 2.0, np.inf], np.float32)
    keys = [O.radix_key(float(v)) for v in vals]
    assert keys == sorted(keys)
    assert O.radix_key(-0.0) == O.radix_key(0.0)


def test_stable_argsort_matches_numpy_stable():
    rng = np.random.default_rng(11)
    for n in (1, 2, 7, 1000, 65537):
        v = rng.normal(size=n).astype(np.float32)
        v[rng.integers(0, n, n // 3)] = 0.25  # many ties
        if n > 3:
            v[0], v[1] = -0.0, 0.0
        order, s = O.stable_argsort(v)
        keys = np.array([O.radix_key(float(x)) for x in v], np.uint64) if n < 2000 else None
        expect = np.argsort(np.where(v == 0, np.float32(0), v), kind="stable")
        assert (order == expect).all()
        assert (s.view(np.uint32) == v[order].view(np.uint32)).all()
        if keys is not None:
            assert (np.diff(keys[order].astype(np.int64)) >= 0).all()
    bad = np.array([0.0, 1.0, np.nan, np.inf], np.float32)
    with pytest.raises(FloatingPointError, match="row 2"):
        O.stable_argsort(bad)


def test_ranks():
    rng = np.random.default_rng(2)
    v = rng.normal(size=500).astype(np.float32)
    order, s = O.stable_argsort(v)
    ranks = O.inverse_ranks(order)
    assert (order[ranks] == np.arange(500)).all()
    q = np.concatenate([rng.normal(size=50), s[:10], [100.0, -100.0]]).astype(np.float32)
    lb = O.lower_bound_ranks(s, q)
    assert (lb == np.minimum(np.searchsorted(s, q, side="left"), 499)).all()


# ------------------------------------------------------------------------ RNG
def test_chacha_round_function_rfc7539_vectors():
    # RFC 7539 section 2.3.2 (ChaCha20 block, counter 1).
    key = np.frombuffer(bytes(range(32)), np.uint32)
    state = np.array([0x61707865, 0x3320646E, 0x79622D32, 0x6B206574, *key, 1, 0x09000000,
                      0x4A000000, 0], np.uint32)
    expect = [0xE4E7F110, 0x15593BD1, 0x1FDD0F50, 0xC47120A3, 0xC7F4D1C7, 0x0368C033,
              0x9AAA2204, 0x4E6CD4C3, 0x466482D2, 0x09AA9F07, 0x05D7C214, 0xA2028BD9,
              0xD19C12B5, 0xB94E16DE, 0xE883D0CB, 0x4E3C50A2]
    assert O.chacha_block(state, 20).tolist() == expect
    # Zero key, zero nonce, block 0: the original ChaCha20 keystream vector.
    zero = np.array([0x61707865, 0x3320646E, 0x79622D32, 0x6B206574] + [0] * 12, np.uint32)
    first = O.chacha_block(zero, 20)
    ks = first.astype("<u4").tobytes()
    assert ks[:16].hex() == "76b8e0ada0f13d90405d6ae55386bd28"


def test_rng_stream_conventions_self_consistent():
    r1, r2 = O.StdRng(42), O.StdRng(42)
    words = [r1.u32() for _ in range(130)]
    lo, hi = r2.u32(), r2.u32()
    r3 = O.StdRng(42)
    assert r3.u64() == (hi << 32) | lo
    # Refill continuity: 130 words cross two 64-word buffers.
    r4 = O.StdRng(42)
    assert [r4.u32() for _ in range(130)] == words
    assert O.StdRng(43).u32() != words[0]
    f = O.StdRng(7).f64()
    assert 0.0 <= f < 1.0


def test_range_and_shuffle_distribution():
    r = O.StdRng(1)
    draws = np.array([r.range_u32(6) for _ in range(60000)])
    counts = np.bincount(draws, minlength=6)
    assert draws.max() < 6 and (abs(counts - 10000) < 500).all()
    r = O.StdRng(9)
    x = np.arange(1000, dtype=np.int32)
    r.shuffle(x)
    assert sorted(x.tolist()) == list(range(1000)) and not (x == np.arange(1000)).all()
    # Position-uniformity of the Durstenfeld shuffle on 4 items (24 perms).
    r = O.StdRng(5)
    seen = {}
    for _ in range(24000):
        a = np.arange(4, dtype=np.int32)
        r.shuffle(a)
        seen[tuple(a)] = seen.get(tuple(a), 0) + 1
    assert len(seen) == 24 and max(seen.values()) < 1300 and min(seen.values()) > 700
    v = [r.range_f32(-10.0, 10.0) for _ in range(1000)]
    assert min(v) >= -10.0 and max(v) < 10.0


def test_projection_direction_unit_and_degenerate():
    for d in (1, 2, 3, 16):
        v = O.StdRng(42).direction(d)
        assert abs(float(np.linalg.norm(v.astype(np.float64))) - 1.0) < 1e-6


def test_fit_curve_default_matches_umap_learn():
    a, b = O.fit_curve(0.1, 1.0)
    # umap-learn's published defaults for min_dist=0.1, spread=1.0.
    assert abs(a - 1.5769434) < 2e-4 and abs(b - 0.8950609) < 2e-4


def test_vectorised_clock_replay_matches_advance_clock():
    from tests.gpu_helpers import advance_graph, random_graph

    g = random_graph(50, 50, 4, 3, 30)
    v = advance_graph(g, 13)
    npos, nneg = g["next_positive"].copy(), g["next_negative"].copy()
    for t in range(13):
        for e in range(len(npos)):
            _, npos[e], nneg[e] = O.advance_clock(npos[e], nneg[e], g["periods"][e],
                                                  g["negative_periods"][e], t)
    assert (npos == v["next_positive"]).all() and (nneg == v["next_negative"]).all()


def test_oracle_cross_entropy_matches_numpy_fsum():
    """The FP64 objective against an independent vectorised restatement
    summed exactly (math.fsum), and its closed form on one edge."""
    import math

    import numpy as np

    from oracle import oracle as O
    from tests.gpu_helpers import random_graph

    g = random_graph(400, 400, 7, 5, 50)
    rng = np.random.default_rng(1)
    head = rng.normal(0, 3, (400, 2)).astype(np.float32)
    head[10] = head[int(g["indices"][g["offsets"][10]])]   # a coincident pair (clipped term)
    w = rng.uniform(0.05, 1.0, len(g["indices"])).astype(np.float32)
    a, b = float(np.float32(1.5769434)), float(np.float32(0.8950608))
    rows = np.repeat(np.arange(400), np.diff(g["offsets"]))
    diff = head[rows].astype(np.float64) - head[g["indices"]].astype(np.float64)
    q = (diff ** 2).sum(1)
    p = np.where(q > 0, a * np.power(q, b, where=q > 0, out=np.zeros_like(q)), 0.0)
    phi, rest = 1 / (1 + p), p / (1 + p)
    w64 = w.astype(np.float64)
    terms = -w64 * np.log(np.maximum(phi, 1e-4)) - (1 - w64) * np.log(np.maximum(rest, 1e-4))
    want = math.fsum(terms.tolist())
    got = O.cross_entropy(head, None, g, w, a, b)
    assert abs(got - want) <= 1e-12 * abs(want)
    # one edge at unit distance with w = 1: -ln(1 / (1 + a))
    one = {"offsets": np.array([0, 1, 1], np.uint32), "indices": np.array([1], np.uint32)}
    pts = np.array([[0, 0], [1, 0]], np.float32)
    assert abs(O.cross_entropy(pts, None, one, np.ones(1, np.float32), a, b) - math.log1p(a)) < 1e-14
