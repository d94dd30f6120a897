"""The oracle's compatibility optimizer against umap-learn's recorded
single-epoch frames (clocks 1e-5, Tausworthe states exact, serial
coordinates 2e-5), plus closed-form checks of the generator."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.classic_probes import A_PROBE, B_PROBE, cases, replay

CASES = cases()


@pytest.mark.parametrize("i", range(len(CASES)))
def test_oracle_classic_epoch_matches_umap_learn_frames(i):
    case, edge_count, frames = CASES[i]

    def run(st, epoch, alpha):
        head = st["head"]
        tail = head if st["shared"] else st["tail"]
        O.classic_epoch(head, tail, st["tails"], st["weights"], st["per"], st["npos"],
                        st["nper"], st["nneg"], st["rng"].reshape(-1), st["offsets"], st["ids"],
                        epoch=epoch, a=A_PROBE, b=B_PROBE, gamma=st["gamma"], alpha=alpha,
                        rate=st["rate"], shared=st["shared"])

    def coords(st, want_head, want_tail):
        # serial head order equals umap-learn's edge order for these probes
        np.testing.assert_allclose(st["head"], want_head, rtol=2e-5, atol=2e-5)
        if not st["shared"]:
            assert (st["tail"] == want_tail).all()

    replay(case, edge_count, frames, run, coords)


def test_tausworthe_components_closed_form():
    # Each component with zero high bits is a 32-bit Tausworthe (L'Ecuyer
    # 1996) step: ((s & mask) << l) ^ (((s << m) ^ s) >> r) within 32 bits.
    rng = np.random.default_rng(3)
    for _ in range(200):
        s = rng.integers(0, 2**32, 3, dtype=np.uint64)
        st = s.astype(np.int64).copy()
        O.taus_next(st)
        def comp(x, mask, l, m, r):
            x = int(x)
            return (((x & mask) << l) & 0xFFFFFFFF) ^ ((((x << m) & 0xFFFFFFFF) ^ x) >> r)
        want = [comp(s[0], 4294967294, 12, 13, 19), comp(s[1], 4294967288, 4, 2, 25),
                comp(s[2], 4294967280, 17, 3, 11)]
        assert list(st) == want
    # negative (high-bit) states shift arithmetically
    st = np.array([-(1 << 40) - 7, -5, -(1 << 62)], np.int64)
    hi = [int(x) for x in st]
    O.taus_next(st)
    def comp_signed(x, mask, l, m, r):
        return ((((x & mask) << l) & 0xFFFFFFFF) ^
                ((((x << m) & 0xFFFFFFFF) ^ x) >> r))  # python >> is arithmetic
    assert list(st) == [comp_signed(hi[0], 4294967294, 12, 13, 19),
                        comp_signed(hi[1], 4294967288, 4, 2, 25),
                        comp_signed(hi[2], 4294967280, 17, 3, 11)]
