"""CPU checks of the C ABI: the library loads, exports every declared symbol,
validates its inputs before touching a device, and its host RNG replay of
the per-epoch orders matches the oracle's."""
import re
from pathlib import Path

import numpy as np
import pytest

from cuscan_amd import _lib, layout
from oracle import oracle as O

HEADER = Path(__file__).resolve().parents[1] / "include" / "umaphip.h"


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(umaphip_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_symbols()
    assert len(names) >= 15
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(_lib.SIGNATURES), "ctypes signatures drifted from the header"
    assert _lib.lib().umaphip_version() >= 100


def test_invalid_params_rejected_before_device_work():
    p = layout.LayoutParams(epochs=10, a=1.0, b=1.0, optimizer="adam")
    c = p.to_c()
    c.negative_selection_range = 0
    with pytest.raises(ValueError, match="negative_selection_range"):
        _lib.check(_lib.lib().umaphip_orders_at_epoch(4, 4, c, 1, 0, None, None))
    with pytest.raises(ValueError, match="optimizer"):
        layout.LayoutParams(epochs=1, a=1, b=1, optimizer="sgd").to_c()


def _oracle_orders(h, n, seed, momentum, shared, hard, epoch):
    rng = O.StdRng(seed)
    src = np.arange(h, dtype=np.int32)
    alias = shared and momentum and not hard
    neg = src if alias else np.arange(n, dtype=np.int32)
    live = not alias and not hard
    for _ in range(epoch):
        rng.shuffle(src)
        if live:
            rng.shuffle(neg)
    if hard and (epoch > 0 or h > 1024):
        neg = np.arange(n, dtype=np.int32)
    return src, (src if alias else neg)


@pytest.mark.parametrize("h,n,shared,momentum,q", [
    (300, 300, True, False, 200_000),     # fit adam, global
    (300, 300, True, True, 200_000),      # fit momentum, aliased
    (70, 500, False, False, 200_000),     # transform, global
    (2000, 2000, True, True, 500),        # fit momentum, projection (retired)
    (1500, 1500, True, False, 200_000),   # 1500 < Q
    (70_000, 70_000, True, False, 200_000),  # spans above 65,535: one draw per element
])
def test_orders_match_oracle_rng(h, n, shared, momentum, q):
    n_eff = h if shared else n
    p = layout.LayoutParams(epochs=10, a=1, b=1, optimizer="momentum" if momentum else "adam",
                            negative_selection_range=q, seed=1234)
    for epoch in (0, 1, 2, 7):
        src, neg = layout.orders_at_epoch(h, n, p, shared=shared, epoch=epoch)
        es, en = _oracle_orders(h, n_eff, 1234, momentum, shared, q < n_eff, epoch)
        assert (src == es).all(), epoch
        assert (neg == en).all(), epoch
