"""Shared builders for the GPU parity tests (graphs, tensors, metrics)."""
import numpy as np

from oracle import oracle as O


def random_graph(n_rows, n_cols, k, seed, epochs, rate=5.0, symmetric=True):
    """A kNN-like positive graph, compacted by the oracle (A1)."""
    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(n_rows), k)
    if symmetric:
        assert n_rows == n_cols
        cols = (rows + rng.integers(1, n_cols, rows.size)) % n_cols
        w = rng.uniform(0.02, 1.0, rows.size).astype(np.float32)
        rows, cols, w = (np.concatenate([rows, cols]), np.concatenate([cols, rows]),
                         np.concatenate([w, w]))
    else:
        cols = rng.integers(0, n_cols, rows.size)
        w = rng.uniform(0.02, 1.0, rows.size).astype(np.float32)
    return O.compact_graph(n_rows, n_cols, rows, cols, w, epochs, rate)


def hub_graph(n, k, seed, epochs, rate=5.0, hubs=(7, 1234), hub_degree=1500, isolated=(3, 4, 2999)):
    """random_graph plus hub rows joined to `hub_degree` others (so a hub's
    CSR row spans many edge units) and rows with no edge at all."""
    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(n), k)
    cols = (rows + rng.integers(1, n, rows.size)) % n
    extra_r = np.concatenate([np.full(hub_degree, h) for h in hubs])
    extra_c = np.concatenate([rng.choice(np.setdiff1d(np.arange(n), [h]), hub_degree, replace=False)
                              for h in hubs])
    rows, cols = np.concatenate([rows, extra_r]), np.concatenate([cols, extra_c])
    keep = ~np.isin(rows, isolated) & ~np.isin(cols, isolated)
    rows, cols = rows[keep], cols[keep]
    w = rng.uniform(0.02, 1.0, rows.size).astype(np.float32)
    rows, cols, w = np.concatenate([rows, cols]), np.concatenate([cols, rows]), np.concatenate([w, w])
    return O.compact_graph(n, n, rows, cols, w, epochs, rate)


def advance_graph(g, epochs_done):
    """Clocks as they stand after `epochs_done` epochs (data independent);
    vectorised restatement of advance_clock (no fused multiply-adds)."""
    npos = g["next_positive"].copy()
    nneg = g["next_negative"].copy()
    per, nper = g["periods"], g["negative_periods"]
    for t in range(epochs_done):
        fire = npos <= float(t)
        k = np.trunc((float(t) - nneg[fire]) / nper[fire])
        npos[fire] = npos[fire] + per[fire]
        nneg[fire] = nneg[fire] + k * nper[fire]
    out = dict(g)
    out["next_positive"], out["next_negative"] = npos, nneg
    return out


def row_relative_error(g, g64):
    """max over rows with ||g64|| > 1e-6 * max of ||g - g64|| / ||g64||."""
    g = np.asarray(g, np.float64)
    n64 = np.linalg.norm(g64, axis=1)
    if n64.max() == 0.0:
        return 0.0 if not np.any(g) else float("inf")
    keep = n64 > 1e-6 * n64.max()
    return float((np.linalg.norm(g - g64, axis=1)[keep] / n64[keep]).max())


def term_relative_error(g, g64, magnitude, rel=1e-5):
    """max over rows of (||g - g64|| - rel * ||g64||) / magnitude, with
    magnitude the row's sum of term norms (FP64 oracle): how far each row is
    from its FP64 value beyond `rel` of it, in units of the row's terms. A
    sum whose every term is within eps relative of its exact value, added
    without loss, stays within eps here whatever cancellation the row has."""
    g = np.asarray(g, np.float64)
    err = np.linalg.norm(g - g64, axis=1) - rel * np.linalg.norm(g64, axis=1)
    mag = np.asarray(magnitude, np.float64)
    if not np.any(mag > 0):
        return 0.0 if not np.any(g) else float("inf")
    if np.any((mag == 0) & (err > 0)):
        return float("inf")
    keep = mag > 0
    return float(max((err[keep] / mag[keep]).max(), 0.0))


def corrections(epoch):
    c1 = float(np.float32(1.0 - float(np.float32(0.9)) ** (epoch + 1)))
    c2 = float(np.float32(1.0 - float(np.float32(0.99)) ** (epoch + 1)))
    return c1, c2
