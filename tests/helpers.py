"""Shared helpers for the parity tests."""
import numpy as np


def rel_close(a, b, rtol=1e-12, scale=None):
    """|a - b| <= rtol * s elementwise, s = max(|a|, |b|) or the given scale."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    s = np.maximum(np.abs(a), np.abs(b)) if scale is None else np.broadcast_to(scale, a.shape)
    bad = ~((np.abs(a - b) <= rtol * s) | ((a == b) | (np.isnan(a) & np.isnan(b))))
    return bad


def assert_rel(a, b, rtol=1e-12, scale=None, what=""):
    bad = rel_close(a, b, rtol, scale)
    if bad.any():
        idx = np.argwhere(bad)[:5]
        a = np.asarray(a)
        b = np.asarray(b)
        raise AssertionError(f"{what}: {bad.sum()} of {bad.size} entries differ beyond {rtol}; first {idx.tolist()}: "
                             f"{[a[tuple(i)] for i in idx]} vs {[b[tuple(i)] for i in idx]}")


def stats_tuple(st):
    return np.array([st.n, st.min, st.max, st.sum, st.sum2, st.n_finite])
