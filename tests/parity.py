"""Comparison helpers shared by the parity tests.

The bar: NaN masks identical; every finite output within
|a-b| / max(1, |a|, |b|) <= tol (the reference corpus comparator,
benchmarks/talib_compare_helpers.py:594-640); and a minimum fraction of
outputs bit-identical to the oracle's float32 result.
"""
import numpy as np


def compare(got, want, tol=1e-5, min_exact=0.999, name=""):
    got = np.asarray(got, dtype=np.float32)
    want = np.asarray(want, dtype=np.float32)
    assert got.shape == want.shape, f"{name}: shape {got.shape} vs {want.shape}"
    gn, wn = np.isnan(got), np.isnan(want)
    if not np.array_equal(gn, wn):
        idx = np.flatnonzero(gn != wn)[:10]
        raise AssertionError(f"{name}: NaN mask differs at {idx.tolist()} "
                             f"got={got.ravel()[idx].tolist()} want={want.ravel()[idx].tolist()}")
    ok = ~wn
    if not ok.any():
        return {"exact": 1.0, "max_rel": 0.0, "n": 0}
    g = got[ok].astype(np.float64)
    w = want[ok].astype(np.float64)
    inf_mismatch = np.isinf(g) != np.isinf(w)
    assert not inf_mismatch.any(), f"{name}: Inf mask differs"
    winf = np.isinf(w)
    assert np.array_equal(g[winf], w[winf]), f"{name}: Inf sign differs"
    fin = np.isfinite(w)
    rel = np.abs(g[fin] - w[fin]) / np.maximum(1.0, np.maximum(np.abs(g[fin]), np.abs(w[fin])))
    exact = float(np.mean(got[ok].view(np.uint32) == want[ok].view(np.uint32)))
    max_rel = float(rel.max()) if rel.size else 0.0
    if max_rel > tol:
        i = int(np.argmax(rel))
        raise AssertionError(f"{name}: max rel err {max_rel:.3g} > {tol} "
                             f"(got {g[fin][i]!r}, want {w[fin][i]!r})")
    if exact < min_exact:
        raise AssertionError(f"{name}: only {exact:.6f} of outputs bit-identical "
                             f"(need {min_exact}); max rel {max_rel:.3g}")
    return {"exact": exact, "max_rel": max_rel, "n": int(ok.sum())}
