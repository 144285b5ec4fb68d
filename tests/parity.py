"""Point-by-point parity checks between HIP results and the FP64 oracle.

The bar is a relative error of at most 1e-5 of the field's largest magnitude
at every face or cell.  The transport's discrete choices (MacCormack donor
sets, the correction's range test, the solid cell look-ups) can go the other
way in FP32 where a trace coordinate lies within rounding distance of its
threshold; ``oracle.decisions`` gives each point's distance to its nearest
threshold.  A point whose margin is at most ``eps`` cells is "fragile": it
must still lie inside the hull of every donor set it could have chosen, and
the fragile points are few (``max_fragile``).  All other points meet the
1e-5 bar.  ``eps`` is derived from the FP32 position error: positions of at
most n cells carry an error below n * 2^-23 ~ 1.2e-7 n cells per operation,
and a trace has fewer than 50 operations, so 1e-4 cells covers n <= 64 with a
margin of over 2x.
"""
import numpy as np

TOL = 1e-5


def to_np(field, shape):
    t = field.tensor if hasattr(field, "tensor") else field
    return t.detach().cpu().double().numpy().reshape(shape)


def pointwise(got, ref, margin=None, eps=1e-4, tol=TOL, hull=None, max_fragile=2e-3, label=""):
    """Assert the bar described in the module docstring; returns a report."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    assert got.shape == ref.shape, (label, got.shape, ref.shape)
    assert np.isfinite(got).all(), label
    scale = max(float(np.abs(ref).max()), 1e-30)
    err = np.abs(got - ref) / scale
    fragile = np.zeros(got.shape, bool) if margin is None else (np.asarray(margin) <= eps)
    firm = ~fragile
    rep = {"label": label, "max_err": float(err.max()), "max_firm": float(err[firm].max()) if firm.any() else 0.0,
           "fragile": int(fragile.sum()), "fragile_frac": float(fragile.mean()),
           "max_fragile_err": float(err[fragile].max()) if fragile.any() else 0.0}
    bad = firm & (err > tol)
    if bad.any():
        idx = np.argwhere(bad)[:5].tolist()
        rep["bad"] = int(bad.sum())
        rep["bad_examples"] = [(i, float(err[tuple(i)]),
                                float(np.asarray(margin)[tuple(i)]) if margin is not None else None) for i in idx]
    assert not bad.any(), rep
    assert rep["fragile_frac"] <= max_fragile, rep
    if hull is not None and fragile.any():
        lo, hi = hull
        slack = tol * scale
        out = fragile & ((got < lo - slack) | (got > hi + slack))
        rep["outside_hull"] = int(out.sum())
        assert not out.any(), rep
    return rep


def gauge_free(p, fluid=None):
    """Pressure with its (fluid) mean removed: the Neumann problem fixes p up
    to a constant."""
    p = np.asarray(p, np.float64)
    if fluid is None:
        return p - p.mean()
    return np.where(fluid, p - p[fluid].mean(), 0.0)
