"""Point-by-point parity checks between HIP results and the FP64 oracle.

The bar is an error of at most 1e-5 of the field's largest magnitude at every
face or cell.  The transport's discrete choices (MacCormack donor sets, the
correction's range test, the solid cell look-ups) can go the other way in FP32
where a trace coordinate lies within rounding distance of its threshold;
``oracle.decisions`` gives every point an allowance: the largest change of
its value that taking the other branch of each such near decision can make
(0 where no near decision changes anything).  A point may differ from the
oracle by its allowance plus the 1e-5 bar, and the points that use their
allowance (those beyond the plain 1e-5 bar) must be few: at most
``max_exempt`` of the field (0.1 %).  How many points could use one (their
allowance exceeds the bar, "decisive") depends on the flow and is reported.

The rounding distance ``eps`` is derived from the FP32 position error: trace
coordinates of at most n cells carry an error below n * 2^-23 ~ 1.2e-7 n
cells per operation, and a trace has fewer than 50 operations, so 1e-4 cells
covers n <= 64 with a margin of over 2x.
"""
import numpy as np

TOL = 1e-5
EPS = 1e-4


def to_np(field, shape):
    t = field.tensor if hasattr(field, "tensor") else field
    return t.detach().cpu().double().numpy().reshape(shape)


def pointwise(got, ref, allowance=None, tol=TOL, max_exempt=1e-3, label=""):
    """Assert the bar described in the module docstring; returns a report."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    assert got.shape == ref.shape, (label, got.shape, ref.shape)
    assert np.isfinite(got).all(), label
    scale = max(float(np.abs(ref).max()), 1e-30)
    err = np.abs(got - ref)
    allow = np.zeros(got.shape) if allowance is None else np.asarray(allowance, np.float64)
    firm = allow == 0
    decisive = allow > tol * scale
    rep = {"label": label, "max_err": float(err.max() / scale),
           "max_firm": float(err[firm].max() / scale) if firm.any() else 0.0,
           "decisive": int(decisive.sum()), "decisive_frac": float(decisive.mean()),
           "max_decisive_err": float(err[decisive].max() / scale) if decisive.any() else 0.0,
           "exempt": int((err > tol * scale).sum())}
    bad = err > allow + tol * scale
    if bad.any():
        idx = np.argwhere(bad)[:5].tolist()
        rep["bad"] = int(bad.sum())
        rep["bad_examples"] = [(i, float(err[tuple(i)] / scale), float(allow[tuple(i)] / scale)) for i in idx]
    assert not bad.any(), rep
    assert rep["exempt"] <= max_exempt * err.size, rep
    return rep


def laplacian_condition(shape):
    """Condition number of the 7-point Neumann Laplacian on a box of
    ``shape`` cells (nonzero spectrum): 4*3 / (4 sin^2(pi / (2 n_max)))."""
    import math
    return 3.0 / math.sin(math.pi / (2 * max(shape))) ** 2


def potential_tol(shape):
    """Bound for a pressure (impulse) field computed in FP32: the forward
    error of any FP32 solve of L q = b is about eps32 * cond(L) relative to q
    (eps32 = 2^-24; 7.4e-5 at 32^3).  The corrected velocity, a gradient of
    q that the smooth error barely moves, keeps the 1e-5 bar."""
    return max(TOL, 2.0 ** -24 * laplacian_condition(shape))


def gauge_free(p, fluid=None):
    """Pressure with its (fluid) mean removed: the Neumann problem fixes p up
    to a constant."""
    p = np.asarray(p, np.float64)
    if fluid is None:
        return p - p.mean()
    return np.where(fluid, p - p[fluid].mean(), 0.0)
