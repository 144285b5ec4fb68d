"""Pins the oracle's MacCormack branches (velocity: predictor, forward trace,
reverse sample, 8-donor clamp and the 1/2 + 1/2 blend with the
semi-Lagrangian value; scalars: the unblended clamped correction) against
cases whose answer is written down by hand.

Reference semantics: shaders/transport_face.glsl:1-20,
shaders/correct_velocity_fused.glsl:5-27, shaders/transport_scalar.glsl:1-19,
shaders/correct_scalar.glsl:1-21.  The reference has no CPU oracle for these
passes, so each case reduces to one dimension where every stage is a linear
interpolation that can be computed by a few lines of independent code:

* U = c constant and W = 0 make every trace move x by exactly c*dt (a
  displacement of s cells) whatever the point;
* V = f(x) depends on x only, so the y part of the trace (driven by V)
  changes no sampled value, and only decides whether the correction's range
  test passes at the y ends.

A wrong blend factor, a reverse sample taken at the departure instead of the
arrival, a clamp over the arrival's donors or a sign error in either trace
changes the values at the faces where the step in f sits.
"""
import math

import numpy as np
import pytest

from oracle import mac, transport as tr

NX, NY, NZ = 12, 6, 5
DT = 0.05


def _lerp_face(f, q):
    """Clamp-to-edge linear sample of a 1-D face row (coordinate clamped)."""
    q = min(max(q, 0.0), len(f) - 1)
    b = math.floor(q)
    w = q - b
    return (1 - w) * f[b] + w * f[min(b + 1, len(f) - 1)]


def _lerp_closed(f, q):
    """Closed-box linear sample of a 1-D cell row (indices clamped)."""
    b = math.floor(q)
    w = q - b
    at = lambda i: f[min(max(i, 0), len(f) - 1)]
    return (1 - w) * at(b) + w * at(b + 1)


def _shear_case(s):
    g = mac.Grid((NX, NY, NZ))
    hx, hy = g.cell_size[0], g.cell_size[1]
    c = s * hx / DT
    f = np.array([0.0, 0.1, 0.2, 1.0, 1.0, 0.9, -0.5, -0.4, 0.3, 0.3, 0.0, -0.2])
    u = np.full(g.face_array_shape(0), c)
    v = np.broadcast_to(f, g.face_array_shape(1)).copy()
    w = np.zeros(g.face_array_shape(2))
    return g, (u, v, w), f, c, hy


@pytest.mark.parametrize("s", [0.3, 0.55, 0.8])
def test_velocity_maccormack_hand_computed_1d(s):
    g, vel, f, c, hy = _shear_case(s)
    pred = tr.face_predictor(vel, g, DT)
    out = tr.advect_velocity(vel, g, DT, True, pred)
    # U constant, W zero: exact, predictor included
    assert np.abs(out[0] - c).max() < 1e-12 and np.abs(out[2]).max() < 1e-12
    assert np.abs(pred[0] - c).max() < 1e-12
    # hand computation for the V faces (x index i, y face index j)
    low = np.array([_lerp_face(f, i - s) for i in range(NX)])
    assert np.abs(pred[1] - low[None, None, :]).max() < 1e-12
    for j in range(NY + 1):
        for i in range(NX):
            # y of the back-trace: V at the midpoint x - s/2; forward: V at x + s/2
            qy_dep = j - DT * _lerp_face(f, i - s / 2) / hy
            qy_arr = j + DT * _lerp_face(f, i + s / 2) / hy
            inside = (0 <= i - s and i + s <= NX - 1 and 0 <= qy_dep <= NY and 0 <= qy_arr <= NY)
            expect = low[i]
            if inside:
                rev = _lerp_face(low, i + s)
                corr = low[i] + 0.5 * (f[i] - rev)
                b = math.floor(i - s)
                donors = (f[max(b, 0)], f[min(b + 1, NX - 1)])
                expect = 0.5 * low[i] + 0.5 * min(max(corr, min(donors)), max(donors))
            got = out[1][:, j, i]
            assert np.abs(got - expect).max() < 1e-12, (s, i, j, got[0], expect)


def test_velocity_maccormack_blend_and_clamp_are_active():
    """The case exercises both halves of the rule: the blended value differs
    from the semi-Lagrangian one, and the clamp cuts the correction."""
    s = 0.3
    g, vel, f, c, hy = _shear_case(s)
    pred = tr.face_predictor(vel, g, DT)
    out = tr.advect_velocity(vel, g, DT, True, pred)[1][2, 3]
    sl = tr.advect_velocity(vel, g, DT, False)[1][2, 3]
    corr_full = []
    for i in range(1, NX - 1):
        low = _lerp_face(f, i - s)
        rev = _lerp_face([_lerp_face(f, k - s) for k in range(NX)], i + s)
        corr_full.append(low + 0.5 * (f[i] - rev))
    clamped = [i for i, cv in zip(range(1, NX - 1), corr_full)
               if not min(f[i - 1], f[i]) <= cv <= max(f[i - 1], f[i])]
    assert clamped, "no face hits the clamp"
    assert np.abs(out - sl).max() > 1e-3


def test_velocity_maccormack_keeps_affine_and_constant_flows():
    """Constant flow and a field linear in the transported direction stay
    exact (trilinear sampling reproduces them, so the correction vanishes),
    and a shear U = y is stationary, with MacCormack on."""
    g = mac.Grid((8, 7, 6), (1.0, 0.9, 1.2))
    const = tuple(np.full(g.face_array_shape(a), v) for a, v in enumerate((.3, -.2, .1)))
    out = tr.advect_velocity(const, g, .05, True, tr.face_predictor(const, g, .05))
    for o, v in zip(out, (.3, -.2, .1)):
        assert np.abs(o - v).max() < 1e-12
    (px, py, pz), _ = tr.face_positions(g, 0)
    shear = (py - .45, np.zeros(g.face_array_shape(1)), np.zeros(g.face_array_shape(2)))
    out = tr.advect_velocity(shear, g, .05, True, tr.face_predictor(shear, g, .05))
    for a, b in zip(shear, out):
        assert np.abs(a - b).max() < 1e-12


def test_scalar_maccormack_hand_computed_1d():
    g = mac.Grid((NX, NY, NZ))
    s = 0.35
    c = s * g.cell_size[0] / DT
    vel = (np.full(g.face_array_shape(0), c), np.zeros(g.face_array_shape(1)), np.zeros(g.face_array_shape(2)))
    rho = np.array([0.0, 0.0, 1.0, 1.0, 0.2, 0.9, 0.1, 0.0, 0.5, 0.5, 0.6, 0.0])
    field = np.broadcast_to(rho, g.array_shape).copy()
    out = tr.advect_scalars([field], vel, g, DT, maccormack=True)[0]
    pred = [_lerp_closed(rho, i - s) for i in range(NX)]
    for i in range(NX):
        expect = pred[i]
        if 0 <= i - s and i + s <= NX - 1:
            rev = _lerp_closed(pred, i + s)
            corr = pred[i] + 0.5 * (rho[i] - rev)
            b = math.floor(i - s)
            donors = (rho[max(b, 0)], rho[min(b + 1, NX - 1)])
            expect = min(max(corr, min(donors)), max(donors))
        assert np.abs(out[..., i] - expect).max() < 1e-12, (i, out[0, 0, i], expect)
    # the predictor pass alone is the semi-Lagrangian value
    assert np.abs(tr.transport_scalar(field, vel, g, DT)[0, 0] - pred).max() < 1e-12
