"""Pins of the oracle's MacCormack advection (advect.py:11-137,
utils/util.py:20-159) by independent computations, not retyped formulas:

* sample_velocity_3d reproduces a linear velocity field (trilinear on the MAC
  staggering) away from the clamped boundary;
* the tricubic interpolation reproduces any polynomial of degree <= 3 per axis
  (it is the 4-point Lagrange interpolant) away from the clamped boundary;
* the RK3 backtrace moves a point by -dt/h U in a constant field and matches an
  independent FP64 RK3 of an analytic linear field;
* sample_minmax is the min / max of the 8 clamped samples at floor(p/h - st),
  gathered here with NumPy;
* MacCormack keeps a constant field (to FP32 rounding of the weights) and,
  with zero velocity, returns the field itself exactly; it equals the semi-Lagrangian value wherever the clamped
  corners coincide (min == max).
"""
import numpy as np

from oracle import oracle as O

F32 = np.float32


def _faces(n, fu, fv, fw):
    i = np.arange(n + 1, dtype=np.float64)
    c = np.arange(n, dtype=np.float64) + 0.5
    u = fu(i[:, None, None], c[None, :, None], c[None, None, :]) * np.ones((n + 1, n, n))
    v = fv(c[:, None, None], i[None, :, None], c[None, None, :]) * np.ones((n, n + 1, n))
    w = fw(c[:, None, None], c[None, :, None], i[None, None, :]) * np.ones((n, n, n + 1))
    return u.astype(F32), v.astype(F32), w.astype(F32)


LIN = (lambda x, y, z: 0.3 + 0.02 * x - 0.01 * y + 0.015 * z,
       lambda x, y, z: -0.2 + 0.01 * x + 0.02 * y - 0.005 * z,
       lambda x, y, z: 0.1 - 0.015 * x + 0.005 * y + 0.01 * z)


def test_sample_velocity_linear_exact():
    n = 16
    u, v, w = _faces(n, *LIN)
    rng = np.random.default_rng(0)
    p = rng.uniform(2.0, n - 2.0, (200, 3)).astype(F32)
    got = O.sample_velocity(u, v, w, p)
    want = np.stack([f(p[:, 0].astype(np.float64), p[:, 1], p[:, 2]) for f in LIN], axis=1)
    np.testing.assert_allclose(got, want, atol=2e-6)


def test_tricubic_reproduces_cubic_polynomials():
    rng = np.random.default_rng(1)
    g = (12, 10, 14)
    cu, cv, cw = (rng.normal(size=4) for _ in range(3))
    poly = lambda c, x: c[0] + c[1] * x + c[2] * x ** 2 / 10 + c[3] * x ** 3 / 100  # noqa: E731
    i, j, k = np.meshgrid(*(np.arange(s, dtype=np.float64) for s in g), indexing="ij")
    q = (poly(cu, i) * poly(cv, j) * poly(cw, k)).astype(F32)
    pts = np.stack([rng.uniform(1.0, s - 3.0, 300) for s in g], axis=1).astype(F32)
    got = O.tricubic(q, pts)
    p64 = pts.astype(np.float64)
    want = poly(cu, p64[:, 0]) * poly(cv, p64[:, 1]) * poly(cw, p64[:, 2])
    np.testing.assert_allclose(got, want, rtol=2e-4, atol=2e-4 * np.abs(q).max())


def test_backtrace_constant_and_linear_fields():
    n = 16
    dt, h = 0.01, 1.0 / n
    u, v, w = _faces(n, lambda x, y, z: 0.7 + 0 * x, lambda x, y, z: -0.4 + 0 * x, lambda x, y, z: 0.25 + 0 * x)
    rng = np.random.default_rng(2)
    p = rng.uniform(3.0, n - 3.0, (100, 3)).astype(F32)
    got = O.backtrace(u, v, w, p, dt, h)
    np.testing.assert_allclose(got, p - dt / h * np.array([0.7, -0.4, 0.25]), atol=2e-5)
    # linear field: FP64 RK3 of the analytic velocity (the interpolation is exact for it)
    u, v, w = _faces(n, *LIN)
    got = O.backtrace(u, v, w, p, dt, h)
    vel = lambda q: np.stack([f(q[:, 0], q[:, 1], q[:, 2]) for f in LIN], axis=1)  # noqa: E731
    p0 = p.astype(np.float64)
    k1 = vel(p0)
    k2 = vel(p0 - 0.5 * dt / h * k1)
    k3 = vel(p0 - 0.75 * dt / h * k2)
    want = p0 - dt / h * (2 / 9 * k1 + 1 / 3 * k2 + 4 / 9 * k3)
    np.testing.assert_allclose(got, want, atol=3e-5)


def test_sample_minmax_gathers_clamped_corners():
    rng = np.random.default_rng(3)
    q = rng.normal(size=(9, 8, 7)).astype(F32)
    h = 1.0 / 8
    st = np.array([0.0, 0.5, 0.5], F32)
    p = rng.uniform(0.0, 1.5, (200, 3)).astype(F32)
    mn, mx = O.sample_minmax(q, p, h, st)
    base = np.floor(p / F32(h) - st).astype(int)
    for t in range(len(p)):
        vals = [q[min(max(base[t, 0] + a, 0), 8), min(max(base[t, 1] + b, 0), 7), min(max(base[t, 2] + c, 0), 6)]
                for a in (0, 1) for b in (0, 1) for c in (0, 1)]
        assert mn[t] == min(vals) and mx[t] == max(vals)


def test_maccormack_invariants():
    n = 12
    rng = np.random.default_rng(4)
    u, v, w = (rng.normal(0, 0.5, s).astype(F32) for s in ((n + 1, n, n), (n, n + 1, n), (n, n, n + 1)))
    st = (0.0, 0.5, 0.5)
    const = np.full((n + 1, n, n), 0.75, F32)
    # the 4-point weights sum to 1 up to FP32 rounding
    np.testing.assert_allclose(O.advect_mc(u, v, w, const, st, 0.01, 1.0 / n), 0.75, rtol=3e-7)
    q = rng.normal(size=(n + 1, n, n)).astype(F32)
    z = [np.zeros_like(a) for a in (u, v, w)]
    assert np.array_equal(O.advect_mc(*z, q, st, 0.01, 1.0 / n), q)
    # away from the low faces the clamped corners coincide: the semi-Lagrangian value
    out = O.advect_mc(u, v, w, q, st, 0.01, 1.0 / n)
    i, j, k = np.meshgrid(*(np.arange(s) for s in q.shape), indexing="ij")
    p = np.stack([i + st[0], j + st[1], k + st[2]], -1).reshape(-1, 3).astype(F32)
    pb = O.backtrace(u, v, w, p, 0.01, 1.0 / n)
    q_sl = O.tricubic(q, pb - np.array(st, F32)).reshape(q.shape)
    mn, mx = O.sample_minmax(q, pb, 1.0 / n, st)
    eq = (mn == mx).reshape(q.shape)
    assert eq.mean() > 0.5
    assert np.array_equal(out[eq], q_sl[eq])
