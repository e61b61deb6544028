"""Checks that pin the oracle's point queries and grid fields
(MEpsilonFast::eps2h / h2, XVelocity, XVorticity, XStreamfunction, XPressure)
by routes independent of its formulas:

* eps2h / h2 with every leaf near (huge far criterion) are the second
  smallest non-zero squared distance to all vortices and the smallest to all
  segment centres (numpy partition);
* the vorticity of Gaussian blobs integrates to their circulation times
  (1 - e^-6), the share inside the exponent cut (quadrature on a fine grid);
* the streamfunction's finite differences are the velocity of the same
  smoothed vortices (u = d psi / dy, v = -d psi / dx), including the free
  stream relative to the reference frame;
* the pressure of moving point-like vortices equals -d phi / dt - |u|^2 / 2,
  with d phi / dt taken by a finite difference in time of the vortices'
  potential sum g / (2 pi) atan2 (unsteady Bernoulli, U = 0);
* XVelocity is velocity(p) minus the reference frame's speed.
"""
import math

import numpy as np
import pytest

from cases import cylinder_corners, oracle_for, random_cluster_case

VX, VY, VORT, STREAM, PRESS = range(5)


def spread_case(n, seed, box=1.0):
    rng = np.random.default_rng(seed)
    objs = np.zeros((n, 6))
    objs[:, :2] = rng.uniform(0, box, (n, 2))
    objs[:, 2] = rng.choice([-1.0, 1.0], n) * rng.uniform(0.5, 1.5, n) * 1e-3
    return objs


def test_eps2h_h2_all_near_brute_force():
    objs = spread_case(400, 1)
    corners = cylinder_corners(60, r=0.1) + [0.5, 0.5]
    objs = objs[np.hypot(objs[:, 0] - 0.5, objs[:, 1] - 0.5) > 0.12]
    o = oracle_for(objs, corners=corners)
    o.tree_build(10 ** 7)
    v = o.vortices()
    segs = o.segments()
    rng = np.random.default_rng(2)
    pts = np.concatenate([rng.uniform(0, 1, (50, 2)), v[:7, :2]])   # some points sit on vortices
    got = o.eps2h(pts)
    got_h2 = o.eps2h(pts, h2=True)
    for p, e, h in zip(pts, got, got_h2):
        d = np.sum((v[:, :2] - p) ** 2, 1)
        d = d[d != 0]
        assert e == np.partition(d, 1)[1]
        assert h == np.min((segs["x"] - p[0]) ** 2 + (segs["y"] - p[1]) ** 2)
    own = o.eps2h(None)
    for i in (0, 5, 100):
        d = np.sum((v[:, :2] - v[i, :2]) ** 2, 1)
        assert own[i] == np.partition(d[d != 0], 1)[1]
    # one vortex alone: its own distance is zero, so no candidate at all
    solo = oracle_for(np.array([[0.3, 0.3, 1e-3, 0, 0, 0]]))
    solo.tree_build(8)
    assert solo.eps2h(None)[0] == np.finfo(np.float64).min
    assert solo.eps2h(np.array([[0.0, 0.0]]))[0] == pytest.approx(0.18, rel=1e-15)


def test_vorticity_integrates_to_circulation():
    objs = spread_case(60, 3, box=0.6)
    objs[:, :2] += 0.2
    o = oracle_for(objs)
    o.tree_build(10 ** 7)
    # the blobs reach sqrt(6 m^2 eps2h) from their centres: the grid covers
    # [-0.4, 1.4]^2, beyond every cut
    h, m = 0.002, 1.0
    n = int(1.8 / h)
    w = o.field(VORT, -0.4, -0.4, h, n, n, eps_mult=m, dl=0.0).astype(np.float64)
    total = w.sum() * h * h
    g = objs[:, 2]
    assert total == pytest.approx((1 - math.exp(-6)) * g.sum(), rel=2e-3, abs=2e-3 * np.abs(g).sum())
    # positive and negative parts separately (the signs do not mix per blob)
    pos = oracle_for(objs[g > 0])
    pos.tree_build(10 ** 7)
    wp = pos.field(VORT, -0.4, -0.4, h, n, n, eps_mult=m, dl=0.0).astype(np.float64)
    assert wp.sum() * h * h == pytest.approx((1 - math.exp(-6)) * g[g > 0].sum(), rel=2e-3)


def test_vorticity_is_zero_inside_a_body_and_has_the_wall_term():
    corners = cylinder_corners(100)
    o = oracle_for(np.zeros((0, 6)), corners=corners)
    dl = o.body(0)["slen"] / 99
    o.tree_build(8, dl * 20)
    w = o.field(VORT, -1.0, -1.0, 0.05, 41, 41, eps_mult=1.0, dl=dl)
    xs = -1.0 + 0.05 * np.arange(41)
    X, Y = np.meshgrid(xs, xs)
    inside = np.hypot(X, Y) < 0.49
    assert np.all(w[inside] == 0)
    # just outside the wall the term is 0.5 (1 - erf(h2 / dl^2)) with h2 the
    # squared distance to the nearest segment centre: between 0 and 0.5
    near = (np.hypot(X, Y) > 0.5) & (np.hypot(X, Y) < 0.52)
    assert np.all(w[near] > 0) and np.all(w[near] <= 0.5)
    far = np.hypot(X, Y) > 0.7
    assert np.all(w[far] == 0)


def lamb_velocity(v, e2, p, U):
    d = p[None, :] - v[:, :2]
    f = v[:, 2] / (np.sum(d * d, 1) + e2) / (2 * math.pi)
    return np.array([U[0] + np.sum(-d[:, 1] * f), U[1] + np.sum(d[:, 0] * f)])


def test_streamfunction_derivatives_are_the_velocity():
    objs = spread_case(120, 4, box=0.5)
    o = oracle_for(objs, inf=(0.7, -0.2))
    o.tree_build(10 ** 7)   # every leaf near: the vortex sum is direct
    v = o.vortices()
    m = 2.0
    # the weights the evaluator uses: e^2 = m^2 max(eps2h / 4, 0) (no body)
    e2 = m * m * np.maximum(o.eps2h(None) * 0.25, 0.0)
    ref = (0.1, 0.05)
    h = 1e-3
    rng = np.random.default_rng(5)
    for p in rng.uniform(0.6, 0.9, (6, 2)):
        psi = o.field(STREAM, p[0] - h, p[1] - h, h, 3, 3, eps_mult=m, dl=0.0, ref=ref).astype(np.float64)
        u = (psi[2, 1] - psi[0, 1]) / (2 * h)
        w = -(psi[1, 2] - psi[1, 0]) / (2 * h)
        want = lamb_velocity(v, e2, p, (0.7 - ref[0], -0.2 - ref[1]))
        assert np.allclose([u, w], want, rtol=0, atol=2e-3 * np.abs(want).max())


def test_pressure_is_unsteady_bernoulli_of_moving_vortices():
    rng = np.random.default_rng(6)
    n = 5
    objs = np.zeros((n, 6))
    objs[:, :2] = rng.uniform(0.4, 0.6, (n, 2))
    objs[:, 2] = rng.uniform(-1, 1, n) * 1e-2
    objs[:, 5] = 1e6   # point-like
    o = oracle_for(objs, inf=(0.0, 0.0))
    o.tree_build(10 ** 7)
    st = o.vortices()
    vel = rng.normal(size=(n, 2)) * 0.3
    st[:, 3:5] = vel   # the vortices' own velocities enter the pressure
    o.set_vortices(st)
    o.tree_build(10 ** 7)
    st = o.vortices()
    pts = rng.uniform(0.0, 1.0, (8, 2))
    pts = pts[np.min(np.hypot(pts[:, None, 0] - st[None, :, 0], pts[:, None, 1] - st[None, :, 1]), 1) > 0.1]

    def phi(p, t):
        r = st[:, :2] + st[:, 3:5] * t
        return np.sum(st[:, 2] / (2 * math.pi) * np.arctan2(p[1] - r[:, 1], p[0] - r[:, 0]))

    for p in pts:
        cp = o.field(PRESS, p[0], p[1], 1.0, 2, 2, eps_mult=1.0, dl=0.0)[0, 0]
        dt = 1e-6
        dphidt = (phi(p, dt) - phi(p, -dt)) / (2 * dt)
        u = lamb_velocity(st, 1e-12, p, (0.0, 0.0))
        want = -dphidt - 0.5 * (u @ u)
        assert cp == pytest.approx(want, rel=1e-4, abs=1e-6)


def test_velocity_field_is_velocity_minus_reference():
    objs = random_cluster_case(500, 7, "mixed")
    o = oracle_for(objs, inf=(1.0, 0.25))
    o.tree_build(10 ** 7)
    o.epsilon(False)
    v = o.vortices()
    ref = (0.3, -0.1)
    fx = o.field(VX, 0.1, 0.2, 0.05, 4, 3, ref=ref)
    fy = o.field(VY, 0.1, 0.2, 0.05, 4, 3, ref=ref)
    e2 = (1.0 / v[:, 5]) ** 2
    for yj in range(3):
        for xi in range(4):
            p = np.array([0.1 + 0.05 * xi, 0.2 + 0.05 * yj])
            want = lamb_velocity(v, e2, p, (1.0, 0.25)) - ref
            assert fx[yj, xi] == pytest.approx(want[0], rel=1e-6, abs=1e-7)
            assert fy[yj, xi] == pytest.approx(want[1], rel=1e-6, abs=1e-7)
