"""Checks that pin the oracle functions the reference build cannot reach
(MConvectiveFast, Space and MFlowmove need Lua/LAPACK, absent here) through
relations the mathematics fixes, computed by a different route than the
oracle's own formula:

* MConvectiveFast::velocity(p) (MConvectiveFast.cpp:20-34, 139-151): the
  all-near limit is the Lamb-kernel direct sum plus U; the far part is the
  exact field of each far node's per-sign centroid monopoles, recomputed here
  from the node's vortex range.
* Space::calc_forces (TSpace.cpp:94-120): friction over a run of segments
  telescopes to the difference of its end corners; its moment over a closed
  contour is twice the shoelace area; Cp is the running sum of gsum.
* The SLAE's speed, hydro and Newton rows (MConvectiveFast.cpp:533-739) for a
  static body: zero speed; the hydrodynamic force equals minus the rate of
  change of the fluid impulse (and angular impulse) of the circulation the
  body creates, measured on the vortex list through Space's
  hydrodynamic_momentum / integral after vortex_shed; Newton's law with zero
  acceleration gives holder force = hydrodynamic force; friction_prev and
  fdt_dead change only the force unknowns.
* MFlowmove's dead force, moment and circulation (MFlowmove.cpp:124-144,
  205-214): equal to the loss of momentum, angular impulse and circulation of
  the vortex list (Space integrals before and after).
"""
import math

import numpy as np
import pytest

from cases import cylinder_corners, oracle_for, random_cluster_case

C_1_2PI = 1.0 / (2.0 * math.pi)


def lamb_field(src, p, eps2=None):
    """Lamb-kernel field of sources src (rows x, y, g, ., ., 1/eps) at p."""
    d = p[None, :] - src[:, :2]
    e2 = (1.0 / src[:, 5]) ** 2 if eps2 is None else eps2
    f = src[:, 2] / (np.sum(d * d, 1) + e2)
    return np.array([np.sum(-d[:, 1] * f), np.sum(d[:, 0] * f)]) * C_1_2PI


def monopole_field(src, p):
    """Field at p of the per-sign centroid monopoles of src (no smoothing)."""
    out = np.zeros(2)
    for sel in (src[:, 2] > 0, src[:, 2] < 0):
        G = src[sel, 2].sum()
        if G == 0:
            continue
        c = (src[sel, 2][:, None] * src[sel, :2]).sum(0) / G
        d = p - c
        out += G * C_1_2PI * np.array([-d[1], d[0]]) / (d @ d)
    return out


def shoelace(corners):
    x, y = corners[:, 0], corners[:, 1]
    return 0.5 * np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y)


# --------------------------------------------------------------- velocity(p)

def test_velocity_points_all_near_is_direct_sum_plus_u():
    objs = random_cluster_case(900, 21, "mixed")
    o = oracle_for(objs, inf=(0.7, -0.3))
    o.tree_build(10 ** 7)
    assert len(o.tree()["far"][1]) == 0
    o.epsilon(False)
    v = o.vortices()
    rng = np.random.default_rng(5)
    pts = rng.uniform(v[:, :2].min(0) - 0.3, v[:, :2].max(0) + 0.3, (60, 2))
    got = o.velocity(pts)
    for p, g in zip(pts, got):
        want = lamb_field(v, p) + [0.7, -0.3]
        assert np.allclose(g, want, rtol=1e-12, atol=1e-15)


def test_velocity_points_far_part_is_centroid_monopole_field():
    # two well separated clusters: for a point in one of them the far list
    # holds nodes of the other; velocity(p) minus U and the near leaves'
    # Lamb sums must equal the exact field of each far node's per-sign
    # centroid monopoles, unsmoothed (the far nodes' _1_eps is +inf).
    rng = np.random.default_rng(7)
    a = np.zeros((600, 6))
    a[:, :2] = rng.normal(size=(600, 2)) * 0.05
    b = np.zeros((600, 6))
    b[:, :2] = [3.0, 1.0] + rng.normal(size=(600, 2)) * 0.05
    objs = np.concatenate([a, b])
    objs[:, 2] = rng.choice([-1.0, 1.0], 1200) * rng.uniform(0.5, 1.5, 1200) * 1e-3
    o = oracle_for(objs, inf=(1.0, 0.0))
    o.tree_build(8)
    o.epsilon(False)
    v = o.vortices()
    T = o.tree()
    leaves = T["leaves"]
    noff, nl = T["near"]
    foff, fl = T["far"]
    ints = T["ints"]
    pts = rng.normal(size=(40, 2)) * 0.04 + np.where(rng.random((40, 1)) < 0.5, [0.0, 0.0], [3.0, 1.0])
    got = o.velocity(pts)
    checked = 0
    for p, g in zip(pts, got):
        leaf = o.find(p[0], p[1])
        li = int(np.nonzero(leaves == leaf)[0][0])
        near_src = np.concatenate([v[ints[n, 0]:ints[n, 1]] for n in nl[noff[li]:noff[li + 1]]])
        rest = g - [1.0, 0.0] - lamb_field(near_src, p)
        far = fl[foff[li]:foff[li + 1]]
        want = sum((monopole_field(v[ints[f, 0]:ints[f, 1]], p) for f in far), np.zeros(2))
        assert len(far) > 0
        assert np.allclose(rest, want, rtol=1e-9, atol=1e-14 * np.abs(want).max())
        # and the far field is not negligible next to what is compared
        assert np.abs(want).max() > 1e-6
        checked += 1
    assert checked == 40


# ---------------------------------------------------------------- calc_forces

def _cylinder(n_seg=120, vort=None, re=600.0, dt=0.005):
    return oracle_for(np.zeros((0, 6)) if vort is None else vort, corners=cylinder_corners(n_seg), re=re,
                      inf=(1.0, 0.0), dt=dt)


def test_calc_forces_closed_forms():
    re, dt, f = 350.0, 0.01, 0.37
    o = _cylinder(120, re=re, dt=dt)
    segs = o.segments()
    corners = np.stack([segs["cx"], segs["cy"]], 1)
    dl = np.hypot(segs["dlx"], segs["dly"])
    rng = np.random.default_rng(11)
    segs["gsum"] = rng.normal(size=len(segs)) * 1e-3
    segs["Cp"] = rng.normal(size=len(segs))
    segs["Fr"] = rng.normal(size=len(segs))
    upper = segs["y"] > 0.1
    segs["fric"] = np.where(upper, f * dl, 0.0)   # friction per unit length f on a run of segments
    o.set_segments(segs)
    o.calc_forces()
    out = o.segments()
    body = o.body(0)
    # Cp accumulates the running sum of gsum along the contour
    assert np.allclose(out["Cp"], segs["Cp"] + np.cumsum(segs["gsum"]), rtol=1e-13, atol=1e-15)
    # Fr += fric dt / (pi Re)
    assert np.allclose(out["Fr"], segs["Fr"] + segs["fric"] * dt / (math.pi * re), rtol=1e-14, atol=0)
    # the run of upper segments is contiguous; sum of dl over it telescopes to
    # (last corner + dl) - first corner
    idx = np.nonzero(upper)[0]
    assert np.all(np.diff(idx) == 1)
    run = (corners[idx[-1]] + [segs["dlx"][idx[-1]], segs["dly"][idx[-1]]]) - corners[idx[0]]
    assert np.allclose(body["friction"][:2], -f / (math.pi * re) * run, rtol=1e-12, atol=1e-16)
    assert body["nusselt"] == 0.0   # no heat particles
    # the whole closed contour: no net force, moment = -(f / (pi Re)) * 2 * area
    segs["fric"] = f * dl
    o.set_segments(segs)
    o.calc_forces()
    body = o.body(0)
    area = shoelace(corners)
    assert np.abs(body["friction"][:2]).max() < 1e-14
    assert body["friction"][2] == pytest.approx(-f / (math.pi * re) * 2 * area, rel=1e-12)


# ------------------------------------------------------------------------ SLAE

def _wake_around_cylinder(n, seed):
    rng = np.random.default_rng(seed)
    th = rng.uniform(0, 2 * np.pi, n)
    rr = 0.6 + rng.exponential(0.3, n)
    v = np.zeros((n, 6))
    v[:, 0], v[:, 1] = rr * np.cos(th), rr * np.sin(th)
    v[:, 2] = rng.normal(size=n) * 1e-2
    return v


def _slae_step(o, n_seg):
    dl = o.body(0)["slen"] / (n_seg - 1)
    o.tree_build(8, 5 * dl, 100 * dl)
    o.calc_circulation()
    o.tree_destroy()


def _impulses(o):
    v = o.vortices()
    g = v[:, 2]
    return np.array([np.sum(g * v[:, 0]), np.sum(g * v[:, 1]), np.sum(g * (v[:, 0] ** 2 + v[:, 1] ** 2))])


def test_slae_static_body_forces_are_impulse_rates():
    n_seg, dt = 160, 0.004
    o = _cylinder(n_seg, _wake_around_cylinder(400, 3), dt=dt)
    o.zero_forces()
    _slae_step(o, n_seg)
    body = o.body(0)
    segs = o.segments()
    # static body: the speed rows pin its speed to zero
    assert np.abs(body["speed_slae"]).max() < 1e-12
    # Newton's law at zero acceleration, no gravity: holder = hydrodynamic force
    fscale = np.abs(body["force_hydro"]).max()
    assert np.allclose(body["force_holder"], body["force_hydro"], rtol=1e-12, atol=1e-12 * fscale)
    # the circulation created this step enters the flow at the corners (shed
    # 1e-4 |dl| off the wall); the fluid impulse P = sum g (y, -x) and angular
    # impulse -1/2 sum g |r|^2 change by it, and the body feels minus the rate
    before = _impulses(o)
    o.vortex_shed()
    after = _impulses(o)
    dM = after - before   # change of (sum g x, sum g y, sum g |r|^2)
    scale = np.abs(segs["g"]).sum()
    assert scale > 1e-4
    fx = -dM[1] / dt
    fy = dM[0] / dt
    mo = dM[2] / (2 * dt)
    tol = 1e-5 * scale / dt
    assert abs(body["force_hydro"][0] - fx) < tol
    assert abs(body["force_hydro"][1] - fy) < tol
    assert abs(body["force_hydro"][2] - mo) < tol
    assert max(abs(fx), abs(fy), abs(mo)) > 100 * tol


def test_slae_friction_prev_and_dead_force_shift_only_the_forces():
    n_seg, dt = 160, 0.004
    vort = _wake_around_cylinder(300, 4)
    base = _cylinder(n_seg, vort.copy(), dt=dt)
    base.zero_forces()
    _slae_step(base, n_seg)
    o = _cylinder(n_seg, vort.copy(), dt=dt)
    o.zero_forces()
    st = o.body(0)
    st["friction_prev"] = [0.3, -0.2, 0.05]
    st["fdt_dead"] = [1e-3, 2e-3, -4e-3]
    o.set_body(0, st)
    _slae_step(o, n_seg)
    b0, b1 = base.body(0), o.body(0)
    # (same matrix; the solve rounds differently with the other right-hand side)
    assert np.allclose(o.segments()["g"], base.segments()["g"], rtol=1e-9, atol=1e-12)
    assert np.abs(b1["speed_slae"]).max() < 1e-12
    # the body's previous friction adds to the hydrodynamic force, the momentum
    # of circulation destroyed inside it (fdt_dead) is taken back per step
    want = b0["force_hydro"] + np.array([0.3, -0.2, 0.05]) - np.array([1e-3 / dt, 2e-3 / dt, -4e-3 / (2 * dt)])
    fscale = np.abs(b1["force_hydro"]).max()
    assert np.allclose(b1["force_hydro"], want, rtol=1e-10, atol=1e-12 * fscale)
    assert np.allclose(b1["force_holder"], b1["force_hydro"], rtol=1e-12, atol=1e-12 * fscale)


# ------------------------------------------------------------------ flow-move

def test_dead_force_moment_and_circulation_are_what_the_list_loses():
    rng = np.random.default_rng(13)
    n_in, n_out = 40, 200
    r_in = 0.45 * np.sqrt(rng.random(n_in))
    th = rng.uniform(0, 2 * np.pi, n_in + n_out)
    r_out = 0.55 + rng.exponential(0.2, n_out)
    rr = np.concatenate([r_in, r_out])
    objs = np.zeros((n_in + n_out, 6))
    objs[:, 0], objs[:, 1] = rr * np.cos(th), rr * np.sin(th)
    objs[:, 2] = rng.normal(size=n_in + n_out) * 1e-3
    o = _cylinder(100, objs)
    segs = o.segments()
    segs["g"] = 1e-3 * (1 + rng.random(len(segs)))   # no slip, none below remove_eps: segments add nothing
    segs["gsum"] = 0.0
    o.set_segments(segs)
    o.zero_forces()
    before = _impulses(o)
    g_before = o.vortices()[:, 2].sum()
    cleaned = o.move_and_clean(True)
    assert cleaned == n_in
    after = _impulses(o)
    g_after = o.vortices()[:, 2].sum()
    body = o.body(0)
    lost = before - after   # (sum g x, sum g y, sum g |r|^2) of the removed vortices
    assert np.allclose(body["fdt_dead"][:2], [-lost[1], lost[0]], rtol=1e-10, atol=1e-16)
    assert body["fdt_dead"][2] == pytest.approx(lost[2], rel=1e-10, abs=1e-16)
    assert body["g_dead"] == pytest.approx(g_before - g_after, rel=1e-10, abs=1e-16)
    # each removed vortex's circulation is taken from one segment's gsum
    assert o.segments()["gsum"].sum() == pytest.approx(-(g_before - g_after), rel=1e-10, abs=1e-16)


def test_dead_terms_of_slip_and_vanishing_segments():
    # segments with slip or |g| < remove_eps shed nothing: their circulation
    # goes to the dead accumulators as if shed at the corner and destroyed
    o = _cylinder(100, np.zeros((0, 6)))
    segs = o.segments()
    rng = np.random.default_rng(17)
    segs["g"] = rng.normal(size=len(segs)) * 1e-3
    dead = np.zeros(len(segs), bool)
    dead[::7] = True
    segs["slip"] = np.where(dead, 1, 0)
    segs["gsum"] = 0.0
    o.set_segments(segs)
    o.zero_forces()
    # shed the same segments as vortices exactly at their corners in a second
    # context and measure their impulse there
    probe = np.zeros((int(dead.sum()), 6))
    probe[:, 0], probe[:, 1], probe[:, 2] = segs["cx"][dead], segs["cy"][dead], segs["g"][dead]
    p = oracle_for(probe)
    imp = _impulses(p)
    o.move_and_clean(True)
    body = o.body(0)
    assert np.allclose(body["fdt_dead"][:2], [-imp[1], imp[0]], rtol=1e-10, atol=1e-16)
    assert body["fdt_dead"][2] == pytest.approx(imp[2], rel=1e-10)
    assert body["g_dead"] == pytest.approx(probe[:, 2].sum(), rel=1e-10)
    assert np.allclose(o.segments()["gsum"], np.where(dead, -segs["g"], 0.0), rtol=0, atol=0)


# ------------------------------------------------------ cylinder run fixture

def test_cylinder_run_fixture_reproduces():
    """tests/golden/cylinder400_oracle.npz (the oracle's 400-step cylinder
    run, the yardstick of the GPU's free-running run) follows from the
    oracle's code: its first steps are reproduced exactly."""
    import os
    import sys
    from cases import GOLDEN
    sys.path.insert(0, GOLDEN)
    from make_cylinder_run import run
    ref = np.load(os.path.join(GOLDEN, "cylinder400_oracle.npz"), allow_pickle=False)
    got = run(12)
    for k in ("n", "merged", "gsum", "gabs", "force"):
        assert np.array_equal(got[k], ref[k][:12]), k
