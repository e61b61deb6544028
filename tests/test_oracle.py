"""Checks that pin the FP64 oracle where the reference's code is not built
here (convective velocity, SLAE, flow-move, integrals) or independently of
it: closed forms, brute force on small inputs, invariants and the TBody
known answers of libvvhd/test/test_TBody.hpp:78-160."""
import math

import numpy as np
import pytest

from cases import DBL_MAX, cylinder_corners, oracle_for, random_cluster_case
from oracle_bind import Oracle, direct_sum

C_1_2PI = 1.0 / (2.0 * math.pi)


def lamb(objs, p):
    """Independent numpy restatement of the Lamb-kernel direct sum."""
    d = p[None, :] - objs[:, :2]
    e2 = (1.0 / objs[:, 5]) ** 2
    f = objs[:, 2] / (np.sum(d * d, 1) + e2)
    return np.array([np.sum(-d[:, 1] * f), np.sum(d[:, 0] * f)]) * C_1_2PI


def test_single_vortex_velocity_closed_form():
    # one source, one target: v = U + g/(2 pi) rotl(r) / (|r|^2 + eps^2)
    objs = np.array([[0.0, 0.0, 0.3, 0, 0, 1 / 0.05], [0.2, 0.1, 0.7, 0, 0, 1 / 0.05]])
    o = oracle_for(objs, inf=(1.0, 0.5))
    o.tree_build(8)
    o.convective()
    v = o.vortices()
    r = np.array([0.2, 0.1])
    exp1 = np.array([1.0, 0.5]) + 0.3 * C_1_2PI * np.array([-r[1], r[0]]) / (r @ r + 0.05 ** 2)
    exp0 = np.array([1.0, 0.5]) + 0.7 * C_1_2PI * np.array([r[1], -r[0]]) / (r @ r + 0.05 ** 2)
    assert np.allclose(v[1, 3:5], exp1, rtol=1e-14, atol=0)
    assert np.allclose(v[0, 3:5], exp0, rtol=1e-14, atol=0)


def test_direct_sum_matches_numpy():
    objs = random_cluster_case(300, 1, "mixed")
    objs[:, 5] = np.random.default_rng(0).uniform(50, 500, len(objs))
    ds = direct_sum(objs)
    ref = np.array([lamb(objs, objs[i, :2]) for i in range(len(objs))])
    assert np.allclose(ds, ref, rtol=1e-12, atol=1e-15)


def test_near_field_only_equals_direct_sum():
    # With a huge far criterion every leaf is near: process_all_lists is the
    # direct sum plus U.
    objs = random_cluster_case(800, 2, "mixed")
    o = oracle_for(objs, inf=(1.0, 0.0))
    o.tree_build(10 ** 6)
    T = o.tree()
    assert len(T["far"][1]) == 0
    o.epsilon(False)
    o.convective()
    v = o.vortices()
    ds = direct_sum(v)
    assert np.allclose(v[:, 3:5] - [1.0, 0.0], ds, rtol=1e-11, atol=1e-13)


def test_far_field_taylor_is_first_order_expansion():
    # Sources far from the targets: the far terms must be the first-order
    # Taylor expansion, about the target leaf centre, of the far nodes'
    # per-sign monopole fields (MConvectiveFast.cpp:48-69).
    rng = np.random.default_rng(3)
    tgt = np.zeros((12, 6))
    tgt[:, :2] = rng.normal(size=(12, 2)) * 0.01
    tgt[:, 2] = 1e-14   # targets with negligible circulation
    src = np.zeros((40, 6))
    src[:, :2] = [5.0, 3.0] + rng.normal(size=(40, 2)) * 0.02
    src[:, 2] = rng.uniform(-1, 1, 40) * 1e-3
    objs = np.concatenate([tgt, src])
    objs[:, 5] = 1e3
    o = oracle_for(objs, inf=(0.0, 0.0))
    o.tree_build(8)
    o.epsilon(False)
    o.convective()
    v = o.vortices()
    T = o.tree()
    leaves = T["leaves"]
    checked = 0
    for li, leaf in enumerate(leaves):
        v0, v1 = T["ints"][leaf, :2]
        foff, flst = T["far"]
        far = flst[foff[li]:foff[li + 1]]
        if v1 - v0 == 0 or len(far) == 0 or v[v0, 2] != 1e-14:
            continue
        cm = T["cm"][far]
        def field(p):
            out = np.zeros(2)
            for (px, py, pg, mx, my, mg) in cm:
                for (cx, cy, g) in ((px, py, pg), (mx, my, mg)):
                    d = p - [cx, cy]
                    out += g * C_1_2PI * np.array([-d[1], d[0]]) / (d @ d)
            return out
        cen = T["geom"][leaf, :2]
        h = 1e-5
        J = np.stack([(field(cen + [h, 0]) - field(cen - [h, 0])) / (2 * h),
                      (field(cen + [0, h]) - field(cen - [0, h])) / (2 * h)], 1)
        for i in range(v0, v1):
            lin = field(cen) + J @ (v[i, :2] - cen)
            assert np.linalg.norm(v[i, 3:5] - lin) < 1e-6 * np.linalg.norm(field(cen))
            checked += 1
    assert checked >= 10


def brute_top2(objs, i, cand):
    best = []
    for j in cand:
        if j == i or objs[j, 2] == 0:
            continue
        d = (objs[i, 0] - objs[j, 0]) ** 2 + (objs[i, 1] - objs[j, 1]) ** 2
        best.append((d, j))
    best.sort()
    return best[:2]


def test_epsilon_brute_force_no_merge():
    objs = random_cluster_case(400, 4, "mixed")
    o = oracle_for(objs)
    o.tree_build(10 ** 6)   # every leaf near every leaf: global 2-NN
    v = o.vortices()
    nn1, nn2, _, m = o.epsilon(False)
    assert m == 0
    w = o.vortices()
    for i in range(len(v)):
        b = brute_top2(v, i, range(len(v)))
        assert nn1[i] == b[0][1] and nn2[i] == b[1][1]
        assert w[i, 5] == 1.0 / math.sqrt(b[1][0])


def test_epsilon_special_cases():
    # one vortex: no neighbour -> eps = DBL_MIN; two: eps = sqrt(d1)
    o = oracle_for(np.array([[0.1, 0.2, 1e-3, 0, 0, 0]]))
    o.tree_build(8)
    o.epsilon(True)
    assert o.vortices()[0, 5] == 1.0 / np.finfo(np.float64).tiny
    o = oracle_for(np.array([[0.0, 0.0, 1e-3, 0, 0, 0], [0.3, 0.4, 1e-3, 0, 0, 0]]))
    o.tree_build(8)
    o.epsilon(True)
    assert np.allclose(o.vortices()[:, 5], 1 / 0.5, rtol=0, atol=0)


def test_merge_conserves_circulation_and_moment():
    objs = random_cluster_case(3000, 6, "mixed")
    o = oracle_for(objs)
    o.tree_build(8)
    before = o.vortices()
    nn1, nn2, merges, m = o.epsilon(True)
    after = o.vortices()
    assert m > 0
    assert abs(after[:, 2].sum() - before[:, 2].sum()) < 1e-15 * np.abs(before[:, 2]).sum()
    # same-sign merges conserve sum g r; check per merge on the recorded pairs
    for (i, j) in merges[:50]:
        assert after[j, 2] == 0.0


def test_diffusive_brute_force():
    objs = random_cluster_case(300, 7, "positive")
    o = oracle_for(objs, re=600.0)
    o.tree_build(10 ** 6)
    o.epsilon(False)
    v = o.vortices()
    o.diffusive()
    w = o.vortices()
    for i in range(0, len(v), 7):
        ie = v[i, 5]
        S2 = np.zeros(2)
        S1 = 0.0
        for j in range(len(v)):
            if np.sign(v[j, 2]) != np.sign(v[i, 2]) or v[j, 2] == 0:
                continue
            dr = v[i, :2] - v[j, :2]
            if abs(dr[0]) + abs(dr[1]) < 1e-10:
                continue
            a = math.hypot(*dr)
            if -a * ie < -8:
                continue
            e = v[j, 2] * math.exp(-a * ie)
            S2 += dr * (e / a)
            S1 += e
        if np.sign(S1) != np.sign(v[i, 2]) or abs(S1) < abs(0.1 * v[i, 2]):
            S1 = 0.1 * v[i, 2]
        dv = ie / (600.0 * S1) * S2
        assert np.allclose(w[i, 3:5] - v[i, 3:5], dv, rtol=1e-10, atol=1e-14)


# ---- TBody known answers (libvvhd/test/test_TBody.hpp:78-160)
def test_body_square_known_answers():
    o = Oracle()
    o.add_body([[1, 1], [1, -1], [-1, -1], [-1, 1]])   # clockwise: external flow
    b = o.body(0)
    assert b["slen"] == 8.0 and b["area"] == 4.0
    assert b["cofm"][0] ** 2 + b["cofm"][1] ** 2 == 0.0
    assert b["moi_cofm"] == 16.0 / 6.0
    assert o.point_invalid(0, 0.9, 0.0) == 0
    assert o.point_invalid(0, 1.1, 0.0) == -1


def test_body_inverted_and_thin_known_answers():
    o = Oracle()
    o.add_body([[2, -2], [2, 2], [-2, 2], [-2, -2]])   # counter-clockwise: internal flow
    b = o.body(0)
    assert b["slen"] == 16.0 and b["area"] == -16.0
    assert b["moi_cofm"] == -256.0 / 6.0
    assert o.point_invalid(0, 1.9, 0.0) == -1
    assert o.point_invalid(0, 2.1, 0.0) == 0
    o = Oracle()
    o.add_body([[0.0, 1.0], [0.5, 1.0], [1.0, 1.0]])
    b = o.body(0)
    assert b["area"] == 0.0 and b["moi_cofm"] == 0.0 and b["slen"] == 2.0
    assert b["cofm"][0] == 0.5 and b["cofm"][1] == 1.0


def _cylinder_slae(n_seg=200, vort=None):
    o = oracle_for(np.zeros((0, 6)) if vort is None else vort, corners=cylinder_corners(n_seg), inf=(1.0, 0.0))
    dl = o.body(0)["slen"] / (n_seg - 1)
    o.tree_build(8, 5 * dl, 100 * dl)
    A, b = o.calc_circulation(want_matrix=True)
    return o, A, b


def test_slae_residual_and_potential_flow():
    o, A, b = _cylinder_slae(200)
    o2, _, _ = _cylinder_slae(200)
    x = np.linalg.solve(A, b)
    segs = o.segments()
    assert np.allclose(segs["g"], x[:200], rtol=1e-9, atol=1e-12)
    # bound circulation of a cylinder in uniform flow U=(1,0): gamma = 2 U sin(theta)
    th = np.arctan2(segs["y"], segs["x"])
    dl = np.hypot(segs["dlx"], segs["dly"])
    gamma = segs["g"] / dl
    assert np.allclose(np.abs(gamma), np.abs(2 * np.sin(th)), atol=0.05)
    assert abs(segs["g"].sum()) < 1e-9   # circulation conservation row
    body = o.body(0)
    assert np.linalg.norm(body["speed_slae"]) < 1e-10


def test_slae_with_vortices_and_inverse_path():
    rng = np.random.default_rng(1)
    n = 500
    th = rng.uniform(0, 2 * np.pi, n)
    rr = 0.6 + rng.exponential(0.2, n)
    v = np.zeros((n, 6))
    v[:, 0], v[:, 1] = rr * np.cos(th), rr * np.sin(th)
    v[:, 2] = rng.normal(size=n) * 1e-2
    v[:, 5] = 100.0
    o, A, b = _cylinder_slae(150, v)
    x = np.linalg.solve(A, b)
    assert np.allclose(o.segments()["g"], x[:150], rtol=1e-8, atol=1e-12)
    assert abs(o.segments()["g"].sum() + v[:, 2].sum()) < 1e-12   # -gsum - inf_g
    # t > 0 uses the cached inverse: same answer
    o.set_params(re=600, dt=0.005, inf=(1.0, 0.0), time=0.005)
    o.tree_build(8, 5 * o.body(0)["slen"] / 149, 100 * o.body(0)["slen"] / 149)
    A2, b2 = o.calc_circulation(want_matrix=True)
    assert np.allclose(o.segments()["g"], np.linalg.solve(A2, b2)[:150], rtol=1e-8, atol=1e-12)


def test_flow_move_closed_form():
    objs = random_cluster_case(200, 8, "mixed")
    objs[:, 3] = np.linspace(-1, 1, 200)
    objs[:, 4] = 0.5
    objs[5, 2] = 1e-12   # removed (|g| < 1e-10)
    o = oracle_for(objs, dt=0.01)
    o.move_and_clean(True)
    w = o.vortices()
    keep = np.abs(objs[:, 2]) >= 1e-10
    assert np.array_equal(w[:, 0], objs[keep, 0] + objs[keep, 3] * 0.01)
    assert np.array_equal(w[:, 1], objs[keep, 1] + objs[keep, 4] * 0.01)
    assert np.all(w[:, 3:5] == 0)


def test_flow_move_removes_vortices_inside_body_and_shed():
    corners = cylinder_corners(100)
    objs = np.array([[0.0, 0.0, 1e-3, 0, 0, 0], [2.0, 0.0, -1e-3, 0, 0, 0], [0.1, 0.2, 2e-3, 0, 0, 0]])
    o = oracle_for(objs, corners=corners)
    assert o.move_and_clean(True) == 2
    assert len(o.vortices()) == 1
    assert o.body(0)["g_dead"] == 3e-3
    segs = o.segments()
    segs["g"] = 1e-3
    o.set_segments(segs)
    o.vortex_shed()
    w = o.vortices()
    assert len(w) == 101
    assert np.allclose(w[1:, 0], segs["cx"] - segs["dly"] * 1e-4)


def test_integrals():
    objs = random_cluster_case(500, 9, "mixed")
    o = oracle_for(objs)
    gs, integ, mx, my, gm = o.integrals()
    assert gs == pytest.approx(objs[:, 2].sum(), rel=1e-14)
    assert integ == pytest.approx((objs[:, 2] * (objs[:, 0] ** 2 + objs[:, 1] ** 2)).sum(), rel=1e-13)
    assert mx == pytest.approx((objs[:, 2] * objs[:, 0]).sum(), rel=1e-13)
    assert abs(gm) == np.abs(objs[:, 2]).max()


def test_find_node_matches_bucketing():
    objs = random_cluster_case(2000, 10, "mixed")
    o = oracle_for(objs)
    o.tree_build(8)
    T = o.tree()
    v = o.vortices()
    for leaf in T["leaves"][::7]:
        v0, v1 = T["ints"][leaf, :2]
        for k in range(v0, v1):
            assert o.find(v[k, 0], v[k, 1]) == leaf
