"""The oracle's heat and streak lists (HeatList, StreakList) against the
reference's own code and against rules that fix them independently:

* tree over vortices, heat and streak particles and a body, CalcEpsilonFast
  (epsv and epsh with merges), process_vort_list and process_heat_list:
  bit-exact against tests/golden/heat_cylinder.npz, written by
  tests/golden/make_golden.py from the reference's TSortedTree.cpp,
  MEpsilonFast.cpp and MDiffusiveFast.cpp;
* process_all_lists on the heat and streak ranges (MConvectiveFast.cpp:
  87-110): a streak or heat copy of every vortex, at the vortex's position,
  leaves the tree unchanged and must get exactly the vortex's velocity;
* flow-move of the heat and streak lists, heat_shed, heat_crop and
  streak_shed (MFlowmove.cpp:107-198, 237-291): small cases whose outcome
  follows from the rules (which particles go, which segment absorbs them,
  the shed particles' positions and strengths) computed here in numpy.
"""
import numpy as np
import pytest

from cases import cylinder_corners, load_golden, oracle_for, params, path_code32

HEAT, STREAK, SOURCE = 1, 2, 3


def golden_oracle(d):
    objs = d["objs"].copy()
    objs[:, 3] = np.arange(len(objs))
    o = oracle_for(objs, segs=d["segs"], re=float(d["re"]))
    o.set_pr(float(d["pr"]))
    for lid, key in ((HEAT, "heat"), (STREAK, "streak")):
        a = d[key].copy()
        a[:, 3] = np.arange(len(a))
        o.set_list(lid, a)
    o.tree_build(*params(d))
    return o


def test_heat_tree_matches_reference():
    d = load_golden("heat_cylinder")
    o = golden_oracle(d)
    assert np.array_equal(o.vortices()[:, 3].astype(np.int32), d["perm"])
    assert np.array_equal(o.get_list(HEAT)[:, 3].astype(np.int32), d["heat_perm"])
    assert np.array_equal(o.get_list(STREAK)[:, 3].astype(np.int32), d["streak_perm"])
    T = o.tree()
    leaves = T["leaves"]
    assert np.array_equal(T["geom"][leaves], d["leaf_geom"])
    assert np.array_equal(T["ints"][leaves][:, :2], d["leaf_ints"][:, :2])
    assert np.array_equal(path_code32(T["path"][leaves]), d["leaf_ints"][:, 3].astype(np.uint32).astype(np.int64))
    assert np.array_equal(o.leaf_hs(), d["leaf_hs"])
    for key in ("near", "far"):
        off, lst = T[key]
        ref = d[key].astype(np.int64)
        assert np.array_equal(T["ints"][lst, 4], ref[:, 0])
        assert np.array_equal(path_code32(T["path"][lst]), ref[:, 1].astype(np.uint32).astype(np.int64))
    # the heat and streak particles change the tree: without them it differs
    bare = oracle_for(d["objs"], segs=d["segs"])
    bare.tree_build(*params(d))
    assert len(bare.tree()["leaves"]) != len(leaves)


def test_heat_epsilon_and_diffusive_match_reference():
    d = load_golden("heat_cylinder")
    o = golden_oracle(d)
    for lid in (0, HEAT, STREAK):
        a = o.vortices() if lid == 0 else o.get_list(lid)
        a[:, 3] = 0.0
        if lid == 0:
            o.set_vortices(a)
        else:
            o.set_list(lid, a)
    o.tree_build(*params(d))
    _, _, _, merged = o.epsilon(True)
    assert merged == int(d["merged"])
    assert np.array_equal(o.vortices()[:, [0, 1, 2, 5]], d["after_eps"])
    h = o.get_list(HEAT)
    assert np.array_equal(h[:, [0, 1, 2, 5]], d["heat_after_eps"])
    # heat merges happened and were of heat particles only
    assert (h[:, 2] == 0).sum() > (d["heat"][:, 2] == 0).sum()
    assert np.sum(h[:, 2]) == pytest.approx(np.sum(d["heat"][:, 2]), rel=1e-12)
    o.diffusive()
    o.diffusive_heat()
    assert np.array_equal(o.get_list(HEAT)[:, 3:5], d["heat_after_diff_v"])


def test_passive_lists_get_the_vortex_velocity():
    rng = np.random.default_rng(4)
    n = 3000
    objs = np.zeros((n, 6))
    th = rng.uniform(0, 2 * np.pi, n)
    rr = 0.5 + 1e-3 + rng.exponential(0.05, n)
    objs[:, 0], objs[:, 1] = rr * np.cos(th), rr * np.sin(th)
    objs[:, 2] = rng.normal(size=n) * 1e-3
    corners = cylinder_corners(200)
    dl = 200 * np.sin(np.pi / 200) / 199
    base = oracle_for(objs, corners=corners, inf=(1.0, 0.2))
    base.tree_build(8, 5 * dl, 100 * dl)
    base.epsilon(False)
    state = base.vortices()
    base.convective()
    want = base.vortices()[:, 3:5]
    for lid in (HEAT, STREAK):
        o = oracle_for(state, corners=corners, inf=(1.0, 0.2))
        o.set_list(lid, state)   # copies at the vortices' positions, same g
        o.tree_build(8, 5 * dl, 100 * dl)
        assert np.array_equal(o.vortices()[:, :3], state[:, :3])
        o.convective()
        assert np.array_equal(o.vortices()[:, 3:5], want)
        got = o.get_list(lid)
        assert np.array_equal(got[:, :2], o.vortices()[:, :2])   # the same permutation
        assert np.array_equal(got[:, 3:5], want)


def test_streak_velocity_ignores_its_own_circulation():
    rng = np.random.default_rng(5)
    objs = np.zeros((500, 6))
    objs[:, :2] = rng.uniform(0, 1, (500, 2))
    objs[:, 2] = rng.normal(size=500) * 1e-3
    objs[:, 5] = 50.0
    st = np.zeros((40, 6))
    st[:, :2] = rng.uniform(0.2, 0.8, (40, 2))
    a = oracle_for(objs, inf=(0.3, 0.0))
    a.set_list(STREAK, st)
    a.tree_build(8)
    a.convective()
    st2 = st.copy()
    st2[:, 2] = 5.0   # a streak particle induces nothing and is moved whatever its g
    b = oracle_for(objs, inf=(0.3, 0.0))
    b.set_list(STREAK, st2)
    b.tree_build(8)
    b.convective()
    assert np.array_equal(a.get_list(STREAK)[:, 3:5], b.get_list(STREAK)[:, 3:5])
    assert np.array_equal(a.vortices()[:, 3:5], b.vortices()[:, 3:5])
    assert np.all(np.abs(a.get_list(STREAK)[:, 3:5]).sum(1) > 0)
    # heat particles with g = 0 are skipped (their velocity stays as it was)
    h = st.copy()
    h[:, 2] = 0.0
    h[:, 3] = 7.0
    c = oracle_for(objs, inf=(0.3, 0.0))
    c.set_list(HEAT, h)
    c.tree_build(8)
    c.convective()
    assert np.all(c.get_list(HEAT)[:, 3] == 7.0)


# ---------------------------------------------------------------- flow-move

def pip(poly, p):
    """Even-odd crossing test of point p against a closed polygon (numpy)."""
    x, y = p
    inside = False
    n = len(poly)
    for i in range(n):
        x1, y1 = poly[i - 1]
        x2, y2 = poly[i]
        if (y1 > y) != (y2 > y):
            xc = x1 + (y - y1) * (x2 - x1) / (y2 - y1)
            if xc > x:
                inside = not inside
    return inside


def test_heat_and_streak_flowmove_rules():
    corners = cylinder_corners(100)
    o = oracle_for(np.zeros((0, 6)), corners=corners, dt=0.1)
    segs = o.segments()
    r = np.stack([segs["x"], segs["y"]], 1)
    heat = np.array([
        [0.0, 0.0, 2e-3, 0, 0, 0],       # inside the body: absorbed by the nearest segment
        [0.3, 0.05, 1e-3, 0, 0, 0],      # inside too
        [1.0, 0.0, 5e-11, 0, 0, 0],      # |g| < 1e-10: removed, nothing absorbed
        [0.9, 0.9, 1e-3, 1.0, -2.0, 0],  # moves to (1.0, 0.7) and stays
        [-2.0, 0.0, 3e-4, 0, 0, 0],      # stays
    ])
    streak = np.array([
        [0.1, 0.1, 0, 0, 0, 0],          # inside: removed
        [0.7, 0.0, 0, -1.0, 0, 0],       # moves to (0.6, 0): outside, stays
        [0.55, 0.0, 0, -1.0, 0, 0],      # moves to (0.45, 0): inside, removed
    ])
    o.set_list(HEAT, heat)
    o.set_list(STREAK, streak)
    o.move_and_clean(True)
    h = o.get_list(HEAT)
    assert np.allclose(h[:, :3], [[1.0, 0.7, 1e-3], [-2.0, 0.0, 3e-4]], rtol=0, atol=1e-15)
    assert np.all(h[:, 3:5] == 0)
    s = o.get_list(STREAK)
    assert np.allclose(s[:, :2], [[0.6, 0.0]], rtol=0, atol=1e-15)
    hsum = o.segments()["hsum"]
    want = np.zeros(len(segs))
    for p, g in (((0.0, 0.0), 2e-3), ((0.3, 0.05), 1e-3)):
        want[np.argmin(np.sum((r - np.array(p)) ** 2, 1))] -= g
    assert np.array_equal(hsum, want)
    assert np.all(o.heat_layer() == -1)   # not a const_t body


def star_corners(n=24, r0=1.0, r1=0.55):
    """A star-shaped body (clockwise, as gen_cylinder's corners): its heat
    layer reaches into the notches, inside the body's bounding disc."""
    k = np.arange(2 * n)
    rad = np.where(k % 2 == 0, r0, r1)
    ang = 2 * np.pi - np.pi * k / n
    return np.stack([rad * np.cos(ang), rad * np.sin(ang)], 1)


def test_heat_layer_of_a_const_t_body():
    corners = star_corners()
    o = oracle_for(np.zeros((0, 6)), corners=corners, dt=1e-9)
    segs = o.segments()
    r = np.stack([segs["x"], segs["y"]], 1)
    layer = r + np.stack([-segs["dly"], segs["dlx"]], 1)
    body = corners
    # points in the layer band outside the body, inside the bounding disc
    rng = np.random.default_rng(8)
    pts = []
    while len(pts) < 40:
        p = rng.uniform(-1, 1, 2)
        if np.hypot(*p) < 0.98 and pip(layer, p) and not pip(body, p):
            pts.append(p)
    pts = np.array(pts)
    g = rng.uniform(1, 2, len(pts)) * 1e-3
    heat = np.zeros((len(pts), 6))
    heat[:, :2], heat[:, 2] = pts, g
    o.set_heat([2], np.zeros(len(segs), np.float32))   # const_t
    o.set_list(HEAT, heat)
    o.move_and_clean(True)
    near = [int(np.argmin(np.sum((r - p) ** 2, 1))) for p in pts]
    keep, lay, hsum = [], -np.ones(len(segs), np.int64), np.zeros(len(segs))
    for k, s in enumerate(near):
        if lay[s] >= 0:
            hsum[s] -= g[k]
        else:
            lay[s] = len(keep)
            keep.append(k)
    assert len(keep) < len(pts)   # some segments saw two particles
    assert np.array_equal(o.get_list(HEAT)[:, :3], heat[keep, :3])
    assert np.array_equal(o.heat_layer(), lay)
    assert np.array_equal(o.segments()["hsum"], hsum)
    # neglect: the same particles all stay
    o2 = oracle_for(np.zeros((0, 6)), corners=corners, dt=1e-9)
    o2.set_list(HEAT, heat)
    o2.move_and_clean(True)
    assert len(o2.get_list(HEAT)) == len(pts)
    # zero_forces clears the layer indices (TSpace.cpp:131)
    o.zero_forces()
    assert np.all(o.heat_layer() == -1)


def test_heat_shed_conditions():
    corners = cylinder_corners(40)
    for cond in (0, 1, 2, 3):
        o = oracle_for(np.zeros((0, 6)), corners=corners)
        segs = o.segments()
        rng = np.random.default_rng(cond)
        hc = rng.uniform(0.5, 2.0, len(segs)).astype(np.float32)
        segs["hsum"] = np.where(np.arange(len(segs)) % 3 == 0, 0.0, rng.normal(size=len(segs)) * 1e-3)
        o.set_segments(segs)
        o.set_heat([cond], hc)
        pre = np.zeros((2, 6))
        pre[:, 2] = [4e-3, 5e-3]
        o.set_list(HEAT, pre)
        lay = -np.ones(len(segs), np.int64)
        lay[5], lay[9] = 1, 0   # two segments own a layer particle (const_t)
        o.set_heat_layer(lay)
        o.heat_shed()
        h = o.get_list(HEAT)
        s2 = o.segments()
        pos = np.stack([segs["x"] - 0.5 * segs["dly"], segs["y"] + 0.5 * segs["dlx"]], 1)
        tg = (segs["dlx"] ** 2 + segs["dly"] ** 2) * hc.astype(np.float64)
        if cond == 0:
            assert len(h) == 2 and np.array_equal(s2["hsum"], segs["hsum"])
        elif cond == 1:   # isolate: every segment with hsum != 0 sheds -hsum
            nz = segs["hsum"] != 0
            assert np.array_equal(h[2:, :2], pos[nz]) and np.array_equal(h[2:, 2], -segs["hsum"][nz])
            assert np.all(s2["hsum"] == 0)
        elif cond == 2:   # const_t: layer particles are reset to tg, the rest shed tg
            fresh = lay < 0
            assert np.array_equal(h[2:, :2], pos[fresh]) and np.array_equal(h[2:, 2], tg[fresh])
            assert h[1, 2] == tg[5] and h[0, 2] == tg[9]
            want = segs["hsum"] + tg
            want[5] = segs["hsum"][5] + (tg[5] - 5e-3)
            want[9] = segs["hsum"][9] + (tg[9] - 4e-3)
            assert np.array_equal(s2["hsum"], want)
        else:             # const_w: every segment sheds tg
            assert np.array_equal(h[2:, :2], pos) and np.array_equal(h[2:, 2], tg)
            assert np.array_equal(s2["hsum"], segs["hsum"] + tg)


def test_heat_crop_and_streak_shed():
    o = oracle_for(np.zeros((0, 6)))
    h = np.zeros((6, 6))
    h[:, 0] = [1.0, 17.0, 16.0, -3.0, 30.0, 15.9]
    h[:, 2] = np.arange(1, 7)
    o.set_list(HEAT, h)
    o.heat_crop(16.0)
    assert np.array_equal(o.get_list(HEAT)[:, 2], [1, 3, 4, 6])
    src = np.array([[0.1, 0.2, 0, 0, 0, 0], [0.3, 0.4, 0, 0, 0, 0]])
    o.set_list(SOURCE, src)
    o.set_list(STREAK, np.array([[5.0, 5.0, 0, 0, 0, 0]]))
    o.streak_shed()
    o.streak_shed()
    assert np.array_equal(o.get_list(STREAK)[:, :2], [[5, 5], [0.1, 0.2], [0.3, 0.4], [0.1, 0.2], [0.3, 0.4]])
