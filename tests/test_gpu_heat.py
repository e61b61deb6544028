"""The heat and streak lists on the GPU (tree ranges, epsh with merges,
convective on the heat and streak targets, process_heat_list, their
flow-move, heat_shed, heat_crop, streak_shed) through the C ABI, against the
reference's own outputs (tests/golden/heat_cylinder.npz) and the FP64
oracle.

Bars: tree permutations of all three lists, leaf ranges, epsh merges and the
heat particles' state after epsilon bit-exact (same FP64 comparisons, the
reference's sequential merge order); flow-move, shedding and cropping
bit-exact (same FP64 operations in the same order).  Velocities: heat and
streak convective targets are summed in FP64 (relative L2 within 1e-12 of
the oracle: only the order of the sums and the device's division differ);
process_heat_list within 1e-10 of the reference (device exp against
glibc's, and the order of the sums).
"""
import numpy as np
import pytest

import vvflow_amd as V
from cases import cylinder_corners, load_golden, oracle_for, params, rel_l2

pytestmark = pytest.mark.gpu

HEAT, STREAK, SOURCE = V.LIST_HEAT, V.LIST_STREAK, V.LIST_STREAK_SOURCE


@pytest.fixture(scope="module")
def ctx():
    c = V.Context(0)
    yield c
    c.close()


def body_record(o, b=0):
    rec = np.zeros(1, dtype=V.BODY_DTYPE)
    ob = o.body(b)
    for f in V.BODY_DTYPE.names:
        rec[f] = ob[f]
    return rec


def mirror(ctx, o, re=600.0, dt=0.005, pr=0.7):
    """The oracle's bodies, vortices, heat and streak lists on the device."""
    ctx.set_params(re=re, dt=dt)
    ctx.set_pr(pr)
    if o.L.vvo_num_bodies(o.h):
        ctx.upload_bodies(np.ascontiguousarray(o.segments(), dtype=V.SEG_DTYPE), body_record(o))
    else:
        ctx.upload_bodies(np.zeros(0, V.SEG_DTYPE), np.zeros(0, V.BODY_DTYPE))
    ctx.upload(o.vortices())
    ctx.upload_list(HEAT, o.get_list(HEAT))
    ctx.upload_list(STREAK, o.get_list(STREAK))


def golden_lists(d):
    lists = []
    for key in ("objs", "heat", "streak"):
        a = d[key].copy()
        a[:, 3] = np.arange(len(a))   # the original index travels in v.x
        lists.append(a)
    return lists


def load_golden_case(ctx, d):
    v, h, s = golden_lists(d)
    segs = d["segs"]
    o = oracle_for(v, segs=segs, re=float(d["re"]))
    ctx.set_params(re=float(d["re"]), dt=0.005)
    ctx.set_pr(float(d["pr"]))
    ctx.upload_bodies(np.ascontiguousarray(o.segments(), dtype=V.SEG_DTYPE), body_record(o))
    ctx.upload(v)
    ctx.upload_list(HEAT, h)
    ctx.upload_list(STREAK, s)


def test_heat_tree_matches_reference(ctx):
    print("case heat_cylinder tree", flush=True)
    d = load_golden("heat_cylinder")
    load_golden_case(ctx, d)
    ctx.tree_build(*params(d))
    assert np.array_equal(ctx.download()[:, 3].astype(np.int32), d["perm"])
    assert np.array_equal(ctx.download_list(HEAT)[:, 3].astype(np.int32), d["heat_perm"])
    assert np.array_equal(ctx.download_list(STREAK)[:, 3].astype(np.int32), d["streak_perm"])
    lv = ctx.tree_leaves()
    assert np.array_equal(lv["geom"], d["leaf_geom"])
    assert np.array_equal(lv["ints"][:, :2], d["leaf_ints"][:, :2])
    assert np.array_equal(ctx.tree_leaf_lists(), d["leaf_hs"])


def test_heat_epsilon_and_diffusive_match_reference(ctx):
    print("case heat_cylinder epsilon + process_heat_list", flush=True)
    d = load_golden("heat_cylinder")
    load_golden_case(ctx, d)
    ctx.tree_build(*params(d))
    # velocities zero before epsilon (as after flow-move)
    v = ctx.download()
    v[:, 3] = 0
    h = ctx.download_list(HEAT)
    h[:, 3] = 0
    st = ctx.download_list(STREAK)
    st[:, 3] = 0
    ctx.upload(v)
    ctx.upload_list(HEAT, h)
    ctx.upload_list(STREAK, st)
    ctx.tree_build(*params(d))
    merged = ctx.epsilon(True)
    assert merged == int(d["merged"])
    assert np.array_equal(ctx.download()[:, [0, 1, 2, 5]], d["after_eps"])
    assert np.array_equal(ctx.download_list(HEAT)[:, [0, 1, 2, 5]], d["heat_after_eps"])
    ctx.diffusive_heat()
    got = ctx.download_list(HEAT)[:, 3:5]
    want = d["heat_after_diff_v"]
    err = rel_l2(got, want)
    print(f"  heat diffusive rel L2 vs reference {err:.2e}", flush=True)
    assert err < 1e-10
    assert np.all((got == 0) == (want == 0))


def heat_case(seed=3, n=3000, nh=800, ns=500):
    rng = np.random.default_rng(seed)
    corners = cylinder_corners(300)
    dl = 300 * np.sin(np.pi / 300) / 299
    th = rng.uniform(0, 2 * np.pi, n)
    rr = 0.5 + 1e-3 + rng.exponential(0.05, n)
    v = np.zeros((n, 6))
    v[:, 0], v[:, 1] = rr * np.cos(th), rr * np.sin(th)
    v[:, 2] = rng.normal(size=n) * 1e-3
    th = rng.uniform(0, 2 * np.pi, nh)
    rr = 0.5 + 1e-3 + rng.exponential(0.03, nh)
    h = np.zeros((nh, 6))
    h[:, 0], h[:, 1] = rr * np.cos(th), rr * np.sin(th)
    h[:, 2] = rng.uniform(0.5, 1.5, nh) * 1e-4
    th = rng.uniform(0, 2 * np.pi, ns)
    rr = 0.5 + rng.uniform(0.01, 1.0, ns)
    s = np.zeros((ns, 6))
    s[:, 0], s[:, 1] = rr * np.cos(th), rr * np.sin(th)
    o = oracle_for(v, corners=corners, inf=(1.0, 0.0))
    o.set_pr(0.7)
    o.set_list(HEAT, h)
    o.set_list(STREAK, s)
    return o, dl


@pytest.mark.parametrize("learned", [False, True])
def test_passive_convective_vs_oracle(ctx, learned):
    print(f"case heat/streak convective (learned={learned})", flush=True)
    o, dl = heat_case()
    o.tree_build(8, 5 * dl, 100 * dl)
    o.epsilon(True)
    o.tree_build(8, 5 * dl, 100 * dl)   # the same tree on both sides after the merges
    mirror(ctx, o)
    ctx.tree_build(8, 5 * dl, 100 * dl)
    assert np.array_equal(ctx.download_list(HEAT), o.get_list(HEAT))
    assert np.array_equal(ctx.download_list(STREAK), o.get_list(STREAK))
    o.convective(threads=8)
    if learned:
        from vvflow_amd.train_farfield import MODEL_PATH, load_model
        ctx.far_model_set(load_model(MODEL_PATH))
        ctx.convective_learned((1.0, 0.0))
    else:
        ctx.convective((1.0, 0.0))
    for lid in (HEAT, STREAK):
        got, want = ctx.download_list(lid)[:, 3:5], o.get_list(lid)[:, 3:5]
        err = rel_l2(got, want)
        print(f"  list {lid}: rel L2 vs oracle {err:.2e}", flush=True)
        assert err < 1e-12


def test_passive_flowmove_vs_oracle(ctx):
    print("case heat/streak flow-move, const_t star body", flush=True)
    k = np.arange(48)
    rad = np.where(k % 2 == 0, 1.0, 0.55)
    ang = 2 * np.pi - np.pi * k / 24
    corners = np.stack([rad * np.cos(ang), rad * np.sin(ang)], 1)
    rng = np.random.default_rng(12)
    nh = 3000
    h = np.zeros((nh, 6))
    h[:, :2] = rng.uniform(-1.2, 1.2, (nh, 2))
    h[:, 2] = rng.uniform(0.5, 2, nh) * 1e-3
    h[::37, 2] = 3e-11   # below remove_eps
    h[:, 3:5] = rng.normal(size=(nh, 2)) * 0.2
    s = np.zeros((1000, 6))
    s[:, :2] = rng.uniform(-1.2, 1.2, (1000, 2))
    s[:, 3:5] = rng.normal(size=(1000, 2)) * 0.2
    o = oracle_for(np.zeros((0, 6)), corners=corners, dt=0.01)
    segs = o.segments()
    hc = rng.uniform(0.5, 1.5, len(segs)).astype(np.float32)
    o.set_heat([V.HC_CONST_T], hc)
    o.set_list(HEAT, h)
    o.set_list(STREAK, s)
    mirror(ctx, o, dt=0.01)
    ctx.set_heat([V.HC_CONST_T], hc)
    for step in range(2):
        o.move_and_clean(True)
        ctx.move_and_clean(True)
        for lid in (HEAT, STREAK):
            assert np.array_equal(ctx.download_list(lid), o.get_list(lid)), (step, lid)
        assert np.array_equal(ctx.heat_layer(), o.heat_layer())
        assert np.array_equal(ctx.download_bodies()[0]["hsum"], o.segments()["hsum"])
        # the heat layer took some particles and absorbed others
        assert (o.heat_layer() >= 0).sum() > 0
        # shed into the layer: const_t resets the layer particles' strengths
        o.heat_shed()
        ctx.heat_shed()
        assert np.array_equal(ctx.download_list(HEAT), o.get_list(HEAT))
        assert np.array_equal(ctx.download_bodies()[0]["hsum"], o.segments()["hsum"])
        h2 = o.get_list(HEAT)
        h2[:, 3:5] = rng.normal(size=(len(h2), 2)) * 0.2
        o.set_list(HEAT, h2)
        ctx.upload_list(HEAT, h2)


@pytest.mark.parametrize("cond", [V.HC_ISOLATE, V.HC_CONST_W])
def test_heat_shed_crop_streak_shed_vs_oracle(ctx, cond):
    print(f"case heat_shed condition {cond}, heat_crop, streak_shed", flush=True)
    o = oracle_for(np.zeros((0, 6)), corners=cylinder_corners(200))
    segs = o.segments()
    rng = np.random.default_rng(cond)
    segs["hsum"] = np.where(np.arange(len(segs)) % 4 == 0, 0.0, rng.normal(size=len(segs)) * 1e-3)
    o.set_segments(segs)
    hc = rng.uniform(0.5, 2.0, len(segs)).astype(np.float32)
    o.set_heat([cond], hc)
    h = np.zeros((50, 6))
    h[:, 0] = rng.uniform(-5, 30, 50)
    h[:, 2] = 1e-3
    o.set_list(HEAT, h)
    src = rng.uniform(-1, 1, (7, 6))
    o.set_list(SOURCE, src)
    mirror(ctx, o)
    ctx.set_heat([cond], hc)
    ctx.upload_list(SOURCE, src)
    o.heat_shed()
    ctx.heat_shed()
    assert np.array_equal(ctx.download_list(HEAT), o.get_list(HEAT))
    assert np.array_equal(ctx.download_bodies()[0]["hsum"], o.segments()["hsum"])
    o.heat_crop(16.0)
    ctx.heat_crop(16.0)
    assert np.array_equal(ctx.download_list(HEAT), o.get_list(HEAT))
    o.streak_shed()
    ctx.streak_shed()
    assert np.array_equal(ctx.download_list(STREAK), o.get_list(STREAK))
    assert np.array_equal(ctx.download_list(SOURCE), src)


def test_heated_cylinder_step_vs_oracle(ctx):
    """One step of the reference's loop with a heated (const_w) cylinder and
    streak sources, from the same state on both sides: heat_shed,
    vortex_shed, streak_shed, tree, epsilon (both lists), convective,
    process_vort_list, process_heat_list, flow-move, heat_crop."""
    print("case heated cylinder step", flush=True)
    o, dl = heat_case(seed=21)
    segs = o.segments()
    segs["g"] = np.random.default_rng(2).normal(size=len(segs)) * 1e-3
    o.set_segments(segs)
    hc = np.full(len(segs), 0.5, np.float32)
    o.set_heat([V.HC_CONST_W], hc)
    src = np.zeros((20, 6))
    src[:, 0] = -1.0
    src[:, 1] = np.linspace(-1, 1, 20)
    o.set_list(SOURCE, src)
    mirror(ctx, o)
    ctx.set_heat([V.HC_CONST_W], hc)
    ctx.upload_list(SOURCE, src)
    for side in (o, ctx):
        side.heat_shed()
        side.vortex_shed()
        side.streak_shed()
    assert np.array_equal(ctx.download_list(HEAT), o.get_list(HEAT))
    assert np.array_equal(ctx.download(), o.vortices())
    o.tree_build(8, 5 * dl, 100 * dl)
    ctx.tree_build(8, 5 * dl, 100 * dl)
    assert o.epsilon(True)[3] == ctx.epsilon(True)
    for lid in (HEAT, STREAK):
        assert np.array_equal(ctx.download_list(lid), o.get_list(lid))
    o.convective(threads=8)
    o.diffusive(threads=8)
    o.diffusive_heat(threads=8)
    ctx.convective((1.0, 0.0))
    ctx.diffusive()
    ctx.diffusive_heat()
    for lid in (HEAT, STREAK):
        err = rel_l2(ctx.download_list(lid)[:, 3:5], o.get_list(lid)[:, 3:5])
        print(f"  list {lid} velocity rel L2 vs oracle {err:.2e}", flush=True)
        assert err < 1e-10
    hv = o.get_list(HEAT)
    o.move_and_clean(True)
    # the device moves the heat list with its own velocities: move the
    # oracle's with the same ones so that removal decisions can be compared
    o2 = oracle_for(np.zeros((0, 6)), corners=cylinder_corners(300))
    hg = ctx.download_list(HEAT)
    ctx.move_and_clean(True)
    o2.set_list(HEAT, hg)
    o2.set_heat([V.HC_CONST_W], hc)
    o2.move_and_clean(True)
    assert np.array_equal(ctx.download_list(HEAT), o2.get_list(HEAT))
    assert len(o.get_list(HEAT)) == len(o2.get_list(HEAT))
    assert np.allclose(o.get_list(HEAT)[:, :2], o2.get_list(HEAT)[:, :2], rtol=0, atol=1e-12)
    assert len(hv) > 800
    o.heat_crop(0.7)
    ctx.heat_crop(0.7)
    assert len(ctx.download_list(HEAT)) == len(o.get_list(HEAT))
