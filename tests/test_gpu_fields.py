"""Point queries and grid fields on the GPU (vvhd_eps2h, vvhd_h2, vvhd_field)
against the FP64 oracle on the same tree and state.

Bars: eps2h and h2 bit-exact (the same FP64 differences and products, and a
top-2 / minimum that does not depend on the order).  Grid maps (float, as
XField::map) within 1e-5 of the oracle's relative to the map's largest value
(device exp / log / erf / atan2 against glibc's, and the order of the sums).
"""
import numpy as np
import pytest

import vvflow_amd as V
from cases import cylinder_corners, layer_case, oracle_for

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = V.Context(0)
    yield c
    c.close()


def body_record(o):
    rec = np.zeros(1, dtype=V.BODY_DTYPE)
    ob = o.body(0)
    for f in V.BODY_DTYPE.names:
        rec[f] = ob[f]
    return rec


def mirror(ctx, o, vort, re=600.0, dt=0.005, inf=(1.0, 0.0)):
    """The oracle's bodies and the given vortex records on the device."""
    ctx.set_params(re=re, dt=dt)
    ctx.upload_bodies(np.ascontiguousarray(o.segments(), dtype=V.SEG_DTYPE), body_record(o))
    ctx.upload(np.ascontiguousarray(vort, dtype=np.float64))


def case(n=6000, seed=11):
    objs, corners = layer_case(n, seed, n_seg=400)
    o = oracle_for(objs, corners=corners, inf=(1.0, 0.0))
    dl = o.body(0)["slen"] / 399
    return objs, o, dl


def assert_map_close(g, o, tol=1e-5):
    scale = max(float(np.abs(o).max()), 1e-30)
    err = float(np.abs(g.astype(np.float64) - o.astype(np.float64)).max()) / scale
    assert err < tol, err
    return err


def test_eps2h_h2_vs_oracle(ctx):
    print("case eps2h/h2 layer n=6000", flush=True)
    objs, o, dl = case()
    o.tree_build(8, dl * 20)
    mirror(ctx, o, objs)
    ctx.tree_build(8, dl * 20)
    assert np.array_equal(ctx.download()[:, :3], o.vortices()[:, :3])
    assert np.array_equal(ctx.eps2h(), o.eps2h(None))
    assert np.array_equal(ctx.h2(), o.eps2h(None, h2=True))
    pts = np.random.default_rng(3).uniform(-1.5, 1.5, (500, 2))
    assert np.array_equal(ctx.eps2h(pts), o.eps2h(pts))
    assert np.array_equal(ctx.h2(pts), o.eps2h(pts, h2=True))
    # over an explicit leaf: the near list of leaf 0 for every point
    lv = ctx.tree_leaves()
    e0 = ctx.eps2h(pts[:20], leaf=np.zeros(20, np.int32))
    T = o.tree()
    near0 = T["near"][1][T["near"][0][0]:T["near"][0][1]]
    v = o.vortices()
    src = np.concatenate([v[T["ints"][m, 0]:T["ints"][m, 1], :2] for m in near0])
    for p, e in zip(pts[:20], e0):
        d = np.sum((src - p) ** 2, 1)
        d = d[d != 0]
        want = np.partition(d, 1)[1] if len(d) > 1 else (d[0] if len(d) else np.finfo(np.float64).min)
        assert e == want
    assert len(lv["geom"]) == len(T["leaves"])


@pytest.mark.parametrize("kind", ["vx", "vy", "vorticity", "streamfunction", "pressure"])
def test_fields_vs_oracle(ctx, kind):
    print(f"case field {kind}", flush=True)
    objs, o, dl = case(4000, 12)
    grid = (-1.2, -1.0, 0.04, 70, 50)
    m = 1.5
    ref = (0.2, -0.1)
    if kind in ("vx", "vy"):
        o.tree_build(8, dl * 20, 0.1)
        o.epsilon(False)
        state = o.vortices()
        mirror(ctx, o, objs)
        ctx.tree_build(8, dl * 20, 0.1)
        ctx.epsilon(False)
        assert np.array_equal(ctx.download()[:, [0, 1, 2, 5]], state[:, [0, 1, 2, 5]])
        k = V.FIELD_VX if kind == "vx" else V.FIELD_VY
        g = ctx.field(k, *grid, inf=(1.0, 0.0), ref=ref)
        w = o.field(0 if kind == "vx" else 1, *grid, ref=ref)
    elif kind == "vorticity":
        segs = o.segments()
        segs["g"] = np.random.default_rng(6).normal(size=len(segs)) * 1e-3
        o.set_segments(segs)
        o.vortex_shed()
        shed = o.vortices()
        o.tree_build(8, dl * 20)
        mirror(ctx, o, shed)
        ctx.tree_build(8, dl * 20)
        g = ctx.field(V.FIELD_VORTICITY, *grid, eps_mult=m, dl=dl)
        w = o.field(2, *grid, eps_mult=m, dl=dl)
    elif kind == "streamfunction":
        segs = o.segments()
        segs["g"] = np.random.default_rng(4).normal(size=len(segs)) * 1e-3
        o.set_segments(segs)
        st = o.body(0)
        st["speed_slae"] = (0.05, -0.02, 0.1)   # a moving body exercises the surface terms
        o.set_body(0, st)
        o.tree_build(8, dl * 10, dl * 20)
        mirror(ctx, o, objs)
        ctx.tree_build(8, dl * 10, dl * 20)
        g = ctx.field(V.FIELD_STREAMFUNCTION, *grid, eps_mult=m, dl=dl, inf=(1.0, 0.0), ref=ref)
        w = o.field(3, *grid, eps_mult=m, dl=dl, ref=ref)
    else:
        segs = o.segments()
        rng = np.random.default_rng(5)
        segs["gsum"] = rng.normal(size=len(segs)) * 1e-3   # the running gsum term
        o.set_segments(segs)
        o.tree_build(8, dl * 20, 0.1)
        o.epsilon(False)
        o.convective(threads=8)
        o.diffusive()
        state = o.vortices()   # the vortices' velocities enter the pressure
        mirror(ctx, o, state)
        ctx.tree_build(8, dl * 20, 0.1)
        ctx.epsilon(False)
        assert np.array_equal(ctx.download(), state)
        g = ctx.field(V.FIELD_PRESSURE, *grid, eps_mult=m, dl=dl, inf=(1.0, 0.0), ref=ref)
        w = o.field(4, *grid, eps_mult=m, dl=dl, ref=ref)
    err = assert_map_close(g, w)
    print(f"  max |gpu - oracle| / max |oracle| = {err:.2e}", flush=True)
    assert np.abs(w).max() > 0
