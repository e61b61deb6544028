"""GPU (gfx950) hot path through the C ABI against the reference's own
outputs (tests/golden) and the FP64 oracle.

Bars: tree permutation, leaf table, near/far lists, monopoles, epsilon
neighbour indices, merges and the post-epsilon vortex state are bit-exact.
Convective velocity: FP32 pair sums with Kahan compensation, relative L2 of
the induced velocity within 2e-6 of the FP64 oracle (plain FP32 within 1e-5).
Diffusive velocity: FP32 pair sums, relative L2 within 1e-5 and pointwise
within 1e-3 (floor 1e-3 rms) of the FP64 reference (exp-cutoff decisions at
the |dr|/eps = 8 boundary may flip in FP32).  Segment friction: FP64, 1e-12.
"""
import os

import numpy as np
import pytest

import vvflow_amd as V
from cases import (DBL_MAX, FIXTURES, layer_case, load_golden, max_rel, oracle_for, params, path_code32,
                   random_cluster_case, rel_l2)
from oracle_bind import direct_sum

pytestmark = pytest.mark.gpu

CONV_TOL_KAHAN = 2e-6
CONV_TOL_PLAIN = 1e-5
DIFF_TOL_L2 = 1e-5
DIFF_TOL_MAX = 1e-3


@pytest.fixture(scope="module")
def ctx():
    c = V.Context(0)
    yield c
    c.close()


def body_records(segs, oracle=None):
    b = np.zeros(1, dtype=V.BODY_DTYPE)
    b["first_seg"] = 0
    b["nseg"] = len(segs)
    if oracle is not None:
        ob = oracle.body(0)
        b["slen"], b["area"] = ob["slen"], ob["area"]
    return b


def load_case(ctx, objs, segs=None, re=600.0):
    ctx.set_params(re=re, dt=0.005)
    if segs is None:
        ctx.upload_bodies(np.zeros(0, V.SEG_DTYPE), np.zeros(0, V.BODY_DTYPE))
    else:
        ctx.upload_bodies(np.ascontiguousarray(segs, dtype=V.SEG_DTYPE), body_records(segs))
    ctx.upload(objs)


# ------------------------------------------------------------ golden (reference)

@pytest.mark.parametrize("name", FIXTURES)
def test_tree_matches_reference(ctx, name):
    d = load_golden(name)
    print(f"case golden {name}", flush=True)
    load_case(ctx, d["objs"], d["segs"], float(d["re"]))
    perm = ctx.tree_build(*params(d), perm=True)
    assert np.array_equal(perm, d["perm"])
    lv = ctx.tree_leaves()
    assert np.array_equal(lv["geom"], d["leaf_geom"])
    assert np.array_equal(lv["ints"][:, :2], d["leaf_ints"][:, :2])
    assert np.array_equal(lv["ints"][:, 2], d["leaf_ints"][:, 2])
    assert np.array_equal(path_code32(lv["path"]), d["leaf_ints"][:, 3].astype(np.uint32).astype(np.int64))
    for far in (False, True):
        g = ctx.tree_lists(far=far)
        ref = d["far" if far else "near"].astype(np.int64)
        assert np.array_equal(g["depth"], ref[:, 0])
        assert np.array_equal(path_code32(g["path"]), ref[:, 1].astype(np.uint32).astype(np.int64))
        if far:
            cm = d["far_cm"]
            assert np.array_equal(g["cm"], cm[:, [0, 1, 2, 4, 5, 6]])


@pytest.mark.parametrize("name", FIXTURES)
def test_epsilon_and_diffusive_match_reference(ctx, name):
    d = load_golden(name)
    print(f"case golden eps/diff {name}", flush=True)
    load_case(ctx, d["objs"], d["segs"], float(d["re"]))
    ctx.tree_build(*params(d))
    merged = ctx.epsilon(bool(d["merge"]))
    assert merged == int(d["merged"])
    st = ctx.download()
    assert np.array_equal(st[:, [0, 1, 2, 5]], d["after_eps"])
    ctx.diffusive()
    v = ctx.download()[:, 3:5]
    assert rel_l2(v, d["after_diff_v"]) < DIFF_TOL_L2
    assert max_rel(v, d["after_diff_v"]) < DIFF_TOL_MAX
    if d["segs"] is not None:
        segs, _ = ctx.download_bodies()
        assert np.allclose(segs["fric"], d["fric"], rtol=1e-12, atol=1e-15 * np.abs(d["fric"]).max())


# ------------------------------------------------------------ random vs oracle

def run_both(ctx, objs, corners=None, fc=8, mn=0.0, mx=DBL_MAX, merge=True, kahan=True, inf=(1.0, 0.0)):
    o = oracle_for(objs, corners=corners, inf=inf)
    segs = o.segments() if corners is not None else None
    load_case(ctx, objs, segs)
    o.tree_build(fc, mn, mx)
    perm = ctx.tree_build(fc, mn, mx, perm=True)
    ov = o.vortices()
    gv = ctx.download()
    assert np.array_equal(gv[:, :3], ov[:, :3])
    nn1o, nn2o, _, mo = o.epsilon(merge)
    nn1, nn2, mg = ctx.epsilon(merge, want_nn=True)
    assert mg == mo
    assert np.array_equal(nn1, nn1o) and np.array_equal(nn2, nn2o)
    ov = o.vortices()
    gv = ctx.download()
    assert np.array_equal(gv[:, [0, 1, 2, 5]], ov[:, [0, 1, 2, 5]], equal_nan=True)
    o.convective(threads=4)
    ctx.convective(inf, kahan=kahan)
    ov = o.vortices()
    gv = ctx.download()
    ind = ov[:, 3:5] - np.asarray(inf)
    conv_err = rel_l2(gv[:, 3:5] - np.asarray(inf), ind)
    o.diffusive()
    ctx.diffusive()
    ov2 = o.vortices()
    gv2 = ctx.download()
    dv_o = ov2[:, 3:5] - ov[:, 3:5]
    dv_g = gv2[:, 3:5] - gv[:, 3:5]
    return conv_err, dv_g, dv_o, o


@pytest.mark.parametrize("seed,signs,n", [(1, "mixed", 4000), (2, "cluster", 6000), (3, "positive", 3000),
                                          (4, "cluster", 12000)])
def test_random_clusters_vs_oracle(ctx, seed, signs, n):
    print(f"case random seed={seed} signs={signs} n={n}", flush=True)
    objs = random_cluster_case(n, seed, signs)
    conv_err, dv_g, dv_o, _ = run_both(ctx, objs)
    assert conv_err < CONV_TOL_KAHAN
    assert rel_l2(dv_g, dv_o) < DIFF_TOL_L2
    assert max_rel(dv_g, dv_o) < DIFF_TOL_MAX


def test_large_leaves_vs_oracle(ctx):
    """Min node 0.02 stops the splits inside the clusters: leaves of up to 63
    vortices, and 115 of the 958 batches of 64 near-list entries hold more
    than 1024 sources, so the near kernel locates sources both through its
    LDS index table and through the prefix-sum search."""
    print("case large leaves", flush=True)
    conv_err, dv_g, dv_o, _ = run_both(ctx, random_cluster_case(8000, 41, "cluster"), mn=0.02)
    assert conv_err < CONV_TOL_KAHAN
    assert rel_l2(dv_g, dv_o) < DIFF_TOL_L2
    assert max_rel(dv_g, dv_o) < DIFF_TOL_MAX


def test_plain_fp32_accumulation(ctx):
    print("case plain fp32", flush=True)
    conv_err, _, _, _ = run_both(ctx, random_cluster_case(5000, 7, "cluster"), kahan=False)
    assert conv_err < CONV_TOL_PLAIN


@pytest.mark.parametrize("fc", [2, 30])
def test_far_criteria_vs_oracle(ctx, fc):
    print(f"case fc={fc}", flush=True)
    conv_err, dv_g, dv_o, _ = run_both(ctx, random_cluster_case(5000, 9, "cluster"), fc=fc)
    assert conv_err < CONV_TOL_KAHAN
    assert rel_l2(dv_g, dv_o) < DIFF_TOL_L2


def test_no_merge_mode(ctx):
    print("case nomerge", flush=True)
    run_both(ctx, random_cluster_case(3000, 11, "mixed"), merge=False)


@pytest.mark.parametrize("seed", [1, 2])
def test_body_layer_vs_oracle(ctx, seed):
    print(f"case body layer seed={seed}", flush=True)
    objs, corners = layer_case(4000, seed)
    dl = np.pi / 499
    conv_err, dv_g, dv_o, o = run_both(ctx, objs, corners=corners, mn=5 * dl, mx=100 * dl)
    assert conv_err < CONV_TOL_KAHAN
    assert rel_l2(dv_g, dv_o) < DIFF_TOL_L2
    segs, _ = ctx.download_bodies()
    fo = o.segments()["fric"]
    assert np.allclose(segs["fric"], fo, rtol=1e-10, atol=1e-13 * np.abs(fo).max())


@pytest.mark.parametrize("n", [1, 2, 3, 17, 40])
def test_tiny_inputs(ctx, n):
    print(f"case tiny n={n}", flush=True)
    run_both(ctx, random_cluster_case(n, 20 + n, "mixed"))


def test_duplicate_positions(ctx):
    print("case duplicates", flush=True)
    base = random_cluster_case(500, 3, "mixed")
    dup = np.concatenate([base, base[:60], base[:7]])
    dup = dup[np.random.default_rng(0).permutation(len(dup))]
    o = oracle_for(dup)
    load_case(ctx, dup)
    o.tree_build(8)
    perm = ctx.tree_build(8, perm=True)
    assert np.array_equal(ctx.download()[:, :3], o.vortices()[:, :3])
    o.epsilon(True)
    ctx.epsilon(True)
    assert np.array_equal(ctx.download()[:, [0, 1, 2, 5]], o.vortices()[:, [0, 1, 2, 5]], equal_nan=True)


def test_velocity_points_vs_oracle(ctx):
    print("case velocity points", flush=True)
    objs = random_cluster_case(5000, 13, "cluster")
    o = oracle_for(objs)
    load_case(ctx, objs)
    o.tree_build(8)
    ctx.tree_build(8)
    o.epsilon(True)
    ctx.epsilon(True)
    pts = np.random.default_rng(2).uniform(-0.1, 1.1, (300, 2))
    vo = o.velocity(pts)
    vg = ctx.velocity(pts, inf=(1.0, 0.0))
    assert np.allclose(vg, vo, rtol=1e-10, atol=1e-12)


def test_tree_find_and_integrals_vs_oracle(ctx):
    """stree::findNode (TSortedTree.cpp:284-303) for vortex positions, random
    points and points far outside the root box (the descent still ends in a
    leaf, as the reference's does), and Space::integral / the impulse
    integrals after merges: the same leaf and the same sums as the oracle."""
    print("case tree find and integrals", flush=True)
    objs = random_cluster_case(20000, 21, "mixed")
    o = oracle_for(objs)
    load_case(ctx, objs)
    o.tree_build(8)
    ctx.tree_build(8)
    leaves = o.tree()["leaves"]
    rng = np.random.default_rng(4)
    v = o.vortices()
    pts = np.concatenate([v[rng.choice(len(v), 500, replace=False), :2],
                          rng.uniform(-0.2, 1.2, (500, 2)),
                          np.array([[50.0, 50.0], [-40.0, 3.0], [0.5, -60.0]])])
    got = ctx.tree_find(pts)
    want = np.array([o.find(x, y) for x, y in pts])
    assert np.all(got >= 0)
    assert np.array_equal(leaves[got], want)
    o.epsilon(True)
    ctx.epsilon(True)
    gi, oi = ctx.integrals(), o.integrals()
    # mixed signs: the sums may cancel, so the floor scales with sum |g|
    assert np.allclose(gi, oi, rtol=1e-12, atol=1e-13 * np.abs(objs[:, 2]).sum()), (gi, oi)
    assert gi[4] == oi[4]


def test_direct_sum_vs_oracle(ctx):
    print("case direct sum", flush=True)
    objs = random_cluster_case(3000, 15, "mixed")
    objs[:, 5] = np.random.default_rng(1).uniform(100, 1000, len(objs))
    load_case(ctx, objs)
    tg = np.arange(0, 3000, 7, dtype=np.int64)
    assert np.allclose(ctx.direct_sum(tg), direct_sum(objs, tg), rtol=1e-12, atol=1e-15)


def test_repeatable(ctx):
    print("case repeatability", flush=True)
    objs = random_cluster_case(8000, 17, "mixed")
    outs = []
    for _ in range(2):
        load_case(ctx, objs)
        ctx.tree_build(8)
        ctx.epsilon(True)
        ctx.convective((1.0, 0.0))
        ctx.diffusive()
        outs.append(ctx.download())
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("learned", [False, True])
def test_diffusive_beside_convective(ctx, learned):
    """A diffusive call right after a convective one runs on a second stream
    beside the convective kernels and adds its increments afterwards (and the
    learned near field beside its far field); any call in between keeps the
    diffusive serial, and vvhd_set_overlap(0) runs every kernel alone.  All
    three give identical velocities."""
    print(f"case diffusive beside convective learned={learned}", flush=True)
    from vvflow_amd.train_farfield import MODEL_PATH, load_model
    if learned and not os.path.exists(MODEL_PATH):
        pytest.skip("no trained far-field model in the tree")
    objs = V.gen_wake(200_000, 0x5EED0002)
    load_case(ctx, objs)
    if learned:
        ctx.far_model_set(load_model(MODEL_PATH))
    conv = (lambda: ctx.convective_learned((1.0, 0.0))) if learned else (lambda: ctx.convective((1.0, 0.0)))
    out = []
    for mode in ("overlap", "serial diffusive", "no overlap"):
        load_case(ctx, objs)
        ctx.set_overlap(mode != "no overlap")
        ctx.tree_build(8)
        ctx.epsilon(True)
        conv()
        if mode == "serial diffusive":
            ctx.phase_times()
        ctx.diffusive()
        out.append(ctx.download())
    ctx.set_overlap(True)
    assert np.array_equal(out[0], out[1], equal_nan=True)
    assert np.array_equal(out[0], out[2], equal_nan=True)


# ------------------------------------------------------------ full size

def test_full_size_tree_repeatable(ctx):
    """Three builds of the 1M-vortex wake give the same tree: permutation,
    every node's geometry (exact bounding boxes) and element range, keyed by
    (depth, path) since device node ids follow allocation order, and the leaf
    table.  Races inside a level (e.g. reading coordinates that a concurrent
    swap rewrites) show up here long before they do on small cases."""
    print("case wake 1M tree repeatability", flush=True)
    objs = V.gen_wake(1_000_000)
    ref = None
    for _ in range(3):
        load_case(ctx, objs)
        perm = ctx.tree_build(8, perm=True)
        nodes = ctx.tree_nodes()
        order = np.lexsort((nodes["path"], nodes["ints"][:, 3]))
        out = (perm, nodes["geom"][order], nodes["ints"][order][:, [0, 1, 3]], nodes["path"][order],
               ctx.tree_leaves()["geom"])
        if ref is None:
            ref = out
        else:
            for a, b in zip(ref, out):
                assert np.array_equal(a, b, equal_nan=True)


def test_full_size_wake_properties(ctx):
    """1M-vortex wake: permutation, leaf partition, circulation conservation
    through merges, agreement with the FP64 direct sum on sampled targets."""
    print("case wake 1M", flush=True)
    n = 1_000_000
    objs = V.gen_wake(n)
    load_case(ctx, objs)
    perm = ctx.tree_build(8, perm=True)
    assert np.array_equal(np.sort(perm), np.arange(n, dtype=np.int32))
    lv = ctx.tree_leaves()
    r = lv["ints"][:, :2]
    assert r[0, 0] == 0 and r[-1, 1] == n and np.all(r[1:, 0] == r[:-1, 1]) and np.all(r[:, 1] - r[:, 0] < 16)
    g0 = objs[:, 2].sum()
    merged = ctx.epsilon(True)
    st = ctx.download()
    assert abs(st[:, 2].sum() - g0) < 1e-13 * np.abs(objs[:, 2]).sum()
    assert np.sum(st[:, 2] == 0) == merged
    ctx.convective((0.0, 0.0))
    v = ctx.download()[:, 3:5]
    tg = np.random.default_rng(0).choice(n, 2000, replace=False).astype(np.int64)
    ref = ctx.direct_sum(tg)
    err = rel_l2(v[tg], ref)
    print(f"  merged={merged} tree-method rel L2 vs direct sum (fc=8) = {err:.3e}", flush=True)
    assert err < 1e-2


# ------------------------------------------------------------ merge commit paths

def _one_flag_chain_case():
    """A single-leaf set (fewer than 16 vortices) where pass 1 flags one
    target and its merge flags another: the parallel commit must give way to
    the sequential pass.  Found by a seeded search checked on the oracle."""
    rng = np.random.default_rng(1)
    for _ in range(2000):
        n = int(rng.integers(5, 15))
        v = np.zeros((n, 6))
        v[:, :2] = rng.uniform(0, 1, (n, 2))
        v[:, 2] = rng.choice([-1.0, 1.0], n) * rng.uniform(0.5, 1.5, n) * 1e-3
        d2 = ((v[:, None, :2] - v[None, :, :2]) ** 2).sum(-1)
        np.fill_diagonal(d2, np.inf)
        order = np.argsort(d2, axis=1, kind="stable")[:, :2]
        s = np.sign(v[:, 2])
        flags = (s[order[:, 0]] == s[order[:, 1]]) & (s[order[:, 0]] != s)
        if flags.sum() != 1:
            continue
        o = oracle_for(v, inf=(0.0, 0.0))
        o.tree_build(8)
        if o.epsilon(True)[3] >= 2:
            return v
    raise AssertionError("no case found")


def test_merge_that_flags_another(ctx):
    print("case one flag whose merge flags another", flush=True)
    v = _one_flag_chain_case()
    run_both(ctx, v, inf=(0.0, 0.0))


def test_wake_merges_vs_oracle(ctx):
    """A 200k wake at far criterion 16 merges a few vortex pairs: neighbour
    indices, merges and the post-epsilon state against the oracle."""
    print("case wake 200000 merges fc=16", flush=True)
    objs = V.gen_wake(200_000, 0x5EED0003)
    o = oracle_for(objs, inf=(0.0, 0.0))
    load_case(ctx, objs)
    o.tree_build(16)
    ctx.tree_build(16)
    nn1o, nn2o, _, mo = o.epsilon(True)
    nn1, nn2, mg = ctx.epsilon(True, want_nn=True)
    print(f"  merged {mg}", flush=True)
    assert mg == mo and mg > 0
    assert np.array_equal(nn1, nn1o) and np.array_equal(nn2, nn2o)
    assert np.array_equal(ctx.download()[:, [0, 1, 2, 5]], o.vortices()[:, [0, 1, 2, 5]], equal_nan=True)
