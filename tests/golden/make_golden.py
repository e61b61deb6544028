"""Writes the golden fixtures of tests/golden/ from the reference's own code.

Runs in the build container only: it loads oracle/_ref/libvvhd_ref.so, which
oracle/Makefile compiles from the reference's unmodified TSortedTree.cpp,
MEpsilonFast.cpp and MDiffusiveFast.cpp.  Each fixture holds the inputs and
the reference's outputs (tree permutation, leaf table, near/far lists with
far-node monopoles, vortex state after CalcEpsilonFast, velocities and
segment friction after process_vort_list; for heat_cylinder also the heat
and streak lists' permutations and leaf ranges, the heat state after epsh
and the heat velocities after process_heat_list) as a compressed .npz of
plain arrays.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import vvflow_amd as V  # noqa: E402  (host-side generators only)
from oracle_bind import SEG_DTYPE, RefHarness  # noqa: E402

DBL_MAX = np.finfo(np.float64).max


def wake_case(n, seed):
    return V.gen_wake(n, seed), None, dict(fc=8, min_node=0.0, max_node=DBL_MAX, re=600.0)


def mixed_case(n, seed):
    """Uniform positions with random signs: many opposite-sign merges."""
    rng = np.random.default_rng(seed)
    objs = np.zeros((n, 6))
    objs[:, 0] = rng.uniform(0, 1, n)
    objs[:, 1] = rng.uniform(0, 1, n)
    objs[:, 2] = rng.choice([-1.0, 1.0], n) * rng.uniform(0.5, 1.5, n) * 1e-3
    return objs, None, dict(fc=8, min_node=0.0, max_node=DBL_MAX, re=600.0)


def cylinder_case(n, seed, n_seg=500):
    """Vortices in a thin layer around a 500-segment cylinder (merges happen)."""
    corners = V.gen_cylinder(0.5, n_seg)
    segs = V.segments_from_corners(corners)
    rng = np.random.default_rng(seed)
    th = rng.uniform(0, 2 * np.pi, n)
    rr = 0.5 + 1e-4 + rng.exponential(0.03, n)
    objs = np.zeros((n, 6))
    objs[:, 0] = rr * np.cos(th)
    objs[:, 1] = rr * np.sin(th)
    objs[:, 2] = rng.normal(size=n) * 1e-3
    slen = np.sqrt(segs["dlx"] ** 2 + segs["dly"] ** 2).sum()
    dl = slen / (n_seg - 1)
    return objs, segs, dict(fc=8, min_node=5 * dl, max_node=100 * dl, re=600.0)


def run(name, objs, segs, p, merge=True):
    r = RefHarness()
    r.set_re(p["re"])
    tagged = objs.copy()
    tagged[:, 3] = np.arange(len(objs))   # original index travels with each TObj
    r.set_vortices(tagged)
    if segs is not None:
        r.add_body(segs)
    r.tree_build(p["fc"], p["min_node"], p["max_node"])
    after_build = r.vortices()
    perm = after_build[:, 3].astype(np.int32)
    lv = r.leaves()
    # epsilon on the built tree (velocities zero, as after flow-move)
    state = after_build.copy()
    state[:, 3] = 0.0
    r.set_vortices(state)
    merged = r.epsilon(merge)
    after_eps = r.vortices()
    r.diffusive()
    after_diff = r.vortices()
    out = dict(
        objs=objs, fc=np.int32(p["fc"]), min_node=np.float64(p["min_node"]),
        max_node=np.float64(p["max_node"]), re=np.float64(p["re"]), merge=np.int32(merge),
        perm=perm, leaf_geom=lv["geom"], leaf_ints=lv["ints"], near=lv["near"], far=lv["far"],
        leaf_segs=lv["segs"], far_cm=lv["far_cm"], merged=np.int32(merged),
        after_eps=after_eps[:, [0, 1, 2, 5]], after_diff_v=after_diff[:, 3:5])
    if segs is not None:
        out["segs"] = np.frombuffer(segs.tobytes(), dtype=np.uint8)
        out["fric"] = r.fric()
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    print(f"{name}: n={len(objs)} leaves={len(lv['geom'])} merged={merged} -> "
          f"{os.path.getsize(path) / 1024:.0f} KiB")


def heat_case(n, nh, ns, seed, n_seg=500):
    """The cylinder layer of cylinder_case plus heat particles in a thinner
    layer (some in close pairs, so that epsh merges them) and streak
    particles around the body."""
    objs, segs, p = cylinder_case(n, seed, n_seg)
    rng = np.random.default_rng(seed + 100)
    th = rng.uniform(0, 2 * np.pi, nh)
    rr = 0.5 + 1e-4 + rng.exponential(0.02, nh)
    heat = np.zeros((nh, 6))
    heat[:, 0] = rr * np.cos(th)
    heat[:, 1] = rr * np.sin(th)
    heat[:, 2] = rng.uniform(0.5, 1.5, nh) * 1e-4
    pairs = nh // 10   # partners 1e-4 .. 1e-3 away from the first tenth
    ang = rng.uniform(0, 2 * np.pi, pairs)
    d = rng.uniform(1e-4, 1e-3, pairs)
    heat[nh - pairs:, 0] = heat[:pairs, 0] + d * np.cos(ang)
    heat[nh - pairs:, 1] = heat[:pairs, 1] + d * np.sin(ang)
    heat[rng.choice(nh, nh // 50, replace=False), 2] = 0.0   # absorbed particles stay in the list
    streak = np.zeros((ns, 6))
    th = rng.uniform(0, 2 * np.pi, ns)
    rr = 0.5 + rng.uniform(0.01, 1.5, ns)
    streak[:, 0] = rr * np.cos(th)
    streak[:, 1] = rr * np.sin(th)
    p = dict(p, pr=0.7)
    return objs, heat, streak, segs, p


def run_heat(name, objs, heat, streak, segs, p):
    """Tree over vortices, heat and streak lists and a body; CalcEpsilonFast
    (epsv and epsh); process_vort_list and process_heat_list."""
    r = RefHarness()
    r.set_re(p["re"])
    r.set_pr(p["pr"])
    lists = [objs.copy(), heat.copy(), streak.copy()]
    for k, a in enumerate(lists):
        a[:, 3] = np.arange(len(a))   # original index travels with each record
    r.set_vortices(lists[0])
    r.set_list(1, lists[1])
    r.set_list(2, lists[2])
    r.add_body(segs)
    r.tree_build(p["fc"], p["min_node"], p["max_node"])
    built = [r.vortices(), r.get_list(1), r.get_list(2)]
    perms = [b[:, 3].astype(np.int32) for b in built]
    lv = r.leaves()
    hs = r.leaves_hs()
    for b in built:
        b[:, 3] = 0.0
    r.set_vortices(built[0])
    r.set_list(1, built[1])
    r.set_list(2, built[2])
    merged = r.epsilon(True)
    after_eps = r.vortices()
    heat_eps = r.get_list(1)
    r.diffusive()
    r.diffusive_heat()
    heat_diff = r.get_list(1)
    out = dict(
        objs=objs, heat=heat, streak=streak, fc=np.int32(p["fc"]), min_node=np.float64(p["min_node"]),
        max_node=np.float64(p["max_node"]), re=np.float64(p["re"]), pr=np.float64(p["pr"]),
        perm=perms[0], heat_perm=perms[1], streak_perm=perms[2], leaf_geom=lv["geom"], leaf_ints=lv["ints"],
        leaf_hs=hs, near=lv["near"], far=lv["far"], leaf_segs=lv["segs"], merged=np.int32(merged),
        after_eps=after_eps[:, [0, 1, 2, 5]], heat_after_eps=heat_eps[:, [0, 1, 2, 5]],
        heat_after_diff_v=heat_diff[:, 3:5], segs=np.frombuffer(segs.tobytes(), dtype=np.uint8))
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    nz = int((np.abs(heat_eps[:, 2]) > 0).sum())
    print(f"{name}: n={len(objs)} heat={len(heat)} ({nz} left after merges) streak={len(streak)} "
          f"leaves={len(lv['geom'])} merged={merged} -> {os.path.getsize(path) / 1024:.0f} KiB")


def main():
    if not RefHarness.available():
        sys.exit("oracle/_ref/libvvhd_ref.so is missing: run `make -C oracle ref` first")
    o, s, p = wake_case(3000, 0x5EED0001)
    run("wake3000", o, s, p, merge=True)
    o, s, p = wake_case(2000, 0x5EED0002)
    run("wake2000_nomerge", o, s, p, merge=False)
    o, s, p = mixed_case(3000, 11)
    run("mixed3000", o, s, p, merge=True)
    o, s, p = cylinder_case(3000, 7)
    run("cylinder3000", o, s, p, merge=True)
    o, h, st, s, p = heat_case(2000, 1500, 800, 9)
    run_heat("heat_cylinder", o, h, st, s, p)


if __name__ == "__main__":
    main()
