"""Launch forms of the convective phase: the near field one wave per leaf or
as a persistent grid (8, 16, 32 waves per compute unit) taking work items
from a device queue, the autotuner's choice, and the plain convective phase
replayed as a captured HIP graph.

Bar: every form, the tuned choice and every graph replay give bit-identical
velocities (each work item runs exactly the arithmetic of one wave per leaf
for its targets).  Graph replays are checked one by one against the eager
result, with eager work of the library between them.
"""
import numpy as np
import pytest

import vvflow_amd as V
from cases import layer_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = V.Context(0)
    yield c
    c.close()


def cases():
    wake = V.gen_wake(200_000, 0x5EED0031)
    layer, corners = layer_case(40_000, 3)   # leaves of many targets near the body
    return [("wake 200k", wake, None, (8, 0.0, np.finfo(np.float64).max)),
            ("cylinder layer 40k", layer, corners, (8, 0.0314, 0.63))]


def load(ctx, objs, corners):
    ctx.set_params(re=600.0, dt=0.005)
    if corners is None:
        ctx.upload_bodies(np.zeros(0, V.SEG_DTYPE), np.zeros(0, V.BODY_DTYPE))
    else:
        segs = V.segments_from_corners(corners)
        b = np.zeros(1, dtype=V.BODY_DTYPE)
        b["nseg"] = len(segs)
        ctx.upload_bodies(np.ascontiguousarray(segs, dtype=V.SEG_DTYPE), b)
    ctx.upload(objs)
    ctx.snapshot()


def velocities(ctx, tree, kahan=True):
    ctx.restore()
    ctx.tree_build(*tree)
    ctx.epsilon(True)
    ctx.convective((1.0, 0.0), kahan=kahan)
    return ctx.download()


@pytest.mark.parametrize("kahan", [True, False])
def test_near_forms_are_bit_identical(ctx, kahan):
    for name, objs, corners, tree in cases():
        print(f"case near forms, {name}, kahan={kahan}", flush=True)
        load(ctx, objs, corners)
        ctx.set_tuning(False, 0)
        ref = velocities(ctx, tree, kahan)
        for form in (8, 16, 32):
            ctx.set_tuning(False, form)
            got = velocities(ctx, tree, kahan)
            assert ctx.tuning()["near_form"] == form
            assert np.array_equal(got, ref), f"{name}: form {form}"
        ctx.set_tuning(True)
        got = velocities(ctx, tree, kahan)
        t = ctx.tuning()
        print(f"  tuned form {t['near_form']}", flush=True)
        assert t["classes_tuned"] >= 1 and np.array_equal(got, ref)
    ctx.set_tuning(True)


def test_graph_replays_match_eager(ctx):
    for name, objs, corners, tree in cases():
        print(f"case convective graph, {name}", flush=True)
        load(ctx, objs, corners)
        ctx.set_graph(False)
        ctx.set_tuning(False, 16)   # the graph needs the persistent form
        ref = velocities(ctx, tree)
        before = ctx.tuning()
        ctx.set_graph(True)
        for rep in range(3):
            got = velocities(ctx, tree)
            assert np.array_equal(got, ref), f"{name}: replay {rep}"
            ctx.set_graph(False)   # eager work between the replays
            assert np.array_equal(velocities(ctx, tree), ref)
            ctx.set_graph(True)
        after = ctx.tuning()
        print(f"  graphs captured {after['graph_captures'] - before['graph_captures']}, "
              f"launched {after['graph_launches'] - before['graph_launches']}", flush=True)
        assert after["graph_launches"] - before["graph_launches"] == 3
        assert after["graph_captures"] - before["graph_captures"] <= 2
        ctx.set_graph(False)
    ctx.set_tuning(True)


def test_graph_whole_cylinder_steps(ctx):
    """The captured convective phase inside whole cylinder steps (the vortex
    arrays swap with their double buffers at every flow-move): the same
    bits as eager steps."""
    print("case graph in whole cylinder steps", flush=True)
    nseg = 300
    corners = V.gen_cylinder(0.5, nseg)
    segs = np.ascontiguousarray(V.segments_from_corners(corners), dtype=V.SEG_DTYPE)
    body = np.zeros(1, dtype=V.BODY_DTYPE)
    body["nseg"] = nseg
    dl = nseg * np.sin(np.pi / nseg) / (nseg - 1)
    rng = np.random.default_rng(7)
    n = 20_000
    th = rng.uniform(0, 2 * np.pi, n)
    rr = 0.5 + 2e-3 + rng.exponential(0.1, n)
    objs = np.zeros((n, 6))
    objs[:, 0], objs[:, 1] = rr * np.cos(th), rr * np.sin(th)
    objs[:, 2] = rng.normal(size=n) * 1e-3

    def run(graph, steps=4):
        ctx.set_graph(graph)
        ctx.set_tuning(False, 16)
        ctx.set_params(re=600.0, dt=0.005, time=0.0)
        ctx.upload_bodies(segs, body)
        ctx.upload(objs)
        out = []
        for k in range(steps):
            ctx.set_params(re=600.0, dt=0.005, time=k * 0.005)
            ctx.tree_build(8, 5 * dl, 100 * dl)
            ctx.calc_circulation((1.0, 0.0))
            ctx.tree_destroy()
            ctx.vortex_shed()
            ctx.tree_build(8, 5 * dl, 100 * dl)
            ctx.epsilon(True)
            ctx.convective((1.0, 0.0))
            ctx.diffusive()
            ctx.tree_destroy()
            ctx.move_and_clean(True)
            out.append(ctx.download())
        return out

    eager = run(False)
    graphed = run(True)
    for k, (a, b) in enumerate(zip(eager, graphed)):
        assert np.array_equal(a, b), f"step {k}"
    ctx.set_graph(False)
    ctx.set_tuning(True)


def test_graph_batched_cylinder_steps(ctx):
    """Four cylinder simulations in one context (lattice states as in the
    batched benchmark, Re 150-2800), whole steps: two eager runs (the tuned
    form) and a run with the captured convective phase (persistent form 16)
    give the same bits after every step."""
    from bench import cylinder_wake
    print("case graph in batched cylinder steps", flush=True)
    nseg = 500
    corners = V.gen_cylinder(0.5, nseg)
    segs1 = np.ascontiguousarray(V.segments_from_corners(corners), dtype=V.SEG_DTYPE)
    body1 = np.zeros(1, dtype=V.BODY_DTYPE)
    body1["nseg"] = nseg
    ns = [20_000, 50_000, 12_000, 35_000]
    res = np.array([150.0, 600.0, 1500.0, 2800.0])
    states = np.concatenate([cylinder_wake(n, 0x5EED2100 + s) for s, n in enumerate(ns)])
    off = np.concatenate([[0], np.cumsum(ns)]).astype(np.int64)
    segs = np.concatenate([segs1] * len(ns))
    bodies = np.concatenate([body1] * len(ns))
    bodies["first_seg"] = np.arange(len(ns)) * nseg
    dl = nseg * np.sin(np.pi / nseg) / (nseg - 1)

    def run(graph, steps=4):
        ctx.set_graph(graph)
        if graph:
            ctx.set_tuning(False, 16)
        else:
            ctx.set_tuning(True)
        ctx.set_params(re=600.0, dt=0.005, time=0.0)
        ctx.upload_bodies(segs, bodies)
        ctx.upload(states)
        ctx.set_batch(off, np.arange(len(ns) + 1, dtype=np.int32), res)
        out = []
        for k in range(steps):
            ctx.set_params(re=600.0, dt=0.005, time=k * 0.005)
            ctx.tree_build(8, 5 * dl, 100 * dl)
            ctx.calc_circulation((1.0, 0.0))
            ctx.tree_destroy()
            ctx.vortex_shed()
            ctx.calc_forces()
            ctx.zero_forces()
            ctx.tree_build(8, 5 * dl, 100 * dl)
            ctx.epsilon(True)
            ctx.convective((1.0, 0.0))
            ctx.diffusive()
            ctx.tree_destroy()
            ctx.move_and_clean(True)
            out.append((ctx.download(), ctx.batch_offsets(), ctx.download_bodies()[0]["g"].copy()))
        return out

    eager = run(False)
    again = run(False)
    graphed = run(True)
    for k in range(len(eager)):
        for name, other in (("eager again", again), ("graph", graphed)):
            assert np.array_equal(eager[k][1], other[k][1]), f"{name}: step {k} offsets"
            assert np.array_equal(eager[k][0], other[k][0]), f"{name}: step {k} vortices"
            assert np.array_equal(eager[k][2], other[k][2]), f"{name}: step {k} segment circulations"
    print(f"  vortices after {len(eager)} steps: {eager[-1][1][-1]}", flush=True)
    ctx.set_graph(False)
    ctx.set_tuning(True)
