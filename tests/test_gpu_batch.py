"""Several independent simulations in one context (vvhd_set_batch): one
tree per simulation, and epsilon, convective (plain and learned) and
diffusive advancing all of them in the same launches, and whole cylinder
steps (boundary system per simulation, shedding, forces, flow-move).

Bar: every simulation's state after tree build, epsilon (with merges),
convective and diffusive, and after whole cylinder steps, is bit-identical
to the same simulation run alone (same kernels, same lists in the same
order, same sums, the same batched LU), and its FP64 direct sum runs over
its own vortices only.  Entry points that act on the whole Space refuse to
run while a batch is declared.
"""
import numpy as np
import pytest

import vvflow_amd as V
from cases import layer_case, random_cluster_case

pytestmark = pytest.mark.gpu

FC, MIN_NODE, MAX_NODE = 8, 0.0314, 0.63


@pytest.fixture(scope="module")
def ctx():
    c = V.Context(0)
    yield c
    c.close()


def body_of(corners):
    segs = V.segments_from_corners(corners)
    b = np.zeros(1, dtype=V.BODY_DTYPE)
    b["nseg"] = len(segs)
    return np.ascontiguousarray(segs, dtype=V.SEG_DTYPE), b


def simulations():
    """(vortices, body or None, Re) of four simulations of uneven size: a
    wake, a vortex layer around a cylinder, an empty one and clusters."""
    wake = V.gen_wake(20_000, 0x5EED0021)
    layer, corners = layer_case(6000, 5)
    clusters = random_cluster_case(8000, 9, "mixed")
    return [(wake, None, 150.0), (layer, body_of(corners), 600.0), (np.zeros((0, 6)), None, 1000.0),
            (clusters, None, 2500.0)]


def step(ctx, learned=False):
    ctx.tree_build(FC, MIN_NODE, MAX_NODE)
    merged = ctx.epsilon(True)
    if learned:
        ctx.convective_learned((1.0, 0.0))
    else:
        ctx.convective((1.0, 0.0))
    ctx.diffusive()
    return merged


def run_alone(ctx, objs, body, re, learned=False):
    ctx.set_params(re=re, dt=0.005)
    if body is None:
        ctx.upload_bodies(np.zeros(0, V.SEG_DTYPE), np.zeros(0, V.BODY_DTYPE))
    else:
        ctx.upload_bodies(*body)
    ctx.upload(objs)
    merged = step(ctx, learned)
    return ctx.download(), merged


def load_batch(ctx, sims):
    objs = np.concatenate([s[0] for s in sims])
    segs, bodies, body_off = [], [], [0]
    nseg = 0
    for _, body, _ in sims:
        if body is not None:
            sg, b = body
            b = b.copy()
            b["first_seg"] = nseg
            nseg += len(sg)
            segs.append(sg)
            bodies.append(b)
        body_off.append(len(bodies))
    ctx.set_params(re=600.0, dt=0.005)
    ctx.upload_bodies(np.concatenate(segs) if segs else np.zeros(0, V.SEG_DTYPE),
                      np.concatenate(bodies) if bodies else np.zeros(0, V.BODY_DTYPE))
    ctx.upload(objs)
    off = np.concatenate([[0], np.cumsum([len(s[0]) for s in sims])]).astype(np.int64)
    ctx.set_batch(off, body_off, [s[2] for s in sims])
    return off


@pytest.mark.parametrize("learned", [False, True])
def test_batch_matches_simulations_alone(ctx, learned):
    print(f"case batch of 4 simulations, learned={learned}", flush=True)
    sims = simulations()
    if learned:
        from vvflow_amd.train_farfield import MODEL_PATH, load_model
        ctx.far_model_set(load_model(MODEL_PATH))
    alone = [run_alone(ctx, *s, learned=learned) for s in sims]
    off = load_batch(ctx, sims)
    merged = step(ctx, learned)
    st = ctx.download()
    assert merged == sum(m for _, m in alone)
    for s, (ref, _) in enumerate(alone):
        got = st[off[s]:off[s + 1]]
        assert np.array_equal(got, ref, equal_nan=True), f"simulation {s}"
    print(f"  merged {merged}", flush=True)


def test_batch_direct_sum_per_simulation(ctx):
    print("case batch direct sum", flush=True)
    sims = simulations()
    rng = np.random.default_rng(4)
    refs = []
    for objs, body, re in sims:
        run_alone(ctx, objs, body, re)
        tg = np.sort(rng.choice(len(objs), min(50, len(objs)), replace=False)) if len(objs) else np.zeros(0, int)
        refs.append((tg, ctx.direct_sum(tg.astype(np.int64)) if len(tg) else np.zeros((0, 2))))
    off = load_batch(ctx, sims)
    step(ctx)
    tg = np.concatenate([off[s] + t for s, (t, _) in enumerate(refs)]).astype(np.int64)
    got = ctx.direct_sum(tg)
    assert np.array_equal(got, np.concatenate([r for _, r in refs]))


def test_batch_refuses_whole_space_calls(ctx):
    sims = simulations()
    load_batch(ctx, sims)
    step(ctx)
    # a boundary system per simulation needs bodies in every simulation;
    # integrals and point queries act on the whole Space
    for call in (lambda: ctx.calc_circulation((1.0, 0.0)), ctx.integrals, lambda: ctx.velocity(np.zeros((1, 2))),
                 lambda: ctx.tree_find(np.zeros((1, 2)))):
        with pytest.raises(V.VVHDError) as e:
            call()
        assert e.value.code == 6
    # a new upload returns to one simulation
    ctx.upload(sims[0][0])
    ctx.upload_bodies(np.zeros(0, V.SEG_DTYPE), np.zeros(0, V.BODY_DTYPE))
    ctx.upload(sims[0][0])
    step(ctx)
    ctx.move_and_clean(True)
    assert len(ctx.batch_offsets()) == 2


def cylinder_sims(nseg=200):
    """Four cylinder simulations (one body of nseg segments each, so their
    boundary systems have one order) with vortex layers of uneven size and
    their own Reynolds numbers."""
    out = []
    for k, (n, re) in enumerate([(3000, 200.0), (0, 800.0), (9000, 1500.0), (1200, 3000.0)]):
        corners = V.gen_cylinder(0.5, nseg)
        rng = np.random.default_rng(40 + k)
        th = rng.uniform(0, 2 * np.pi, n)
        rr = 0.5 + 2e-3 + rng.exponential(0.1, n)
        objs = np.zeros((n, 6))
        objs[:, 0], objs[:, 1] = rr * np.cos(th), rr * np.sin(th)
        objs[:, 2] = rng.normal(size=n) * 1e-3
        out.append((objs, body_of(corners), re))
    return out


def cylinder_step(ctx, k, dl, dt=0.005):
    """One step of the reference's loop (utils/vvflow/vvflow.cpp:214-262)."""
    ctx.tree_build(8, 5 * dl, 100 * dl)
    ctx.calc_circulation((1.0, 0.0))
    ctx.tree_destroy()
    ctx.vortex_shed()
    ctx.calc_forces()
    forces = ctx.download_bodies()[1][["force_hydro", "friction"]].copy()
    ctx.zero_forces()
    ctx.tree_build(8, 5 * dl, 100 * dl)
    merged = ctx.epsilon(True)
    ctx.convective((1.0, 0.0))
    ctx.diffusive()
    ctx.tree_destroy()
    removed = ctx.move_and_clean(True)
    return forces, merged, removed


def test_batch_cylinder_steps_match_simulations_alone(ctx):
    """Whole cylinder steps (boundary system, shedding, forces, flow-move)
    of four simulations in one context against each run alone: vortex lists,
    segments and bodies bit-identical after every step."""
    steps, nseg, dt = 4, 200, 0.005
    dl = nseg * np.sin(np.pi / nseg) / (nseg - 1)
    print(f"case batch of 4 cylinder simulations, {steps} whole steps", flush=True)
    sims = cylinder_sims(nseg)
    alone = []
    for objs, body, re in sims:
        ctx.set_params(re=re, dt=dt, time=0.0)
        ctx.upload_bodies(*body)
        ctx.upload(objs)
        hist = []
        for k in range(steps):
            ctx.set_params(re=re, dt=dt, time=k * dt)
            f, m, r = cylinder_step(ctx, k, dl)
            hist.append((ctx.download(), ctx.download_bodies(), f, m, r))
        alone.append(hist)
    load_batch(ctx, sims)
    for k in range(steps):
        ctx.set_params(re=600.0, dt=dt, time=k * dt)
        f, m, r = cylinder_step(ctx, k, dl)
        off = ctx.batch_offsets()
        st = ctx.download()
        segs, bodies = ctx.download_bodies()
        assert m == sum(a[k][3] for a in alone)
        assert r == sum(a[k][4] for a in alone)
        for s in range(len(sims)):
            ref_v, (ref_segs, ref_bodies), ref_f, _, _ = alone[s][k]
            assert np.array_equal(st[off[s]:off[s + 1]], ref_v), f"step {k} simulation {s} vortices"
            assert np.array_equal(segs[s * nseg:(s + 1) * nseg], ref_segs), f"step {k} simulation {s} segments"
            ref_b = ref_bodies.copy()
            ref_b["first_seg"] = s * nseg
            assert np.array_equal(bodies[s:s + 1], ref_b), f"step {k} simulation {s} body"
            assert np.array_equal(f[s:s + 1], ref_f), f"step {k} simulation {s} forces"
    print(f"  vortices after {steps} steps: {np.diff(off).tolist()}", flush=True)
