"""The C++ module classes (include/vvhd_modules.hpp) drive the reference's
time loop on the GPU: the cylinder driver built on them (vvflow_cylinder)
against the Python driver over the C ABI, step by step.

Bar: identical vortex counts, forces within 1e-9 relative (the segment
friction is accumulated with FP64 atomics, so its last bits depend on the
order in which the device adds them)."""
import os
import subprocess

import numpy as np
import pytest

import vvflow_amd as V
from vvflow_amd.simulate import CylinderRun

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "vvflow_amd", "vvflow_cylinder")


@pytest.mark.skipif(not os.path.exists(EXE), reason="vvflow_cylinder not built")
def test_cpp_driver_matches_python_driver():
    steps = 25
    print(f"case C++ cylinder driver, {steps} steps", flush=True)
    r = subprocess.run([EXE, "--steps", str(steps)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    rows = np.array([[float(x) for x in ln.split()] for ln in r.stdout.splitlines() if not ln.startswith("#")])
    assert rows.shape == (steps, 10)
    ctx = V.Context(0)
    run = CylinderRun(ctx)
    counts = []
    for _ in range(steps):
        run.step()
        counts.append(ctx.n)
    ctx.close()
    f = np.array(run.forces)
    assert np.array_equal(rows[:, 2].astype(int), np.array(counts))
    assert np.allclose(rows[:, 3:9], f, rtol=1e-9, atol=1e-12 * np.abs(f).max())
    print(f"  n after {steps} steps: {counts[-1]}; force_hydro x {f[-1, 0]:.6f}", flush=True)


def test_cpp_driver_heat_and_streak_match_python_driver():
    """A heated cylinder (const_w, heat_const 0.5, Pr 0.7) with 11 streak
    sources shed every 4 steps: the module classes (heat_shed, streak_shed,
    process_heat_list, heat_crop and the lists carried by Space) against the
    Python driver's direct ABI calls in the reference's order."""
    steps = 20
    print(f"case C++ cylinder driver with heat and streak, {steps} steps", flush=True)
    r = subprocess.run([EXE, "--steps", str(steps), "--heat", "3:0.5:0.7", "--streak", "11:0.02"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    heat_line = [ln for ln in r.stdout.splitlines() if ln.startswith("# heat")][0].split()
    ctx = V.Context(0)
    run = CylinderRun(ctx, heat=(V.HC_CONST_W, 0.5, 0.7), streak=(11, 0.02))
    for _ in range(steps):
        run.step()
    h = ctx.download_list(V.LIST_HEAT)
    st = ctx.download_list(V.LIST_STREAK)
    hsum = ctx.download_bodies()[0]["hsum"].sum()
    ctx.close()
    print(f"  heat {len(h)} (sum {h[:, 2].sum():.6e}), streak {len(st)}", flush=True)
    assert int(heat_line[2]) == len(h)
    assert float(heat_line[3]) == pytest.approx(h[:, 2].sum(), rel=1e-12)
    assert float(heat_line[5]) == pytest.approx(hsum, rel=1e-12, abs=1e-15)
    assert int(heat_line[7]) == len(st)
    assert len(st) == 11 * 5   # shed at t = 0, 0.02, 0.04, 0.06, 0.08
    # const_w sheds one particle per segment and step; epsh merges the
    # layers shed at the same places and the wall absorbs some of them
    assert 500 * 2 < len(h) < 500 * steps


def test_cpp_driver_snapshot_save_and_load(tmp_path):
    """Space::save after ten steps of the C++ driver holds the state the
    Python driver reaches with the same calls (vortices and the body's
    corners, bound circulation and gsum, bit for bit); Space::load restarts
    the driver from it (every vortex with _1_eps = 0, as the reference's
    reader leaves them)."""
    from vvflow_amd import hdf
    print("case C++ driver snapshot, 10 steps", flush=True)
    path = str(tmp_path / "cyl.h5")
    r = subprocess.run([EXE, "--steps", "10", "--save", path], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    sp, lists, bodies = hdf.load(path)
    ctx = V.Context(0)
    run = CylinderRun(ctx)
    for _ in range(10):
        run.step()
    v = ctx.download()
    segs = ctx.download_bodies()[0]
    ctx.close()
    assert sp["caption"] == "cylinder" and sp["re"] == 600.0 and sp["dt"] == 0.005
    assert sp["time"] == pytest.approx(0.05, abs=1e-12)
    assert np.array_equal(lists["vort"][:, :3], v[:, :3])
    assert len(bodies) == 1 and bodies[0]["nseg"] == 500
    seg4 = bodies[0]["seg4"]
    assert np.array_equal(seg4[:, 0], segs["cx"]) and np.array_equal(seg4[:, 1], segs["cy"])
    assert np.array_equal(seg4[:, 2], segs["g"]) and np.array_equal(seg4[:, 3], segs["gsum"])
    r2 = subprocess.run([EXE, "--load", path, "--steps", "3"], capture_output=True, text=True, timeout=300)
    assert r2.returncode == 0, r2.stderr
    rows = [ln.split() for ln in r2.stdout.splitlines() if not ln.startswith("#")]
    assert len(rows) == 3 and float(rows[0][1]) == pytest.approx(0.055, abs=1e-12)
    assert int(rows[-1][2]) > len(v)
    print(f"  snapshot of {len(v)} vortices; restarted for 3 steps: {rows[-1][2]} vortices", flush=True)


def h5_values(path, dset, tmp_path):
    """A dataset's values through h5dump's binary output (float32)."""
    import shutil
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    out = tmp_path / (dset.strip("/").replace("/", "_") + ".bin")
    subprocess.run([h5dump, "-d", dset, "-b", "LE", "-o", str(out), path], capture_output=True, check=True)
    return np.fromfile(out, dtype="<f4")


@pytest.mark.skipif(not os.path.exists(EXE), reason="vvflow_cylinder not built")
def test_cpp_driver_stepdata(tmp_path):
    """Stepdata (MStepdata.cpp) in the C++ driver's loop: one row per step
    after the forces, holding the force_hydro and friction the driver prints
    (as float); the friction profile Fr / dt of every segment reproduces the
    friction force, F = -sum dl (Fr / dt) / |dl| (TSpace.cpp:94-120), which
    holds only if the accumulators are reset after every row; a restart from
    a snapshot continues the series from the snapshot's time."""
    print("case C++ driver step data, 8 + 4 steps", flush=True)
    sd, snap = str(tmp_path / "stepdata_cylinder.h5"), str(tmp_path / "cyl.h5")
    r = subprocess.run([EXE, "--steps", "8", "--stepdata", sd, "--profile", "--save", snap], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    rows = np.array([[float(x) for x in ln.split()] for ln in r.stdout.splitlines() if not ln.startswith("#")])
    t = h5_values(sd, "/time", tmp_path)
    assert np.array_equal(t, (0.005 * np.arange(8)).astype(np.float32))
    fh = h5_values(sd, "/body00/force_hydro", tmp_path).reshape(8, 3)
    ff = h5_values(sd, "/body00/force_friction", tmp_path).reshape(8, 3)
    assert np.array_equal(fh, rows[:, 3:6].astype(np.float32))
    assert np.array_equal(ff, rows[:, 6:9].astype(np.float32))
    fr = h5_values(sd, "/body00/friction", tmp_path).reshape(8, 500).astype(np.float64)
    cp = h5_values(sd, "/body00/pressure", tmp_path).reshape(8, 500)
    assert np.all(np.isfinite(cp)) and np.all(np.isfinite(fr))
    c = V.gen_cylinder(0.5, 500)
    dl = np.roll(c, -1, axis=0) - c
    fx = -(dl[:, 0] * fr / np.hypot(dl[:, 0], dl[:, 1])).sum(1)
    fy = -(dl[:, 1] * fr / np.hypot(dl[:, 0], dl[:, 1])).sum(1)
    scale = np.abs(rows[:, 6:8]).max()
    assert np.abs(fx - rows[:, 6]).max() < 1e-4 * scale and np.abs(fy - rows[:, 7]).max() < 1e-4 * scale
    # restart at the snapshot's time: eight rows kept, four appended
    r2 = subprocess.run([EXE, "--load", snap, "--steps", "4", "--stepdata", sd], capture_output=True, text=True,
                        timeout=300)
    assert r2.returncode == 0, r2.stderr
    t = h5_values(sd, "/time", tmp_path)
    assert np.array_equal(t, (0.005 * np.arange(12)).astype(np.float32))
    assert h5_values(sd, "/body00/pressure", tmp_path).size == 8 * 500   # cut at the restart, no profiles after
    print(f"  {len(t)} rows; friction force from the profile within "
          f"{np.abs(fx - rows[:, 6]).max() / scale:.1e} of the force", flush=True)


def read_xfield(path):
    """XField's binary matrix (XField.cpp:124-152) -> (xs, ys, map)."""
    a = np.fromfile(path, dtype=np.float32)
    n = int(a[0])
    xs = a[1:1 + n]
    rows = a[1 + n:].reshape(-1, n + 1)
    return xs, rows[:, 0], rows[:, 1:]


def read_state(path):
    """The state vvflow_cylinder --fields writes before evaluating: vortex
    records, segment records (16 doubles in vvhd_seg order) and the body's
    kinematic triples and g_dead."""
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw, np.int64, 1, 0)[0])
    o = 8
    vort = np.frombuffer(raw, np.float64, 6 * n, o).reshape(n, 6).copy()
    o += 48 * n
    ns = int(np.frombuffer(raw, np.int64, 1, o)[0])
    o += 8
    rec = np.frombuffer(raw, np.float64, 16 * ns, o).reshape(ns, 16)
    o += 128 * ns
    segs = np.zeros(ns, dtype=V.SEG_DTYPE)
    for k, name in enumerate(V.SEG_DTYPE.names):
        segs[name] = rec[:, k]
    body = np.zeros(1, dtype=V.BODY_DTYPE)
    body["nseg"] = ns
    tri = np.frombuffer(raw, np.float64, 18, o).reshape(6, 3)
    for k, name in enumerate(("holder", "dpos", "speed_slae", "speed_slae_prev", "friction_prev", "fdt_dead")):
        body[name] = tri[k]
    body["g_dead"] = np.frombuffer(raw, np.float64, 1, o + 144)[0]
    return vort, segs, body


def field_copy(state, kind, run):
    """The preparation of the reference's evaluators on a copy of the state
    (XVelocity.cpp:47-52, XVorticity.cpp:35-44, XStreamfunction.cpp:69-78,
    XPressure.cpp:55-66), then vvhd_field on the driver's grid (its float
    origin and step, as XField keeps them)."""
    vort, segs, body = state
    c = V.Context(0)
    try:
        c.set_params(re=run.re, dt=run.dt, time=run.time)
        c.upload_bodies(segs, body)
        c.upload(vort)
        dl = c.download_bodies()[1][0]["slen"] / (len(segs) - 1)
        f = float(np.float32(-1.0)), float(np.float32(0.05))
        grid = (f[0], f[0], f[1], 80, 40)
        if kind == "velocity":
            c.tree_build(8, dl * 20, 0.1)
            c.epsilon(False)
            return c.field(V.FIELD_VX, *grid, inf=run.inf, ref=(0.0, 0.0))
        if kind == "vorticity":
            c.vortex_shed()
            c.tree_build(8, dl * 20)
            return c.field(V.FIELD_VORTICITY, *grid, eps_mult=1.0, dl=dl)
        if kind == "streamfunction":
            c.tree_build(8, dl * 10, dl * 20)
            return c.field(V.FIELD_STREAMFUNCTION, *grid, eps_mult=1.0, dl=dl, inf=run.inf, ref=(0.0, 0.0))
        c.vortex_shed()
        c.tree_build(8, dl * 20, 0.1)
        c.epsilon(False)
        c.convective(run.inf)
        c.diffusive()
        return c.field(V.FIELD_PRESSURE, *grid, eps_mult=1.0, dl=dl, inf=run.inf, ref=None)
    finally:
        c.close()


@pytest.mark.skipif(not os.path.exists(EXE), reason="vvflow_cylinder not built")
def test_cpp_sensors_fields_and_node_views(tmp_path):
    """The module classes' sensors, XField evaluators, node views and
    MEpsilonFast::eps2h/h2 against the same computations through the C ABI
    on the Python driver's state."""
    steps = 12
    print(f"case C++ sensors and fields after {steps} steps", flush=True)
    pts = np.array([[0.8, 0.05], [1.2, -0.3], [2.0, 0.4], [0.6, -0.6]])
    pfile = tmp_path / "points.txt"
    pfile.write_text("".join(f"{x} {y}\n" for x, y in pts))
    sout = tmp_path / "sensors.txt"
    prefix = str(tmp_path / "f")
    r = subprocess.run([EXE, "--steps", str(steps), "--sensors", str(pfile), "--sensors-out", str(sout),
                        "--fields", prefix], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    ctx = V.Context(0)
    try:
        run = CylinderRun(ctx)
        run.sensor_points = pts
        for _ in range(steps):
            run.step()
        # sensors: time then (u, v) per point, every step, printed with %lg
        # (six significant digits) as sensors::output does
        rows = np.array([[float(x) for x in ln.split()] for ln in sout.read_text().splitlines()])
        assert rows.shape == (steps, 1 + 2 * len(pts))
        want = np.array([s.reshape(-1) for s in run.sensor_rows])
        assert np.allclose(rows[:, 1:], want, rtol=1e-5, atol=1e-6)
        assert np.allclose(rows[:, 0], 0.005 * np.arange(steps), rtol=1e-5, atol=1e-9)
        # fields through the classes == fields through the C ABI from the
        # same state (the driver's, written before its evaluators ran)
        state = read_state(f"{prefix}_state.bin")
        assert len(state[0]) == ctx.n
        for kind in ("velocity", "vorticity", "streamfunction", "pressure"):
            xs, ys, m = read_xfield(f"{prefix}_{kind}.bin")
            assert m.shape == (40, 80)
            assert np.allclose(xs, np.float32(-1.0) + np.float32(0.05) * np.arange(80, dtype=np.float32))
            g = field_copy(state, kind, run)
            scale = float(np.abs(g).max())
            err = float(np.abs(m - g).max()) / scale
            print(f"  {kind}: max |class - abi| / max = {err:.2e}", flush=True)
            assert scale > 0 and err < 1e-6
        # node views and the static eps2h / h2 at the driver's probe points
        ctx.tree_build(8, run.min_node, run.max_node)
        lv = ctx.tree_leaves()
        nl = ctx.tree_lists(False)
        fl = ctx.tree_lists(True)
        probe = np.array([[0.7, 0.1], [1.5, -0.2], [0.0, 0.55]])
        leaf = ctx.tree_find(probe)
        e2 = ctx.eps2h(probe, leaf=leaf)
        h2 = ctx.h2(probe, leaf=leaf)
        lines = [ln.split() for ln in r.stdout.splitlines() if ln.startswith("# node")]
        assert len(lines) == 3 and len(lv["geom"]) > 0
        for k, ln in enumerate(lines):
            assert float(ln[5]) == pytest.approx(e2[k], rel=1e-9)
            assert float(ln[7]) == pytest.approx(h2[k], rel=1e-9)
            assert int(ln[9]) == nl["off"][leaf[k] + 1] - nl["off"][leaf[k]]
            assert int(ln[11]) == fl["off"][leaf[k] + 1] - fl["off"][leaf[k]]
        gs, integ, mx, my, _ = ctx.integrals()
        ln = [x for x in r.stdout.splitlines() if x.startswith("# integrals")][0].split()
        assert float(ln[3]) == pytest.approx(gs, rel=1e-9, abs=1e-12)
        assert float(ln[5]) == pytest.approx(integ, rel=1e-9)
        assert float(ln[7]) == pytest.approx(mx, rel=1e-9) and float(ln[8]) == pytest.approx(my, rel=1e-9)
    finally:
        ctx.close()
