"""Single-GPU benchmark of the hot path: learned vs plain far field at a
matched convective error.

Workload: the 1M-vortex Karman-street wake (two staggered rows of 200
Gaussian clusters, sigma 0.05, SplitMix64 seed 0x5EED0001), Re = 600,
U = (1, 0), merging on.  One step = tree build + epsilon (with merges) +
convective + diffusive velocity from a device-resident copy of the input
(merges mutate it, so every step restores the input first; the restore is a
device-to-device copy inside the timed region).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N]
                    [--learned-fc F] [--plain-fcs 512,1024,2048] [--no-plain]
                    [--no-cpu] [--error-sample S] [--model PATH]

value: particle-steps/s of the learned path (far criterion --learned-fc,
default 8 as in the reference, learned far-field correction, FP32 compensated pair
sums) = N * K / (time of K steps).  The plain tree path (the reference's method) is timed at the
smallest far criterion of --plain-fcs whose convective error reaches 1e-5,
and at the reference's far criterion 8.  Errors: relative L2 of the induced
convective velocity against the FP64 direct sum on --error-sample targets.
roofline: the kernel with the largest device time in the learned step --
the near-field pair kernel (11 FP32 flops per near pair,
MConvectiveFast.cpp:116-120, against the FP32 vector peak 157.3 TFLOP/s) or
the learned far-field network kernel (9984 matrix flops per network pass,
layer 1 in bf16 and layers 2-3 in fp16, three passes -- value and two
forward-mode tangents -- per (leaf, far node) term, against the dense
bf16/fp16 peak 2.5 PFLOP/s) -- measured with HIP events on
the library's stream, in an untimed pass with the step's kernels run one at a
time (the timed step overlaps the near field, the far field and the
diffusive on separate streams; phases_ms then overlap too).
cpu_baseline: the FP64 oracle (the reference's algorithm and OpenMP structure,
plain path) on this host's cores: at far criterion 8 the whole step (tree and
epsilon serial as in the reference, convective and diffusive over the host
threads); at the plain path's matched criterion the full tree build, and
epsilon, convective and diffusive on a contiguous block of leaves, scaled by
the vortices of the block.  Threads: the CPUs this process may run on,
limited by OMP_NUM_THREADS when that is set (the host's share for one GPU).
roofline_other also carries the tree build and flow-move against the HBM
bandwidth (algorithmic bytes: 72 B per vortex and tree level -- an 8-byte key
and two 32-byte records moved per level; 96 B per vortex for flow-move -- six
FP64 fields read and written once).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0
FLOPS_PER_NEAR_PAIR = 11
FLOPS_PER_NET_PASS = 2 * (12 * 64 + 64 * 64 + 64 * 2)
TARGET_ERR = 1e-5
HBM_PEAK_GBS = 8000.0
TREE_BYTES_PER_VORTEX_LEVEL = 8 + 2 * 32
FLOWMOVE_BYTES_PER_VORTEX = 6 * 8 * 2
TRAFFIC_FILES = {"k_near": os.path.join(ROOT, "profiles", "near_kernel_traffic_wake1M.txt"),
                 "k_far_learned": os.path.join(ROOT, "profiles", "far_learned_traffic_wake1M.txt"),
                 "tree": os.path.join(ROOT, "profiles", "tree_traffic_wake1M.txt"),
                 "flowmove": os.path.join(ROOT, "profiles", "flowmove_traffic_wake1M.txt")}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--learned-fc", type=int, default=8)
    ap.add_argument("--plain-fcs", default="512,1024,2048,4096")
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--plain-only", action="store_true", help="time the plain path at --learned-fc instead")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--far-gradient", choices=("network", "geometric"), default="network",
                    help="gradient of the learned correction: the network's own forward-mode derivative "
                         "(default) or the geometric terms only, with the network output held fixed")
    ap.add_argument("--error-sample", type=int, default=4000)
    ap.add_argument("--cpu-fraction", type=float, default=0.1,
                    help="leaf block sampled by the CPU baseline at the plain path's matched criterion")
    ap.add_argument("--model", default=os.path.join(ROOT, "vvflow_amd", "farfield_mlp.npy"))
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 256 MiB device copy (counter calibration for rocprofv3 --pmc)")
    ap.add_argument("--precision-sweep", action="store_true",
                    help="4M-vortex wake: FP32 pair sums plain vs Kahan against the FP64 direct sum, and the "
                         "2000-segment body SLAE filled and solved in FP64 (one JSON line)")
    ap.add_argument("--sweep-fcs", default="8,64,256")
    ap.add_argument("--batched", action="store_true",
                    help="64 independent cylinder simulations of uneven size, whole steps, in one context vs one at "
                         "a time (one JSON line)")
    ap.add_argument("--nsim", type=int, default=64)
    ap.add_argument("--graph", action="store_true",
                    help="batched: run the plain convective phase as a captured HIP graph")
    ap.add_argument("--near-form", type=int, default=-1, choices=(-1, 0, 8, 16, 32),
                    help="near-field launch form: -1 autotuned (default), 0 one wave per leaf, 8/16/32 persistent "
                         "waves per compute unit (fixes the form for counter runs, whose timings would steer the "
                         "autotuner elsewhere)")
    return ap.parse_args()


def copy_calibration():
    """A plain device-to-device copy kernel of known size: FETCH_SIZE and
    WRITE_SIZE of this kernel calibrate the same counters on the kernels."""
    import torch
    nbytes = 256 << 20
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_kernel_bytes_read": nbytes, "copy_kernel_bytes_written": nbytes, "launches": 5}))


def traffic_bytes(kernel):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, calibrated)."""
    try:
        for line in open(TRAFFIC_FILES[kernel]):
            if line.startswith("traffic_bytes_per_launch"):
                return float(line.split("=")[1].split()[0])
    except (OSError, KeyError):
        pass
    return None


def host_cpu():
    """(model, nproc, os.cpu_count(), threads used by the CPU baseline)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count() or 1
    threads = nproc
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        threads = min(threads, int(omp))
    return model, nproc, os.cpu_count() or 1, threads


def _leaf_vortex_ranges(o):
    """Vortex range [v0, v1) of every leaf of the oracle's tree, DFS order."""
    nn = o.L.vvo_tree_num_nodes(o.h)
    nl = o.L.vvo_tree_num_leaves(o.h)
    geom = np.zeros((nn, 4))
    ints = np.zeros((nn, 6), dtype=np.int32)
    cm = np.zeros((nn, 6))
    cme = np.zeros((nn, 2))
    o.L.vvo_tree_nodes(o.h, geom, ints, cm, cme)
    leaves = np.zeros(nl, dtype=np.int32)
    o.L.vvo_tree_leaves(o.h, leaves)
    return ints[leaves, 0], ints[leaves, 1]


def cpu_baseline(objs, fc, fraction, threads):
    """FP64 oracle on the host, the plain path at far criterion fc: returns
    (particle-steps/s, seconds of CPU work, sample text).  fraction 1 runs
    the whole step; below 1, epsilon, convective and diffusive run on a
    contiguous block of that share of the leaves (middle of the DFS order)
    and their time is scaled by the vortex count over the block's."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bind import Oracle

    n = len(objs)
    o = Oracle()
    o.set_params(re=600.0, dt=0.005, inf=(1.0, 0.0))
    o.set_vortices(objs)
    t0 = time.perf_counter()
    o.tree_build(fc)
    t_tree = time.perf_counter() - t0
    if fraction >= 1:
        t0 = time.perf_counter()
        o.epsilon(True)
        t_eps = time.perf_counter() - t0
        t0 = time.perf_counter()
        o.convective(threads=threads)
        o.diffusive(threads=threads)
        t_cd = time.perf_counter() - t0
        t_step = t_tree + t_eps + t_cd
        sample = (f"whole step at far criterion {fc}: tree build {t_tree:.2f} s and epsilon {t_eps:.2f} s "
                  f"(serial, as the reference), convective + diffusive {t_cd:.2f} s on {threads} threads")
        return n / t_step, t_tree + t_eps + t_cd, sample
    v0, v1 = _leaf_vortex_ranges(o)
    nl = len(v0)
    cnt = max(1, int(nl * fraction))
    lb = (nl - cnt) // 2
    le = lb + cnt
    block = int(v1[le - 1] - v0[lb])
    t0 = time.perf_counter()
    o.epsilon_leaves(True, lb, le)
    t_eps = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.convective(threads=threads, leaf_begin=lb, leaf_end=le)
    o.diffusive(threads=threads, leaf_begin=lb, leaf_end=le)
    t_cd = time.perf_counter() - t0
    scale = n / max(block, 1)
    t_step = t_tree + (t_eps + t_cd) * scale
    sample = (f"far criterion {fc}: full tree build {t_tree:.2f} s (serial); epsilon (serial, {t_eps:.2f} s) and "
              f"convective + diffusive ({t_cd:.2f} s on {threads} threads) on {cnt} of {nl} leaves "
              f"({block} vortices), scaled by {scale:.2f}")
    return n / t_step, t_tree + t_eps + t_cd, sample


class Bench:
    def __init__(self, V, ctx, n, inf=(1.0, 0.0)):
        self.V, self.ctx, self.n, self.inf = V, ctx, n, inf

    def step(self, fc, learned):
        c = self.ctx
        c.restore()
        c.tree_build(fc)
        c.epsilon(True)
        if learned:
            c.convective_learned(self.inf, kahan=True)
        else:
            c.convective(self.inf, kahan=True)
        c.diffusive()

    def time(self, fc, learned, steps, warmup):
        for _ in range(warmup):
            self.step(fc, learned)
        self.ctx.phase_times()   # synchronises
        phases = np.zeros(8)
        t0 = time.perf_counter()
        for _ in range(steps):
            self.step(fc, learned)
            phases += self.ctx.phase_times()
        el = time.perf_counter() - t0
        return el, phases / steps

    def error(self, fc, learned, nsample):
        """Convective (induced, U = 0) error against the FP64 direct sum of
        the same post-epsilon state, and the pair counts of this tree."""
        c = self.ctx
        c.restore()
        c.tree_build(fc)
        merged = c.epsilon(True)
        if learned:
            c.convective_learned((0.0, 0.0), kahan=True)
        else:
            c.convective((0.0, 0.0), kahan=True)
        st = c.download()
        v = st[:, 3:5]
        live = np.nonzero(st[:, 2] != 0)[0]   # merge partners (g = 0) are skipped by the method
        tg = np.random.default_rng(0).choice(live, min(nsample, len(live)), replace=False).astype(np.int64)
        ref = c.direct_sum(tg)
        err = float(np.linalg.norm(v[tg] - ref) / np.linalg.norm(ref))
        nr = np.linalg.norm(ref, axis=1)
        err_max = float(np.max(np.linalg.norm(v[tg] - ref, axis=1) / np.maximum(nr, 1e-3 * np.sqrt(np.mean(nr ** 2)))))
        work = c.work_counts()
        return dict(rel_l2=err, max_rel=err_max, merged=int(merged), near_pairs=int(work[0]),
                    far_pairs=int(work[2]), net_passes=int(work[3]), levels=int(work[4]))

    def flowmove(self, fc, learned, reps):
        """MFlowmove::move_and_clean after a whole step (velocities set):
        mean device time (HIP events) over reps and the vortex count."""
        c = self.ctx
        ms = 0.0
        n0 = 0
        for _ in range(reps):
            self.step(fc, learned)
            c.tree_destroy()
            n0 = c.n
            c.move_and_clean(True)
            ms += float(c.phase_times()[4])
        return ms / reps, n0


ROOFLINE_STEPS = 3
ALONE_NOTE = ("kernel_ms: HIP events around the kernel, mean of %d steps run with the step's kernels one at a time "
              "(vvhd_set_overlap(0)); kernel_ms_in_timed_step: the same kernel in the timed steps, where it shares "
              "the GPU with the step's concurrent kernels" % ROOFLINE_STEPS)


def roofline(name, bound, achieved, peak, unit, traffic, extra):
    out = {"kernel": name, "bound": bound, "achieved": achieved, "peak": peak, "unit": unit,
           "frac": achieved / peak, "traffic": traffic}
    out.update(extra)
    return out


def precision_sweep(a):
    """4M-vortex wake (seed 0x5EED0004): convective error of the plain tree
    path with FP32 pair sums, plain and Kahan-compensated, against the FP64
    direct sum at several far criteria, with the near kernel's FP32 rate;
    then the boundary SLAE of a 2000-segment cylinder (n = 2009) filled and
    solved in FP64: the first call (with the solver library's set-up), a
    second call at t = 0 (fill, LU, solve) and a later step (fill compared
    with the factorised matrix, right-hand side and solve with the factors
    kept), and the relative residual."""
    import vvflow_amd as V
    n = 4_000_000
    objs = V.gen_wake(n, 0x5EED0004)
    ctx = V.Context(0)
    ctx.set_params(re=600.0, dt=0.005)
    ctx.upload_bodies(np.zeros(0, V.SEG_DTYPE), np.zeros(0, V.BODY_DTYPE))
    ctx.upload(objs)
    ctx.snapshot()
    rows = []
    tg = None
    for fc in [int(f) for f in a.sweep_fcs.split(",")]:
        for kahan in (False, True):
            ctx.restore()
            ctx.tree_build(fc)
            ctx.epsilon(True)
            before = ctx.download()[:, 3:5]
            ctx.convective((0.0, 0.0), kahan=kahan)
            ph = ctx.phase_times()
            st = ctx.download()
            if tg is None:
                live = np.nonzero(st[:, 2] != 0)[0]
                tg = np.random.default_rng(0).choice(live, min(a.error_sample, len(live)),
                                                     replace=False).astype(np.int64)
                ref = ctx.direct_sum(tg)
            v = st[tg, 3:5] - before[tg]
            nr = np.linalg.norm(ref, axis=1)
            work = ctx.work_counts()
            near_ms = float(ph[6])
            rows.append({"far_criteria": fc, "accumulation": "kahan" if kahan else "plain",
                         "rel_l2": float(np.linalg.norm(v - ref) / np.linalg.norm(ref)),
                         "rms_rel": float(np.sqrt(np.mean((np.linalg.norm(v - ref, axis=1) /
                                                           np.maximum(nr, 1e-3 * np.sqrt(np.mean(nr ** 2)))) ** 2))),
                         "max_rel": float(np.max(np.linalg.norm(v - ref, axis=1) /
                                                 np.maximum(nr, 1e-3 * np.sqrt(np.mean(nr ** 2))))),
                         "near_pairs": int(work[0]), "near_kernel_ms": near_ms,
                         "near_tflops": work[0] * FLOPS_PER_NEAR_PAIR / (near_ms * 1e-3) / 1e12})
            print(f"fc={fc} {rows[-1]['accumulation']}: rel L2 {rows[-1]['rel_l2']:.3e}", file=sys.stderr, flush=True)
    # the 2000-segment cylinder, shifted upstream of the wake
    corners = V.gen_cylinder(0.5, 2000)
    corners[:, 0] -= 2.0
    segs = V.segments_from_corners(corners)
    body = np.zeros(1, dtype=V.BODY_DTYPE)
    body["nseg"] = len(segs)
    ctx.restore()
    ctx.set_params(re=600.0, dt=0.005, time=0.0)
    ctx.upload_bodies(np.ascontiguousarray(segs, dtype=V.SEG_DTYPE), body)
    dl = ctx.download_bodies()[1][0]["slen"] / (len(segs) - 1)   # Space::average_segment_length
    ctx.tree_build(8, 5 * dl, 100 * dl)
    ctx.calc_circulation((1.0, 0.0))
    first_ms = float(ctx.phase_times()[5])   # includes the solver library's first-use set-up
    ctx.calc_circulation((1.0, 0.0))
    refactor_ms = float(ctx.phase_times()[5])   # t = 0: fill, LU and solve again
    ctx.tree_destroy()
    ctx.set_params(re=600.0, dt=0.005, time=0.005)
    ctx.tree_build(8, 5 * dl, 100 * dl)
    ctx.calc_circulation((1.0, 0.0))
    cached_ms = float(ctx.phase_times()[5])
    A, b = ctx.slae_system()
    sg, bd = ctx.download_bodies()
    x = np.concatenate([sg["g"], bd[0]["speed_slae"], bd[0]["force_hydro"], bd[0]["force_holder"]])
    resid = float(np.linalg.norm(A @ x - b) / np.linalg.norm(b))
    ctx.close()
    out = {"metric": "relative convective velocity error vs FP64 direct sum, FP32 plain vs Kahan pair sums",
           "workload": f"karman-street wake, {n} vortices (seed 0x5EED0004), plain tree path, merging epsilon",
           "sample_targets": int(len(tg)), "sweep": rows,
           "slae": {"segments": int(len(segs)), "n": int(A.shape[0]), "dtype": "fp64",
                    "first_call_ms": first_ms, "fill_factor_solve_ms": refactor_ms,
                    "factor_reuse_ms": cached_ms, "relative_residual": resid}}
    print(json.dumps(out), flush=True)


def cylinder_wake(n, seed):
    """A cylinder simulation's state of n vortices (TObj records) near the
    balance the method keeps by itself: the merge distance of epsilon grows
    with the distance d from the body as 0.4 |dl| sqrt(1 + d)
    (MEpsilonFast.cpp:29-43), so the vortices sit on jittered lattices of
    spacing 0.006 sqrt(1 + d) -- rings round the cylinder (R = 0.5; 10 % of
    them) and columns behind it across |y| <= w, w = max(1, n / 1e5) -- with
    circulation -sign(y) U(0.5, 1.5) 1e-3 (negative above the axis, as the
    upper side sheds).  Few of them merge on the first step (about 0.1 %
    against the FP64 oracle at 100k and 400k)."""
    rng = np.random.default_rng(seed)
    c, jit = 0.006, 0.2
    pts = []
    nb = int(0.1 * n)
    r, cnt = 0.502, 0
    while cnt < nb:
        h = c * np.sqrt(1.0 + (r - 0.5))
        m = int(2 * np.pi * r / h)
        th = (np.arange(m) + rng.uniform(-jit, jit, m)) * (2 * np.pi / m)
        rr = r + rng.uniform(-jit, jit, m) * h
        pts.append(np.stack([rr * np.cos(th), rr * np.sin(th)], 1))
        cnt += m
        r += h
    layer = np.concatenate(pts)[:nb]
    w = max(1.0, n / 1e5)
    pts, x, cnt = [], max(1.0, r + 0.1), 0
    while cnt < n - nb:
        h = c * np.sqrt(1.0 + x)
        m = int(2 * w / h)
        y = -w + (np.arange(m) + 0.5 + rng.uniform(-jit, jit, m)) * (2 * w / m)
        pts.append(np.stack([x + rng.uniform(-jit, jit, m) * h, y], 1))
        cnt += m
        x += h
    out = np.zeros((n, 6))
    out[:, :2] = np.concatenate([layer, np.concatenate(pts)[:n - nb]])
    out[:, 2] = np.where(out[:, 1] > 0, -1.0, 1.0) * rng.uniform(0.5, 1.5, n) * 1e-3
    return out


def batched(a):
    """Config 4: nsim independent cylinder simulations (R = 0.5, 500
    segments, U = (1, 0), dt = 0.005) with Reynolds numbers log-uniform in
    [100, 3000] and n_s vortices each, log-uniform in [10k, 400k] and scaled
    to about 8M in all (cylinder_wake lattices, seeds 0x5EED2000 + s).  One
    step is the reference's whole loop (utils/vvflow/vvflow.cpp:214-262):
    tree, boundary system (FP64), shedding, forces, tree, epsilon (merging),
    convective (plain, FP32 Kahan), diffusive, flow-move.
    Sequential: each simulation alone in the context, W + K steps from its
    state.  Batched: all simulations in one context (vvhd_set_batch), W + K
    steps; every simulation evolves bit-identically to its sequential run,
    so both do the same work.  Per-simulation figures: the time of its own
    step (sequential: its mean step; batched: the step of all) and the time
    at which its step is done when every simulation takes one step
    (sequential: one after another, in order; batched: the step of all)."""
    import vvflow_amd as V
    rng = np.random.default_rng(0x5EED0040)
    raw = np.exp(rng.uniform(np.log(10_000), np.log(400_000), a.nsim))
    ns = np.clip(raw * (8_000_000 / raw.sum()), 10_000, 400_000).astype(np.int64)
    res = np.exp(rng.uniform(np.log(100.0), np.log(3000.0), a.nsim))
    nseg, dt = 500, 0.005
    corners = V.gen_cylinder(0.5, nseg)
    segs1 = np.ascontiguousarray(V.segments_from_corners(corners), dtype=V.SEG_DTYPE)
    body1 = np.zeros(1, dtype=V.BODY_DTYPE)
    body1["nseg"] = nseg
    states = [cylinder_wake(int(ns[s]), 0x5EED2000 + s) for s in range(a.nsim)]
    ctx = V.Context(0)
    ctx.set_graph(a.graph)
    ctx.set_params(re=600.0, dt=dt)
    ctx.upload_bodies(segs1, body1)
    dl = ctx.download_bodies()[1][0]["slen"] / (nseg - 1)
    tree_args = (8, 5 * dl, 100 * dl)

    def step(k, re):
        ctx.set_params(re=re, dt=dt, time=k * dt)
        ctx.tree_build(*tree_args)
        ctx.calc_circulation((1.0, 0.0))
        ctx.tree_destroy()
        ctx.vortex_shed()
        ctx.calc_forces()
        ctx.zero_forces()
        ctx.tree_build(*tree_args)
        ctx.epsilon(True)
        ctx.convective((1.0, 0.0), kahan=True)
        ctx.diffusive()
        ctx.tree_destroy()
        ctx.move_and_clean(True)

    def run(re, w, k):
        """w untimed and k timed steps; particles stepped in the timed ones."""
        for j in range(w):
            step(j, re)
        ctx.phase_times()   # synchronises
        parts = 0
        t0 = time.perf_counter()
        for j in range(w, w + k):
            parts += ctx.n
            step(j, re)
        ctx.phase_times()
        return (time.perf_counter() - t0) / k, parts

    seq, seq_parts = [], 0
    for s in range(a.nsim):
        ctx.upload_bodies(segs1, body1)
        ctx.upload(states[s])
        t, p = run(float(res[s]), a.warmup, a.steps)
        seq.append(t)
        seq_parts += p
        if (s + 1) % 8 == 0:
            print(f"sequential: {s + 1}/{a.nsim} simulations", file=sys.stderr, flush=True)
    seq = np.array(seq)
    segs = np.concatenate([segs1] * a.nsim)
    bodies = np.concatenate([body1] * a.nsim)
    bodies["first_seg"] = np.arange(a.nsim) * nseg
    ctx.set_params(re=600.0, dt=dt)
    ctx.upload_bodies(segs, bodies)
    ctx.upload(np.concatenate(states))
    off = np.concatenate([[0], np.cumsum(ns)]).astype(np.int64)
    ctx.set_batch(off, np.arange(a.nsim + 1, dtype=np.int32), res)
    t_batch, batch_parts = run(600.0, a.warmup, a.steps)
    print("batched: done", file=sys.stderr, flush=True)
    ph = ctx.phase_times()
    final = np.diff(ctx.batch_offsets())
    tuning = ctx.tuning()
    ctx.close()
    done_seq = np.cumsum(seq)   # each simulation's step done, one after another
    value = batch_parts / (t_batch * a.steps)
    out = {"metric": "aggregate particle-steps/s, independent cylinder simulations advanced together",
           "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": t_batch * 1e3, "higher_is_better": True, "dtype": "fp32",
           "data": "synthetic",
           "config": {"workload": f"{a.nsim} cylinder simulations (500-segment cylinder, vortex lattices round "
                                  f"and behind it) of {int(ns.min())}-{int(ns.max())} vortices ({int(ns.sum())} "
                                  f"in all), Re {res.min():.0f}-{res.max():.0f}; one step = the whole loop: "
                                  "tree, boundary system (FP64), shedding, forces, tree, epsilon (merge), "
                                  "convective (plain, Kahan FP32), diffusive, flow-move",
                      "nsim": a.nsim, "far_criteria": 8, "vortices_after": int(final.sum()),
                      "near_form": tuning["near_form"], "graph": bool(a.graph),
                      "graph_launches": tuning["graph_launches"]},
           "batched": {"ms_per_step": t_batch * 1e3,
                       "per_simulation_step_ms_p50": t_batch * 1e3, "per_simulation_step_ms_p99": t_batch * 1e3,
                       "step_done_ms_p50": t_batch * 1e3, "step_done_ms_p99": t_batch * 1e3},
           "sequential": {"ms_all_simulations": float(seq.sum() * 1e3),
                          "value": seq_parts / (float(seq.sum()) * a.steps),
                          "per_simulation_step_ms_p50": float(np.percentile(seq, 50) * 1e3),
                          "per_simulation_step_ms_p99": float(np.percentile(seq, 99) * 1e3),
                          "step_done_ms_p50": float(np.percentile(done_seq, 50) * 1e3),
                          "step_done_ms_p99": float(np.percentile(done_seq, 99) * 1e3)},
           "speedup_vs_sequential": float(seq.sum() / t_batch),
           "phases_ms_batched": {"tree": float(ph[0]), "epsilon": float(ph[1]), "convective": float(ph[2]),
                                 "diffusive": float(ph[3]), "flow_move": float(ph[4]),
                                 "boundary_system": float(ph[5])}}
    print(json.dumps(out), flush=True)


def main():
    a = parse()
    if a.batched:
        batched(a)
        return
    if a.copy_calibration:
        copy_calibration()
        return
    if a.precision_sweep:
        precision_sweep(a)
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU: the hot path runs on a single MI355X (asked for --gpus {a.gpus}); "
              "across GPUs it would shard by slabs of tree leaves with a halo exchange of the near region, "
              "which is not built here", file=sys.stderr)
        sys.exit(2)
    import vvflow_amd as V

    n = a.n
    objs = V.gen_wake(n, 0x5EED0001)
    ctx = V.Context(0)
    ctx.set_params(re=600.0, dt=0.005)
    ctx.upload_bodies(np.zeros(0, V.SEG_DTYPE), np.zeros(0, V.BODY_DTYPE))
    ctx.upload(objs)
    ctx.snapshot()
    if a.near_form >= 0:
        ctx.set_tuning(False, a.near_form)
    B = Bench(V, ctx, n)
    learned = not a.plain_only
    if learned:
        from vvflow_amd.train_farfield import load_model
        ctx.far_model_set(load_model(a.model))
        ctx.set_far_gradient(a.far_gradient == "network")
    fc = a.learned_fc

    elapsed, ph = B.time(fc, learned, a.steps, a.warmup)
    value = n * a.steps / elapsed
    # The timed step runs kernels concurrently (near field beside the far
    # field, diffusive beside both): a kernel's own rate is measured in a
    # second, untimed pass with the kernels one at a time (same results).
    # (the near field in the form the timed steps used)
    timed_form = int(ctx.tuning()["near_form"])
    ctx.set_overlap(False)
    ctx.set_tuning(False, timed_form)
    _, ph_alone = B.time(fc, learned, ROOFLINE_STEPS, 1)
    ctx.set_overlap(True)
    if a.near_form < 0:
        ctx.set_tuning(True)
    acc = B.error(fc, learned, a.error_sample)
    near_ms, far_ms = float(ph_alone[6]), float(ph_alone[7]) if learned else 0.0
    near_tf = acc["near_pairs"] * FLOPS_PER_NEAR_PAIR / (near_ms * 1e-3) / 1e12
    near_name = "k_near" if timed_form == 0 else f"k_near_persist ({timed_form} waves per compute unit)"
    rl_near = roofline(f"{near_name} (convective near field, the autotuned form, FP32 VALU)", "fp32_valu", near_tf,
                       FP32_PEAK_TFLOPS, "TFLOP/s",
                       traffic_bytes("k_near"),
                       {"compute_unit": "FP32 VALU (no matrix instructions); peak is the published FP32 figure",
                        "algorithmic_flops": acc["near_pairs"] * FLOPS_PER_NEAR_PAIR, "near_pairs": acc["near_pairs"],
                        "kernel_ms": near_ms, "kernel_ms_in_timed_step": float(ph[6]),
                        "timing": ALONE_NOTE})
    rls = [rl_near]
    if learned:
        far_tf = acc["net_passes"] * FLOPS_PER_NET_PASS / (far_ms * 1e-3) / 1e12
        rls.append(roofline("k_far_learned (learned far-field network, bf16/fp16 MFMA)", "mfma", far_tf, BF16_PEAK_TFLOPS,
                            "TFLOP/s", traffic_bytes("k_far_learned"),
                            {"compute_unit": "bf16/fp16 matrix cores, dense peak (the same for both)",
                             "algorithmic_flops": acc["net_passes"] * FLOPS_PER_NET_PASS,
                             "network_passes": acc["net_passes"], "far_pairs": acc["far_pairs"],
                             "kernel_ms": far_ms, "kernel_ms_in_timed_step": float(ph[7]),
                             "timing": ALONE_NOTE}))
    rls.sort(key=lambda r: -r["kernel_ms"])
    # tree build and flow-move against the HBM bandwidth
    tree_ms = float(ph_alone[0])
    tree_bytes = n * acc["levels"] * TREE_BYTES_PER_VORTEX_LEVEL
    hbm = [roofline("tree build (level kernels, monopoles, interaction lists)", "hbm",
                    tree_bytes / (tree_ms * 1e-3) / 1e9, HBM_PEAK_GBS, "GB/s", traffic_bytes("tree"),
                    {"algorithmic_bytes": tree_bytes, "levels": acc["levels"], "phase_ms": tree_ms,
                     "timing": "HIP events around the tree build, steps run with kernels one at a time"})]
    fm_ms, fm_n = B.flowmove(fc, learned, ROOFLINE_STEPS)
    fm_bytes = fm_n * FLOWMOVE_BYTES_PER_VORTEX
    hbm.append(roofline("flow-move (advect, remove, compact)", "hbm", fm_bytes / (fm_ms * 1e-3) / 1e9,
                        HBM_PEAK_GBS, "GB/s", traffic_bytes("flowmove"),
                        {"algorithmic_bytes": fm_bytes, "vortices": fm_n, "phase_ms": fm_ms,
                         "timing": "HIP events around MFlowmove::move_and_clean after a step, mean of %d" %
                                   ROOFLINE_STEPS}))

    plain = None
    cpu_fc = [8]
    if learned and not a.no_plain:
        plain = {"reference_fc8": None, "at_target": None, "sweep": []}
        e8 = B.error(8, False, a.error_sample)
        el8, _ = B.time(8, False, a.plain_steps, 1)
        plain["reference_fc8"] = {"far_criteria": 8, "value": n * a.plain_steps / el8, "rel_l2": e8["rel_l2"],
                                  "ms_per_step": el8 / a.plain_steps * 1e3}
        errs = {}

        def err_at(f):
            if f not in errs:
                e = B.error(f, False, a.error_sample)
                errs[f] = e
                plain["sweep"].append({"far_criteria": f, "rel_l2": e["rel_l2"], "near_pairs": e["near_pairs"]})
                print(f"plain fc={f}: rel L2 {e['rel_l2']:.4e}", file=sys.stderr, flush=True)
            return errs[f]["rel_l2"]

        lo, hi = None, None
        for f in [int(x) for x in a.plain_fcs.split(",")]:
            if err_at(f) <= TARGET_ERR:
                hi = f
                break
            lo = f
        if hi is not None:
            # smallest integer criterion in (lo, hi] that reaches the target
            lo = lo if lo is not None else 0
            while hi - lo > 1:
                mid = (lo + hi) // 2
                if err_at(mid) <= TARGET_ERR:
                    hi = mid
                else:
                    lo = mid
            elp, _ = B.time(hi, False, a.plain_steps, 1)
            rec = {"far_criteria": hi, "rel_l2": errs[hi]["rel_l2"], "near_pairs": errs[hi]["near_pairs"],
                   "value": n * a.plain_steps / elp, "ms_per_step": elp / a.plain_steps * 1e3,
                   "search": "smallest integer far criterion with rel L2 <= 1e-5: sweep %s, then bisection "
                             "(error taken as non-increasing in the criterion between the bracketing points)" %
                             a.plain_fcs}
            plain["at_target"] = rec
            cpu_fc.append(hi)
        plain["sweep"].sort(key=lambda r: r["far_criteria"])

    cpu = None
    if not a.no_cpu:
        model, nproc, ncpu, threads = host_cpu()
        cv, cs, sample = cpu_baseline(objs, 8, 1.0, threads)
        cpu = {"value": cv, "unit": "particle-steps/s", "cores": threads, "kind": "port", "sample": sample,
               "cpu_seconds": cs, "far_criteria": 8, "cpu_model": model, "nproc": nproc, "os_cpu_count": ncpu,
               "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}
        if len(cpu_fc) > 1:
            cv2, cs2, sample2 = cpu_baseline(objs, cpu_fc[1], a.cpu_fraction, threads)
            cpu["at_plain_matched_criterion"] = {"far_criteria": cpu_fc[1], "value": cv2, "cpu_seconds": cs2,
                                                 "sample": sample2}
        print(f"cpu baseline: {cv:.4g} particle-steps/s at fc 8 on {threads} threads", file=sys.stderr, flush=True)

    out = {
        "metric": "particle-steps/s at N=1M (conv+eps+diff), learned vs plain at <=1e-5 rel err",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": f"karman-street wake, {n} vortices, one step: tree + epsilon (merge) + convective + "
                        f"diffusive, {'learned far field' if learned else 'plain tree path'} at far criterion {fc}",
            "n": n, "far_criteria": fc, "re": 600.0, "accumulation": "kahan", "pair_dtype": "fp32",
            "tree_epsilon_dtype": "fp64", "far_field": "learned (bf16/fp16 MLP)" if learned else "plain",
            "far_gradient": a.far_gradient if learned else None,
        },
        "phases_ms": {"tree": float(ph[0]), "epsilon": float(ph[1]), "convective": float(ph[2]),
                      "diffusive": float(ph[3]), "near_kernel": near_ms, "far_learned_kernel": far_ms},
        "accuracy": {"rel_l2_vs_fp64_direct_sum": acc["rel_l2"], "max_rel": acc["max_rel"],
                     "sample_targets": min(a.error_sample, n), "merged": acc["merged"]},
        "roofline": rls[0],
        "roofline_other": rls[1:] + hbm,
        "plain": plain,
        "speedup_vs_plain_at_target": (value / plain["at_target"]["value"]
                                       if plain and plain["at_target"] and acc["rel_l2"] <= TARGET_ERR else None),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
