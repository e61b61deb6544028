"""Single-step benchmark of the hot path on one MI355X.

Workload: the 1M-vortex Karman-street wake (two staggered rows of 200
Gaussian clusters, sigma 0.05, SplitMix64 seed 0x5EED0001), Re = 600,
U = (1, 0), merging on.  One step = tree build + epsilon (with merges) +
convective + diffusive velocity, starting from a device-resident copy of the
input (merges mutate it, so every step restores the input first; the
restore is a device-to-device copy inside the timed region).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N]
                    [--far-criteria F] [--plain] [--no-cpu] [--error-sample S]

Prints one JSON line.  value = N * K / (time of K steps) in particle-steps/s.
roofline: the near-field convective kernel (dominant), algorithmic FP32
flops = 11 per near pair (MConvectiveFast.cpp:116-120) x the near pairs of the
built tree, over that kernel's duration measured with HIP events on the
library's stream, against the MI355X FP32 peak of 157.3 TFLOP/s.
cpu_baseline: the FP64 oracle (the reference's algorithm) on this host:
full tree build and epsilon (serial, as the reference), convective and
diffusive on a deterministic subset of leaves spread over all host cores,
scaled to the full step.
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
HBM_PEAK_GBS = 8000.0
FLOPS_PER_NEAR_PAIR = 11
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "near_kernel_traffic_wake1M.txt")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--far-criteria", type=int, default=8)
    ap.add_argument("--plain", action="store_true", help="plain FP32 accumulation (default: Kahan)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--error-sample", type=int, default=4000)
    ap.add_argument("--cpu-fraction", type=float, default=0.05)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 256 MiB device copy (counter calibration for rocprofv3 --pmc)")
    return ap.parse_args()


def copy_calibration():
    """A plain device-to-device copy kernel of known size: FETCH_SIZE and
    WRITE_SIZE of this kernel calibrate the same counters on k_near."""
    import torch
    nbytes = 256 << 20
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_kernel_bytes_read": nbytes, "copy_kernel_bytes_written": nbytes, "launches": 5}))


def traffic_bytes():
    """HBM bytes per launch of the near-field kernel, from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, calibrated)."""
    try:
        for line in open(TRAFFIC_FILE):
            if line.startswith("traffic_bytes_per_launch"):
                return float(line.split("=")[1].split()[0])
    except OSError:
        pass
    return None


def cpu_baseline(objs, fc, fraction):
    """FP64 oracle on the host: returns (particle-steps/s, threads, sample text)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bind import Oracle

    threads = min(16, os.cpu_count() or 1)
    n = len(objs)
    o = Oracle()
    o.set_params(re=600.0, dt=0.005, inf=(1.0, 0.0))
    o.set_vortices(objs)
    t0 = time.perf_counter()
    o.tree_build(fc)
    t1 = time.perf_counter()
    o.epsilon(True)
    t2 = time.perf_counter()
    nl = o.L.vvo_tree_num_leaves(o.h)
    # a contiguous block of leaves in the middle of the DFS order
    cnt = max(1, int(nl * fraction))
    lb = (nl - cnt) // 2
    le = lb + cnt
    before = o.vortices()
    t3 = time.perf_counter()
    o.convective(threads=threads, leaf_begin=lb, leaf_end=le)
    o.diffusive(threads=threads, leaf_begin=lb, leaf_end=le)
    t4 = time.perf_counter()
    after = o.vortices()
    touched = int(np.sum(np.any(after[:, 3:5] != before[:, 3:5], axis=1)))
    touched = max(touched, 1)
    t_step = (t1 - t0) + (t2 - t1) + (t4 - t3) * (n / touched)
    sample = (f"full tree build ({t1 - t0:.2f} s) and epsilon ({t2 - t1:.2f} s), serial; convective + diffusive "
              f"on {cnt} of {nl} leaves ({touched} targets, {t4 - t3:.2f} s on {threads} threads) scaled by "
              f"{n / touched:.1f}")
    return n / t_step, threads, sample


def main():
    a = parse()
    if a.copy_calibration:
        copy_calibration()
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU: the hot path runs on a single MI355X (asked for --gpus {a.gpus}); "
              "see DESIGN.md for how it would shard", file=sys.stderr)
        sys.exit(2)
    import vvflow_amd as V

    n = a.n
    objs = V.gen_wake(n, 0x5EED0001)
    ctx = V.Context(0)
    ctx.set_params(re=600.0, dt=0.005)
    ctx.upload_bodies(np.zeros(0, V.SEG_DTYPE), np.zeros(0, V.BODY_DTYPE))
    ctx.upload(objs)
    ctx.snapshot()
    kahan = not a.plain
    inf = (1.0, 0.0)

    def step():
        ctx.restore()
        ctx.tree_build(a.far_criteria)
        ctx.epsilon(True)
        ctx.convective(inf, kahan=kahan)
        ctx.diffusive()

    for _ in range(a.warmup):
        step()
    ctx.phase_times()   # synchronises
    near_ms, phases = [], np.zeros(8)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
        pt = ctx.phase_times()
        phases += pt
        near_ms.append(pt[6])
    t1 = time.perf_counter()
    elapsed = t1 - t0
    value = n * a.steps / elapsed
    phases /= a.steps

    work = ctx.work_counts()
    near_pairs = int(work[0])
    flops = near_pairs * FLOPS_PER_NEAR_PAIR
    near_avg_ms = float(np.mean(near_ms))
    achieved = flops / (near_avg_ms * 1e-3) / 1e12

    # accuracy of this step against the FP64 direct sum (induced velocity)
    ctx.restore()
    ctx.tree_build(a.far_criteria)
    merged = ctx.epsilon(True)
    ctx.convective((0.0, 0.0), kahan=kahan)
    v = ctx.download()[:, 3:5]
    tg = np.random.default_rng(0).choice(n, min(a.error_sample, n), replace=False).astype(np.int64)
    ref = ctx.direct_sum(tg)
    err_l2 = float(np.linalg.norm(v[tg] - ref) / np.linalg.norm(ref))
    nr = np.linalg.norm(ref, axis=1)
    err_max = float(np.max(np.linalg.norm(v[tg] - ref, axis=1) / np.maximum(nr, 1e-3 * np.sqrt(np.mean(nr ** 2)))))

    cpu = None
    if not a.no_cpu:
        cv, cores, sample = cpu_baseline(objs, a.far_criteria, a.cpu_fraction)
        cpu = {"value": cv, "unit": "particle-steps/s", "cores": cores, "kind": "port", "sample": sample}

    tb = traffic_bytes()
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
                        f"diffusive, plain tree path",
            "n": n, "far_criteria": a.far_criteria, "re": 600.0, "accumulation": "kahan" if kahan else "plain",
            "pair_dtype": "fp32", "tree_epsilon_dtype": "fp64",
        },
        "phases_ms": {"tree": float(phases[0]), "epsilon": float(phases[1]), "convective": float(phases[2]),
                      "diffusive": float(phases[3]), "near_kernel": float(phases[6])},
        "accuracy": {"rel_l2_vs_fp64_direct_sum": err_l2, "max_rel": err_max, "sample_targets": int(len(tg)),
                     "merged": int(merged)},
        "roofline": {
            "kernel": "k_near (convective near field)",
            "bound": "mfma",
            "compute_unit": "FP32 VALU (no matrix instructions); peak is the published FP32 figure",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": tb,
            "algorithmic_flops": flops,
            "near_pairs": near_pairs,
            "kernel_ms": near_avg_ms,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
