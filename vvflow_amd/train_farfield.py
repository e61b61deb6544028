"""train_farfield.py -- trains the learned far-field correction on the GPU.

    python -m vvflow_amd.train_farfield [--out PATH] [--pairs 20000000]
        [--steps 300000] [--batch 8192] [--re 200,400,800,1400,2000]
        [--sim-steps 400] [--keep-every 20] [--fc 8,10,12,16]
        [--wake 250000:0x5EED0011,500000:0x5EED0012,1000000:0x5EED0013]
        [--wake-pairs 20000000] [--weight s2|s|one] [--lr0 2e-3]

1. Snapshots: impulsively started cylinders (R = 0.5, 500 segments,
   dt = 0.005) run on the device at each Reynolds number; the vortex list is
   kept every `keep_every` steps.  These wakes stay small (a few thousand
   vortices after 400 steps), so dense Karman-street clouds from the wake
   generator are added (sizes and seeds of --wake; the benchmark's seed is
   not among them).
2. Pairs: every snapshot is rebuilt as a free vortex cloud with the far
   criteria listed (no body, unbounded node sizes, as for the wake
   benchmark), the epsilon pass runs, and (vortex, far node) pairs are
   sampled with their FP64 direct-sum residual labels (vvhd_far_sample),
   about --pairs from the cylinders and --wake-pairs from the wakes.
3. Training: feature means / deviations and output scales from the pairs,
   then Adam on batches drawn on the device, cosine learning-rate decay, with
   the device's bf16 MFMA forward and backward kernels.  The loss weights
   each pair by S^2 (mean 1), S being the pair's velocity scale: the error
   of a far sum is the sum of S (y - Y) over its pairs, so pairs close to
   their node or of strong nodes count as much as they do in the velocity.
   Against unweighted training this halves the velocity error of the learned
   path at a given far criterion.  A held-out 5 % of
   the pairs gives the relative error of the predicted correction.
The model (VVHD_FAR_PARAMS float32 values) is written with numpy.save.
"""
import argparse
import json
import os
import time

import numpy as np

from . import FAR_PARAMS, Context, gen_wake
from .simulate import CylinderRun

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_PATH = os.path.join(HERE, "farfield_mlp.npy")

N_IN, N_HID = 12, 64


def init_params(rng, mu, sd, ysd):
    p = np.zeros(FAR_PARAMS, dtype=np.float32)
    o = 0
    p[o:o + N_HID * N_IN] = rng.normal(0, np.sqrt(2.0 / N_IN), N_HID * N_IN)
    o += N_HID * N_IN + N_HID
    p[o:o + N_HID * N_HID] = rng.normal(0, np.sqrt(2.0 / N_HID), N_HID * N_HID)
    o += N_HID * N_HID + N_HID
    p[o:o + 2 * N_HID] = rng.normal(0, np.sqrt(1.0 / N_HID), 2 * N_HID)
    o += 2 * N_HID + 2
    p[o:o + N_IN] = mu
    p[o + N_IN:o + 2 * N_IN] = sd
    p[o + 2 * N_IN:o + 2 * N_IN + 2] = ysd
    return p


def load_model(path=MODEL_PATH):
    p = np.load(path, allow_pickle=False)
    if p.shape != (FAR_PARAMS,) or p.dtype != np.float32:
        raise ValueError(f"{path}: expected {FAR_PARAMS} float32 parameters")
    return p


def snapshots(ctx, re_list, sim_steps, keep_every, progress=100):
    out = []
    for re in re_list:
        run = CylinderRun(ctx, re=re)
        for v in run.run(sim_steps, keep_every=keep_every, progress=progress):
            out.append((re, v))
    return out


def sample_pairs(ctx, snaps, fcs, target_pairs, seed=1):
    from . import BODY_DTYPE, SEG_DTYPE
    ctx.upload_bodies(np.zeros(0, SEG_DTYPE), np.zeros(0, BODY_DTYPE))
    # rate from a first count so that the total lands near target_pairs
    counts = []
    for k, (re, v) in enumerate(snaps):
        ctx.set_params(re=re, dt=0.005)
        ctx.upload(v)
        ctx.tree_build(fcs[k % len(fcs)])
        ctx.epsilon(True)
        counts.append(ctx.far_pair_count())
    rate = min(1.0, target_pairs / max(1, sum(counts)))
    Xs, Ys, Ss = [], [], []
    for k, (re, v) in enumerate(snaps):
        ctx.set_params(re=re, dt=0.005)
        ctx.upload(v)
        ctx.tree_build(fcs[k % len(fcs)])
        ctx.epsilon(True)
        s = ctx.far_sample(rate, seed=seed + k)
        ok = s["scale"] > 0
        Xs.append(s["X"][ok])
        Ys.append(s["Y"][ok])
        Ss.append(s["scale"][ok])
    return np.concatenate(Xs), np.concatenate(Ys), np.concatenate(Ss), rate, sum(counts)


def correction_error(ctx, X, Y, S):
    """Relative L2 error of the predicted velocity correction (scale-weighted)."""
    p = ctx.far_model_get()
    ysd = p[-2:]
    y = ctx.far_forward(X) * ysd
    num = np.sqrt(np.sum((S[:, None] * (y - Y)) ** 2))
    den = np.sqrt(np.sum((S[:, None] * Y) ** 2))
    return float(num / den)


def loss_weights(S, kind):
    """Per-pair loss weights, mean 1: 'one', or the pair's velocity scale S
    ('s') or S^2 ('s2').  The far sum's error at a target is the sum over its
    far nodes of S (y - Y), so 's2' weights the squared error as it enters
    the velocity."""
    if kind == "one":
        w = np.ones(len(S))
    elif kind == "s":
        w = np.abs(S).astype(np.float64)
    elif kind == "s2":
        w = np.square(S.astype(np.float64))
    else:
        raise ValueError(f"unknown loss weighting {kind!r}")
    return (w / w.mean()).astype(np.float32)


def train(ctx, X, Y, steps, batch, lr0=2e-3, lr1=2e-5, seed=7, log_every=2000, W=None):
    rng = np.random.default_rng(seed)
    mu = X.mean(0)
    sd = X.std(0) + 1e-6
    ysd = Y.std(0) + 1e-12
    ctx.far_model_set(init_params(rng, mu, sd, ysd))
    ctx.far_train_data(X, Y, np.ones(len(X), np.float32) if W is None else W)
    k = np.arange(steps)
    lr = (lr1 + 0.5 * (lr0 - lr1) * (1 + np.cos(np.pi * k / max(1, steps - 1)))).astype(np.float32)
    losses = []
    for a in range(0, steps, log_every):
        b = min(steps, a + log_every)
        ls = ctx.far_train(lr[a:b], batch=batch, seed=seed * 1000003 + a)
        losses.append(float(np.mean(ls)))
        print(f"  train steps {b}/{steps} batch loss {losses[-1]:.4e} lr {lr[b - 1]:.2e}", flush=True)
    return losses


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=MODEL_PATH)
    ap.add_argument("--pairs", type=int, default=20_000_000)
    ap.add_argument("--steps", type=int, default=300_000)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--re", default="200,400,800,1400,2000")
    ap.add_argument("--sim-steps", type=int, default=400)
    ap.add_argument("--keep-every", type=int, default=20)
    ap.add_argument("--fc", default="8,10,12,16")
    ap.add_argument("--wake", default="250000:0x5EED0011,500000:0x5EED0012,1000000:0x5EED0013")
    ap.add_argument("--wake-pairs", type=int, default=20_000_000)
    ap.add_argument("--weight", default="s2", choices=["s2", "s", "one"])
    ap.add_argument("--lr0", type=float, default=2e-3)
    a = ap.parse_args()
    t0 = time.perf_counter()
    ctx = Context(0)
    re_list = [float(r) for r in a.re.split(",")]
    snaps = snapshots(ctx, re_list, a.sim_steps, a.keep_every)
    print(f"snapshots: {len(snaps)}, vortices {[len(v) for _, v in snaps]} ({time.perf_counter() - t0:.0f} s)",
          flush=True)
    fcs = [int(f) for f in a.fc.split(",")]
    X, Y, S, rate, total = sample_pairs(ctx, snaps, fcs, a.pairs)
    print(f"cylinder pairs: {len(X)} of {total} (rate {rate:.4f}) ({time.perf_counter() - t0:.0f} s)", flush=True)
    wakes = []
    for item in filter(None, a.wake.split(",")):
        n, seed = item.split(":")
        wakes.append((600.0, gen_wake(int(n), int(seed, 0))))
    if wakes:
        Xw, Yw, Sw, rate_w, total_w = sample_pairs(ctx, wakes * len(fcs), [f for f in fcs for _ in wakes],
                                                   a.wake_pairs, seed=1001)
        print(f"wake pairs: {len(Xw)} of {total_w} (rate {rate_w:.5f}) ({time.perf_counter() - t0:.0f} s)",
              flush=True)
        X, Y, S = np.concatenate([X, Xw]), np.concatenate([Y, Yw]), np.concatenate([S, Sw])
        total += total_w
    perm = np.random.default_rng(0).permutation(len(X))
    nval = max(1, len(X) // 20)
    va, tr = perm[:nval], perm[nval:]
    losses = train(ctx, X[tr], Y[tr], a.steps, a.batch, lr0=a.lr0, W=loss_weights(S[tr], a.weight))
    err_tr = correction_error(ctx, X[tr[:200000]], Y[tr[:200000]], S[tr[:200000]])
    err_va = correction_error(ctx, X[va], Y[va], S[va])
    p = ctx.far_model_get()
    np.save(a.out, p, allow_pickle=False)
    info = dict(pairs=int(len(X)), total_pairs=int(total), rate=rate, steps=a.steps, batch=a.batch,
                re=re_list, fc=fcs, sim_steps=a.sim_steps, keep_every=a.keep_every, wake=a.wake,
                weight=a.weight, lr0=a.lr0,
                correction_rel_err_train=err_tr, correction_rel_err_heldout=err_va, final_loss=losses[-1],
                seconds=time.perf_counter() - t0)
    with open(os.path.splitext(a.out)[0] + ".json", "w") as f:
        json.dump(info, f, indent=1)
    print(json.dumps(info), flush=True)


if __name__ == "__main__":
    main()
