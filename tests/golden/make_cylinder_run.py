"""Writes tests/golden/cylinder400_oracle.npz: the FP64 oracle's run of the
impulsively started cylinder (R = 0.5, 500 segments, Re = 600, U = (1, 0),
dt = 0.005) for 400 steps of the reference's loop (utils/vvflow/vvflow.cpp:
214-262: tree, calc_circulation, shed, forces, tree, epsilon with merges,
convective, diffusive, flow-move).

Per step it keeps what a statistical comparison of a free-running GPU run
needs: the vortex count after flow-move, the merges of the epsilon pass, the
total and absolute circulation, and the hydrodynamic force solved by the
boundary system.  The oracle is deterministic (its OpenMP loops write
disjoint targets), so the file depends only on the oracle's code.

    python tests/golden/make_cylinder_run.py [steps]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from cases import cylinder_corners, oracle_for  # noqa: E402

DL = 500 * np.sin(np.pi / 500) / 499       # Space::average_segment_length: slen / (N - 1)
MIN_NODE, MAX_NODE = 5 * DL, 100 * DL   # utils/vvflow/vvflow.cpp:200-202
DT = 0.005


def oracle_step(o, k, threads=8):
    """One step of the reference's loop on the oracle; returns (merges,
    force_hydro) of the step."""
    o.set_params(re=600.0, dt=DT, inf=(1.0, 0.0), time=k * DT)
    o.tree_build(8, MIN_NODE, MAX_NODE)
    o.calc_circulation()
    o.tree_destroy()
    o.vortex_shed()
    o.calc_forces()
    force = np.array(o.body(0)["force_hydro"], dtype=np.float64)
    o.zero_forces()
    o.tree_build(8, MIN_NODE, MAX_NODE)
    _, _, _, merged = o.epsilon(True)
    o.convective(threads=threads)
    o.diffusive(threads=threads)
    o.tree_destroy()
    o.move_and_clean(True)
    return merged, force


def run(steps, threads=8):
    o = oracle_for(np.zeros((0, 6)), corners=cylinder_corners(), inf=(1.0, 0.0), dt=DT)
    rec = {k: [] for k in ("n", "merged", "gsum", "gabs", "force")}
    for k in range(steps):
        merged, force = oracle_step(o, k, threads)
        v = o.vortices()
        rec["n"].append(len(v))
        rec["merged"].append(merged)
        rec["gsum"].append(v[:, 2].sum())
        rec["gabs"].append(np.abs(v[:, 2]).sum())
        rec["force"].append(force)
    return {k: np.array(x) for k, x in rec.items()}


if __name__ == "__main__":
    steps = int(sys.argv[1]) if len(sys.argv) > 1 else 400
    out = run(steps)
    np.savez_compressed(os.path.join(HERE, "cylinder400_oracle.npz"), dt=DT, **out)
    print(f"{steps} steps: final n = {out['n'][-1]}, merges in the last step = {out['merged'][-1]}")
