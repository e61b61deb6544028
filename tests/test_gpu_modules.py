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
