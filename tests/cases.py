"""Shared test inputs and helpers (fixtures, synthetic cases, comparisons)."""
import os

import numpy as np

from oracle_bind import SEG_DTYPE, Oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
DBL_MAX = np.finfo(np.float64).max
FIXTURES = ["wake3000", "wake2000_nomerge", "mixed3000", "cylinder3000"]


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    if "segs" in d:
        d["segs"] = np.frombuffer(d["segs"].tobytes(), dtype=SEG_DTYPE).copy()
    else:
        d["segs"] = None
    return d


def params(d):
    return int(d["fc"]), float(d["min_node"]), float(d["max_node"])


def oracle_for(objs, segs=None, re=600.0, inf=(1.0, 0.0), dt=0.005, corners=None):
    """Oracle context with vortices and (optionally) one body.  A body is
    given by corners (doUpdateSegments runs in the oracle) or by segments
    whose corners are used."""
    o = Oracle()
    o.set_params(re=re, dt=dt, inf=inf)
    o.set_vortices(objs)
    if corners is not None:
        o.add_body(corners)
    elif segs is not None:
        o.add_body(np.stack([segs["cx"], segs["cy"]], 1))
    return o


def path_code32(path):
    """The reference's snode::j (path code in a 32-bit int, wrapping)."""
    return (np.asarray(path).astype(np.uint64) & np.uint64(0xFFFFFFFF)).astype(np.uint32).astype(np.int64)


def random_cluster_case(n, seed, signs="mixed", spread=1.0):
    rng = np.random.default_rng(seed)
    objs = np.zeros((n, 6))
    k = max(1, n // 200)
    centres = rng.uniform(0, spread, (k, 2))
    lab = rng.integers(0, k, n)
    objs[:, :2] = centres[lab] + rng.normal(size=(n, 2)) * 0.02 * spread
    mag = rng.uniform(0.5, 1.5, n) * 1e-3
    if signs == "mixed":
        objs[:, 2] = rng.choice([-1.0, 1.0], n) * mag
    elif signs == "cluster":
        objs[:, 2] = np.where(lab % 2 == 0, 1.0, -1.0) * mag
    else:
        objs[:, 2] = mag
    return objs


def cylinder_corners(n=500, r=0.5):
    """gen_cylinder corners (utils/vvflow/gen_body.cpp:74-88 with a1=2pi, a2=0)."""
    i = np.arange(n, dtype=np.float64)
    a = 2 * np.pi + (0 - 2 * np.pi) * i / n
    return np.stack([r * np.cos(a), r * np.sin(a)], 1)


def layer_case(n, seed, n_seg=500):
    """Vortices in a thin layer around a cylinder (merges and segment terms)."""
    rng = np.random.default_rng(seed)
    th = rng.uniform(0, 2 * np.pi, n)
    rr = 0.5 + 1e-4 + rng.exponential(0.03, n)
    objs = np.zeros((n, 6))
    objs[:, 0] = rr * np.cos(th)
    objs[:, 1] = rr * np.sin(th)
    objs[:, 2] = rng.normal(size=n) * 1e-3
    return objs, cylinder_corners(n_seg)


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def max_rel(a, b, floor_frac=1e-3):
    """max |a-b| / max(|b|, floor_frac * rms|b|) over rows of 2-vectors."""
    a = np.asarray(a).reshape(-1, 2)
    b = np.asarray(b).reshape(-1, 2)
    nb = np.linalg.norm(b, axis=1)
    floor = floor_frac * np.sqrt(np.mean(nb ** 2)) if len(nb) else 0.0
    return float(np.max(np.linalg.norm(a - b, axis=1) / np.maximum(nb, max(floor, 1e-300)))) if len(nb) else 0.0
