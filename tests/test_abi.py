"""The C ABI library: loads on a machine without a GPU, exports every entry
point declared in include/vvhd_gpu.h, record layouts agree, and the host-side
input generators follow their specification."""
import ctypes as C
import math
import os
import re

import numpy as np

import vvflow_amd as V

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "vvhd_gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(vvhd_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 30
    L = V.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(V.ABI) == names


def test_record_layouts():
    assert V.SEG_DTYPE.itemsize == 14 * 8 + 8
    assert V.BODY_DTYPE.itemsize == 27 * 8 + 7 * 8 + 4 * 4


def test_create_without_device_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        return
    h = C.c_void_p()
    rc = V.lib().vvhd_create(0, C.byref(h))
    assert rc in (V.VVHD_EHIP, V.VVHD_EINVAL) and not h.value


def _splitmix(seed):
    s = seed
    M = (1 << 64) - 1
    while True:
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        yield z ^ (z >> 31)


def test_wake_generator_follows_spec():
    # independent restatement of the documented generator for the first
    # clusters: SplitMix64, 53-bit uniforms, Box-Muller, g = s (0.5 + u) 1e-3
    n = 4000
    objs = V.gen_wake(n, 0x5EED0001)
    rng = _splitmix(0x5EED0001)
    u = lambda: (next(rng) >> 11) * 2.0 ** -53
    per = n // 400
    for c in range(3):
        for b in range(per):
            u1 = 1.0 - u()
            u2 = u()
            r = math.sqrt(-2.0 * math.log(u1))
            x = 0.3 * c + 0.05 * (r * math.cos(2.0 * math.pi * u2))
            y = 0.2 + 0.05 * (r * math.sin(2.0 * math.pi * u2))
            g = -1.0 * (0.5 + u()) * 1e-3
            assert tuple(objs[c * per + b, :3]) == (x, y, g)
    # rows and signs
    top = objs[: n // 2]
    bot = objs[n // 2:]
    assert np.all(top[:, 2] < 0) and np.all(bot[:, 2] > 0)
    assert abs(top[:, 1].mean() - 0.2) < 0.01 and abs(bot[:, 1].mean() + 0.2) < 0.01
    assert np.all(np.abs(objs[:, 2]) >= 0.5e-3) and np.all(np.abs(objs[:, 2]) <= 1.5e-3)


def test_cylinder_generator_matches_gen_arc():
    # libm cos/sin as gen_arc_N uses them; a compiler may fuse the pair into
    # sincos, which can differ from separate calls in the last bit
    c = V.gen_cylinder(0.5, 500)
    for i in range(500):
        a = 2 * math.pi + (0 - 2 * math.pi) * float(i) / 500.0
        ex = (0.5 * math.cos(a), 0.5 * math.sin(a))
        for k in range(2):
            assert abs(c[i, k] - ex[k]) <= np.spacing(abs(ex[k]))


def test_segments_from_corners_match_oracle():
    from oracle_bind import Oracle
    corners = V.gen_cylinder(0.5, 64)
    s = V.segments_from_corners(corners)
    o = Oracle()
    o.add_body(corners)
    so = o.segments()
    for f in ("x", "y", "dlx", "dly", "inv_eps", "cx", "cy"):
        assert np.array_equal(s[f], so[f]), f
