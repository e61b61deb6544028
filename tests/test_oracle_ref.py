"""The FP64 oracle against the reference's own tree/epsilon/diffusive code,
compiled from /root/reference into oracle/_ref (skipped where that build is
absent, e.g. on a machine without the reference sources)."""
import numpy as np
import pytest

from cases import DBL_MAX, layer_case, oracle_for, path_code32, random_cluster_case
from oracle_bind import RefHarness

pytestmark = pytest.mark.skipif(not RefHarness.available(), reason="oracle/_ref not built (no reference sources)")


def _compare(objs, corners=None, fc=8, mn=0.0, mx=DBL_MAX, merge=True):
    objs = objs.copy()
    objs[:, 3] = np.arange(len(objs))
    o = oracle_for(objs, corners=corners)
    r = RefHarness()
    r.set_re(600.0)
    r.set_vortices(objs)
    if corners is not None:
        r.add_body(o.segments())
    o.tree_build(fc, mn, mx)
    r.tree_build(fc, mn, mx)
    assert np.array_equal(o.vortices(), r.vortices(), equal_nan=True)
    T = o.tree()
    R = r.leaves()
    lv = T["leaves"]
    assert np.array_equal(T["geom"][lv], R["geom"])
    assert np.array_equal(T["ints"][lv][:, :2], R["ints"][:, :2])
    for key in ("near", "far"):
        off, lst = T[key]
        ref = R[key].astype(np.int64)
        assert np.array_equal(T["ints"][lst, 4], ref[:, 0])
        assert np.array_equal(path_code32(T["path"][lst]), ref[:, 1].astype(np.uint32).astype(np.int64))
    v = o.vortices()
    v[:, 3] = 0
    o.set_vortices(v)
    r.set_vortices(v)
    o.tree_build(fc, mn, mx)
    r.tree_build(fc, mn, mx)
    _, _, _, m = o.epsilon(merge)
    assert m == r.epsilon(merge)
    assert np.array_equal(o.vortices(), r.vortices(), equal_nan=True)
    o.diffusive()
    r.diffusive()
    assert np.array_equal(o.vortices(), r.vortices(), equal_nan=True)
    if corners is not None:
        assert np.array_equal(o.segments()["fric"], r.fric())
    return m


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("signs", ["mixed", "cluster", "positive"])
def test_random_clusters(seed, signs):
    _compare(random_cluster_case(1500 + 700 * seed, seed, signs))


@pytest.mark.parametrize("fc", [2, 8, 30])
def test_far_criteria(fc):
    _compare(random_cluster_case(2500, 5, "cluster"), fc=fc)


def test_no_merge_mode():
    _compare(random_cluster_case(2000, 9, "mixed"), merge=False)


@pytest.mark.parametrize("seed", [4, 5])
def test_cylinder_layer(seed):
    objs, corners = layer_case(2500, seed)
    dl = np.pi / 499
    assert _compare(objs, corners=corners, mn=5 * dl, mx=100 * dl) > 0


def test_duplicates_and_tiny():
    rng = np.random.default_rng(0)
    base = random_cluster_case(300, 3, "mixed")
    dup = np.concatenate([base, base[:50], base[:5]])   # exact duplicate positions
    _compare(dup[rng.permutation(len(dup))])
    for n in (1, 2, 3, 17):
        _compare(random_cluster_case(n, 10 + n, "mixed"))
