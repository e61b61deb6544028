"""The FP64 oracle against the reference's own outputs (tests/golden/*.npz,
written by tests/golden/make_golden.py from TSortedTree.cpp, MEpsilonFast.cpp
and MDiffusiveFast.cpp): bit-exact tree, lists, epsilon/merges, diffusive."""
import numpy as np
import pytest

from cases import FIXTURES, load_golden, oracle_for, params, path_code32


def _oracle_tree(d):
    objs = d["objs"].copy()
    objs[:, 3] = np.arange(len(objs))
    o = oracle_for(objs, segs=d["segs"], re=float(d["re"]))
    o.tree_build(*params(d))
    return o


@pytest.mark.parametrize("name", FIXTURES)
def test_tree_matches_reference(name):
    d = load_golden(name)
    o = _oracle_tree(d)
    v = o.vortices()
    assert np.array_equal(v[:, 3].astype(np.int32), d["perm"])
    T = o.tree()
    leaves = T["leaves"]
    assert np.array_equal(T["geom"][leaves], d["leaf_geom"])
    assert np.array_equal(T["ints"][leaves][:, :2], d["leaf_ints"][:, :2])
    assert np.array_equal(T["ints"][leaves][:, 4], d["leaf_ints"][:, 2])
    assert np.array_equal(path_code32(T["path"][leaves]), d["leaf_ints"][:, 3].astype(np.uint32).astype(np.int64))
    for key in ("near", "far"):
        off, lst = T[key]
        ref = d[key].astype(np.int64)
        assert np.array_equal(T["ints"][lst, 4], ref[:, 0])
        assert np.array_equal(path_code32(T["path"][lst]), ref[:, 1].astype(np.uint32).astype(np.int64))
    off, lst = T["far"]
    cm = np.concatenate([T["cm"][lst][:, 0:3], T["cm_inv_eps"][lst][:, 0:1],
                         T["cm"][lst][:, 3:6], T["cm_inv_eps"][lst][:, 1:2]], 1)
    assert np.array_equal(cm, d["far_cm"])
    if d["segs"] is not None:
        off, sg = T["segs"]
        assert np.array_equal(sg, d["leaf_segs"])


@pytest.mark.parametrize("name", FIXTURES)
def test_epsilon_and_diffusive_match_reference(name):
    d = load_golden(name)
    o = _oracle_tree(d)
    v = o.vortices()
    v[:, 3] = 0.0
    o.set_vortices(v)
    o.tree_build(*params(d))
    nn1, nn2, mg, merged = o.epsilon(bool(d["merge"]))
    assert merged == int(d["merged"])
    assert np.array_equal(o.vortices()[:, [0, 1, 2, 5]], d["after_eps"])
    o.diffusive()
    assert np.array_equal(o.vortices()[:, 3:5], d["after_diff_v"])
    if d["segs"] is not None:
        assert np.array_equal(o.segments()["fric"], d["fric"])
