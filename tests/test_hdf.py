"""The HDF5 snapshot format (vvflow_amd.hdf over libvvhd_hdf.so): round trips
and the file's structure as h5dump (the HDF5 tools of the image, independent
of the library's reader) shows it, against what the reference's writer puts
there (TSpace_save_hdf.cpp:11-221, vvhdf.cpp, vvhdf_h5t.cpp): the committed
types vec_t, vec3d_t, fraction_t and string_t, N x 3 list datasets, N x 4
body datasets with their attributes, times as fractions.  The reference's
writer itself cannot be built here (its Space needs the Lua-backed TEval and
generated version files), so the format is pinned by these structural
checks: parity unpinned against files the reference wrote (none ship with
it)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import vvflow_amd as V
from vvflow_amd import hdf

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)


def space_case():
    rng = np.random.default_rng(3)
    lists = {}
    for name, n in (("vort", 1500), ("heat", 300), ("ink", 40), ("ink_source", 5)):
        a = rng.normal(size=(n, 6))
        lists[name] = a
    corners = V.gen_cylinder(0.5, 64)
    segs = V.segments_from_corners(corners)
    seg4 = np.stack([segs["cx"], segs["cy"], rng.normal(size=64) * 1e-3, rng.normal(size=64) * 1e-2], 1)
    b0 = dict(label="cylinder", speed_x="0", speed_y="0", speed_o="0.1*t", holder=[0.1, 0.2, 0.3],
              dpos=[0.0, -0.5, 1.5], speed_slae=[1.0, 2.0, 3.0], kspring=[-1.0, -1.0, -1.0], density=2.5,
              force_hydro=[0.5, -0.25, 0.125], g_dead=-3e-4, area=0.785, cofm=[0.0, 0.0], moi_cofm=0.1,
              special_segment_no=3, boundary_condition=1, heat_condition=2, heat_const=0.75, seg4=seg4,
              slip=np.zeros(64, np.uint32))
    seg4b = seg4[:20] + 3.0
    slip_b = (np.arange(20) % 3 == 0).astype(np.uint32)   # not uniform: stored as an array
    b1 = dict(label="plate", root_body="body00", seg4=seg4b, slip=slip_b, heat_const=0.0)
    space = dict(caption="cyl_re600", time=0.375, dt=0.005, dt_save=0.5, dt_streak=0.02, dt_profile=0.0, re=600.0,
                 pr=0.7, inf_g=0.1, finish=100.0, inf_marker=[1.5, 0.0], gravity=[0.0, -9.8],
                 inf_speed_x="1", inf_speed_y="0.1*sin(t)")
    return space, lists, [b0, b1]


def test_round_trip(tmp_path):
    space, lists, bodies = space_case()
    path = str(tmp_path / "snap.h5")
    hdf.save(path, space, lists, bodies)
    sp, ls, bs = hdf.load(path)
    for k, v in space.items():
        if isinstance(v, list):
            assert sp[k] == v, k
        elif isinstance(v, str):
            assert sp[k] == v, k
        else:
            assert sp[k] == v, k
    for name, a in lists.items():
        got = ls[name]
        assert got.shape == a.shape
        assert np.array_equal(got[:, :3], a[:, :3]), name
        assert np.all(got[:, 3:] == 0), name   # v and _1_eps come back zero (strided read)
    assert ls["source"].shape == (0, 6)
    assert len(bs) == 2
    for b, ref in zip(bs, bodies):
        assert np.array_equal(b["seg4"], ref["seg4"])
        assert np.array_equal(b["slip"], ref["slip"])
        for k, v in ref.items():
            if k in ("seg4", "slip"):
                continue
            assert b[k] == v, k
    assert bs[1]["boundary_condition"] == 0 and bs[1]["heat_condition"] == 0


def test_empty_lists_and_fractions(tmp_path):
    path = str(tmp_path / "empty.h5")
    hdf.save(path, dict(time=7 / 200, dt=1 / 3, re=float("inf")), {}, [])
    sp, ls, bs = hdf.load(path)
    assert sp["time"] == 7 / 200 and sp["dt"] == 1 / 3 and sp["re"] == float("inf")
    assert all(len(v) == 0 for v in ls.values()) and bs == []


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not in the image")
def test_file_structure_as_h5dump_sees_it(tmp_path):
    space, lists, bodies = space_case()
    path = str(tmp_path / "snap.h5")
    hdf.save(path, space, lists, bodies)
    out = subprocess.run([H5DUMP, "-H", path], capture_output=True, text=True, check=True).stdout
    for t in ('DATATYPE "vec_t"', 'DATATYPE "vec3d_t"', 'DATATYPE "fraction_t"', 'DATATYPE "string_t"'):
        assert t in out, t
    for name, n in (("vort", 1500), ("heat", 300), ("ink", 40), ("ink_source", 5)):
        block = out[out.index(f'DATASET "{name}"'):]
        assert f"( {n}, 3 )" in block[:400], name
    assert 'DATASET "source"' not in out   # empty lists are not written
    body = out[out.index('DATASET "body00"'):]
    assert "( 64, 4 )" in body[:300]
    for a in ("general_slip", "general_heat_const", "simplified_dataset", "holder_position", "delta_position",
              "speed_slae", "spring_const", "special_segment_no", "boundary_condition", "heat_condition", "label",
              "moi_axis", "collision_min"):
        assert f'ATTRIBUTE "{a}"' in body, a
    body1 = out[out.index('DATASET "body01"'):]
    assert 'ATTRIBUTE "slip"' in body1 and 'ATTRIBUTE "root_body"' in body1
    # a time attribute: fraction_t {value, timescale}
    full = subprocess.run([H5DUMP, "-a", "/dt", path], capture_output=True, text=True, check=True).stdout
    assert "200" in full and '"value"' in full and '"timescale"' in full
    full = subprocess.run([H5DUMP, "-a", "/time", path], capture_output=True, text=True, check=True).stdout
    assert "3," in full and "8" in full    # 0.375 = 3 / 8, in lowest terms
    # chunked and deflated like the reference's datasets
    vort = subprocess.run([H5DUMP, "-p", "-H", "-d", "/vort", path], capture_output=True, text=True,
                          check=True).stdout
    assert "CHUNKED ( 512, 3 )" in vort and "DEFLATE { LEVEL 9 }" in vort


def dump(path, dset, tmp_path):
    """A dataset's values through h5dump's binary output (float32, little endian)."""
    out = tmp_path / (dset.replace("/", "_") + ".bin")
    subprocess.run([H5DUMP, "-d", dset, "-b", "LE", "-o", str(out), path], capture_output=True, check=True)
    return np.fromfile(out, dtype="<f4")


def step_bodies(k, nb=2):
    return [dict(force_hydro=[k + 0.5, -k, 0.25 * b], force_holder=[k, b, 1.0], friction=[1e-3 * k, 2e-3, b],
                 holder=[b, 0.0, 0.1 * k], dpos=[0.0, 0.0, 0.0], speed_slae=[0.0, 0.0, 0.01 * k]) for b in range(nb)]


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not in the image")
def test_stepdata_rows_restart_and_structure(tmp_path):
    """Stepdata (MStepdata.cpp): one row per write; reopening at a time keeps
    the rows written before it (a restart) and cuts every dataset there;
    float datasets chunked by 128 rows, deflate 9, NaN fill, unlimited rows."""
    path = str(tmp_path / "stepdata_cyl.h5")
    nseg = [64, 20]
    rng = np.random.default_rng(5)
    cps = [rng.normal(size=84) for _ in range(5)]
    with hdf.Stepdata(path, "cyl", 0.0, nseg, profile=True) as sd:
        for k in range(5):
            sd.write(0.005 * k, step_bodies(k), pressure=cps[k], friction=-cps[k])
    t = dump(path, "/time", tmp_path)
    assert np.array_equal(t, (0.005 * np.arange(5)).astype(np.float32))
    fh = dump(path, "/body01/force_hydro", tmp_path).reshape(5, 3)
    assert np.array_equal(fh, np.array([b["force_hydro"] for b in (step_bodies(k)[1] for k in range(5))],
                                       dtype=np.float32))
    pr = dump(path, "/body00/pressure", tmp_path).reshape(5, 64)
    assert np.array_equal(pr, np.array([c[:64] for c in cps], dtype=np.float32))
    fr = dump(path, "/body01/friction", tmp_path).reshape(5, 20)
    assert np.array_equal(fr, np.array([-c[64:] for c in cps], dtype=np.float32))
    # restart at t = 0.0125: rows 0.000, 0.005, 0.010 stay; then two new rows
    with hdf.Stepdata(path, "cyl", 0.0125, nseg, profile=False) as sd:
        for k in (7, 8):
            sd.write(0.005 * k, step_bodies(k))
    t = dump(path, "/time", tmp_path)
    assert np.array_equal(t, np.float32([0.0, 0.005, 0.01, 0.035, 0.04]))
    sp = dump(path, "/body00/speed_slae", tmp_path).reshape(5, 3)
    assert np.array_equal(sp[:, 2], np.float32([0.0, 0.01, 0.02, 0.07, 0.08]))
    # without profiles the profile datasets were cut at the restart row
    assert dump(path, "/body00/pressure", tmp_path).size == 3 * 64
    head = subprocess.run([H5DUMP, "-H", "-p", path], capture_output=True, text=True, check=True).stdout
    assert 'DATATYPE "string_t"' in head and 'ATTRIBUTE "caption"' in head and 'ATTRIBUTE "git_rev"' in head
    blk = head[head.index('DATASET "force_holder"'):][:900]
    assert "H5T_IEEE_F32LE" in blk and "H5S_UNLIMITED" in blk and "CHUNKED ( 128, 3 )" in blk
    assert "DEFLATE { LEVEL 9 }" in blk and "nan" in blk.lower()
    for g in ("body00", "body01"):
        assert f'GROUP "{g}"' in head
    for dset in ("force_hydro", "force_holder", "force_friction", "holder_position", "delta_position", "speed_slae"):
        assert f'DATASET "{dset}"' in head


def test_stepdata_bad_arguments(tmp_path):
    with pytest.raises(RuntimeError):
        hdf.Stepdata(str(tmp_path / "no_such_dir" / "s.h5"), "x", 0.0, [4])
    with hdf.Stepdata(str(tmp_path / "s.h5"), "x", 0.0, [4], profile=True) as sd:
        with pytest.raises(RuntimeError):
            sd.write(0.0, step_bodies(0, 1))   # profiles without pressure / friction
