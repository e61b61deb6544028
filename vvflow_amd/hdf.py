"""hdf.py -- the reference's HDF5 formats through libvvhd_hdf.so
(include/vvhd_hdf.h): snapshots (Space::save_hdf / load_hdf,
libvvhd/src/TSpace_save_hdf.cpp, TSpace_load_hdf.cpp) and the per-step time
series (Stepdata, libvvhd/src/MStepdata.cpp).  Host code: no device is
needed.

    save(path, space, lists, bodies)
    space, lists, bodies = load(path)
    with Stepdata(path, caption, time, nseg=[500]) as sd:
        sd.write(time, bodies)      # one row per step

space: dict of the header fields of vvhd_hdf_space (caption, time, dt,
dt_save, dt_streak, dt_profile, re, pr, inf_g, finish, inf_marker, gravity,
inf_speed_x, inf_speed_y); lists: dict name -> (n, 6) TObj array for
"vort", "heat", "ink", "ink_source", "source" (missing names are empty);
bodies: list of dicts with the vvhd_hdf_body fields plus "seg4" ((nseg, 4):
corner x, corner y, g, gsum) and "slip" ((nseg,) uint32).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvvhd_hdf.so")
LISTS = ("vort", "heat", "ink", "ink_source", "source")
STR = 256

_V3 = C.c_double * 3


class Space(C.Structure):
    _fields_ = [("caption", C.c_char * STR), ("time", C.c_double), ("dt", C.c_double), ("dt_save", C.c_double),
                ("dt_streak", C.c_double), ("dt_profile", C.c_double), ("re", C.c_double), ("pr", C.c_double),
                ("inf_g", C.c_double), ("finish", C.c_double), ("inf_marker", C.c_double * 2),
                ("gravity", C.c_double * 2), ("inf_speed_x", C.c_char * STR), ("inf_speed_y", C.c_char * STR),
                ("frac_value", C.c_int32 * 5), ("frac_scale", C.c_uint32 * 5)]


class Body(C.Structure):
    _fields_ = [("label", C.c_char * STR), ("root_body", C.c_char * STR), ("speed_x", C.c_char * STR),
                ("speed_y", C.c_char * STR), ("speed_o", C.c_char * STR), ("force_o", C.c_char * STR),
                ("holder", _V3), ("dpos", _V3), ("speed_slae", _V3), ("speed_slae_prev", _V3), ("kspring", _V3),
                ("damping", _V3), ("force_hydro", _V3), ("force_holder", _V3), ("friction_prev", _V3),
                ("fdt_dead", _V3), ("collision_min", _V3), ("collision_max", _V3), ("density", C.c_double),
                ("g_dead", C.c_double), ("bounce", C.c_double), ("area", C.c_double), ("cofm", C.c_double * 2),
                ("moi_cofm", C.c_double), ("moi_axis", C.c_double), ("heat_const", C.c_double),
                ("special_segment_no", C.c_int32), ("boundary_condition", C.c_uint32),
                ("heat_condition", C.c_uint32), ("nseg", C.c_int32)]


class StepBody(C.Structure):
    _fields_ = [("force_hydro", _V3), ("force_holder", _V3), ("friction", _V3), ("holder", _V3), ("dpos", _V3),
                ("speed_slae", _V3)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.vvhd_hdf_save.restype = C.c_int
        L.vvhd_hdf_save.argtypes = [C.c_char_p, C.POINTER(Space), C.POINTER(C.c_void_p), C.POINTER(C.c_int64),
                                    C.c_int32, C.POINTER(Body), C.c_void_p, C.c_void_p]
        L.vvhd_hdf_sizes.restype = C.c_int
        L.vvhd_hdf_sizes.argtypes = [C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.vvhd_hdf_load.restype = C.c_int
        L.vvhd_hdf_load.argtypes = [C.c_char_p, C.POINTER(Space), C.POINTER(C.c_void_p), C.POINTER(Body),
                                    C.c_void_p, C.c_void_p]
        L.vvhd_hdf_error.restype = C.c_char_p
        L.vvhd_hdf_error.argtypes = []
        L.vvhd_stepdata_open.restype = C.c_void_p
        L.vvhd_stepdata_open.argtypes = [C.c_char_p, C.c_char_p, C.c_double, C.c_int32, C.POINTER(C.c_int32),
                                         C.c_int32]
        L.vvhd_stepdata_write.restype = C.c_int
        L.vvhd_stepdata_write.argtypes = [C.c_void_p, C.c_double, C.POINTER(StepBody), C.c_void_p, C.c_void_p]
        L.vvhd_stepdata_flush.restype = C.c_int
        L.vvhd_stepdata_flush.argtypes = [C.c_void_p]
        L.vvhd_stepdata_close.restype = C.c_int
        L.vvhd_stepdata_close.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().vvhd_hdf_error().decode())


def _set(struct, d):
    for name, _ in struct._fields_:
        if name not in d:
            continue
        v = d[name]
        cur = getattr(struct, name)
        if isinstance(v, str):
            setattr(struct, name, v.encode())
        elif isinstance(cur, C.Array):
            for k, x in enumerate(v):
                cur[k] = x
        else:
            setattr(struct, name, v)


def _get(struct):
    out = {}
    for name, _ in struct._fields_:
        v = getattr(struct, name)
        if isinstance(v, bytes):
            out[name] = v.decode()
        elif isinstance(v, C.Array):
            out[name] = list(v)
        else:
            out[name] = v
    return out


def save(path, space, lists=None, bodies=()):
    sp = Space()
    _set(sp, space)
    arrays = [np.ascontiguousarray((lists or {}).get(k, np.zeros((0, 6))), dtype=np.float64).reshape(-1, 6)
              for k in LISTS]
    ptrs = (C.c_void_p * 5)(*[a.ctypes.data for a in arrays])
    counts = (C.c_int64 * 5)(*[len(a) for a in arrays])
    nb = len(bodies)
    bs = (Body * max(nb, 1))()
    seg4 = np.concatenate([np.asarray(b["seg4"], dtype=np.float64).reshape(-1, 4) for b in bodies]) if nb else \
        np.zeros((0, 4))
    slip = np.concatenate([np.asarray(b.get("slip", np.zeros(len(b["seg4"]))), dtype=np.uint32) for b in bodies]) \
        if nb else np.zeros(0, np.uint32)
    seg4 = np.ascontiguousarray(seg4)
    slip = np.ascontiguousarray(slip, dtype=np.uint32)
    for k, b in enumerate(bodies):
        _set(bs[k], {**b, "nseg": len(b["seg4"])})
    _check(lib().vvhd_hdf_save(os.fsencode(path), C.byref(sp), ptrs, counts, nb, bs, seg4.ctypes.data,
                               slip.ctypes.data))


def load(path):
    counts = (C.c_int64 * 5)()
    nb, ns = C.c_int32(0), C.c_int32(0)
    _check(lib().vvhd_hdf_sizes(os.fsencode(path), counts, C.byref(nb), C.byref(ns)))
    arrays = [np.zeros((max(counts[k], 1), 6)) for k in range(5)]
    ptrs = (C.c_void_p * 5)(*[a.ctypes.data for a in arrays])
    bs = (Body * max(nb.value, 1))()
    seg4 = np.zeros((max(ns.value, 1), 4))
    slip = np.zeros(max(ns.value, 1), np.uint32)
    sp = Space()
    _check(lib().vvhd_hdf_load(os.fsencode(path), C.byref(sp), ptrs, bs, seg4.ctypes.data, slip.ctypes.data))
    lists = {k: arrays[i][:counts[i]] for i, k in enumerate(LISTS)}
    bodies, s = [], 0
    for k in range(nb.value):
        b = _get(bs[k])
        n = b["nseg"]
        b["seg4"] = seg4[s:s + n].copy()
        b["slip"] = slip[s:s + n].copy()
        s += n
        bodies.append(b)
    return _get(sp), lists, bodies


class Stepdata:
    """The reference's per-step time series (Stepdata, MStepdata.cpp): one
    row per write of the time and of every body's forces, holder position,
    spring displacement and SLAE speed, and with profile=True the
    per-segment pressure and friction.  Opening an existing file keeps the
    rows before `time` (a restart).  bodies: dicts with force_hydro,
    force_holder, friction, holder, dpos, speed_slae (3 numbers each);
    pressure / friction: one array of all bodies' segments."""

    def __init__(self, path, caption="", time=0.0, nseg=(), profile=False):
        self.nb = len(nseg)
        ns = (C.c_int32 * max(self.nb, 1))(*nseg)
        self.h = lib().vvhd_stepdata_open(os.fsencode(path), caption.encode(), float(time), self.nb, ns, int(profile))
        if not self.h:
            raise RuntimeError(lib().vvhd_hdf_error().decode())

    def write(self, time, bodies, pressure=None, friction=None):
        bs = (StepBody * max(self.nb, 1))()
        for k, b in enumerate(bodies):
            _set(bs[k], b)
        p = None if pressure is None else np.ascontiguousarray(pressure, dtype=np.float64)
        f = None if friction is None else np.ascontiguousarray(friction, dtype=np.float64)
        _check(lib().vvhd_stepdata_write(self.h, float(time), bs, None if p is None else p.ctypes.data,
                                         None if f is None else f.ctypes.data))

    def flush(self):
        lib().vvhd_stepdata_flush(self.h)

    def close(self):
        if self.h:
            lib().vvhd_stepdata_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
