"""vvflow_amd -- MI355X (gfx950) implementation of libvvhd's per-step particle
hot path, driven through the C ABI of ``include/vvhd_gpu.h``.

``Context`` owns one device context (vortex list, bodies, tree) and mirrors
the reference modules' calls:

=====================================  ======================================
reference (vvflow/vvflow libvvhd)      Context method
=====================================  ======================================
stree::build (TSortedTree.cpp:232)     tree_build(far_criteria, min, max)
MEpsilonFast::CalcEpsilonFast (:11)    epsilon(merge)
MConvectiveFast::process_all_lists     convective(inf, kahan)
MConvectiveFast::velocity (:20)        velocity(points, inf)
MDiffusiveFast::process_vort_list      diffusive()
MFlowmove::move_and_clean / shed       move_and_clean(remove) / vortex_shed()
MConvectiveFast::calc_circulation      calc_circulation(inf)
Space::calc_forces / zero_forces       calc_forces() / zero_forces()
=====================================  ======================================

The library is built in-tree by ``__graft_entry__.build()`` (hipcc,
--offload-arch=gfx950).  There is no fallback: if the library or a GPU is
missing, constructing a Context raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvvhd_gpu.so")

VVHD_OK, VVHD_EINVAL, VVHD_ESTATE, VVHD_EHIP, VVHD_ENOMEM, VVHD_ESINGULAR, VVHD_EUNSUPPORTED = range(7)
FAR_TRAINABLE, FAR_PARAMS = 5122, 5148   # VVHD_FAR_TRAINABLE, VVHD_FAR_PARAMS
NOALLOC_SCRATCH_BYTES = 64 << 20          # VVHD_NOALLOC_SCRATCH_BYTES
FIELD_VX, FIELD_VY, FIELD_VORTICITY, FIELD_STREAMFUNCTION, FIELD_PRESSURE = range(5)   # VVHD_FIELD_*
LIST_VORTEX, LIST_HEAT, LIST_STREAK, LIST_STREAK_SOURCE = range(4)                        # VVHD_LIST_*
HC_NEGLECT, HC_ISOLATE, HC_CONST_T, HC_CONST_W = range(4)                                 # VVHD_HC_*

SEG_DTYPE = np.dtype([(n, np.float64) for n in
                      ("x", "y", "g", "inv_eps", "cx", "cy", "dlx", "dly",
                       "gsum", "hsum", "fric", "Cp", "Fr", "Nu")]
                     + [("slip", np.uint32), ("eq_no", np.uint32)])
BODY_DTYPE = np.dtype([
    ("holder", np.float64, 3), ("dpos", np.float64, 3),
    ("speed_slae", np.float64, 3), ("speed_slae_prev", np.float64, 3),
    ("force_hydro", np.float64, 3), ("force_holder", np.float64, 3),
    ("friction", np.float64, 3), ("friction_prev", np.float64, 3),
    ("fdt_dead", np.float64, 3), ("g_dead", np.float64), ("nusselt", np.float64),
    ("slen", np.float64), ("area", np.float64), ("cofm", np.float64, 2),
    ("moi_cofm", np.float64), ("first_seg", np.int32), ("nseg", np.int32),
    ("special_segment_no", np.int32), ("kutta", np.int32)])

# every entry point of include/vvhd_gpu.h with its ctypes signature
_P, _I, _I32, _I64, _U64, _D = C.c_void_p, C.c_int, C.c_int32, C.c_int64, C.c_uint64, C.c_double
_VP = C.c_void_p
ABI = {
    "vvhd_create": (_I, [_I, C.POINTER(_P)]),
    "vvhd_destroy": (None, [_P]),
    "vvhd_last_error": (C.c_char_p, [_P]),
    "vvhd_stream": (_P, [_P]),
    "vvhd_set_params": (_I, [_P, _D, _D, _D, _D]),
    "vvhd_upload_vortices": (_I, [_P, _VP, _I64]),
    "vvhd_download_vortices": (_I, [_P, _VP, _I64]),
    "vvhd_num_vortices": (_I64, [_P]),
    "vvhd_snapshot": (_I, [_P]),
    "vvhd_restore": (_I, [_P]),
    "vvhd_upload_bodies": (_I, [_P, _VP, _I32, _VP, _I32]),
    "vvhd_set_batch": (_I, [_P, _I32, _VP, _VP, _VP]),
    "vvhd_batch_offsets": (_I, [_P, _VP, _VP]),
    "vvhd_set_tuning": (_I, [_P, _I, _I]),
    "vvhd_set_graph": (_I, [_P, _I]),
    "vvhd_tuning": (_I, [_P, _VP]),
    "vvhd_download_bodies": (_I, [_P, _VP, _I32, _VP, _I32]),
    "vvhd_tree_build": (_I, [_P, _I, _D, _D, _VP]),
    "vvhd_tree_destroy": (_I, [_P]),
    "vvhd_tree_leaves": (_I, [_P, _VP, _VP, _VP, _VP]),
    "vvhd_tree_lists": (_I, [_P, _I, _VP, _VP, _VP, _VP, _VP]),
    "vvhd_tree_find": (_I, [_P, _VP, _I64, _VP]),
    "vvhd_tree_nodes": (_I, [_P, _VP, _VP, _VP, _VP, _VP]),
    "vvhd_epsilon": (_I, [_P, _I, _VP, _VP, _VP]),
    "vvhd_convective": (_I, [_P, _D, _D, _I]),
    "vvhd_convective_noalloc": (_I, [_P, _D, _D, _I]),
    "vvhd_velocity_points": (_I, [_P, _VP, _I64, _D, _D, _VP]),
    "vvhd_diffusive": (_I, [_P]),
    "vvhd_move_and_clean": (_I, [_P, _I, _VP]),
    "vvhd_vortex_shed": (_I, [_P]),
    "vvhd_calc_circulation": (_I, [_P, _D, _D]),
    "vvhd_slae_system": (_I, [_P, _VP, _VP, _VP]),
    "vvhd_calc_forces": (_I, [_P]),
    "vvhd_zero_forces": (_I, [_P]),
    "vvhd_zero_profiles": (_I, [_P]),
    "vvhd_far_model_set": (_I, [_P, _VP, _I32]),
    "vvhd_far_model_get": (_I, [_P, _VP, _I32]),
    "vvhd_convective_learned": (_I, [_P, _D, _D, _I]),
    "vvhd_far_moments": (_I, [_P, _VP, _VP]),
    "vvhd_far_sample": (_I, [_P, _D, _U64, _I64, _VP, _VP, _VP, _VP, _VP, _VP]),
    "vvhd_far_forward": (_I, [_P, _VP, _I64, _VP]),
    "vvhd_far_gradients": (_I, [_P, _VP, _VP, _VP, _I64, _VP, _VP]),
    "vvhd_far_train_data": (_I, [_P, _VP, _VP, _VP, _I64]),
    "vvhd_far_train": (_I, [_P, _I32, _I32, _VP, _U64, _VP]),
    "vvhd_eps2h": (_I, [_P, _VP, _VP, _I64, _VP]),
    "vvhd_h2": (_I, [_P, _VP, _VP, _I64, _VP]),
    "vvhd_field": (_I, [_P, _I, _VP, _D, _D, _D, _D, _I, _D, _D, _VP]),
    "vvhd_integrals": (_I, [_P, _VP]),
    "vvhd_direct_sum": (_I, [_P, _VP, _I64, _VP]),
    "vvhd_phase_times": (_I, [_P, _VP]),
    "vvhd_set_overlap": (_I, [_P, _I32]),
    "vvhd_set_far_gradient": (_I, [_P, _I32]),
    "vvhd_work_counts": (_I, [_P, _VP]),
    "vvhd_gen_wake": (_I, [_I64, _U64, _VP]),
    "vvhd_gen_cylinder": (_I, [_D, _I32, _VP]),
    "vvhd_upload_list": (_I, [_P, _I, _VP, _I64]),
    "vvhd_download_list": (_I, [_P, _I, _VP, _I64]),
    "vvhd_list_size": (_I64, [_P, _I]),
    "vvhd_set_pr": (_I, [_P, _D]),
    "vvhd_set_heat": (_I, [_P, _VP, _VP]),
    "vvhd_heat_layer": (_I, [_P, _VP, _I32]),
    "vvhd_set_heat_layer": (_I, [_P, _VP, _I32]),
    "vvhd_diffusive_heat": (_I, [_P]),
    "vvhd_heat_shed": (_I, [_P]),
    "vvhd_heat_crop": (_I, [_P, _D]),
    "vvhd_streak_shed": (_I, [_P]),
    "vvhd_tree_leaf_lists": (_I, [_P, _VP]),
}

_lib = None


class VVHDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"vvhd error {code}: {msg}")
        self.code = code


def lib():
    """Loads libvvhd_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in ABI.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def gen_wake(n, seed=0x5EED0001):
    """Synthetic Karman-street wake (TObj records, n x 6)."""
    out = np.zeros((n, 6))
    rc = lib().vvhd_gen_wake(n, seed, _ptr(out))
    if rc:
        raise VVHDError(rc, "gen_wake")
    return out


def gen_cylinder(radius=0.5, n=500):
    """Corners of gen_cylinder{R, N} (n x 2)."""
    out = np.zeros((n, 2))
    rc = lib().vvhd_gen_cylinder(radius, n, _ptr(out))
    if rc:
        raise VVHDError(rc, "gen_cylinder")
    return out


def segments_from_corners(corners, slip=None, body_first=0):
    """Segment records of one body (TBody::doUpdateSegments, TBody.cpp:200-215)."""
    corners = np.asarray(corners, dtype=np.float64).reshape(-1, 2)
    n = len(corners)
    s = np.zeros(n, dtype=SEG_DTYPE)
    nxt = np.roll(corners, -1, axis=0)
    s["cx"], s["cy"] = corners[:, 0], corners[:, 1]
    s["dlx"] = nxt[:, 0] - corners[:, 0]
    s["dly"] = nxt[:, 1] - corners[:, 1]
    s["inv_eps"] = 3.0 / np.sqrt(s["dlx"] * s["dlx"] + s["dly"] * s["dly"])
    s["x"] = (nxt[:, 0] + corners[:, 0]) * 0.5
    s["y"] = (nxt[:, 1] + corners[:, 1]) * 0.5
    if slip is not None:
        s["slip"] = slip
    s["eq_no"] = body_first + np.arange(n)
    return s


class Context:
    """One device context of libvvhd_gpu (see module docstring)."""

    def __init__(self, device=0):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.vvhd_create(device, C.byref(h))
        if rc:
            raise VVHDError(rc, "vvhd_create failed (no usable GPU?)")
        self.h = h
        self.nseg = 0
        self.nbody = 0

    def close(self):
        if getattr(self, "h", None):
            self.L.vvhd_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc:
            raise VVHDError(rc, self.L.vvhd_last_error(self.h).decode())

    # ---- Space
    def set_params(self, re=float("inf"), dt=1.0, inf_g=0.0, time=0.0):
        self._check(self.L.vvhd_set_params(self.h, re, dt, inf_g, time))

    def upload(self, objs):
        objs = np.ascontiguousarray(objs, dtype=np.float64).reshape(-1, 6)
        self._check(self.L.vvhd_upload_vortices(self.h, _ptr(objs), objs.shape[0]))

    def set_far_gradient(self, network):
        """Learned far field: differentiate the correction through the network
        (True, default) or hold its output fixed over a leaf (False)."""
        self._check(self.L.vvhd_set_far_gradient(self.h, int(bool(network))))

    def set_overlap(self, enable):
        """Concurrent kernels within a step (default on); off, each kernel
        runs alone (bit-identical results), for per-kernel measurement."""
        self._check(self.L.vvhd_set_overlap(self.h, int(bool(enable))))

    def set_batch(self, vortex_off, body_off, re):
        """Declares the uploaded vortices and bodies as len(re) independent
        simulations: simulation s owns vortices vortex_off[s]:vortex_off[s+1]
        and bodies body_off[s]:body_off[s+1], Reynolds number re[s]."""
        vo = np.ascontiguousarray(vortex_off, dtype=np.int64)
        bo = np.ascontiguousarray(body_off, dtype=np.int32)
        r = np.ascontiguousarray(re, dtype=np.float64)
        self._check(self.L.vvhd_set_batch(self.h, len(r), _ptr(vo), _ptr(bo), _ptr(r)))

    def set_tuning(self, autotune=True, near_form=0):
        """Autotuned near-field form (default), or a fixed one: 0 one wave per
        leaf, 8/16/32 persistent waves per compute unit."""
        self._check(self.L.vvhd_set_tuning(self.h, int(autotune), int(near_form)))

    def set_graph(self, enable):
        """Run the plain convective phase as a captured HIP graph (persistent form)."""
        self._check(self.L.vvhd_set_graph(self.h, int(enable)))

    def tuning(self):
        """{near_form, classes_tuned, graph_captures, graph_launches}."""
        out = np.zeros(4, dtype=np.int64)
        self._check(self.L.vvhd_tuning(self.h, _ptr(out)))
        return dict(zip(("near_form", "classes_tuned", "graph_captures", "graph_launches"), out.tolist()))

    def batch_offsets(self):
        """Current vortex ranges of the simulations sharing the context."""
        ns = C.c_int32(0)
        self._check(self.L.vvhd_batch_offsets(self.h, C.addressof(ns), None))
        out = np.zeros(ns.value + 1, dtype=np.int64)
        self._check(self.L.vvhd_batch_offsets(self.h, C.addressof(ns), _ptr(out)))
        return out

    @property
    def n(self):
        return self.L.vvhd_num_vortices(self.h)

    def download(self):
        n = self.n
        out = np.zeros((n, 6))
        self._check(self.L.vvhd_download_vortices(self.h, _ptr(out), n))
        return out

    def snapshot(self):
        self._check(self.L.vvhd_snapshot(self.h))

    def restore(self):
        self._check(self.L.vvhd_restore(self.h))

    def upload_bodies(self, segs, bodies):
        segs = np.ascontiguousarray(segs, dtype=SEG_DTYPE)
        bodies = np.ascontiguousarray(bodies, dtype=BODY_DTYPE)
        self._check(self.L.vvhd_upload_bodies(self.h, _ptr(segs), len(segs), _ptr(bodies), len(bodies)))
        self.nseg, self.nbody = len(segs), len(bodies)

    def download_bodies(self):
        segs = np.zeros(self.nseg, dtype=SEG_DTYPE)
        bodies = np.zeros(self.nbody, dtype=BODY_DTYPE)
        self._check(self.L.vvhd_download_bodies(self.h, _ptr(segs), self.nseg, _ptr(bodies), self.nbody))
        return segs, bodies

    # ---- tree
    def tree_build(self, far_criteria=8, min_node=0.0, max_node=np.finfo(np.float64).max, perm=False):
        p = np.zeros(self.n, dtype=np.int32) if perm else None
        self._check(self.L.vvhd_tree_build(self.h, far_criteria, min_node, max_node, _ptr(p)))
        return p

    def tree_destroy(self):
        self._check(self.L.vvhd_tree_destroy(self.h))

    def tree_leaves(self):
        nl = C.c_int32(0)
        self._check(self.L.vvhd_tree_leaves(self.h, C.addressof(nl), None, None, None))
        n = nl.value
        geom = np.zeros((n, 4))
        ints = np.zeros((n, 4), dtype=np.int32)
        path = np.zeros(n, dtype=np.uint64)
        self._check(self.L.vvhd_tree_leaves(self.h, C.addressof(nl), _ptr(geom), _ptr(ints), _ptr(path)))
        return dict(geom=geom, ints=ints, path=path)

    def tree_lists(self, far=False):
        cnt = C.c_int64(0)
        self._check(self.L.vvhd_tree_lists(self.h, int(far), C.addressof(cnt), None, None, None, None))
        nl = len(self.tree_leaves()["path"])
        tot = cnt.value
        off = np.zeros(nl + 1, dtype=np.int64)
        depth = np.zeros(max(tot, 1), dtype=np.int32)
        path = np.zeros(max(tot, 1), dtype=np.uint64)
        cm = np.zeros((max(tot, 1), 6)) if far else None
        self._check(self.L.vvhd_tree_lists(self.h, int(far), C.addressof(cnt), _ptr(off), _ptr(depth),
                                           _ptr(path), _ptr(cm)))
        out = dict(off=off, depth=depth[:tot], path=path[:tot])
        if far:
            out["cm"] = cm[:tot]
        return out

    def tree_nodes(self):
        nn = C.c_int32(0)
        self._check(self.L.vvhd_tree_nodes(self.h, C.addressof(nn), None, None, None, None))
        n = nn.value
        geom = np.zeros((n, 4))
        ints = np.zeros((n, 4), dtype=np.int32)
        path = np.zeros(n, dtype=np.uint64)
        cm = np.zeros((n, 6))
        self._check(self.L.vvhd_tree_nodes(self.h, C.addressof(nn), _ptr(geom), _ptr(ints), _ptr(path), _ptr(cm)))
        return dict(geom=geom, ints=ints, path=path, cm=cm)

    def tree_find(self, pts):
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 2)
        out = np.zeros(len(pts), dtype=np.int32)
        self._check(self.L.vvhd_tree_find(self.h, _ptr(pts), len(pts), _ptr(out)))
        return out

    # ---- physics
    def epsilon(self, merge=True, want_nn=False):
        n = self.n
        nn1 = np.zeros(n, dtype=np.int32) if want_nn else None
        nn2 = np.zeros(n, dtype=np.int32) if want_nn else None
        merged = C.c_int32(0)
        self._check(self.L.vvhd_epsilon(self.h, int(merge), _ptr(nn1), _ptr(nn2), C.addressof(merged)))
        return (nn1, nn2, merged.value) if want_nn else merged.value

    def convective(self, inf=(0.0, 0.0), kahan=True):
        self._check(self.L.vvhd_convective(self.h, inf[0], inf[1], int(kahan)))

    def convective_noalloc(self, inf=(0.0, 0.0), kahan=True):
        """vvhd_convective with its scratch in the library's fixed device
        array (no device allocation)."""
        self._check(self.L.vvhd_convective_noalloc(self.h, inf[0], inf[1], int(kahan)))

    def velocity(self, pts, inf=(0.0, 0.0)):
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 2)
        out = np.zeros_like(pts)
        self._check(self.L.vvhd_velocity_points(self.h, _ptr(pts), len(pts), inf[0], inf[1], _ptr(out)))
        return out

    def diffusive(self):
        self._check(self.L.vvhd_diffusive(self.h))

    def move_and_clean(self, remove=True):
        cl = C.c_int64(0)
        self._check(self.L.vvhd_move_and_clean(self.h, int(remove), C.addressof(cl)))
        return cl.value

    def vortex_shed(self):
        self._check(self.L.vvhd_vortex_shed(self.h))

    # ---- heat and streak lists (LIST_HEAT, LIST_STREAK, LIST_STREAK_SOURCE)
    def upload_list(self, list_id, objs):
        objs = np.ascontiguousarray(objs, dtype=np.float64).reshape(-1, 6)
        self._check(self.L.vvhd_upload_list(self.h, list_id, _ptr(objs), objs.shape[0]))

    def download_list(self, list_id):
        n = self.L.vvhd_list_size(self.h, list_id)
        out = np.zeros((n, 6))
        self._check(self.L.vvhd_download_list(self.h, list_id, _ptr(out), n))
        return out

    def set_pr(self, pr):
        self._check(self.L.vvhd_set_pr(self.h, pr))

    def set_heat(self, conditions, heat_const):
        cond = np.ascontiguousarray(conditions, dtype=np.int32)
        hc = np.ascontiguousarray(heat_const, dtype=np.float32)
        self._check(self.L.vvhd_set_heat(self.h, _ptr(cond), _ptr(hc)))

    def heat_layer(self):
        out = np.zeros(self.nseg, dtype=np.int64)
        self._check(self.L.vvhd_heat_layer(self.h, _ptr(out), self.nseg))
        return out

    def set_heat_layer(self, v):
        v = np.ascontiguousarray(v, dtype=np.int64)
        self._check(self.L.vvhd_set_heat_layer(self.h, _ptr(v), len(v)))

    def diffusive_heat(self):
        self._check(self.L.vvhd_diffusive_heat(self.h))

    def heat_shed(self):
        self._check(self.L.vvhd_heat_shed(self.h))

    def heat_crop(self, scale=16.0):
        self._check(self.L.vvhd_heat_crop(self.h, scale))

    def streak_shed(self):
        self._check(self.L.vvhd_streak_shed(self.h))

    def tree_leaf_lists(self):
        """hRange and sRange of every leaf as list indices {h0, h1, s0, s1}."""
        nl = self.tree_leaves()["geom"].shape[0]
        out = np.zeros((max(nl, 1), 4), dtype=np.int32)
        self._check(self.L.vvhd_tree_leaf_lists(self.h, _ptr(out)))
        return out[:nl]

    def calc_circulation(self, inf=(0.0, 0.0)):
        self._check(self.L.vvhd_calc_circulation(self.h, inf[0], inf[1]))

    def slae_system(self):
        """(A row-major n x n, rhs) of the last filled boundary-condition system."""
        n = C.c_int32(0)
        self._check(self.L.vvhd_slae_system(self.h, C.addressof(n), None, None))
        A = np.empty((n.value, n.value), dtype=np.float64)
        b = np.empty(n.value, dtype=np.float64)
        self._check(self.L.vvhd_slae_system(self.h, C.addressof(n), _ptr(A), _ptr(b)))
        return A, b

    def calc_forces(self):
        self._check(self.L.vvhd_calc_forces(self.h))

    def zero_forces(self):
        self._check(self.L.vvhd_zero_forces(self.h))

    def zero_profiles(self):
        """Resets every segment's Cp, Fr and Nu accumulators (Stepdata::write
        after writing the profiles)."""
        self._check(self.L.vvhd_zero_profiles(self.h))

    # ---- learned far field
    def far_model_set(self, params):
        params = np.ascontiguousarray(params, dtype=np.float32)
        self._check(self.L.vvhd_far_model_set(self.h, _ptr(params), len(params)))

    def far_model_get(self):
        out = np.zeros(FAR_PARAMS, dtype=np.float32)
        self._check(self.L.vvhd_far_model_get(self.h, _ptr(out), FAR_PARAMS))
        return out

    def convective_learned(self, inf=(0.0, 0.0), kahan=True):
        self._check(self.L.vvhd_convective_learned(self.h, inf[0], inf[1], int(kahan)))

    def far_moments(self):
        nn = C.c_int32(0)
        self._check(self.L.vvhd_far_moments(self.h, C.addressof(nn), None))
        out = np.zeros((12, nn.value))
        self._check(self.L.vvhd_far_moments(self.h, C.addressof(nn), _ptr(out)))
        return out

    def far_sample(self, rate, seed=1):
        """Training pairs of the built tree: dict of X (n,12) float32, Y (n,2),
        scale (n,), target (n,) tree-order vortex, node (n,) device node id."""
        npairs = C.c_int64(0)
        self._check(self.L.vvhd_far_sample(self.h, rate, seed, 0, C.addressof(npairs), None, None, None, None,
                                           None))
        n = npairs.value
        X = np.zeros((n, 12), dtype=np.float32)
        Y = np.zeros((n, 2), dtype=np.float32)
        sc = np.zeros(n, dtype=np.float32)
        tg = np.zeros(n, dtype=np.int32)
        nd = np.zeros(n, dtype=np.int32)
        self._check(self.L.vvhd_far_sample(self.h, rate, seed, n, C.addressof(npairs), _ptr(X), _ptr(Y), _ptr(sc),
                                           _ptr(tg), _ptr(nd)))
        return dict(X=X, Y=Y, scale=sc, target=tg, node=nd)

    def far_pair_count(self):
        """Number of (vortex with g != 0, far node of its leaf) pairs of the built tree."""
        npairs = C.c_int64(0)
        self._check(self.L.vvhd_far_sample(self.h, 1.0, 0, 0, C.addressof(npairs), None, None, None, None, None))
        return npairs.value

    def far_forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 12)
        Y = np.zeros((len(X), 2), dtype=np.float32)
        self._check(self.L.vvhd_far_forward(self.h, _ptr(X), len(X), _ptr(Y)))
        return Y

    def far_gradients(self, X, Y, W):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 12)
        Y = np.ascontiguousarray(Y, dtype=np.float32).reshape(-1, 2)
        W = np.ascontiguousarray(W, dtype=np.float32).reshape(-1)
        g = np.zeros(FAR_TRAINABLE, dtype=np.float32)
        loss = C.c_float(0)
        self._check(self.L.vvhd_far_gradients(self.h, _ptr(X), _ptr(Y), _ptr(W), len(X), _ptr(g),
                                              C.addressof(loss)))
        return loss.value, g

    def far_train_data(self, X, Y, W):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 12)
        Y = np.ascontiguousarray(Y, dtype=np.float32).reshape(-1, 2)
        W = np.ascontiguousarray(W, dtype=np.float32).reshape(-1)
        self._check(self.L.vvhd_far_train_data(self.h, _ptr(X), _ptr(Y), _ptr(W), len(X)))

    def far_train(self, lr, batch=8192, seed=1):
        lr = np.ascontiguousarray(lr, dtype=np.float32).reshape(-1)
        loss = np.zeros(len(lr), dtype=np.float32)
        self._check(self.L.vvhd_far_train(self.h, len(lr), batch, _ptr(lr), seed, _ptr(loss)))
        return loss

    def eps2h(self, pts=None, leaf=None):
        """MEpsilonFast::eps2h at points (over the near list of leaf, or of
        the leaf findNode gives) or, with pts None, at every vortex with its
        own leaf."""
        return self._eps2h_h2(self.L.vvhd_eps2h, pts, leaf)

    def h2(self, pts=None, leaf=None):
        """MEpsilonFast::h2 at points or at every vortex."""
        return self._eps2h_h2(self.L.vvhd_h2, pts, leaf)

    def _eps2h_h2(self, fn, pts, leaf):
        if pts is None:
            out = np.zeros(self.n)
            self._check(fn(self.h, None, None, len(out), _ptr(out)))
            return out
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 2)
        lf = None if leaf is None else np.ascontiguousarray(leaf, dtype=np.int32).reshape(-1)
        out = np.zeros(len(pts))
        self._check(fn(self.h, _ptr(pts), _ptr(lf), len(pts), _ptr(out)))
        return out

    def field(self, kind, xmin, ymin, dxdy, xres, yres, eps_mult=1.0, dl=0.0, inf=(0.0, 0.0), ref=None):
        """vvhd_field: kind one of FIELD_*; ref None is ref_frame 's' (pressure)
        or a zero reference; returns the (yres, xres) float map."""
        g = (C.c_double * 3)(xmin, ymin, dxdy)
        grid = (C.c_byte * 32)()
        C.memmove(grid, g, 24)
        C.memmove(C.addressof(grid) + 24, (C.c_int32 * 2)(xres, yres), 8)
        out = np.zeros((yres, xres), dtype=np.float32)
        r = (0.0, 0.0) if ref is None else ref
        self._check(self.L.vvhd_field(self.h, int(kind), C.addressof(grid), eps_mult, dl, inf[0], inf[1],
                                      0 if ref is None else 1, r[0], r[1], _ptr(out)))
        return out

    def integrals(self):
        out = np.zeros(5)
        self._check(self.L.vvhd_integrals(self.h, _ptr(out)))
        return out

    def direct_sum(self, targets=None):
        if targets is None:
            nt, tp = self.n, None
        else:
            targets = np.ascontiguousarray(targets, dtype=np.int64)
            nt, tp = len(targets), targets
        out = np.zeros((nt, 2))
        self._check(self.L.vvhd_direct_sum(self.h, _ptr(tp), nt, _ptr(out)))
        return out

    def phase_times(self):
        """{tree, epsilon, convective, diffusive, move_and_clean, slae,
        near-field kernel, reserved} in ms (-1: not run)."""
        out = np.zeros(8, dtype=np.float32)
        self._check(self.L.vvhd_phase_times(self.h, _ptr(out)))
        return out

    def work_counts(self):
        out = np.zeros(7, dtype=np.int64)
        self._check(self.L.vvhd_work_counts(self.h, _ptr(out)))
        return out
