"""ctypes bindings of the checker libraries (test infrastructure only).

* ``Oracle`` wraps oracle/libvvhd_oracle.so, the FP64 restatement of the
  reference's hot path.
* ``RefHarness`` wraps oracle/_ref/libvvhd_ref.so, the reference's own
  TSortedTree.cpp / MEpsilonFast.cpp / MDiffusiveFast.cpp (built only where
  the reference sources are present; tests that need it skip otherwise).
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "libvvhd_oracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libvvhd_ref.so")

P = C.c_void_p
I32 = C.c_int32
I64 = C.c_int64
F64 = C.c_double
dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")

SEG_DTYPE = np.dtype([(n, np.float64) for n in
                      ("x", "y", "g", "inv_eps", "cx", "cy", "dlx", "dly",
                       "gsum", "hsum", "fric", "Cp", "Fr", "Nu")]
                     + [("slip", np.uint32), ("eq_no", np.uint32)])
CORNER_DTYPE = np.dtype([("x", np.float64), ("y", np.float64),
                         ("slip", np.uint32), ("pad", np.uint32)])
BODY_DTYPE = np.dtype([
    ("holder", np.float64, 3), ("dpos", np.float64, 3),
    ("speed_slae", np.float64, 3), ("speed_slae_prev", np.float64, 3),
    ("force_hydro", np.float64, 3), ("force_holder", np.float64, 3),
    ("friction", np.float64, 3), ("friction_prev", np.float64, 3),
    ("fdt_dead", np.float64, 3), ("g_dead", np.float64), ("nusselt", np.float64),
    ("slen", np.float64), ("area", np.float64), ("cofm", np.float64, 2),
    ("moi_cofm", np.float64), ("first_seg", np.int32), ("nseg", np.int32),
    ("special_segment_no", np.int32), ("kutta", np.int32)])


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


class Oracle:
    """One oracle context (a Space plus its tree)."""

    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            if not os.path.exists(ORACLE_SO):
                raise RuntimeError("oracle library missing: run __graft_entry__.build()")
            L = C.CDLL(ORACLE_SO)
            _sig(L, "vvo_create", P, [])
            _sig(L, "vvo_destroy", None, [P])
            _sig(L, "vvo_last_error", C.c_char_p, [P])
            _sig(L, "vvo_set_params", None, [P, F64, F64, F64, F64, F64, F64])
            _sig(L, "vvo_set_vortices", None, [P, dp, I64])
            _sig(L, "vvo_num_vortices", I64, [P])
            _sig(L, "vvo_get_vortices", None, [P, dp])
            _sig(L, "vvo_add_body", C.c_int, [P, C.c_void_p, I32])
            _sig(L, "vvo_num_bodies", I32, [P])
            _sig(L, "vvo_num_segments", I32, [P])
            _sig(L, "vvo_get_segments", None, [P, C.c_void_p])
            _sig(L, "vvo_set_segments", None, [P, C.c_void_p])
            _sig(L, "vvo_get_body", None, [P, I32, C.c_void_p])
            _sig(L, "vvo_set_body", None, [P, I32, C.c_void_p])
            _sig(L, "vvo_point_invalid", I32, [P, I32, F64, F64])
            _sig(L, "vvo_tree_build", C.c_int, [P, C.c_int, F64, F64])
            _sig(L, "vvo_tree_destroy", None, [P])
            _sig(L, "vvo_tree_num_nodes", I32, [P])
            _sig(L, "vvo_tree_num_leaves", I32, [P])
            _sig(L, "vvo_tree_nodes", None, [P, dp, ip, dp, dp])
            _sig(L, "vvo_tree_leaves", None, [P, ip])
            _sig(L, "vvo_tree_list_size", I64, [P, C.c_int])
            _sig(L, "vvo_tree_lists", None, [P, C.c_int, lp, ip])
            _sig(L, "vvo_tree_leaf_segs", None, [P, lp, ip])
            _sig(L, "vvo_tree_find", I32, [P, F64, F64])
            _sig(L, "vvo_epsilon", C.c_int, [P, C.c_int, C.c_void_p, C.c_void_p,
                                             C.c_void_p, I32, C.POINTER(I32)])
            _sig(L, "vvo_epsilon_leaves", C.c_int, [P, C.c_int, I64, I64, C.c_void_p])
            _sig(L, "vvo_eps2h_h2", C.c_int, [P, C.c_int, C.c_void_p, I64, dp])
            _sig(L, "vvo_field", C.c_int, [P, C.c_int, F64, F64, F64, C.c_int, C.c_int, F64, F64, C.c_int, F64, F64,
                                           C.c_void_p])
            _sig(L, "vvo_convective", C.c_int, [P, C.c_int, I64, I64])
            _sig(L, "vvo_velocity", C.c_int, [P, dp, I64, dp])
            _sig(L, "vvo_diffusive", C.c_int, [P, C.c_int, I64, I64])
            _sig(L, "vvo_direct_sum", C.c_int, [dp, I64, C.c_void_p, I64, dp, C.c_int])
            _sig(L, "vvo_move_and_clean", C.c_int, [P, C.c_int, C.POINTER(I64)])
            _sig(L, "vvo_vortex_shed", C.c_int, [P])
            _sig(L, "vvo_slae_size", C.c_int, [P])
            _sig(L, "vvo_calc_circulation", C.c_int, [P, C.c_void_p, C.c_void_p])
            _sig(L, "vvo_calc_forces", None, [P])
            _sig(L, "vvo_zero_forces", None, [P])
            _sig(L, "vvo_integrals", None, [P, dp])
            _sig(L, "vvo_far_moments", I32, [P, C.c_void_p])
            _sig(L, "vvo_far_pairs", C.c_int, [P, ip, ip, I64, dp, dp, dp, C.c_void_p, C.c_void_p])
            _sig(L, "vvo_far_point_features", C.c_int, [P, dp, dp, ip, I64, dp, dp, dp])
            _sig(L, "vvo_set_list", C.c_int, [P, C.c_int, C.c_void_p, I64])
            _sig(L, "vvo_list_size", I64, [P, C.c_int])
            _sig(L, "vvo_get_list", C.c_int, [P, C.c_int, C.c_void_p])
            _sig(L, "vvo_set_pr", None, [P, F64])
            _sig(L, "vvo_set_heat", C.c_int, [P, C.c_void_p, C.c_void_p])
            _sig(L, "vvo_get_heat_layer", None, [P, C.c_void_p])
            _sig(L, "vvo_set_heat_layer", None, [P, C.c_void_p])
            _sig(L, "vvo_heat_shed", C.c_int, [P])
            _sig(L, "vvo_heat_crop", C.c_int, [P, F64])
            _sig(L, "vvo_streak_shed", C.c_int, [P])
            _sig(L, "vvo_diffusive_heat", C.c_int, [P, C.c_int])
            _sig(L, "vvo_tree_leaf_hs", None, [P, ip])
            cls._lib = L
        return cls._lib

    def __init__(self):
        self.L = self.lib()
        self.h = self.L.vvo_create()

    def __del__(self):
        if getattr(self, "h", None):
            self.L.vvo_destroy(self.h)
            self.h = None

    def _check(self, rc):
        if rc:
            raise RuntimeError(self.L.vvo_last_error(self.h).decode())

    # ---- Space
    def set_params(self, re=float("inf"), dt=1.0, inf=(0.0, 0.0), inf_g=0.0, time=0.0):
        self.L.vvo_set_params(self.h, re, dt, inf[0], inf[1], inf_g, time)

    def set_vortices(self, objs):
        objs = np.ascontiguousarray(objs, dtype=np.float64).reshape(-1, 6)
        self.L.vvo_set_vortices(self.h, objs, objs.shape[0])

    def vortices(self):
        n = self.L.vvo_num_vortices(self.h)
        out = np.zeros((n, 6))
        if n:
            self.L.vvo_get_vortices(self.h, out)
        return out

    def add_body(self, corners_xy, slip=None):
        corners_xy = np.asarray(corners_xy, dtype=np.float64).reshape(-1, 2)
        c = np.zeros(len(corners_xy), dtype=CORNER_DTYPE)
        c["x"] = corners_xy[:, 0]
        c["y"] = corners_xy[:, 1]
        if slip is not None:
            c["slip"] = slip
        self._check(self.L.vvo_add_body(self.h, c.ctypes.data, len(c)))

    def segments(self):
        n = self.L.vvo_num_segments(self.h)
        out = np.zeros(n, dtype=SEG_DTYPE)
        if n:
            self.L.vvo_get_segments(self.h, out.ctypes.data)
        return out

    def set_segments(self, segs):
        segs = np.ascontiguousarray(segs, dtype=SEG_DTYPE)
        self.L.vvo_set_segments(self.h, segs.ctypes.data)

    def body(self, b=0):
        out = np.zeros(1, dtype=BODY_DTYPE)
        self.L.vvo_get_body(self.h, b, out.ctypes.data)
        return out[0]

    def set_body(self, b, state):
        arr = np.ascontiguousarray(np.asarray(state, dtype=BODY_DTYPE).reshape(1))
        self.L.vvo_set_body(self.h, b, arr.ctypes.data)

    def point_invalid(self, body, x, y):
        return self.L.vvo_point_invalid(self.h, body, x, y)

    # ---- tree
    def tree_build(self, far_criteria=8, min_node=0.0, max_node=np.finfo(np.float64).max):
        self._check(self.L.vvo_tree_build(self.h, far_criteria, min_node, max_node))

    def tree_destroy(self):
        self.L.vvo_tree_destroy(self.h)

    def tree(self):
        """Returns a dict describing the tree (node table, leaves, lists)."""
        nn = self.L.vvo_tree_num_nodes(self.h)
        nl = self.L.vvo_tree_num_leaves(self.h)
        geom = np.zeros((nn, 4))
        ints = np.zeros((nn, 6), dtype=np.int32)
        cm = np.zeros((nn, 6))
        cme = np.zeros((nn, 2))
        self.L.vvo_tree_nodes(self.h, geom, ints, cm, cme)
        leaves = np.zeros(nl, dtype=np.int32)
        self.L.vvo_tree_leaves(self.h, leaves)
        out = dict(geom=geom, ints=ints, cm=cm, cm_inv_eps=cme, leaves=leaves)
        for far in (0, 1):
            sz = self.L.vvo_tree_list_size(self.h, far)
            off = np.zeros(nl + 1, dtype=np.int64)
            lst = np.zeros(max(sz, 1), dtype=np.int32)
            self.L.vvo_tree_lists(self.h, far, off, lst)
            out["far" if far else "near"] = (off, lst[:sz])
        off = np.zeros(nl + 1, dtype=np.int64)
        segs = np.zeros(max(int(ints[:, 5].sum()), 1), dtype=np.int32)
        self.L.vvo_tree_leaf_segs(self.h, off, segs)
        out["segs"] = (off, segs[:off[-1]])
        # path codes (the reference's snode::j) and depth (snode::i)
        path = np.zeros(nn, dtype=np.int64)
        for i in range(1, nn):
            p = ints[i, 3]
            side = 0 if ints[p, 2] == i else 1
            path[i] = path[p] * 2 + side
        out["path"] = path
        return out

    def find(self, x, y):
        return self.L.vvo_tree_find(self.h, x, y)

    # ---- physics
    def epsilon(self, merge=True, max_merges=1 << 20):
        n = self.L.vvo_num_vortices(self.h)
        nn1 = np.zeros(n, dtype=np.int32)
        nn2 = np.zeros(n, dtype=np.int32)
        mg = np.zeros(2 * max_merges, dtype=np.int32)
        merged = I32(0)
        self._check(self.L.vvo_epsilon(self.h, int(merge), nn1.ctypes.data, nn2.ctypes.data,
                                       mg.ctypes.data, max_merges, C.byref(merged)))
        m = merged.value
        return nn1, nn2, mg[:2 * min(m, max_merges)].reshape(-1, 2), m

    def epsilon_leaves(self, merge=True, leaf_begin=0, leaf_end=-1):
        merged = I32(0)
        self._check(self.L.vvo_epsilon_leaves(self.h, int(merge), leaf_begin, leaf_end, C.byref(merged)))
        return merged.value

    def eps2h(self, pts=None, h2=False):
        n = self.L.vvo_num_vortices(self.h) if pts is None else len(pts)
        out = np.zeros(n)
        xy = None if pts is None else np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 2)
        self._check(self.L.vvo_eps2h_h2(self.h, int(h2), None if xy is None else xy.ctypes.data, n, out))
        return out

    def field(self, kind, xmin, ymin, dxdy, xres, yres, eps_mult=1.0, dl=0.0, ref=None):
        out = np.zeros((yres, xres), dtype=np.float32)
        r = (0.0, 0.0) if ref is None else ref
        self._check(self.L.vvo_field(self.h, kind, xmin, ymin, dxdy, xres, yres, eps_mult, dl,
                                     0 if ref is None else 1, r[0], r[1], out.ctypes.data))
        return out

    def convective(self, threads=0, leaf_begin=0, leaf_end=-1):
        self._check(self.L.vvo_convective(self.h, threads, leaf_begin, leaf_end))

    def velocity(self, pts):
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 2)
        out = np.zeros_like(pts)
        self._check(self.L.vvo_velocity(self.h, pts, pts.shape[0], out))
        return out

    def diffusive(self, threads=0, leaf_begin=0, leaf_end=-1):
        self._check(self.L.vvo_diffusive(self.h, threads, leaf_begin, leaf_end))

    # ---- heat and streak lists: 1 HeatList, 2 StreakList, 3 StreakSourceList
    HEAT, STREAK, STREAK_SOURCE = 1, 2, 3

    def set_list(self, list_id, objs):
        objs = np.ascontiguousarray(objs, dtype=np.float64).reshape(-1, 6)
        self._check(self.L.vvo_set_list(self.h, list_id, objs.ctypes.data, objs.shape[0]))

    def get_list(self, list_id):
        n = self.L.vvo_list_size(self.h, list_id)
        out = np.zeros((n, 6))
        if n:
            self._check(self.L.vvo_get_list(self.h, list_id, out.ctypes.data))
        return out

    def set_pr(self, pr):
        self.L.vvo_set_pr(self.h, pr)

    def set_heat(self, conditions, heat_const):
        cond = np.ascontiguousarray(conditions, dtype=np.int32)
        hc = np.ascontiguousarray(heat_const, dtype=np.float32)
        self._check(self.L.vvo_set_heat(self.h, cond.ctypes.data, hc.ctypes.data))

    def heat_layer(self):
        out = np.zeros(self.L.vvo_num_segments(self.h), dtype=np.int64)
        self.L.vvo_get_heat_layer(self.h, out.ctypes.data)
        return out

    def set_heat_layer(self, v):
        v = np.ascontiguousarray(v, dtype=np.int64)
        self.L.vvo_set_heat_layer(self.h, v.ctypes.data)

    def heat_shed(self):
        self._check(self.L.vvo_heat_shed(self.h))

    def heat_crop(self, scale=16.0):
        self._check(self.L.vvo_heat_crop(self.h, scale))

    def streak_shed(self):
        self._check(self.L.vvo_streak_shed(self.h))

    def diffusive_heat(self, threads=0):
        self._check(self.L.vvo_diffusive_heat(self.h, threads))

    def leaf_hs(self):
        nl = self.L.vvo_tree_num_leaves(self.h)
        out = np.zeros((nl, 4), dtype=np.int32)
        self.L.vvo_tree_leaf_hs(self.h, out)
        return out

    def move_and_clean(self, remove=True):
        cl = I64(0)
        self._check(self.L.vvo_move_and_clean(self.h, int(remove), C.byref(cl)))
        return cl.value

    def vortex_shed(self):
        self._check(self.L.vvo_vortex_shed(self.h))

    def slae_size(self):
        return self.L.vvo_slae_size(self.h)

    def calc_circulation(self, want_matrix=False):
        n = self.slae_size()
        A = np.zeros((n, n)) if want_matrix else None
        b = np.zeros(n) if want_matrix else None
        self._check(self.L.vvo_calc_circulation(
            self.h, A.ctypes.data if want_matrix else None, b.ctypes.data if want_matrix else None))
        return A, b

    def calc_forces(self):
        self.L.vvo_calc_forces(self.h)

    def zero_forces(self):
        self.L.vvo_zero_forces(self.h)

    def far_moments(self):
        nn = self.L.vvo_far_moments(self.h, None)
        out = np.zeros((12, nn))
        self.L.vvo_far_moments(self.h, out.ctypes.data)
        return out

    def far_pairs(self, target, node):
        """Features, labels, scale, monopole field and exact field of pairs
        (tree-order vortex, oracle node id)."""
        target = np.ascontiguousarray(target, dtype=np.int32)
        node = np.ascontiguousarray(node, dtype=np.int32)
        n = len(target)
        X = np.zeros((n, 12)); Y = np.zeros((n, 2)); sc = np.zeros(n)
        base = np.zeros((n, 2)); exact = np.zeros((n, 2))
        self._check(self.L.vvo_far_pairs(self.h, target, node, n, X, Y, sc, base.ctypes.data, exact.ctypes.data))
        return X, Y, sc, base, exact

    def far_point_features(self, px, py, node):
        """Features, scale and unit direction (ux, uy) of (point, oracle node)
        pairs for arbitrary points."""
        px = np.ascontiguousarray(px, dtype=np.float64)
        py = np.ascontiguousarray(py, dtype=np.float64)
        node = np.ascontiguousarray(node, dtype=np.int32)
        n = len(node)
        X = np.zeros((n, 12)); sc = np.zeros(n); u = np.zeros((n, 2))
        self._check(self.L.vvo_far_point_features(self.h, px, py, node, n, X, sc, u))
        return X, sc, u

    def integrals(self):
        out = np.zeros(5)
        self.L.vvo_integrals(self.h, out)
        return out


def direct_sum(objs, targets=None, threads=8):
    """FP64 direct sum of the induced Lamb-kernel velocity (no free stream)."""
    L = Oracle.lib()
    objs = np.ascontiguousarray(objs, dtype=np.float64).reshape(-1, 6)
    if targets is None:
        nt = objs.shape[0]
        tp = None
    else:
        targets = np.ascontiguousarray(targets, dtype=np.int64)
        nt = targets.shape[0]
        tp = targets.ctypes.data
    out = np.zeros((nt, 2))
    L.vvo_direct_sum(objs, objs.shape[0], tp, nt, out, threads)
    return out


class RefHarness:
    """The reference's own tree / epsilon / diffusive code (oracle/_ref)."""

    _lib = None

    @classmethod
    def available(cls):
        return os.path.exists(REF_SO)

    @classmethod
    def lib(cls):
        if cls._lib is None:
            L = C.CDLL(REF_SO)
            _sig(L, "vvr_create", P, [])
            _sig(L, "vvr_destroy", None, [P])
            _sig(L, "vvr_set_re", None, [P, F64])
            _sig(L, "vvr_set_vortices", None, [P, dp, I64])
            _sig(L, "vvr_get_vortices", None, [P, dp])
            _sig(L, "vvr_add_body", None, [P, C.c_void_p, I32])
            _sig(L, "vvr_get_fric", None, [P, dp])
            _sig(L, "vvr_tree_build", None, [P, C.c_int, F64, F64])
            _sig(L, "vvr_num_leaves", I32, [P])
            _sig(L, "vvr_leaves", None, [P, dp, ip])
            _sig(L, "vvr_lists", None, [P, ip, ip, ip])
            _sig(L, "vvr_far_cm", None, [P, dp])
            _sig(L, "vvr_epsilon", I32, [P, C.c_int])
            _sig(L, "vvr_diffusive", None, [P])
            _sig(L, "vvr_set_pr", None, [P, F64])
            _sig(L, "vvr_set_list", C.c_int, [P, C.c_int, dp, I64])
            _sig(L, "vvr_list_size", I64, [P, C.c_int])
            _sig(L, "vvr_get_list", C.c_int, [P, C.c_int, dp])
            _sig(L, "vvr_leaves_hs", None, [P, ip])
            _sig(L, "vvr_diffusive_heat", None, [P])
            cls._lib = L
        return cls._lib

    def __init__(self):
        self.L = self.lib()
        self.h = self.L.vvr_create()
        self.n = 0
        self.nseg = 0

    def __del__(self):
        if getattr(self, "h", None):
            self.L.vvr_destroy(self.h)
            self.h = None

    def set_re(self, re):
        self.L.vvr_set_re(self.h, re)

    def set_vortices(self, objs):
        objs = np.ascontiguousarray(objs, dtype=np.float64).reshape(-1, 6)
        self.n = objs.shape[0]
        self.L.vvr_set_vortices(self.h, objs, self.n)

    def vortices(self):
        out = np.zeros((self.n, 6))
        if self.n:
            self.L.vvr_get_vortices(self.h, out)
        return out

    def add_body(self, segs):
        segs = np.ascontiguousarray(segs, dtype=SEG_DTYPE)
        self.nseg += len(segs)
        self.L.vvr_add_body(self.h, segs.ctypes.data, len(segs))

    def fric(self):
        out = np.zeros(max(self.nseg, 1))
        self.L.vvr_get_fric(self.h, out)
        return out[:self.nseg]

    def tree_build(self, far_criteria=8, min_node=0.0, max_node=np.finfo(np.float64).max):
        self.L.vvr_tree_build(self.h, far_criteria, min_node, max_node)

    def leaves(self):
        nl = self.L.vvr_num_leaves(self.h)
        geom = np.zeros((nl, 4))
        ints = np.zeros((nl, 7), dtype=np.int32)
        self.L.vvr_leaves(self.h, geom, ints)
        nnear, nfar, nseg = ints[:, 5].sum(), ints[:, 6].sum(), ints[:, 4].sum()
        near = np.zeros((max(nnear, 1), 2), dtype=np.int32)
        far = np.zeros((max(nfar, 1), 2), dtype=np.int32)
        segs = np.zeros(max(nseg, 1), dtype=np.int32)
        self.L.vvr_lists(self.h, near, far, segs)
        cm = np.zeros((max(nfar, 1), 8))
        self.L.vvr_far_cm(self.h, cm)
        return dict(geom=geom, ints=ints, near=near[:nnear], far=far[:nfar],
                    segs=segs[:nseg], far_cm=cm[:nfar])

    def epsilon(self, merge=True):
        return self.L.vvr_epsilon(self.h, int(merge))

    def diffusive(self):
        self.L.vvr_diffusive(self.h)

    def set_pr(self, pr):
        self.L.vvr_set_pr(self.h, pr)

    def set_list(self, list_id, objs):
        """list_id 1: HeatList, 2: StreakList."""
        objs = np.ascontiguousarray(objs, dtype=np.float64).reshape(-1, 6)
        self.L.vvr_set_list(self.h, list_id, objs, objs.shape[0])

    def get_list(self, list_id):
        n = self.L.vvr_list_size(self.h, list_id)
        out = np.zeros((max(n, 1), 6))
        self.L.vvr_get_list(self.h, list_id, out)
        return out[:n]

    def leaves_hs(self):
        nl = self.L.vvr_num_leaves(self.h)
        out = np.zeros((max(nl, 1), 4), dtype=np.int32)
        self.L.vvr_leaves_hs(self.h, out)
        return out[:nl]

    def diffusive_heat(self):
        self.L.vvr_diffusive_heat(self.h)
