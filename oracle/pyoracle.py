"""ctypes binding of the CPU oracle (oracle/rs_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker for the MI355X path.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "librs_oracle.so")
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.ro_create.restype = ctypes.c_void_p
        L.ro_create.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_double, ctypes.c_double, dp]
        L.ro_create_error.restype = ctypes.c_char_p
        L.ro_last_error.restype = ctypes.c_char_p
        L.ro_last_error.argtypes = [ctypes.c_void_p]
        L.ro_destroy.argtypes = [ctypes.c_void_p]
        L.ro_set_threads.argtypes = [ctypes.c_int]
        L.ro_info.argtypes = [ctypes.c_void_p, ip, ip, ip, ip, ip, dp]
        L.ro_leaf_box.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
        L.ro_leaf_points.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, ip]
        L.ro_factor.argtypes = [ctypes.c_void_p, dp]
        L.ro_solve.argtypes = [ctypes.c_void_p, dp, dp, dp]
        L.ro_num_nodes.argtypes = [ctypes.c_void_p]
        L.ro_node_rtr.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, ip]
        L.ro_apply_invm.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp]
        L.ro_node_iface.argtypes = [ctypes.c_void_p, ctypes.c_int, ip, ip, ip]
        L.ro_node_ptbox.argtypes = [ctypes.c_void_p, ctypes.c_int, ip]
        L.ro_level_node_id.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.ro_leaf_matrix.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def standard_bndry_params(kb):
    """robin_tree%bndry_params as example.f90:64-65 sets it, (2,2) complex."""
    return np.array([[1j * kb, 1.0], [-1j * kb, 1.0]], dtype=np.complex128)


class OracleTree:
    """Reference-algorithm tree: geometry at construction, factor, solve."""

    def __init__(self, d, L, domain, h_tgt, kb, bndry_params=None, threads=0):
        L_ = lib()
        L_.ro_set_threads(int(threads))
        dom = np.asarray(domain, dtype=np.float64).reshape(3, 2)  # rows: axis; cols: lo, hi
        dom_cm = np.ascontiguousarray(dom.reshape(-1))  # (2,3) column-major == [lo,hi] per axis
        bp = standard_bndry_params(kb) if bndry_params is None else np.asarray(bndry_params, np.complex128)
        bp_cm = np.ascontiguousarray(bp.T.reshape(-1)).view(np.float64)  # column-major
        self._bp = bp_cm
        self.d, self.L, self.kb = d, L, kb
        self.bndry_params = bp
        self.h = L_.ro_create(d, L, _dp(dom_cm), float(h_tgt), float(kb), _dp(bp_cm))
        if not self.h:
            raise RuntimeError(L_.ro_create_error().decode())
        nl, m, ni, nb = (ctypes.c_int() for _ in range(4))
        pb = (ctypes.c_int * 3)()
        hh = np.zeros(3)
        L_.ro_info(self.h, ctypes.byref(nl), ctypes.byref(m), ctypes.byref(ni), ctypes.byref(nb), pb, _dp(hh))
        self.n_leaves, self.m, self.n_int, self.nb = nl.value, m.value, ni.value, nb.value
        self.ptbox = tuple(pb)
        self.grid_h = hh

    def close(self):
        if self.h:
            lib().ro_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def err(self):
        return lib().ro_last_error(self.h).decode()

    def leaf_box(self, leaf):
        b = np.zeros(6)
        lib().ro_leaf_box(self.h, leaf, _dp(b))
        return b.reshape(3, 2)

    def leaf_points(self, leaf):
        xyz = np.zeros((self.m, 3))
        face = np.zeros(self.m, dtype=np.int32)
        lib().ro_leaf_points(self.h, leaf, _dp(xyz), face.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return xyz, face

    def all_points(self):
        pts = np.zeros((self.n_leaves, self.m, 3))
        faces = None
        for q in range(self.n_leaves):
            pts[q], faces = self.leaf_points(q)
        return pts, faces

    def factor(self, k_int):
        k = np.ascontiguousarray(np.asarray(k_int, np.complex128).reshape(self.n_leaves, self.n_int))
        if lib().ro_factor(self.h, _dp(k.view(np.float64))):
            raise RuntimeError(self.err())

    def solve(self, f_int, g_bnd):
        f = np.ascontiguousarray(np.asarray(f_int, np.complex128).reshape(self.n_leaves, self.n_int))
        g = np.ascontiguousarray(np.asarray(g_bnd, np.complex128).reshape(self.n_leaves, self.nb))
        u = np.zeros((self.n_leaves, self.m), np.complex128)
        if lib().ro_solve(self.h, _dp(f.view(np.float64)), _dp(g.view(np.float64)), _dp(u.view(np.float64))):
            raise RuntimeError(self.err())
        return u

    def num_nodes(self):
        return lib().ro_num_nodes(self.h)

    def node_rtr(self, node_id):
        P = ctypes.c_int()
        if lib().ro_node_rtr(self.h, node_id, None, ctypes.byref(P)):
            raise RuntimeError("no RtR for node %d" % node_id)
        flat = np.zeros(P.value * P.value, np.complex128)
        lib().ro_node_rtr(self.h, node_id, _dp(flat.view(np.float64)), ctypes.byref(P))
        return flat.reshape(P.value, P.value, order="F")

    def node_ptbox(self, node_id):
        pb = (ctypes.c_int * 3)()
        if lib().ro_node_ptbox(self.h, node_id, pb):
            raise RuntimeError("bad node id")
        return tuple(pb)

    def node_face_offsets(self, node_id):
        """Start offset of faces 1..2d in the node's face-ordered boundary."""
        pb = self.node_ptbox(node_id)
        d = self.d
        prod = int(np.prod(pb[:d]))
        pf = [prod // pb[j] for j in range(d) for _ in range(2)]
        return np.concatenate([[0], np.cumsum(pf)])[:-1]

    def level_node_id(self, lvl, idx):
        return lib().ro_level_node_id(self.h, lvl, idx)

    def apply_invm(self, node_id, x):
        x = np.ascontiguousarray(np.asarray(x, np.complex128))
        y = np.zeros_like(x)
        if lib().ro_apply_invm(self.h, node_id, _dp(x.view(np.float64)), _dp(y.view(np.float64))):
            raise RuntimeError("apply_invm failed")
        return y

    def node_iface(self, node_id):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        if lib().ro_node_iface(self.h, node_id, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)):
            raise RuntimeError("leaf")
        return a.value, b.value, c.value

    def leaf_matrix(self, leaf, k_leaf):
        k = np.ascontiguousarray(np.asarray(k_leaf, np.complex128))
        A = np.zeros(self.m * self.m, np.complex128)
        lib().ro_leaf_matrix(self.h, leaf, _dp(k.view(np.float64)), _dp(A.view(np.float64)))
        return A.reshape(self.m, self.m, order="F")
