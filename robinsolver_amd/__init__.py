"""robinsolver_amd -- MI355X-native HPS factor and solve for variable-potential
Helmholtz problems, behind robinsolver's own interface.

Two layers:
  * ``GpuTree``: thin ctypes binding of the C ABI (include/robinsolver.h),
    host- or device-resident data.
  * ``robin_tree`` / ``elliptic_operator`` / ``solve_opts`` / ``elliptic_rhs`` /
    ``get_solution``: the reference's Fortran API (robin_tree_module.f90:55-233,
    derived_type_module.f90:72-320, robin_output_module.f90:32-75) restated in
    Python, with the callbacks k, f and g_face sampled on the host at the
    reference's points.

The library is the HIP build in this directory (librobinsolver.so); nothing
here falls back to a CPU path: without the library or a GPU, calls raise.
"""
import ctypes
import os
import struct
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librobinsolver.so")
_lib = None

RS_FP64 = 0
RS_MIXED = 1
RS_MIXED_MERGE = 2   # rs_check_gemm: complex64 products of complex128 data
RS_RETAIN_MAPS, RS_RETAIN_FP64, RS_RETAIN_SCRATCH = 1, 2, 4   # rs_opts.retain


class RobinError(RuntimeError):
    pass


class RsOpts(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("skip_root_rtr", ctypes.c_int), ("n_refine", ctypes.c_int),
                ("verbose", ctypes.c_int), ("hbm_budget", ctypes.c_longlong), ("autotune", ctypes.c_int),
                ("graphs", ctypes.c_int), ("merge_fp64_levels", ctypes.c_int), ("retain", ctypes.c_int)]


def lib():
    """Load librobinsolver.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RobinError("librobinsolver.so is not built; run robinsolver_amd.build.build()")
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        vp = ctypes.c_void_p
        L.rs_opts_default.argtypes = [ctypes.POINTER(RsOpts)]
        L.rs_tree_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(RsOpts)]
        L.rs_tree_destroy.argtypes = [vp]
        L.rs_plan.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp, ctypes.c_double, ctypes.c_double, dp,
                              ip, ip, ip, ip]
        L.rs_leaf_points.argtypes = [vp, ctypes.c_int, dp, ip]
        L.rs_leaf_box.argtypes = [vp, ctypes.c_int, dp]
        L.rs_factor.argtypes = [vp, dp]
        L.rs_solve.argtypes = [vp, ctypes.c_int, dp, dp, ctypes.c_int]
        L.rs_get_solution.argtypes = [vp, ctypes.c_int, dp]
        L.rs_solve_device.argtypes = [vp, ctypes.c_int, vp, vp, vp, ctypes.c_int, vp]
        L.rs_timings.argtypes = [vp, dp, dp]
        L.rs_last_error.argtypes = [vp]
        L.rs_last_error.restype = ctypes.c_char_p
        L.rs_inspect.argtypes = [vp, ip, ip, ctypes.POINTER(ctypes.c_longlong)]
        L.rs_node_rtr.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp, ip]
        L.rs_set_refine.argtypes = [vp, ctypes.c_int]
        L.rs_level_info.argtypes = [vp, ctypes.c_int, ip, ip, ip, ip]
        L.rs_profile.argtypes = [vp, ctypes.c_int]
        L.rs_profile_name.restype = ctypes.c_char_p
        L.rs_profile_name.argtypes = [ctypes.c_int]
        L.rs_profile_read.argtypes = [vp, ctypes.c_int, dp, ip, dp, dp]
        llp = ctypes.POINTER(ctypes.c_longlong)
        L.rs_memory_estimate.argtypes = [vp, ctypes.c_int, llp, llp, llp]
        L.rs_memory_used.argtypes = [vp, llp, llp]
        L.rs_residual_norms.argtypes = [vp, ctypes.c_int, vp, vp, vp, dp]
        L.rs_tile_choices.argtypes = [vp, ctypes.c_int, ip]
        L.rs_leaf_grid.argtypes = [vp, ip, dp]
        L.rs_points_device.argtypes = [vp, vp, vp, vp]
        L.rs_factor_device.argtypes = [vp, vp]
        L.rs_tree_save.argtypes = [vp, ctypes.c_char_p]
        L.rs_tree_load.argtypes = [vp, ctypes.c_char_p]
        from . import serve as _serve
        _serve.bind(L)
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _c128(a, shape):
    return np.ascontiguousarray(np.asarray(a, dtype=np.complex128).reshape(shape))


_SAVED_HEADER = struct.Struct("<8sI11i16d")   # rs_tree_save's header (persist.cpp)


def read_saved_header(path):
    """The plan parameters and flags of a file written by rs_tree_save."""
    with open(path, "rb") as fh:
        raw = fh.read(_SAVED_HEADER.size)
    if len(raw) < _SAVED_HEADER.size:
        raise RobinError("Error: %s is not a saved tree." % path)
    v = _SAVED_HEADER.unpack(raw)
    if v[0] != b"RSTREE\x00\x01":
        raise RobinError("Error: %s is not a saved tree." % path)
    d, L, prec, h64, h32, mc64, nl, m, ni, nb = v[2:12]
    dbl = v[13:]
    return {"version": v[1], "d": d, "L": L, "precision": prec, "have64": h64, "have32": h32,
            "merge_c64": mc64, "n_leaves": nl, "m": m, "n_int": ni, "nb": nb, "domain": list(dbl[0:6]),
            "h_tgt": dbl[6], "kb": dbl[7],
            "bndry_params": [complex(dbl[8 + 2 * q], dbl[9 + 2 * q]) for q in range(4)]}


class GpuTree:
    """One factored tree resident on the GPU (C ABI handle)."""

    def __init__(self, d, L, domain, h_tgt, kb, bndry_params=None, precision=RS_FP64, skip_root_rtr=True,
                 n_refine=2, hbm_budget=0, autotune=True, graphs=True, merge_fp64_levels=None, retain=0):
        L_ = lib()
        o = RsOpts()
        L_.rs_opts_default(ctypes.byref(o))
        o.precision, o.skip_root_rtr, o.n_refine, o.hbm_budget = precision, int(skip_root_rtr), n_refine, hbm_budget
        o.autotune = int(autotune)
        o.graphs = int(graphs)
        o.retain = int(retain)
        if merge_fp64_levels is not None:
            o.merge_fp64_levels = int(merge_fp64_levels)
        h = ctypes.c_void_p()
        if L_.rs_tree_create(ctypes.byref(h), ctypes.byref(o)):
            raise RobinError(L_.rs_last_error(None).decode())
        self.h = h
        dom = np.ascontiguousarray(np.asarray(domain, np.float64).reshape(3, 2).reshape(-1))
        if bndry_params is None:
            bndry_params = np.array([[1j * kb, 1.0], [-1j * kb, 1.0]])
        bp = np.ascontiguousarray(np.asarray(bndry_params, np.complex128).T.reshape(-1)).view(np.float64)
        nl, m, ni, nb = (ctypes.c_int() for _ in range(4))
        self._check(L_.rs_plan(self.h, d, L, _dp(dom), float(h_tgt), float(kb), _dp(bp), ctypes.byref(nl),
                               ctypes.byref(m), ctypes.byref(ni), ctypes.byref(nb)))
        self.d, self.L, self.kb = d, L, kb
        self.n_leaves, self.m, self.n_int, self.nb = nl.value, m.value, ni.value, nb.value
        self._nrhs = 0

    def _check(self, rc):
        if rc:
            raise RobinError(lib().rs_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            lib().rs_tree_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def leaf_points(self, leaf):
        xyz = np.zeros((self.m, 3))
        face = np.zeros(self.m, np.int32)
        self._check(lib().rs_leaf_points(self.h, leaf, _dp(xyz), face.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return xyz, face

    def all_points(self):
        pts = np.zeros((self.n_leaves, self.m, 3))
        face = None
        for q in range(self.n_leaves):
            pts[q], face = self.leaf_points(q)
        return pts, face

    def points_device(self, xyz_ptr, face_ptr=0, stream=None):
        """Every point's coordinates into device memory at xyz_ptr ([n_leaves][m][3]
        float64) and, when face_ptr is given, the face ids ([n_leaves][m] int32)."""
        self._check(lib().rs_points_device(self.h, ctypes.c_void_p(xyz_ptr), ctypes.c_void_p(face_ptr or 0),
                                           ctypes.c_void_p(stream or 0)))

    def factor_device(self, k_ptr):
        """rs_factor from a device array of k at the interior points ([n_leaves][n_int] complex128)."""
        self._check(lib().rs_factor_device(self.h, ctypes.c_void_p(k_ptr)))

    def leaf_box(self, leaf):
        b = np.zeros(6)
        self._check(lib().rs_leaf_box(self.h, leaf, _dp(b)))
        return b.reshape(3, 2)

    def factor(self, k_int):
        k = _c128(k_int, (self.n_leaves, self.n_int))
        self._check(lib().rs_factor(self.h, _dp(k.view(np.float64))))

    def solve(self, f_int, g_bnd, precision=-1):
        """f_int [nrhs][n_leaves][n_int], g_bnd [nrhs][n_leaves][nb] (or without
        the leading nrhs axis).  Returns u [nrhs][n_leaves][m] (squeezed)."""
        f = np.asarray(f_int, np.complex128)
        single = f.ndim == 2
        nrhs = 1 if single else f.shape[0]
        f = _c128(f, (nrhs, self.n_leaves, self.n_int))
        g = _c128(g_bnd, (nrhs, self.n_leaves, self.nb))
        self._check(lib().rs_solve(self.h, nrhs, _dp(f.view(np.float64)), _dp(g.view(np.float64)), precision))
        self._nrhs = nrhs
        u = np.zeros((nrhs, self.n_leaves, self.m), np.complex128)
        for r in range(nrhs):
            self._check(lib().rs_get_solution(self.h, r, _dp(u[r].view(np.float64))))
        return u[0] if single else u

    def solve_device(self, nrhs, f_ptr, g_ptr, u_ptr, precision=-1, stream=None):
        """Device-resident solve: raw device pointers (e.g. torch tensor.data_ptr())."""
        self._check(lib().rs_solve_device(self.h, nrhs, ctypes.c_void_p(f_ptr), ctypes.c_void_p(g_ptr),
                                          ctypes.c_void_p(u_ptr), precision, ctypes.c_void_p(stream or 0)))

    def set_refine(self, n):
        self._check(lib().rs_set_refine(self.h, int(n)))

    def timings(self):
        f, s = ctypes.c_double(), ctypes.c_double()
        lib().rs_timings(self.h, ctypes.byref(f), ctypes.byref(s))
        return f.value, s.value

    def save(self, path):
        """Write the factored tree to a file (rs_tree_save)."""
        self._check(lib().rs_tree_save(self.h, os.fsencode(path)))

    @classmethod
    def load(cls, path, **opts):
        """A tree read from a file written by save(): planned from the file's
        parameters and factored, with the saved precision (rs_tree_load).  The
        keyword options are the constructor's other than the geometry and the
        precision."""
        hdr = read_saved_header(path)
        bp = np.array(hdr["bndry_params"]).reshape(2, 2, order="F")
        t = cls(hdr["d"], hdr["L"], np.array(hdr["domain"]).reshape(3, 2), hdr["h_tgt"], hdr["kb"],
                bndry_params=bp, precision=hdr["precision"], **opts)
        t._check(lib().rs_tree_load(t.h, os.fsencode(path)))
        return t

    def residual_norms(self, nrhs, f_ptr, g_ptr, u_ptr):
        """||r|| / ||(f, g)|| of the global FD system for each of nrhs device-resident solutions."""
        out = np.zeros(nrhs)
        self._check(lib().rs_residual_norms(self.h, nrhs, ctypes.c_void_p(f_ptr), ctypes.c_void_p(g_ptr),
                                            ctypes.c_void_p(u_ptr), _dp(out)))
        return out

    def leaf_grid(self):
        """Points per axis of every leaf (d entries)."""
        pb = np.zeros(3, np.int32)
        h = np.zeros(3)
        self._check(lib().rs_leaf_grid(self.h, pb.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dp(h)))
        return tuple(int(x) for x in pb[:self.d])

    def tile_choices(self, precision):
        """GEMM tile chosen per level and product of the solve workspace: [L][4] (-1: shape rule / none)."""
        out = np.zeros(4 * self.L, np.int32)
        self._check(lib().rs_tile_choices(self.h, precision, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return out.reshape(self.L, 4)

    def memory_estimate(self, ncols=1):
        """(factor resident bytes, factor peak bytes, solve workspace bytes of ncols RHS)."""
        r, p, w = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
        self._check(lib().rs_memory_estimate(self.h, int(ncols), ctypes.byref(r), ctypes.byref(p), ctypes.byref(w)))
        return r.value, p.value, w.value

    def memory_used(self):
        """(device bytes held now, most held)."""
        h, p = ctypes.c_longlong(), ctypes.c_longlong()
        lib().rs_memory_used(self.h, ctypes.byref(h), ctypes.byref(p))
        return h.value, p.value

    def inspect(self):
        fa, nl, db = ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong()
        lib().rs_inspect(self.h, ctypes.byref(fa), ctypes.byref(nl), ctypes.byref(db))
        return bool(fa.value), nl.value, db.value

    def level_info(self, lvl):
        """(nodes, P, interface size, child P) of level lvl."""
        v = [ctypes.c_int() for _ in range(4)]
        self._check(lib().rs_level_info(self.h, lvl, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def algorithmic_counts(self):
        """Reference-algorithm counts (dense leaf LU, D/S LU solves): factor flops,
        solve flops per RHS, stored solve data entries."""
        m, nb = self.m, self.nb
        nl = self.n_leaves
        ff = nl * ((8.0 / 3.0) * m ** 3 + 8.0 * m * m * nb)
        fs = nl * 16.0 * m * m
        for l in range(self.L):
            nodes, P, i, b = self.level_info(l)
            ff += nodes * ((32.0 / 3.0) * i ** 3 + 40.0 * i * i * (b - i) + 8.0 * P * P * i)
            fs += nodes * (80.0 * i * i + 32.0 * i * (b - i))
        return ff, fs

    def profile(self, enable=True):
        self._check(lib().rs_profile(self.h, int(enable)))

    def profile_read(self):
        """{category: (ms, launches, bytes, flops)} since the last read."""
        out = {}
        L_ = lib()
        for c in range(L_.rs_profile_ncat()):
            ms, n, b, f = ctypes.c_double(), ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
            self._check(L_.rs_profile_read(self.h, c, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(b),
                                           ctypes.byref(f)))
            out[L_.rs_profile_name(c).decode()] = (ms.value, n.value, b.value, f.value)
        return out

    def node_rtr(self, lvl, idx):
        P = ctypes.c_int()
        self._check(lib().rs_node_rtr(self.h, lvl, idx, None, ctypes.byref(P)))
        flat = np.zeros(P.value * P.value, np.complex128)
        self._check(lib().rs_node_rtr(self.h, lvl, idx, _dp(flat.view(np.float64)), ctypes.byref(P)))
        return flat.reshape(P.value, P.value, order="F")


# ---------------------------------------------------------------------------
# The reference's interface (Fortran derived types and type-bound procedures)
# ---------------------------------------------------------------------------

class elliptic_operator:
    """type(elliptic_operator) (derived_type_module.f90:72-84).  k(x) is a
    callable of a 3-vector returning a complex; coeff and robin_coeff are kept
    for interface parity and, as in the reference, not read."""

    def __init__(self, d=2, domain=None, k=None, coeff=None, robin_coeff=None):
        self.d = d
        self.domain = np.zeros((2, 3)) if domain is None else np.asarray(domain, np.float64)
        self.k = k
        self.coeff = coeff
        self.robin_coeff = robin_coeff


class debug_opts:
    def __init__(self):
        self.txt_output = False
        self.no_delete = False


class solve_opts:
    """type(solve_opts) (derived_type_module.f90:309-320); `precision` selects
    the MI355X path (RS_FP64 or RS_MIXED)."""

    def __init__(self, h_tgt=0.1, kb=1.0, precision=RS_FP64):
        self.h_tgt = h_tgt
        self.h = np.zeros(3)
        self.ordering = 1
        self.disc = "fd"
        self.local_mode = "LU"
        self.R_mode = "LU"
        self.kb = kb
        self.debug = debug_opts()
        self.precision = precision


class elliptic_rhs:
    """type(elliptic_rhs) (derived_type_module.f90:109-115): f(x) and g[0..5]
    face callables (g_face receives the packed tangential coordinates)."""

    def __init__(self, d=2, f=None, g=None):
        self.d = d
        self.f = f
        self.g = list(g) if g is not None else [None] * 6


class leaf_solution:
    """type(leaf_solution) (robin_output_module.f90:32-37)."""

    def __init__(self, box, u):
        self.box = box
        self.u = u


def _packed(xyz, face, d):
    # fd_make_bndry_vec coordinate packing (fd_module.f90:645-651)
    ax = (face - 1) // 2
    x = np.zeros(3)
    tang = [xyz[j] for j in range(d) if j != ax]
    x[:len(tang)] = tang
    return x


class robin_tree:
    """type(robin_tree) with factor / solve / inspect (robin_tree_module.f90:55-77)."""

    def __init__(self):
        self.nlvls = 0
        self.ell_op = None
        self.opts = None
        self.bndry_params = np.zeros((2, 2), np.complex128)
        self.factored = False
        self.f_time = 0.0
        self.s_time = 0.0
        self._t = None

    def factor(self, L, ell_op, opts):
        t0 = time.perf_counter()
        self.nlvls, self.ell_op, self.opts = L, ell_op, opts
        dom = np.asarray(ell_op.domain, np.float64)  # (2,3): rows lo/hi
        self._t = GpuTree(ell_op.d, L, dom.T, opts.h_tgt, opts.kb, self.bndry_params,
                          precision=getattr(opts, "precision", RS_FP64))
        t = self._t
        pts, face = t.all_points()
        self._pts, self._face = pts, face
        kk = np.array([[ell_op.k(pts[q, p]) for p in range(t.nb, t.m)] for q in range(t.n_leaves)],
                      np.complex128)
        t.factor(kk)
        self.factored = True
        self.f_time = time.perf_counter() - t0

    def solve(self, rhs):
        if not self.factored:
            raise RobinError("Error: Robin Tree must be factored first.")
        if self.ell_op.d != rhs.d:
            raise RobinError("Error: Dimensions of Robin Tree and RHS do not match.")
        t0 = time.perf_counter()
        t, pts, face = self._t, self._pts, self._face
        f = np.array([[rhs.f(pts[q, p]) for p in range(t.nb, t.m)] for q in range(t.n_leaves)], np.complex128)
        # g_face is called at the physical boundary points only, as the reference
        # does (apply_bc, robin_tree_module.f90:1186-1237)
        g = np.zeros((t.n_leaves, t.nb), np.complex128)
        dom = np.asarray(self.ell_op.domain, np.float64)
        for p in range(t.nb):
            fc = int(face[p])
            gf = rhs.g[fc - 1]
            if gf is None:
                continue
            ax, side = (fc - 1) // 2, (fc - 1) % 2
            for q in range(t.n_leaves):
                if pts[q, p, ax] == dom[side, ax]:
                    g[q, p] = gf(_packed(pts[q, p], fc, t.d))
        self._u = t.solve(f, g)
        self.s_time = time.perf_counter() - t0

    def inspect(self):
        print("Factored: %s" % ("T" if self.factored else "F"))
        if self.factored:
            print("Number of levels: %d" % self.nlvls)


def get_solution(T):
    """get_solution (robin_output_module.f90:48-75): leaves in DFS order."""
    if not T.factored:
        raise RobinError("Error in GET_SOLUTION: Robin Tree not factored.")
    t = T._t
    return [leaf_solution(t.leaf_box(q).T.copy(), T._u[q].copy()) for q in range(t.n_leaves)]
