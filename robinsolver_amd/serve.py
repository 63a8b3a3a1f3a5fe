"""Serving loop: ctypes binding of the rs_server_* entry points.

A ``Server`` keeps factored ``GpuTree`` objects resident and runs queued solve
and factor requests on several streams, coalescing each tree's queued
right-hand sides into multi-RHS passes (continuous batching) and admitting
work by free device memory.  The scheduling loop itself is native (C++,
csrc/server.cpp); this module only binds it.

Request arrays are device memory (for instance torch tensors on ``cuda:0``,
passed by ``data_ptr()``) and must stay alive until the request completes.
"""
import ctypes

RS_REQ_QUEUED, RS_REQ_RUNNING, RS_REQ_DONE, RS_REQ_FAILED = 0, 1, 2, 3


class ServerOpts(ctypes.Structure):
    _fields_ = [("max_cols", ctypes.c_int), ("lanes", ctypes.c_int), ("hbm_reserve", ctypes.c_longlong),
                ("hbm_limit", ctypes.c_longlong), ("simulate", ctypes.c_int), ("sim_pass_s", ctypes.c_double),
                ("sim_col_s", ctypes.c_double), ("sim_factor_s", ctypes.c_double), ("group", ctypes.c_int)]


class ServerStats(ctypes.Structure):
    _fields_ = [("requests_done", ctypes.c_longlong), ("rhs_done", ctypes.c_longlong),
                ("passes", ctypes.c_longlong), ("factors_done", ctypes.c_longlong),
                ("admission_waits", ctypes.c_longlong), ("evictions", ctypes.c_longlong),
                ("failed", ctypes.c_longlong), ("busy_s", ctypes.c_double), ("lat_p50_s", ctypes.c_double),
                ("lat_p99_s", ctypes.c_double), ("lat_max_s", ctypes.c_double), ("mean_cols", ctypes.c_double),
                ("held_bytes", ctypes.c_longlong), ("mean_trees", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def bind(L):
    """Declare the server entry points on the loaded library."""
    vp, ll, llp, dp = ctypes.c_void_p, ctypes.c_longlong, ctypes.POINTER(ctypes.c_longlong), \
        ctypes.POINTER(ctypes.c_double)
    L.rs_server_opts_default.argtypes = [ctypes.POINTER(ServerOpts)]
    L.rs_server_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(ServerOpts)]
    L.rs_server_destroy.argtypes = [vp]
    L.rs_server_add_tree.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_int)]
    L.rs_server_factor.argtypes = [vp, ctypes.c_int, vp, llp]
    L.rs_server_solve.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp, vp, llp]
    L.rs_server_test.argtypes = [vp, ll, ctypes.POINTER(ctypes.c_int)]
    L.rs_server_wait.argtypes = [vp, ll]
    L.rs_server_drain.argtypes = [vp]
    L.rs_server_request_info.argtypes = [vp, ll, dp, dp, dp, llp]
    L.rs_server_release.argtypes = [vp, ll]
    L.rs_server_get_stats.argtypes = [vp, ctypes.POINTER(ServerStats)]
    L.rs_server_reset_stats.argtypes = [vp]
    L.rs_server_error.argtypes = [vp, ll]
    L.rs_server_error.restype = ctypes.c_char_p
    L.rs_server_lane_bytes.argtypes = [vp, ctypes.c_int, llp, llp]


def lane_bytes(tree, max_cols):
    """(lane arena bytes tree needs, the tree's descriptor bytes per lane) for a
    server with this max_cols: a lane's arena is the largest over its trees."""
    from . import lib, RobinError
    a, d = ctypes.c_longlong(), ctypes.c_longlong()
    if lib().rs_server_lane_bytes(tree.h, int(max_cols), ctypes.byref(a), ctypes.byref(d)):
        raise RobinError("rs_server_lane_bytes: the tree is not planned, or max_cols < 1.")
    return a.value, d.value


class Server:
    """Request loop over resident trees (see include/robinsolver.h, rs_server_*)."""

    def __init__(self, max_cols=256, lanes=2, hbm_reserve=2 << 30, hbm_limit=0, simulate=False,
                 sim_pass_s=1e-3, sim_col_s=1e-5, sim_factor_s=5e-3, group=1):
        from . import lib, RobinError
        self._L = lib()
        self._err = RobinError
        o = ServerOpts()
        self._L.rs_server_opts_default(ctypes.byref(o))
        o.max_cols, o.lanes, o.hbm_reserve, o.hbm_limit = max_cols, lanes, hbm_reserve, hbm_limit
        o.simulate, o.sim_pass_s, o.sim_col_s, o.sim_factor_s = int(simulate), sim_pass_s, sim_col_s, sim_factor_s
        o.group = group
        h = ctypes.c_void_p()
        if self._L.rs_server_create(ctypes.byref(h), ctypes.byref(o)):
            raise RobinError(self._L.rs_server_error(None, -1).decode())
        self.h = h
        self.max_cols = max_cols
        self.trees = []

    def _check(self, rc, req=-1):
        if rc:
            raise self._err(self._L.rs_server_error(self.h, req).decode())

    def add_tree(self, tree):
        """Register a planned GpuTree (factored or not); returns its id."""
        tid = ctypes.c_int()
        self._check(self._L.rs_server_add_tree(self.h, tree.h, ctypes.byref(tid)))
        self.trees.append(tree)
        return tid.value

    def factor(self, tree_id, k_ptr):
        """Queue a (re)factorization; k_ptr: device [n_leaves][n_int] complex128."""
        r = ctypes.c_longlong()
        self._check(self._L.rs_server_factor(self.h, tree_id, ctypes.c_void_p(k_ptr), ctypes.byref(r)))
        return r.value

    def solve(self, tree_id, nrhs, f_ptr, g_ptr, u_ptr):
        """Queue a solve of nrhs right-hand sides (device arrays); returns the request id."""
        r = ctypes.c_longlong()
        self._check(self._L.rs_server_solve(self.h, tree_id, nrhs, ctypes.c_void_p(f_ptr), ctypes.c_void_p(g_ptr),
                                            ctypes.c_void_p(u_ptr), ctypes.byref(r)))
        return r.value

    def state(self, req):
        st = ctypes.c_int()
        self._check(self._L.rs_server_test(self.h, req, ctypes.byref(st)))
        return st.value

    def wait(self, req):
        """Block until req completes; raises with the request's message if it failed."""
        self._check(self._L.rs_server_wait(self.h, req), req)

    def error(self, req):
        return self._L.rs_server_error(self.h, req).decode()

    def drain(self):
        self._check(self._L.rs_server_drain(self.h))

    def info(self, req):
        """(t_submit, t_start, t_done, pass id) of a request, seconds on one monotonic clock."""
        a, b, c, p = ctypes.c_double(), ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
        self._check(self._L.rs_server_request_info(self.h, req, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                                   ctypes.byref(p)))
        return a.value, b.value, c.value, p.value

    def release(self, req):
        self._check(self._L.rs_server_release(self.h, req))

    def stats(self):
        st = ServerStats()
        self._check(self._L.rs_server_get_stats(self.h, ctypes.byref(st)))
        return st.as_dict()

    def reset_stats(self):
        self._check(self._L.rs_server_reset_stats(self.h))

    def close(self):
        if getattr(self, "h", None):
            self._L.rs_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
