"""ctypes driver of the test host around the reference-side binding
(tests/glue/host.c + tandav_amd/glue/tandav_glue.c)."""
import ctypes
import os

import numpy as np

from tandav_amd.engine import NODE_DTYPE, TOP_DTYPE

import build_host

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p

# Tandav's reference-named exports the binding must provide
EXPORTS = [
    "Compute_Acceleration", "Sort_Particles_By_Peano_Key",
    "Reverse_Peano_Keys", "cmp_peanoKeys", "Peano_Key", "Reversed_Peano_Key",
    "Reverse_Peano_Key", "Short_Peano_Key", "Reversed_Short_Peano_Key",
    "Test_Peanokey", "Setup_Gravity_Tree", "Gravity_Tree_Build",
    "Gravity_Tree_Acceleration", "Gravity_Tree_Periodic",
    "Gravity_Tree_Update_Kicks", "Gravity_Tree_Update_Drift",
    "Gravity_Tree_Free", "Node_Size", "Level", "Node_Is", "Node_Set",
    "Node_Clear"]
EXPORTS_PERIODIC = EXPORTS + ["Gravity_Periodic_Init", "Ewald_Correction",
                              "Gravity_Forcetest"]


def _d(a):
    return a.ctypes.data_as(_dp)


def load(periodic, path=None):
    """The host library; `path` loads a copy of it, which gets its own
    globals and statics (a fresh program)."""
    if path is None:
        path = build_host.lib_path(periodic)
        if not os.path.exists(path):
            build_host.build()
    L = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
    L.host_set_verbose.argtypes = [ctypes.c_int]
    L.host_init.argtypes = [ctypes.c_int] + [_dp] * 7 + [_vp] + \
        [ctypes.c_double] * 5 + [ctypes.c_int]
    L.host_set_domain.argtypes = [ctypes.c_double, _dp]
    L.host_set_max_active_bin.argtypes = [ctypes.c_int]
    L.host_set_top_nodes.argtypes = [_vp, ctypes.c_int]
    L.host_compute_acceleration.argtypes = [ctypes.c_int]
    L.host_tree_acceleration.argtypes = [ctypes.c_int]
    L.host_tree_kicks.argtypes = [ctypes.c_double]
    L.host_drift.argtypes = [ctypes.c_double]
    L.host_column.argtypes = [ctypes.c_char_p, ctypes.c_int]
    L.host_column.restype = _vp
    L.host_tree.argtypes = [ctypes.POINTER(ctypes.c_uint32)]
    L.host_tree.restype = _vp
    L.host_top_nodes.argtypes = [ctypes.POINTER(ctypes.c_uint32)]
    L.host_top_nodes.restype = _vp
    L.host_node_size.argtypes = [ctypes.c_int]
    L.host_node_size.restype = ctypes.c_double
    L.host_level.argtypes = [ctypes.c_int]
    L.host_peano_keys.argtypes = [ctypes.c_int, _dp, _dp, _dp, _vp, _vp,
                                  _vp, _vp]
    L.host_reverse_peano_key.argtypes = [ctypes.c_int, _vp, _vp]
    L.host_cmp_keys.argtypes = [_vp, _vp]
    if periodic:
        L.host_ewald_correction.argtypes = [ctypes.c_int, _dp, _dp]
    return L


_TYPES = {"Type": np.int32, "Time_Bin": np.int32, "It_Drift_Pos": np.uint32,
          "It_Kick_Pos": np.uint32, "ID": np.uint32, "Cost": np.float64,
          "Mass": np.float64, "Last_Acc_Mag": np.float64,
          "Tree_Parent": np.int32}


class Host:
    """One host program (one process-wide set of Tandav globals)."""

    def __init__(self, periodic, path=None):
        self.L = load(periodic, path)
        self.periodic = periodic
        self.n = 0

    def init(self, pos, vel, mass, boxsize=1.0, gravity=1.0, softening=0.01,
             theta_bh=0.1, alpha_rel=0.005, time_bin=None, threads=4):
        self.n = n = len(mass)
        cols = [np.ascontiguousarray(a, dtype=np.float64) for a in
                (pos[:, 0], pos[:, 1], pos[:, 2], vel[:, 0], vel[:, 1],
                 vel[:, 2], mass)]
        tb = None if time_bin is None else np.ascontiguousarray(time_bin,
                                                                np.int32)
        self._keep = (cols, tb)
        self.L.host_init(n, *[_d(c) for c in cols],
                         tb.ctypes.data_as(_vp) if tb is not None else None,
                         boxsize, gravity, softening, theta_bh, alpha_rel,
                         threads)

    def column(self, name, comp=0):
        p = self.L.host_column(name.encode(), comp)
        if name == "Key":
            buf = (ctypes.c_uint64 * (2 * self.n)).from_address(p)
            return np.ctypeslib.as_array(buf).reshape(self.n, 2).copy()
        if name in ("Pos", "Vel", "Acc", "Grav_Acc"):
            buf = (ctypes.c_double * self.n).from_address(p)
            return np.ctypeslib.as_array(buf).copy()
        dt = np.dtype(_TYPES[name])
        buf = (ctypes.c_char * (dt.itemsize * self.n)).from_address(p)
        return np.frombuffer(buf, dtype=dt).copy()

    def vec(self, name):
        return np.stack([self.column(name, d) for d in range(3)], axis=1)

    def tree(self):
        n = ctypes.c_uint32()
        p = self.L.host_tree(ctypes.byref(n))
        buf = (ctypes.c_char * (96 * n.value)).from_address(p)
        return np.frombuffer(buf, dtype=NODE_DTYPE).copy()

    def top_nodes(self):
        n = ctypes.c_uint32()
        p = self.L.host_top_nodes(ctypes.byref(n))
        buf = (ctypes.c_char * (64 * n.value)).from_address(p)
        return np.frombuffer(buf, dtype=TOP_DTYPE).copy()

    def set_top_nodes(self, rec):
        rec = np.ascontiguousarray(rec, dtype=TOP_DTYPE)
        assert self.L.host_set_top_nodes(rec.ctypes.data_as(_vp),
                                         len(rec)) == 0

    def set_domain(self, size, origin):
        o = np.ascontiguousarray(origin, dtype=np.float64)
        self.L.host_set_domain(float(size), _d(o))

    def decomposition(self, size, origin, top_nodes):
        """Domain_Decomposition with the given domain and top nodes."""
        self.set_domain(size, origin)
        self.L.host_sort()
        self.set_top_nodes(top_nodes)
        self.L.host_finish_decomposition()

    def peano_keys(self, x, y, z):
        x, y, z = (np.ascontiguousarray(a, np.float64) for a in (x, y, z))
        n = x.size
        std = np.zeros((n, 2), np.uint64)
        rev = np.zeros((n, 2), np.uint64)
        sh = np.zeros(n, np.uint64)
        rsh = np.zeros(n, np.uint64)
        self.L.host_peano_keys(n, _d(x), _d(y), _d(z),
                               *[a.ctypes.data_as(_vp)
                                 for a in (std, rev, sh, rsh)])
        return std, rev, sh, rsh

    def reverse_peano_key(self, keys2):
        k = np.ascontiguousarray(keys2, np.uint64)
        out = np.zeros_like(k)
        self.L.host_reverse_peano_key(len(k), k.ctypes.data_as(_vp),
                                      out.ctypes.data_as(_vp))
        return out

    def cmp_keys(self, a2, b2):
        a = np.ascontiguousarray(a2, np.uint64)
        b = np.ascontiguousarray(b2, np.uint64)
        return self.L.host_cmp_keys(a.ctypes.data_as(_vp),
                                    b.ctypes.data_as(_vp))

    def ewald_correction(self, dr):
        dr = np.ascontiguousarray(dr, np.float64).reshape(-1, 3)
        f = np.zeros_like(dr)
        self.L.host_ewald_correction(len(dr), _d(dr), _d(f))
        return f
