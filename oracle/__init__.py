"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package ``tandav_amd``.
The C sources restate the reference force path (see oracle.h for the
reference file:line of every function and how each one is pinned).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
NE1 = 65


def build(force=False):
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s"] + (["-B"] if force else [])
                       + ["-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    """The oracle library; ORACLE_LIB names another build of the same
    sources (the mutation tests load deliberately broken copies)."""
    global _lib
    if _lib is None:
        path = os.environ.get("ORACLE_LIB") or _LIB_PATH
        if path == _LIB_PATH and not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(path)
        _declare(_lib)
    return _lib


class OrcParams(ctypes.Structure):
    _fields_ = [("periodic", ctypes.c_int), ("boxsize", ctypes.c_double),
                ("gravity", ctypes.c_double), ("softening", ctypes.c_double),
                ("theta_bh", ctypes.c_double), ("alpha_rel", ctypes.c_double),
                ("vector_size", ctypes.c_int), ("nthreads", ctypes.c_int)]


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p


class OrcParticles(ctypes.Structure):
    _fields_ = [("pos", _dp * 3), ("vel", _dp * 3), ("acc", _dp * 3),
                ("grav_acc", _dp * 3), ("mass", _dp), ("last_acc_mag", _dp),
                ("cost", _dp), ("tree_parent", _ip), ("key", _vp),
                ("id", _vp), ("time_bin", _ip)]


class OrcNode(ctypes.Structure):
    _fields_ = [("DNext", ctypes.c_int), ("Bitfield", ctypes.c_uint32),
                ("DUp", ctypes.c_int), ("Npart", ctypes.c_int),
                ("Pos", ctypes.c_double * 3), ("Mass", ctypes.c_double),
                ("CoM", ctypes.c_double * 3), ("Dp", ctypes.c_double * 3)]


class OrcTopNode(ctypes.Structure):
    _fields_ = [("Key", ctypes.c_uint64), ("Target", ctypes.c_int),
                ("Level", ctypes.c_int), ("First_Part", ctypes.c_int),
                ("Npart", ctypes.c_int), ("Pos", ctypes.c_float * 3),
                ("Mass", ctypes.c_float), ("CoM", ctypes.c_float * 3),
                ("Dp", ctypes.c_float * 3)]


NODE_DTYPE = np.dtype([("DNext", "<i4"), ("Bitfield", "<u4"), ("DUp", "<i4"),
                       ("Npart", "<i4"), ("Pos", "<f8", 3), ("Mass", "<f8"),
                       ("CoM", "<f8", 3), ("Dp", "<f8", 3)])
TOP_DTYPE = np.dtype([("Key", "<u8"), ("Target", "<i4"), ("Level", "<i4"),
                      ("First_Part", "<i4"), ("Npart", "<i4"),
                      ("Pos", "<f4", 3), ("Mass", "<f4"), ("CoM", "<f4", 3),
                      ("Dp", "<f4", 3)])
assert NODE_DTYPE.itemsize == ctypes.sizeof(OrcNode) == 96
assert TOP_DTYPE.itemsize == ctypes.sizeof(OrcTopNode) == 64


def _declare(L):
    L.orc_peano_keys.argtypes = [ctypes.c_size_t, _dp, _dp, _dp, _dp,
                                 ctypes.c_double, _vp]
    L.orc_reversed_peano_keys.argtypes = L.orc_peano_keys.argtypes
    L.orc_qsort_index.argtypes = [ctypes.c_size_t, _vp, _vp, ctypes.c_int]
    L.orc_select.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
    L.orc_select.restype = ctypes.c_double
    L.orc_domain_nonperiodic.argtypes = [ctypes.c_size_t, _dp, _dp, _dp, _dp,
                                         _dp, _dp]
    L.orc_ewald_table.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_int]
    L.orc_sim_create.argtypes = [ctypes.POINTER(OrcParams), ctypes.c_size_t,
                                 ctypes.POINTER(OrcParticles)]
    L.orc_sim_create.restype = _vp
    L.orc_sim_free.argtypes = [_vp]
    L.orc_sim_set_ewald_table.argtypes = [_vp, _dp]
    L.orc_sim_set_active_bin.argtypes = [_vp, ctypes.c_int]
    L.orc_sim_nactive.argtypes = [_vp]
    L.orc_sim_nactive.restype = ctypes.c_int
    L.orc_sim_active.argtypes = [_vp]
    L.orc_sim_active.restype = _ip
    for f in ("orc_sim_setup_domain", "orc_sim_domain_decomposition",
              "orc_sim_tree_build"):
        getattr(L, f).argtypes = [_vp]
    L.orc_sim_set_top_nodes.argtypes = [_vp, _vp, ctypes.c_int]
    L.orc_sim_tree_acceleration.argtypes = [_vp, ctypes.c_int]
    L.orc_sim_compute_acceleration.argtypes = [_vp, ctypes.c_int]
    L.orc_sim_ewald_correction.argtypes = [_vp, _dp, _dp]
    L.orc_sim_periodic_nearest.argtypes = [_vp, _dp]
    L.orc_sim_domain.argtypes = [_vp, _dp, _dp, _dp]
    L.orc_sim_ntop.argtypes = [_vp]
    L.orc_sim_ntop.restype = ctypes.c_int
    L.orc_sim_top.argtypes = [_vp]
    L.orc_sim_top.restype = ctypes.POINTER(OrcTopNode)
    L.orc_sim_nnodes.argtypes = [_vp]
    L.orc_sim_nnodes.restype = ctypes.c_int
    L.orc_sim_tree.argtypes = [_vp]
    L.orc_sim_tree.restype = ctypes.POINTER(OrcNode)
    L.orc_sim_ndumped.argtypes = [_vp]
    L.orc_sim_ndumped.restype = ctypes.c_int
    L.orc_sim_epsilon.argtypes = [_vp]
    L.orc_sim_epsilon.restype = ctypes.c_double
    L.orc_direct_sum.argtypes = [_vp, ctypes.c_int, _ip, _dp]
    L.orc_domain_targets.argtypes = [_vp, _dp, ctypes.c_size_t, ctypes.c_int,
                                     ctypes.c_double, ctypes.c_double, _vp,
                                     ctypes.c_int, _ip, _ip]
    L.orc_domain_targets.restype = ctypes.c_int
    L.orc_sim_domain_decomposition_targets.argtypes = [
        _vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _ip, _ip]
    L.orc_sim_domain_decomposition_targets.restype = ctypes.c_int


def _d(a):
    return a.ctypes.data_as(_dp)


def keys_to_u128(k2):
    """(n,2) uint64 [lo, hi] -> list of python ints (for small checks)."""
    return [int(hi) << 64 | int(lo) for lo, hi in k2]


def peano_keys(x, y, z, origin, size, reversed_=False):
    x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
    o = np.ascontiguousarray(origin, dtype=np.float64)
    out = np.zeros((x.size, 2), dtype=np.uint64)
    f = lib().orc_reversed_peano_keys if reversed_ else lib().orc_peano_keys
    f(x.size, _d(x), _d(y), _d(z), _d(o), float(size),
      out.ctypes.data_as(_vp))
    return out


def qsort_index(keys2, nthreads=1):
    keys2 = np.ascontiguousarray(keys2, dtype=np.uint64)
    perm = np.zeros(keys2.shape[0], dtype=np.uint64)
    lib().orc_qsort_index(keys2.shape[0], keys2.ctypes.data_as(_vp),
                          perm.ctypes.data_as(_vp), nthreads)
    return perm.astype(np.int64)


def select(k, data):
    buf = np.ascontiguousarray(data, dtype=np.float64).copy()
    return lib().orc_select(int(k), buf.size, _d(buf))


def domain_nonperiodic(x, y, z):
    x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
    c = np.zeros(3)
    s = np.zeros(1)
    o = np.zeros(3)
    lib().orc_domain_nonperiodic(x.size, _d(x), _d(y), _d(z), _d(c), _d(s),
                                 _d(o))
    return c, float(s[0]), o


def domain_targets(short_keys, cost, ntarget, part_alloc_factor=1.1,
                   imbal_ceil=0.05):
    """Domain_Decomposition with NTarget = ntarget (domain.c:67-145):
    short keys (upper 64 bits, ascending) and P.Cost in -> top-node records
    in descending key order (Target = -task - 1, Pos 0) and the loop's
    iteration count.  Raises on the oracle's error returns."""
    k = np.ascontiguousarray(short_keys, dtype=np.uint64)
    c = np.ascontiguousarray(cost, dtype=np.float64)
    out = np.zeros(65536, dtype=TOP_DTYPE)
    nt = ctypes.c_int(0)
    it = ctypes.c_int(0)
    rc = lib().orc_domain_targets(k.ctypes.data_as(_vp), _d(c), k.size,
                                  int(ntarget), float(part_alloc_factor),
                                  float(imbal_ceil), out.ctypes.data_as(_vp),
                                  out.size, ctypes.byref(nt), ctypes.byref(it))
    if rc != 0:
        raise ValueError("orc_domain_targets returned %d" % rc)
    return out[:nt.value].copy(), it.value


def ewald_table(nthreads=None):
    """4 x 65^3 table (Fx, Fy, Fz, Fp) in box units, unnormalised."""
    nthreads = nthreads or os.cpu_count() or 1
    t = np.zeros((4, NE1, NE1, NE1))
    lib().orc_ewald_table(_d(t[0]), _d(t[1]), _d(t[2]), _d(t[3]), nthreads)
    return t


def cached_ewald_table(path=None):
    """The table cached on disk next to the library (ewald_tables.dat's
    role in Gravity/periodic.c:303-342)."""
    path = path or os.path.join(_HERE, "_ewald_table_f64.npy")
    if os.path.exists(path):
        t = np.load(path)
        if t.shape == (4, NE1, NE1, NE1):
            return t
    t = ewald_table()
    np.save(path, t)
    return t


class Sim:
    """The reference's global particle state and force path, on the CPU."""

    def __init__(self, pos, mass, vel=None, periodic=False, boxsize=1.0,
                 gravity=1.0, softening=0.01, theta_bh=0.1, alpha_rel=0.02,
                 vector_size=64, nthreads=None, time_bin=None,
                 ewald_table=None):
        n = pos.shape[0]
        self.n = n
        self.pos = [np.ascontiguousarray(pos[:, d], dtype=np.float64).copy()
                    for d in range(3)]
        v = np.zeros((n, 3)) if vel is None else vel
        self.vel = [np.ascontiguousarray(v[:, d], dtype=np.float64).copy()
                    for d in range(3)]
        self.acc = [np.zeros(n) for _ in range(3)]
        self.grav_acc = [np.zeros(n) for _ in range(3)]
        self.mass = np.ascontiguousarray(mass, dtype=np.float64).copy()
        self.last_acc_mag = np.zeros(n)
        self.cost = np.zeros(n)
        self.tree_parent = np.zeros(n, dtype=np.int32)
        self.key = np.zeros((n, 2), dtype=np.uint64)
        self.id = np.arange(n, dtype=np.uint32)
        self.time_bin = (np.zeros(n, dtype=np.int32) if time_bin is None
                         else np.ascontiguousarray(time_bin, dtype=np.int32))
        P = OrcParticles()
        for d in range(3):
            P.pos[d] = _d(self.pos[d])
            P.vel[d] = _d(self.vel[d])
            P.acc[d] = _d(self.acc[d])
            P.grav_acc[d] = _d(self.grav_acc[d])
        P.mass = _d(self.mass)
        P.last_acc_mag = _d(self.last_acc_mag)
        P.cost = _d(self.cost)
        P.tree_parent = self.tree_parent.ctypes.data_as(_ip)
        P.key = self.key.ctypes.data_as(_vp)
        P.id = self.id.ctypes.data_as(_vp)
        P.time_bin = self.time_bin.ctypes.data_as(_ip)
        self._P = P
        self.params = OrcParams(int(periodic), boxsize, gravity, softening,
                                theta_bh, alpha_rel, vector_size,
                                nthreads or os.cpu_count() or 1)
        self.periodic = periodic
        self._s = lib().orc_sim_create(ctypes.byref(self.params), n,
                                       ctypes.byref(P))
        if periodic:
            t = cached_ewald_table() if ewald_table is None else ewald_table
            t = np.ascontiguousarray(t, dtype=np.float64)
            lib().orc_sim_set_ewald_table(self._s, _d(t))

    def __del__(self):
        if getattr(self, "_s", None):
            lib().orc_sim_free(self._s)
            self._s = None

    # reference-named steps
    def setup_domain(self):
        lib().orc_sim_setup_domain(self._s)

    def domain_decomposition(self):
        lib().orc_sim_domain_decomposition(self._s)

    def domain_decomposition_targets(self, ntarget, part_alloc_factor=1.1,
                                     imbal_ceil=0.05):
        """Domain_Decomposition with NTarget = ntarget; returns (tasks per
        top node, loop passes)."""
        tasks = np.zeros(65536, dtype=np.int32)
        it = ctypes.c_int(0)
        rc = lib().orc_sim_domain_decomposition_targets(
            self._s, int(ntarget), float(part_alloc_factor),
            float(imbal_ceil), tasks.ctypes.data_as(_ip), ctypes.byref(it))
        if rc != 0:
            raise ValueError("orc_sim_domain_decomposition_targets: %d" % rc)
        return tasks[:lib().orc_sim_ntop(self._s)].copy(), it.value

    def tree_build(self):
        lib().orc_sim_tree_build(self._s)

    def set_top_nodes(self, rec):
        rec = np.ascontiguousarray(rec, dtype=TOP_DTYPE)
        lib().orc_sim_set_top_nodes(self._s, rec.ctypes.data_as(_vp), len(rec))

    def tree_acceleration(self, use_bh):
        lib().orc_sim_tree_acceleration(self._s, int(use_bh))

    def compute_acceleration(self, prepare_step=True):
        lib().orc_sim_compute_acceleration(self._s, int(prepare_step))

    def set_active_bin(self, b):
        lib().orc_sim_set_active_bin(self._s, int(b))

    @property
    def active(self):
        n = lib().orc_sim_nactive(self._s)
        p = lib().orc_sim_active(self._s)
        return np.ctypeslib.as_array(p, shape=(n,)).copy()

    # views
    @property
    def domain(self):
        s = np.zeros(1)
        o = np.zeros(3)
        c = np.zeros(3)
        lib().orc_sim_domain(self._s, _d(s), _d(o), _d(c))
        return float(s[0]), o, c

    @property
    def top(self):
        n = lib().orc_sim_ntop(self._s)
        p = lib().orc_sim_top(self._s)
        buf = ctypes.string_at(p, n * TOP_DTYPE.itemsize)
        return np.frombuffer(buf, dtype=TOP_DTYPE).copy()

    @property
    def tree(self):
        n = lib().orc_sim_nnodes(self._s)
        p = lib().orc_sim_tree(self._s)
        if n == 0:
            return np.zeros(0, dtype=NODE_DTYPE)
        buf = ctypes.string_at(p, n * NODE_DTYPE.itemsize)
        return np.frombuffer(buf, dtype=NODE_DTYPE).copy()

    @property
    def ndumped(self):
        return lib().orc_sim_ndumped(self._s)

    @property
    def epsilon(self):
        return lib().orc_sim_epsilon(self._s)

    def positions(self):
        return np.stack(self.pos, axis=1)

    def accelerations(self):
        return np.stack(self.acc, axis=1)

    def ewald_correction(self, dr):
        dr = np.ascontiguousarray(dr, dtype=np.float64)
        f = np.zeros(3)
        lib().orc_sim_ewald_correction(self._s, _d(dr), _d(f))
        return f

    def direct_sum(self, idx):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        out = np.zeros((idx.size, 3))
        lib().orc_direct_sum(self._s, idx.size, idx.ctypes.data_as(_ip),
                             _d(out))
        return out
