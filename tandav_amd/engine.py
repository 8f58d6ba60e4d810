"""ctypes binding of libtandav_gpu.so and the reference-shaped host API.

Method names follow the reference functions they replace; each docstring
cites the reference file:line.  Arrays cross the boundary as plain host
pointers and sizes (numpy float64 / int32 / uint32 / uint64 buffers).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtandav_gpu.so")

NODE_DTYPE = np.dtype([("DNext", "<i4"), ("Bitfield", "<u4"), ("DUp", "<i4"),
                       ("Npart", "<i4"), ("Pos", "<f8", 3), ("Mass", "<f8"),
                       ("CoM", "<f8", 3), ("Dp", "<f8", 3)])
TOP_DTYPE = np.dtype([("Key", "<u8"), ("Target", "<i4"), ("Level", "<i4"),
                      ("First_Part", "<i4"), ("Npart", "<i4"),
                      ("Pos", "<f4", 3), ("Mass", "<f4"), ("CoM", "<f4", 3),
                      ("Dp", "<f4", 3)])

STAGES = ("keys", "sort", "reorder", "top_nodes", "build", "walk", "ewald",
          "kick")


class TDVError(RuntimeError):
    pass


class TdvParams(ctypes.Structure):
    _fields_ = [("periodic", ctypes.c_int), ("boxsize", ctypes.c_double),
                ("gravity", ctypes.c_double), ("softening", ctypes.c_double),
                ("theta_bh", ctypes.c_double), ("alpha_rel", ctypes.c_double),
                ("vector_size", ctypes.c_int), ("compat", ctypes.c_int),
                ("device", ctypes.c_int), ("precision", ctypes.c_int),
                ("ewald_size", ctypes.c_double),
                ("ewald_accuracy", ctypes.c_double),
                ("fp32_order", ctypes.c_int),
                ("fp32_criterion", ctypes.c_int),
                ("fp32_carry", ctypes.c_int),
                ("fp32_waves", ctypes.c_int),
                ("fp32_queue", ctypes.c_int),
                ("fp32_split", ctypes.c_int),
                ("fp32_bundle_min", ctypes.c_int)]


class TdvTimeline(ctypes.Structure):
    """tdv_timeline: Time.Begin/End and the comoving background."""
    _fields_ = [("comoving", ctypes.c_int), ("time_begin", ctypes.c_double),
                ("time_end", ctypes.c_double), ("hubble", ctypes.c_double),
                ("omega_0", ctypes.c_double),
                ("omega_matter", ctypes.c_double),
                ("omega_lambda", ctypes.c_double),
                ("omega_rad", ctypes.c_double)]


class TdvTimestepParams(ctypes.Structure):
    """tdv_timestep_params: inputs of set_new_particle_timebins."""
    _fields_ = [("time_int_accuracy", ctypes.c_double),
                ("softening", ctypes.c_double),
                ("dt_max_global", ctypes.c_double),
                ("max_active_bin", ctypes.c_int),
                ("expansion_factor", ctypes.c_double),
                ("hubble_parameter", ctypes.c_double),
                ("grav_accel_factor", ctypes.c_double)]


def timeline(comoving=False, time_begin=0.0, time_end=1.0, hubble=1.0,
             omega_matter=0.3, omega_lambda=0.7, omega_rad=0.0,
             omega_0=None):
    if omega_0 is None:
        omega_0 = omega_matter + omega_lambda + omega_rad
    return TdvTimeline(int(comoving), float(time_begin), float(time_end),
                       float(hubble), float(omega_0), float(omega_matter),
                       float(omega_lambda), float(omega_rad))


def comoving_tables(tl):
    """setup_kick_drift_factors (comoving.c:109-153) on the host: the 512
    knots and the kick and drift integrals (no GPU needed)."""
    L = load_library()
    a, k, d = (np.zeros(512) for _ in range(3))
    rc = L.tdv_comoving_tables(ctypes.byref(tl), _d(a), _d(k), _d(d))
    if rc:
        raise TDVError("tdv_comoving_tables: error %d" % rc)
    return a, k, d


_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_ctx = ctypes.c_void_p
_lib = None


def load_library(path=LIB_PATH):
    """Load the HIP engine; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise TDVError(
            "libtandav_gpu.so is missing: build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` "
            "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    L = ctypes.CDLL(path)
    L.tdv_create.argtypes = [ctypes.POINTER(TdvParams), ctypes.POINTER(_ctx)]
    L.tdv_destroy.argtypes = [_ctx]
    L.tdv_destroy.restype = None
    L.tdv_last_error.argtypes = [_ctx]
    L.tdv_last_error.restype = ctypes.c_char_p
    L.tdv_set_particles.argtypes = [_ctx, _sz, ctypes.POINTER(_dp),
                                    ctypes.POINTER(_dp), _dp,
                                    ctypes.POINTER(_dp), _dp, _vp, _vp]
    L.tdv_get_particles.argtypes = [_ctx, ctypes.POINTER(_dp),
                                    ctypes.POINTER(_dp), _dp,
                                    ctypes.POINTER(_dp), ctypes.POINTER(_dp),
                                    _dp, _dp, _vp, _vp, _vp, _vp]
    for f in ("tdv_setup_domain", "tdv_domain_decomposition",
              "tdv_tree_build"):
        getattr(L, f).argtypes = [_ctx]
    L.tdv_set_domain.argtypes = [_ctx, ctypes.c_double, _dp]
    L.tdv_sort_particles.argtypes = [_ctx]
    L.tdv_set_top_nodes.argtypes = [_ctx, _vp, ctypes.c_int]
    _ip = ctypes.POINTER(ctypes.c_int)
    L.tdv_domain_decomposition_targets.argtypes = [
        _ctx, ctypes.c_int, ctypes.c_double, ctypes.c_double, _ip,
        ctypes.c_int, _ip, _ip]
    L.tdv_tree_acceleration.argtypes = [_ctx, ctypes.c_int]
    L.tdv_compute_acceleration.argtypes = [_ctx, ctypes.c_int]
    L.tdv_forcetest.argtypes = [_ctx, ctypes.c_int, _vp, _dp]
    L.tdv_forcetest_exact.argtypes = [_ctx, ctypes.c_int, _vp, _dp]
    L.tdv_forcetest_exact_pairs.argtypes = [_ctx, ctypes.c_int, _vp, _dp]
    L.tdv_peano_keys.argtypes = [_ctx, _sz, _dp, _dp, _dp, _dp,
                                 ctypes.c_double, ctypes.c_int, _vp]
    L.tdv_sort_keys.argtypes = [_ctx, _sz, _vp, _vp]
    L.tdv_sort_keys_noalloc.argtypes = [_sz, _vp, _vp, _vp]
    L.tdv_ewald_correction.argtypes = [_ctx, _sz, _dp, _dp]
    L.tdv_peano_keys_host.argtypes = [_sz, _dp, _dp, _dp, _dp,
                                      ctypes.c_double, ctypes.c_int, _vp]
    L.tdv_ewald_correction_host.argtypes = [_ctx, _sz, _dp, _dp]
    L.tdv_get_ewald_table.argtypes = [_ctx, _dp]
    L.tdv_set_ewald_table.argtypes = [_ctx, _dp]
    L.tdv_get_domain.argtypes = [_ctx, _dp, _dp, _dp]
    L.tdv_get_top_nodes.argtypes = [_ctx, _vp, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_int)]
    L.tdv_get_tree.argtypes = [_ctx, _vp, _sz, ctypes.POINTER(_sz)]
    L.tdv_stage_times.argtypes = [_ctx, ctypes.POINTER(ctypes.c_float)]
    L.tdv_interaction_counts.argtypes = [_ctx,
                                         ctypes.POINTER(ctypes.c_uint64)]
    L.tdv_interaction_kinds.argtypes = [_ctx,
                                        ctypes.POINTER(ctypes.c_uint64)]
    L.tdv_get_keys.argtypes = [_ctx, _vp, ctypes.c_int]
    L.tdv_permute_host_columns.argtypes = [_ctx, ctypes.c_int,
                                           ctypes.POINTER(_vp),
                                           ctypes.POINTER(ctypes.c_int)]
    L.tdv_update_particles.argtypes = [_ctx, ctypes.POINTER(_dp), _dp,
                                       ctypes.POINTER(_dp), _dp]
    L.tdv_set_active.argtypes = [_ctx, ctypes.c_int, _vp]
    L.tdv_set_tree_moments.argtypes = [_ctx, _vp, _sz, _vp, ctypes.c_int]
    L.tdv_get_ewald_derivs.argtypes = [_ctx, _dp]
    L.tdv_get_ewald_jacobian.argtypes = [_ctx, _dp]
    L.tdv_get_walk32_quadrupoles.argtypes = [_ctx, _vp, _vp, _vp]
    L.tdv_kick_drift.argtypes = [_ctx, ctypes.c_double, ctypes.c_double]
    L.tdv_make_active_list.argtypes = [_ctx, ctypes.c_int, _vp]
    L.tdv_set_fp32_options.argtypes = [_ctx, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int]
    L.tdv_capture_walk.argtypes = [_ctx]
    L.tdv_capture_substep_walk.argtypes = [_ctx]
    L.tdv_capture_step.argtypes = [_ctx, ctypes.c_double, ctypes.c_double]
    L.tdv_launch_step.argtypes = [_ctx]
    L.tdv_set_timeline.argtypes = [_ctx, ctypes.POINTER(TdvTimeline)]
    L.tdv_comoving_tables.argtypes = [ctypes.POINTER(TdvTimeline), _dp, _dp,
                                      _dp]
    L.tdv_set_integer_times.argtypes = [_ctx, _vp, _vp]
    L.tdv_get_integer_times.argtypes = [_ctx, _vp, _vp]
    L.tdv_kick_halfstep.argtypes = [_ctx, ctypes.c_int]
    L.tdv_tree_update_drift.argtypes = [_ctx, ctypes.c_double]
    L.tdv_drift_to_sync_point.argtypes = [_ctx, ctypes.c_uint32]
    L.tdv_set_new_timebins.argtypes = [_ctx,
                                       ctypes.POINTER(TdvTimestepParams),
                                       ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_int)]
    L.tdv_launch_walk.argtypes = [_ctx]
    L.tdv_write_snapshot.argtypes = [_ctx, ctypes.c_char_p,
                                     ctypes.POINTER(SnapshotHeader),
                                     ctypes.c_uint]
    L.tdv_read_snapshot.argtypes = [_ctx, ctypes.c_char_p,
                                    ctypes.POINTER(SnapshotHeader)]
    L.tdv_num_particles.argtypes = [_ctx]
    L.tdv_num_particles.restype = _sz
    _lib = L
    return L


def _d(a):
    return a.ctypes.data_as(_dp)


def peano_keys_host(x, y, z, origin, size, reversed_=False, lib=None):
    """Peano_Key / Reversed_Peano_Key (peano.c:122-276) computed on the host
    by the library's own transform (tdv_peano_keys_host): no device, no
    context.  Returns (n, 2) uint64 words (low, high)."""
    L = lib or load_library()
    x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
    o = np.ascontiguousarray(origin, dtype=np.float64)
    out = np.zeros((x.size, 2), np.uint64)
    rc = L.tdv_peano_keys_host(x.size, _d(x), _d(y), _d(z), _d(o),
                               float(size), int(reversed_),
                               out.ctypes.data_as(ctypes.c_void_p))
    if rc != 0:
        raise TDVError("tdv_peano_keys_host: error %d" % rc)
    return out


class SnapshotHeader(ctypes.Structure):
    """struct gadget_header (IO/io.h:23-42), 256 bytes."""
    _fields_ = [("Npart", ctypes.c_uint32 * 6),
                ("Massarr", ctypes.c_double * 6),
                ("Time", ctypes.c_double), ("Redshift", ctypes.c_double),
                ("Flag_Sfr", ctypes.c_int32), ("Flag_Feedback", ctypes.c_int32),
                ("Nall", ctypes.c_uint32 * 6),
                ("Flag_Cooling", ctypes.c_int32), ("Num_Files", ctypes.c_int32),
                ("Boxsize", ctypes.c_double), ("Omega0", ctypes.c_double),
                ("Omega_Lambda", ctypes.c_double),
                ("Hubble_Param", ctypes.c_double),
                ("Flag_Age", ctypes.c_int32), ("Flag_Metals", ctypes.c_int32),
                ("Nall_High_Word", ctypes.c_uint32 * 6),
                ("fill_bytes", ctypes.c_char * 64)]


assert ctypes.sizeof(SnapshotHeader) == 256


def _col3(cols):
    arr = (_dp * 3)()
    for d in range(3):
        arr[d] = _d(cols[d]) if cols is not None else None
    return arr


class Engine:
    """One GPU's force path with the reference's calling sequence."""

    def __init__(self, periodic=False, boxsize=1.0, gravity=1.0,
                 softening=0.01, theta_bh=0.1, alpha_rel=0.02, vector_size=64,
                 compat=True, device=0, ewald_table=None, precision="fp64",
                 ewald_size=0.125, ewald_accuracy=0.1, fp32_order=2,
                 fp32_criterion=2, fp32_carry=16, fp32_waves=0,
                 fp32_queue=0, fp32_split=0, fp32_bundle_min=0):
        self.lib = load_library()
        prec = {"fp64": 0, "fp32": 1}[precision]
        self.params = TdvParams(int(periodic), float(boxsize), float(gravity),
                                float(softening), float(theta_bh),
                                float(alpha_rel), int(vector_size),
                                int(compat), int(device), prec,
                                float(ewald_size), float(ewald_accuracy),
                                int(fp32_order), int(fp32_criterion),
                                int(fp32_carry), int(fp32_waves),
                                int(fp32_queue), int(fp32_split),
                                int(fp32_bundle_min))
        self._c = _ctx()
        rc = self.lib.tdv_create(ctypes.byref(self.params),
                                 ctypes.byref(self._c))
        if rc:
            raise TDVError("tdv_create failed (%d)" % rc)
        self.periodic = periodic
        self.n = 0
        if ewald_table is not None:
            self.set_ewald_table(ewald_table)

    def close(self):
        if getattr(self, "_c", None) and self._c.value:
            self.lib.tdv_destroy(self._c)
            self._c = _ctx()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ok(self, rc, what):
        if rc:
            msg = self.lib.tdv_last_error(self._c).decode()
            raise TDVError("%s: error %d: %s" % (what, rc, msg))

    # -- particle columns (particles.h:17-37) --------------------------------
    def set_particles(self, pos, mass, vel=None, acc=None, last_acc_mag=None,
                      time_bin=None, ids=None):
        n = pos.shape[0]
        p = [np.ascontiguousarray(pos[:, d], dtype=np.float64)
             for d in range(3)]
        v = None if vel is None else [
            np.ascontiguousarray(vel[:, d], dtype=np.float64)
            for d in range(3)]
        a = None if acc is None else [
            np.ascontiguousarray(acc[:, d], dtype=np.float64)
            for d in range(3)]
        m = np.ascontiguousarray(mass, dtype=np.float64)
        la = (None if last_acc_mag is None
              else np.ascontiguousarray(last_acc_mag, dtype=np.float64))
        tb = (None if time_bin is None
              else np.ascontiguousarray(time_bin, dtype=np.int32))
        ii = None if ids is None else np.ascontiguousarray(ids,
                                                           dtype=np.uint32)
        rc = self.lib.tdv_set_particles(
            self._c, n, _col3(p), _col3(v) if v else None, _d(m),
            _col3(a) if a else None, _d(la) if la is not None else None,
            tb.ctypes.data_as(_vp) if tb is not None else None,
            ii.ctypes.data_as(_vp) if ii is not None else None)
        self._ok(rc, "tdv_set_particles")
        self.n = n

    def get(self, *names):
        """Columns by name: pos vel mass acc grav_acc cost last_acc_mag
        tree_parent key id time_bin."""
        n = self.n
        want = set(names) if names else {
            "pos", "vel", "mass", "acc", "grav_acc", "cost", "last_acc_mag",
            "tree_parent", "key", "id", "time_bin"}
        out = {}
        cols = {}
        for k in ("pos", "vel", "acc", "grav_acc"):
            if k in want:
                cols[k] = [np.zeros(n) for _ in range(3)]
        one = {}
        for k in ("mass", "cost", "last_acc_mag"):
            if k in want:
                one[k] = np.zeros(n)
        tp = np.zeros(n, np.int32) if "tree_parent" in want else None
        key = np.zeros((n, 2), np.uint64) if "key" in want else None
        ids = np.zeros(n, np.uint32) if "id" in want else None
        tb = np.zeros(n, np.int32) if "time_bin" in want else None
        rc = self.lib.tdv_get_particles(
            self._c, _col3(cols["pos"]) if "pos" in cols else None,
            _col3(cols["vel"]) if "vel" in cols else None,
            _d(one["mass"]) if "mass" in one else None,
            _col3(cols["acc"]) if "acc" in cols else None,
            _col3(cols["grav_acc"]) if "grav_acc" in cols else None,
            _d(one["cost"]) if "cost" in one else None,
            _d(one["last_acc_mag"]) if "last_acc_mag" in one else None,
            tp.ctypes.data_as(_vp) if tp is not None else None,
            key.ctypes.data_as(_vp) if key is not None else None,
            ids.ctypes.data_as(_vp) if ids is not None else None,
            tb.ctypes.data_as(_vp) if tb is not None else None)
        self._ok(rc, "tdv_get_particles")
        for k, v in cols.items():
            out[k] = np.stack(v, 1)
        out.update(one)
        if tp is not None:
            out["tree_parent"] = tp
        if key is not None:
            out["key"] = key
        if ids is not None:
            out["id"] = ids
        if tb is not None:
            out["time_bin"] = tb
        return out

    # -- the force path --------------------------------------------------------
    def setup_domain(self):
        """Setup_Domain_Decomposition (domain.c:151-189)."""
        self._ok(self.lib.tdv_setup_domain(self._c), "tdv_setup_domain")

    def domain_decomposition(self):
        """Domain_Decomposition (domain.c:67-145)."""
        self._ok(self.lib.tdv_domain_decomposition(self._c),
                 "tdv_domain_decomposition")

    def domain_decomposition_targets(self, ntarget, part_alloc_factor=1.1,
                                     imbal_ceil=0.05):
        """Domain_Decomposition (domain.c:67-145) with NTarget = ntarget:
        top nodes refined until every target's share of P.Cost is within
        imbal_ceil (DOMAIN_IMBAL_CEIL) of the mean.  Returns (tasks per top
        node, loop passes)."""
        tasks = np.zeros(65536, dtype=np.int32)
        nt = ctypes.c_int(0)
        it = ctypes.c_int(0)
        ip = ctypes.POINTER(ctypes.c_int)
        self._ok(self.lib.tdv_domain_decomposition_targets(
            self._c, int(ntarget), float(part_alloc_factor),
            float(imbal_ceil), tasks.ctypes.data_as(ip), tasks.size,
            ctypes.byref(nt), ctypes.byref(it)),
            "tdv_domain_decomposition_targets")
        return tasks[:nt.value].copy(), it.value

    def write_snapshot(self, path, npart=None, massarr=(0.0,) * 6, time=0.0,
                       redshift=0.0, boxsize=0.0, omega0=0.0,
                       omega_lambda=0.0, hubble_param=0.0, acc=True,
                       gacc=True):
        """Write_Snapshot (IO/write_snapshot.c) for one file: HEAD, POS,
        VEL, ID, MASS, and ACC / GACC (OUTPUT_TOTAL_ACCELERATION,
        OUTPUT_PARTIAL_ACCELERATIONS).  npart: particles per type (default
        all of type 1)."""
        h = SnapshotHeader()
        counts = [0, self.n, 0, 0, 0, 0] if npart is None else list(npart)
        for t in range(6):
            h.Npart[t] = int(counts[t])
            h.Massarr[t] = float(massarr[t])
        h.Time, h.Redshift, h.Boxsize = float(time), float(redshift), \
            float(boxsize)
        h.Omega0, h.Omega_Lambda, h.Hubble_Param = float(omega0), \
            float(omega_lambda), float(hubble_param)
        blocks = (1 if acc else 0) | (2 if gacc else 0)
        self._ok(self.lib.tdv_write_snapshot(self._c, os.fsencode(path),
                                             ctypes.byref(h), blocks),
                 "tdv_write_snapshot")

    def read_snapshot(self, path):
        """Read_Snapshot (IO/read_snapshot.c) into this engine; returns the
        (first file's) header as a dict."""
        h = SnapshotHeader()
        self._ok(self.lib.tdv_read_snapshot(self._c, os.fsencode(path),
                                            ctypes.byref(h)),
                 "tdv_read_snapshot")
        self.n = int(self.lib.tdv_num_particles(self._c))
        out = {}
        for name, _ in SnapshotHeader._fields_:
            if name == "fill_bytes":
                continue
            v = getattr(h, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out

    def set_domain(self, size, origin):
        """Domain (domain.h:58-63) computed by the caller."""
        o = np.ascontiguousarray(origin, dtype=np.float64)
        self._ok(self.lib.tdv_set_domain(self._c, float(size), _d(o)),
                 "tdv_set_domain")

    def sort_particles(self):
        """Sort_Particles_By_Peano_Key (peano.c:20-48) in the set domain."""
        self._ok(self.lib.tdv_sort_particles(self._c), "tdv_sort_particles")

    def set_top_nodes(self, rec):
        """D[].TNode from the caller's decomposition (domain.c:220-250)."""
        rec = np.ascontiguousarray(rec, dtype=TOP_DTYPE)
        self._ok(self.lib.tdv_set_top_nodes(self._c, rec.ctypes.data_as(_vp),
                                            len(rec)), "tdv_set_top_nodes")

    def tree_build(self):
        """Gravity_Tree_Build (Gravity/tree_build.c:47-122)."""
        self._ok(self.lib.tdv_tree_build(self._c), "tdv_tree_build")

    def tree_acceleration(self, use_bh):
        """Gravity_Tree_Acceleration (+ Gravity_Tree_Periodic)."""
        self._ok(self.lib.tdv_tree_acceleration(self._c, int(use_bh)),
                 "tdv_tree_acceleration")

    def compute_acceleration(self, prepare_step=True):
        """Compute_Acceleration (accel.c:15-71)."""
        self._ok(self.lib.tdv_compute_acceleration(self._c, int(prepare_step)),
                 "tdv_compute_acceleration")

    def capture_walk(self):
        """Capture the regular step's FP32 walk as a HIP graph."""
        self._ok(self.lib.tdv_capture_walk(self._c), "tdv_capture_walk")

    def capture_substep_walk(self):
        """Capture the FP32 walk of substeps that keep the tree: the walk
        records refreshed, the quadrupoles of the last eager walk kept."""
        self._ok(self.lib.tdv_capture_substep_walk(self._c),
                 "tdv_capture_substep_walk")

    def launch_walk(self):
        """Replay the captured walk."""
        self._ok(self.lib.tdv_launch_walk(self._c), "tdv_launch_walk")

    def capture_step(self, kick=0.0, drift=0.0):
        """Capture the whole regular step (decomposition, build, FP32 walk,
        kick-drift with these factors) as one HIP graph."""
        self._ok(self.lib.tdv_capture_step(self._c, float(kick),
                                           float(drift)), "tdv_capture_step")

    def launch_step(self):
        """Replay the captured step."""
        self._ok(self.lib.tdv_launch_step(self._c), "tdv_launch_step")

    def forcetest_exact(self, idx, pairs=False):
        """forcetest with the exact periodic force (the Ewald lattice sum
        the table samples) instead of the table's interpolation; pairs=True
        evaluates the reference's sums per pair (slow, the pin of the
        split form)."""
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        out = np.zeros((idx.size, 3))
        f = (self.lib.tdv_forcetest_exact_pairs if pairs
             else self.lib.tdv_forcetest_exact)
        self._ok(f(self._c, idx.size, idx.ctypes.data_as(_vp), _d(out)),
                 "tdv_forcetest_exact")
        return out

    def forcetest(self, idx):
        """Gravity_Forcetest (Gravity/forcetest.c:24-174): FP64 direct sum
        (+ Ewald correction per pair when periodic) for particles idx in the
        current (PH) order."""
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        out = np.zeros((idx.size, 3))
        self._ok(self.lib.tdv_forcetest(self._c, idx.size,
                                        idx.ctypes.data_as(_vp), _d(out)),
                 "tdv_forcetest")
        return out

    def keys(self, reversed_=True):
        """P.Key of the last sort: standard (reversed_=False) or reversed
        (Reverse_Peano_Keys, peano.c:84-97), (n, 2) uint64 words (lo, hi)."""
        out = np.zeros((self.n, 2), np.uint64)
        self._ok(self.lib.tdv_get_keys(self._c, out.ctypes.data_as(_vp),
                                       int(reversed_)), "tdv_get_keys")
        return out

    def permute_columns(self, *cols):
        """Apply the last sort's permutation to host columns (4- or 8-byte
        elements) in place (reorder_collisionless_particles, peano.c:50-82)."""
        cs = [np.ascontiguousarray(c) for c in cols]
        for c, orig in zip(cs, cols):
            if c is not orig or c.itemsize not in (4, 8) or c.size != self.n:
                raise TDVError("permute_columns: contiguous 4/8-byte columns "
                               "of n elements")
        ptrs = (_vp * len(cs))(*[c.ctypes.data_as(_vp) for c in cs])
        sizes = (ctypes.c_int * len(cs))(*[c.itemsize for c in cs])
        self._ok(self.lib.tdv_permute_host_columns(self._c, len(cs), ptrs,
                                                   sizes),
                 "tdv_permute_host_columns")

    def update_particles(self, pos, mass=None, acc=None, last_acc_mag=None):
        """Drifted positions (and masses, accelerations, |a_prev|) in the
        current particle order; keys, top nodes and tree are kept."""
        p = [np.ascontiguousarray(pos[:, d], dtype=np.float64)
             for d in range(3)]
        a = None if acc is None else [
            np.ascontiguousarray(acc[:, d], dtype=np.float64)
            for d in range(3)]
        m = None if mass is None else np.ascontiguousarray(mass, np.float64)
        la = (None if last_acc_mag is None
              else np.ascontiguousarray(last_acc_mag, dtype=np.float64))
        self._ok(self.lib.tdv_update_particles(
            self._c, _col3(p), _d(m) if m is not None else None,
            _col3(a) if a else None, _d(la) if la is not None else None),
            "tdv_update_particles")

    def set_active(self, idx):
        """Active_Particle_List (timestep.c:280-299)."""
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        self._ok(self.lib.tdv_set_active(self._c, idx.size,
                                         idx.ctypes.data_as(_vp)),
                 "tdv_set_active")

    def set_tree_moments(self, tree, top):
        """Tree[] / D[].TNode CoM and Mass moved on the host
        (Gravity/tree_update.c:12-104)."""
        tree = np.ascontiguousarray(tree, dtype=NODE_DTYPE)
        top = np.ascontiguousarray(top, dtype=TOP_DTYPE)
        self._ok(self.lib.tdv_set_tree_moments(
            self._c, tree.ctypes.data_as(_vp), len(tree),
            top.ctypes.data_as(_vp), len(top)), "tdv_set_tree_moments")

    # -- element-wise functions -------------------------------------------------
    def peano_keys(self, x, y, z, origin, size, reversed_=False):
        """Peano_Key / Reversed_Peano_Key (peano.c:122-276)."""
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
        o = np.ascontiguousarray(origin, dtype=np.float64)
        out = np.zeros((x.size, 2), np.uint64)
        self._ok(self.lib.tdv_peano_keys(self._c, x.size, _d(x), _d(y), _d(z),
                                         _d(o), float(size), int(reversed_),
                                         out.ctypes.data_as(_vp)),
                 "tdv_peano_keys")
        return out

    def peano_keys_host(self, x, y, z, origin, size, reversed_=False):
        """The same keys computed on the host (tdv_peano_keys_host)."""
        return peano_keys_host(x, y, z, origin, size, reversed_, self.lib)

    def sort_keys(self, keys2):
        """Qsort_Index(perm, keys, n, 16, cmp_peanoKeys) (sort.c:90-116)."""
        k = np.ascontiguousarray(keys2, dtype=np.uint64)
        perm = np.zeros(k.shape[0], np.uint32)
        self._ok(self.lib.tdv_sort_keys(self._c, k.shape[0],
                                        k.ctypes.data_as(_vp),
                                        perm.ctypes.data_as(_vp)),
                 "tdv_sort_keys")
        return perm

    def ewald_correction(self, dr):
        """Ewald_Correction (Gravity/periodic.c:35-113), batched."""
        dr = np.ascontiguousarray(dr, dtype=np.float64).reshape(-1, 3)
        f = np.zeros_like(dr)
        self._ok(self.lib.tdv_ewald_correction(self._c, dr.shape[0], _d(dr),
                                               _d(f)), "tdv_ewald_correction")
        return f

    def ewald_correction_host(self, dr):
        """The same lookup on the host (tdv_ewald_correction_host)."""
        dr = np.ascontiguousarray(dr, dtype=np.float64).reshape(-1, 3)
        f = np.zeros_like(dr)
        rc = self.lib.tdv_ewald_correction_host(self._c, dr.shape[0], _d(dr),
                                                _d(f))
        if rc != 0:
            raise TDVError("tdv_ewald_correction_host: error %d (no table)"
                           % rc)
        return f

    def ewald_table(self):
        t = np.zeros((4, 65, 65, 65))
        self._ok(self.lib.tdv_get_ewald_table(self._c, _d(t)),
                 "tdv_get_ewald_table")
        return t

    def ewald_derivs(self):
        """Second derivatives d_b d_c E_a of the Ewald correction at the
        table nodes, (10, 65, 65, 65), per unit box, components
        xxx yyy zzz xxy xxz xyy yyz xzz yzz xyz."""
        s = np.zeros((10, 65, 65, 65))
        self._ok(self.lib.tdv_get_ewald_derivs(self._c, _d(s)),
                 "tdv_get_ewald_derivs")
        return s

    def ewald_jacobian(self):
        """First derivatives d_b E_a of the Ewald correction at the table
        nodes, (6, 65, 65, 65), per unit box, components xx yy zz xy xz yz."""
        j = np.zeros((6, 65, 65, 65))
        self._ok(self.lib.tdv_get_ewald_jacobian(self._c, _d(j)),
                 "tdv_get_ewald_jacobian")
        return j

    def set_ewald_table(self, t):
        t = np.ascontiguousarray(t, dtype=np.float64)
        self._ok(self.lib.tdv_set_ewald_table(self._c, _d(t)),
                 "tdv_set_ewald_table")

    # -- views --------------------------------------------------------------------
    @property
    def domain(self):
        s = np.zeros(1)
        o = np.zeros(3)
        c = np.zeros(3)
        self._ok(self.lib.tdv_get_domain(self._c, _d(s), _d(o), _d(c)),
                 "tdv_get_domain")
        return float(s[0]), o, c

    def top_nodes(self):
        n = ctypes.c_int()
        self._ok(self.lib.tdv_get_top_nodes(self._c, None, 0, ctypes.byref(n)),
                 "tdv_get_top_nodes")
        buf = np.zeros(n.value, dtype=TOP_DTYPE)
        self._ok(self.lib.tdv_get_top_nodes(self._c, buf.ctypes.data_as(_vp),
                                            n.value, ctypes.byref(n)),
                 "tdv_get_top_nodes")
        return buf

    def num_nodes(self):
        """Tree nodes of the last build (Tree_Node records, terminators
        included), without copying them."""
        n = _sz()
        self._ok(self.lib.tdv_get_tree(self._c, None, 0, ctypes.byref(n)),
                 "tdv_get_tree")
        return n.value

    def tree(self):
        n = _sz()
        self._ok(self.lib.tdv_get_tree(self._c, None, 0, ctypes.byref(n)),
                 "tdv_get_tree")
        buf = np.zeros(n.value, dtype=NODE_DTYPE)
        self._ok(self.lib.tdv_get_tree(self._c, buf.ctypes.data_as(_vp),
                                       n.value, ctypes.byref(n)),
                 "tdv_get_tree")
        return buf

    def make_active_list(self, max_active_bin, time_bin=None):
        """Make_Active_Particle_List on the device: Time_Bin <= max bin
        (after replacing the Time_Bin column with `time_bin`, if given)."""
        tb = (None if time_bin is None
              else np.ascontiguousarray(time_bin, dtype=np.int32))
        self._ok(self.lib.tdv_make_active_list(
            self._c, int(max_active_bin),
            tb.ctypes.data_as(_vp) if tb is not None else None),
            "tdv_make_active_list")

    def set_fp32_options(self, waves=None, queue=None, split=None):
        """Change the FP32 walk's launch options (register budget, work-queue
        kernel, parts per group); None keeps the current value."""
        p = self.params
        w = p.fp32_waves if waves is None else int(waves)
        q = p.fp32_queue if queue is None else int(queue)
        s = p.fp32_split if split is None else int(split)
        self._ok(self.lib.tdv_set_fp32_options(self._c, w, q, s),
                 "tdv_set_fp32_options")
        p.fp32_waves, p.fp32_queue, p.fp32_split = w, q, s

    def kick_drift(self, kick, drift):
        """The KDK step of a run on one global time step (every particle
        active): v += kick a, then x += drift v (wrapped into the box when
        periodic), on the device."""
        self._ok(self.lib.tdv_kick_drift(self._c, float(kick), float(drift)),
                 "tdv_kick_drift")

    # -- integer timeline (kick.c, drift.c, timestep.c, comoving.c) ----------
    def set_timeline(self, tl):
        """Setup_Time_Integration + Setup_Comoving for a tdv_timeline."""
        self._ok(self.lib.tdv_set_timeline(self._c, ctypes.byref(tl)),
                 "tdv_set_timeline")

    def set_integer_times(self, it_drift=None, it_kick=None):
        """P.It_Drift_Pos / P.It_Kick_Pos in the current particle order."""
        a = (None if it_drift is None
             else np.ascontiguousarray(it_drift, dtype=np.uint32))
        b = (None if it_kick is None
             else np.ascontiguousarray(it_kick, dtype=np.uint32))
        self._ok(self.lib.tdv_set_integer_times(
            self._c, a.ctypes.data_as(_vp) if a is not None else None,
            b.ctypes.data_as(_vp) if b is not None else None),
            "tdv_set_integer_times")

    def integer_times(self):
        a = np.zeros(self.n, np.uint32)
        b = np.zeros(self.n, np.uint32)
        self._ok(self.lib.tdv_get_integer_times(
            self._c, a.ctypes.data_as(_vp), b.ctypes.data_as(_vp)),
            "tdv_get_integer_times")
        return a, b

    def kick_halfstep(self, update_tree=False):
        """Kick_First/Second_Halfstep (kick.c:9-70) on the device; with
        update_tree the tree takes the kicks (Gravity_Tree_Update_Kicks,
        tree_update.c:12-57)."""
        self._ok(self.lib.tdv_kick_halfstep(self._c, int(bool(update_tree))),
                 "tdv_kick_halfstep")

    def tree_update_drift(self, dt):
        """Gravity_Tree_Update_Drift (tree_update.c:64-104) on the device."""
        self._ok(self.lib.tdv_tree_update_drift(self._c, float(dt)),
                 "tdv_tree_update_drift")

    def drift_to_sync_point(self, it_next):
        """Drift_To_Sync_Point (drift.c:12-64) to Int_Time.Next = it_next."""
        self._ok(self.lib.tdv_drift_to_sync_point(self._c, int(it_next)),
                 "tdv_drift_to_sync_point")

    def set_new_timebins(self, time_int_accuracy, softening, dt_max_global,
                         max_active_bin, expansion_factor=1.0,
                         hubble_parameter=1.0, grav_accel_factor=1.0):
        """set_new_particle_timebins (timestep.c:171-201) on the device;
        returns (Time_Bin_Min, Time_Bin_Max)."""
        p = TdvTimestepParams(float(time_int_accuracy), float(softening),
                              float(dt_max_global), int(max_active_bin),
                              float(expansion_factor),
                              float(hubble_parameter),
                              float(grav_accel_factor))
        lo, hi = ctypes.c_int(), ctypes.c_int()
        self._ok(self.lib.tdv_set_new_timebins(self._c, ctypes.byref(p),
                                               ctypes.byref(lo),
                                               ctypes.byref(hi)),
                 "tdv_set_new_timebins")
        return lo.value, hi.value

    def walk32_quadrupoles(self):
        """The FP32 walk's quadrupole records of the last FP32 walk: (nodes
        (nnodes, 8) float32, top nodes (ntop, 8) float32, each node's first
        particle (-1 for terminators))."""
        nn = len(self.tree())
        ntop = len(self.top_nodes())
        q = np.zeros((nn, 8), np.float32)
        tq = np.zeros((ntop, 8), np.float32)
        first = np.zeros(nn, np.int32)
        self._ok(self.lib.tdv_get_walk32_quadrupoles(
            self._c, q.ctypes.data_as(_vp), tq.ctypes.data_as(_vp),
            first.ctypes.data_as(_vp)), "tdv_get_walk32_quadrupoles")
        return q, tq, first

    def interaction_counts(self):
        """(tree-walk interactions, Ewald interactions) of the last walk."""
        c = (ctypes.c_uint64 * 2)()
        self._ok(self.lib.tdv_interaction_counts(self._c, c),
                 "tdv_interaction_counts")
        return int(c[0]), int(c[1])

    def interaction_kinds(self):
        """(tree interactions, Ewald corrections, of those: multipole
        nodes, corrections with quadrupoles) of the last walk."""
        c = (ctypes.c_uint64 * 4)()
        self._ok(self.lib.tdv_interaction_kinds(self._c, c),
                 "tdv_interaction_kinds")
        return tuple(int(x) for x in c)

    def stage_times(self):
        ms = (ctypes.c_float * 8)()
        self._ok(self.lib.tdv_stage_times(self._c, ms), "tdv_stage_times")
        return dict(zip(STAGES, list(ms)))
