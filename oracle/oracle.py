"""ORACLE -- test infrastructure only.

ctypes front end of the CPU restatement (``liboracle.so``, built from
rd_oracle*.c) and, where it was built, of the reference harness
(``_ref/librdref.so``, the RUMDEED reference modules compiled from their own
sources).  Only tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg import this module, and only as the checker.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_int, c_int32, c_uint64

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "librdref.so")

GEOM_PLANAR = 1
GEOM_TIP = 2


class Geom(ctypes.Structure):
    """Mirror of ``rdo_geom`` (oracle/rd_oracle.h)."""

    _fields_ = [
        ("kind", c_int32), ("image_charge", c_int32), ("n_ic_max", c_int32), ("pad0", c_int32),
        ("d", c_double), ("V", c_double), ("E_z", c_double),
        ("a_foci", c_double), ("shift_z", c_double), ("pre_fac_E_tip", c_double),
        ("pre_fac_E_tip_unit_voltage", c_double),
        ("h_tip", c_double), ("r_tip", c_double), ("eta_1", c_double), ("max_xi", c_double),
        ("d_tip", c_double), ("R_base", c_double), ("theta_tip", c_double),
    ]


class Consts(ctypes.Structure):
    _fields_ = [(n, c_double) for n in (
        "pi", "h", "k_b", "c", "mu_0", "epsilon_0", "h_bar", "m_0", "q_0", "q_02", "m_u",
        "m_N2", "m_N2p", "div_fac_c", "length_scale", "time_scale", "eps_soft",
        "a_FN", "b_FN", "l_const")]


_dp = POINTER(c_double)
_ip = POINTER(c_int)


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


_lib = None


def default_threads() -> int:
    """OpenMP threads for the oracle: OMP_NUM_THREADS if set, at most 16."""
    env = os.environ.get("OMP_NUM_THREADS")
    n = int(env) if env and env.isdigit() else (os.cpu_count() or 1)
    return max(1, min(n, 16))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle library missing: {ORACLE_SO} (run make -C oracle)")
        _lib = ctypes.CDLL(ORACLE_SO)
        L = _lib
        L.rdo_log.restype = c_double
        L.rdo_log.argtypes = [c_double]
        L.rdo_exp.restype = c_double
        L.rdo_exp.argtypes = [c_double]
        L.rdo_sum_chunk.restype = c_int
    return _lib


def consts() -> Consts:
    c = Consts()
    lib().rdo_get_consts(ctypes.byref(c))
    return c


def geom_planar(d, V, image_charge=True, n_ic_max=0) -> Geom:
    g = Geom()
    lib().rdo_geom_planar(ctypes.byref(g), c_double(d), c_double(V), c_int(int(image_charge)),
                          c_int(n_ic_max))
    return g


def geom_tip(V, d_tip, R_base, h_tip, image_charge=True) -> Geom:
    g = Geom()
    lib().rdo_geom_tip(ctypes.byref(g), c_double(V), c_double(d_tip), c_double(R_base),
                       c_double(h_tip), c_int(int(image_charge)))
    return g


def accel(g: Geom, pos, q, m, nthreads=None, long_double=False):
    """Accelerations (n,3) of particles pos (n,3), charge q, mass m."""
    pos = _f64(pos)
    n = pos.shape[0]
    q = _f64(q)
    m = _f64(m)
    out = np.zeros((n, 3))
    nt = nthreads or default_threads()
    fn = lib().rdo_accel_ld if long_double else lib().rdo_accel
    fn(ctypes.byref(g), c_int(n), _d(pos), _d(q), _d(m), _d(out), c_int(nt))
    return out


def accel_tts(g: Geom, pos, q, m, species, K, step, acc=None, nthreads=None):
    """Two-time-step accelerations (rdo_accel_tts) of step `step` with
    atom_time_interval K; acc: the accelerations the particles hold (an ion
    adds onto its own), zeros by default."""
    pos = _f64(pos)
    n = pos.shape[0]
    out = np.zeros((n, 3)) if acc is None else np.array(acc, dtype=np.float64).reshape(n, 3)
    sp = np.ascontiguousarray(species, np.int32)
    L = lib()
    L.rdo_accel_tts.restype = None
    L.rdo_accel_tts(ctypes.byref(g), c_int(n), _d(pos), _d(_f64(q)), _d(_f64(m)), _i(sp),
                    c_int(K), c_int(step), _d(out), c_int(nthreads or default_threads()))
    return out


def accel_subset(g: Geom, pos, q, m, targets, nthreads=None, long_double=False):
    """Accelerations (len(targets),3) of the listed particles (0-based) over
    all sources; row k equals row targets[k] of accel()."""
    pos = _f64(pos)
    n = pos.shape[0]
    t = np.ascontiguousarray(targets, dtype=np.int32)
    assert t.size == 0 or (t.min() >= 0 and t.max() < n)
    out = np.zeros((t.size, 3))
    nt = nthreads or default_threads()
    fn = lib().rdo_accel_subset_ld if long_double else lib().rdo_accel_subset
    fn(ctypes.byref(g), c_int(n), _d(pos), _d(_f64(q)), _d(_f64(m)), c_int(t.size), _i(t),
       _d(out), c_int(nt))
    return out


def field(g: Geom, pos, q, pts, nthreads=None, long_double=False):
    """Field (M,3) at points pts (M,3) from particles pos (n,3) with charge q."""
    pos = _f64(pos).reshape(-1, 3)
    q = _f64(q)
    pts = _f64(pts).reshape(-1, 3)
    out = np.zeros((pts.shape[0], 3))
    nt = nthreads or default_threads()
    fn = lib().rdo_field_ld if long_double else lib().rdo_field
    fn(ctypes.byref(g), c_int(pos.shape[0]), _d(pos), _d(q), c_int(pts.shape[0]), _d(pts),
       _d(out), c_int(nt))
    return out


def field_scale(g: Geom, pos, q, pts, nthreads=None):
    """Per point: sum over sources of the norm of each source's field term
    plus the norm of the vacuum field."""
    pos = _f64(pos).reshape(-1, 3)
    pts = _f64(pts).reshape(-1, 3)
    out = np.zeros(pts.shape[0])
    lib().rdo_field_scale(ctypes.byref(g), c_int(pos.shape[0]), _d(pos), _d(_f64(q)),
                          c_int(pts.shape[0]), _d(pts), _d(out), c_int(nthreads or default_threads()))
    return out


def accel_scale(g: Geom, pos, q, m, targets, nthreads=None):
    """Per listed target: sum over sources of the norm of each pair's
    acceleration term plus the norm of the vacuum term."""
    pos = _f64(pos)
    t = np.ascontiguousarray(targets, dtype=np.int32)
    out = np.zeros(t.size)
    lib().rdo_accel_scale(ctypes.byref(g), c_int(pos.shape[0]), _d(pos), _d(_f64(q)),
                          _d(_f64(m)), c_int(t.size), _i(t), _d(out),
                          c_int(nthreads or default_threads()))
    return out


def vacuum_field(g: Geom, p):
    p = _f64(p)
    out = np.zeros(3)
    lib().rdo_vacuum_field(ctypes.byref(g), _d(p), _d(out))
    return out


def log(x):
    return lib().rdo_log(float(x))


def exp(x):
    return lib().rdo_exp(float(x))


def philox4x64(ctr, key):
    out = (c_uint64 * 4)()
    lib().rdo_philox4x64((c_uint64 * 4)(*ctr), (c_uint64 * 2)(*key), out)
    return [int(v) for v in out]


# ---------------------------------------------------------------------------
# Reference harness (the reference's own routines, built by `make -C oracle ref`)
class Ref:
    """Front end of oracle/_ref/librdref.so.  The reference keeps its state
    in module variables, so there is one instance per process."""

    _inst = None

    def __init__(self, cap=1 << 16):
        if not os.path.exists(REF_SO):
            raise RuntimeError(f"reference harness not built: {REF_SO}")
        self.L = ctypes.CDLL(REF_SO)
        self.L.ref_jgtf.restype = c_double
        self.L.ref_jgtf.argtypes = [c_double, c_double, c_double]
        self.L.ref_init(c_int(cap))
        self.cap = cap

    @classmethod
    def get(cls, cap=1 << 16):
        if cls._inst is None:
            cls._inst = Ref(cap)
        return cls._inst

    @staticmethod
    def available():
        return os.path.exists(REF_SO)

    def threads(self, n):
        self.L.ref_set_threads(c_int(n))

    def consts(self):
        out = np.zeros(12)
        self.L.ref_consts(_d(out))
        return out

    def setup_planar(self, d, V, dt=0.25e-15, image_charge=True, n_ic_max=0):
        self.L.ref_setup_planar(c_double(d), c_double(V), c_double(dt),
                                c_int(int(image_charge)), c_int(n_ic_max))

    def setup_tip(self, V, d_tip, R_base, h_tip, dt=0.25e-15, image_charge=True):
        self.L.ref_setup_tip(c_double(V), c_double(d_tip), c_double(R_base), c_double(h_tip),
                             c_double(dt), c_int(int(image_charge)))

    def tip_params(self):
        out = np.zeros(8)
        self.L.ref_tip_params(_d(out))
        return out

    def clear(self):
        self.L.ref_clear()

    def put_snapshot(self, pos, q, m, species=None):
        pos = _f64(pos)
        n = pos.shape[0]
        assert n <= self.cap
        sp = np.ones(n, dtype=np.int32) if species is None else np.ascontiguousarray(species, np.int32)
        self.L.ref_put_snapshot(c_int(n), _d(pos), _d(_f64(q)), _d(_f64(m)), _i(sp))

    def add_particle(self, pos, vel, species=1, step=1, emit=1, sec=1):
        self.L.ref_add_particle(_d(_f64(pos)), _d(_f64(vel)), c_int(species), c_int(step),
                                c_int(emit), c_int(sec))

    def accel(self, which, n):
        """which: 0 = planar-specialized pair loop, 1 = generic, 2 = gather (ACC form)."""
        out = np.zeros((n, 3))
        self.L.ref_accel(c_int(which), c_int(n), _d(out))
        return out

    def field_at(self, pts):
        pts = _f64(pts).reshape(-1, 3)
        out = np.zeros_like(pts)
        self.L.ref_field_at(c_int(pts.shape[0]), _d(pts), _d(out))
        return out

    def update_position(self, step):
        self.L.ref_update_position(c_int(step))

    def set_two_time_step(self, atom_time_interval):
        self.L.ref_set_two_time_step(c_int(atom_time_interval))

    def remove_particles(self, step):
        self.L.ref_remove_particles(c_int(step))

    def mark_remove(self, i1, reason):
        self.L.ref_mark_remove(c_int(i1), c_int(reason))

    def counts(self):
        out = np.zeros(16, dtype=np.int32)
        self.L.ref_counts(_i(out))
        keys = ["nrPart", "nrElec", "nrIon", "nrPart_remove", "nrElec_remove", "nrIon_remove",
                "nrPart_remove_top", "nrPart_remove_bot", "nrElec_remove_top",
                "nrElec_remove_bot", "nrIon_remove_top", "nrIon_remove_bot", "nrID",
                "charge_rev", "nrElec_remove_top_emit1", "nrPart_dropped"]
        return dict(zip(keys, [int(v) for v in out]))

    def state(self, n):
        arrs = [np.zeros((n, 3)) for _ in range(6)]
        q = np.zeros(n)
        m = np.zeros(n)
        ints = np.zeros((n, 6), dtype=np.int32)
        self.L.ref_get_state(c_int(n), *[_d(a) for a in arrs], _d(q), _d(m), _i(ints))
        keys = ["pos", "prev_pos", "vel", "acc", "acc_prev", "acc_prev2"]
        st = dict(zip(keys, arrs))
        st.update(q=q, m=m, species=ints[:, 0], step=ints[:, 1], emitter=ints[:, 2],
                  section=ints[:, 3], id=ints[:, 4], mask=ints[:, 5])
        return st

    def ramo(self):
        out = np.zeros(4)
        self.L.ref_ramo(_d(out))
        return out

    def life_time(self):
        out = np.zeros((3, 1000), dtype=np.int32)
        self.L.ref_life_time(_i(out))
        return out.T

    def jgtf(self, F, T, w):
        return self.L.ref_jgtf(F, T, w)

    def read_work(self, x0, y0, lx, ly):
        self.L.ref_read_work(c_double(x0), c_double(y0), c_double(lx), c_double(ly))

    def tip_xyz(self, xi, eta, phi):
        o = np.zeros(3)
        self.L.ref_tip_xyz(c_double(xi), c_double(eta), c_double(phi), _d(o))
        return o

    def tip_coords(self, p):
        o = np.zeros(3)
        self.L.ref_tip_coords(_d(_f64(p)), _d(o))
        return o

    def tip_normal(self, p):
        o = np.zeros(3)
        self.L.ref_tip_normal(_d(_f64(p)), _d(o))
        return o

    def field_normal(self, p, E):
        self.L.ref_field_normal.restype = c_double
        return self.L.ref_field_normal(_d(_f64(p)), _d(_f64(E)))

    def tip_area(self, xi1, xi2, phi1, phi2):
        self.L.ref_tip_area.restype = c_double
        self.L.ref_tip_area.argtypes = [c_double] * 4
        return self.L.ref_tip_area(xi1, xi2, phi1, phi2)

    def tip_vacuum(self, p):
        o = np.zeros(3)
        self.L.ref_tip_vacuum(_d(_f64(p)), _d(o))
        return o

    def mb_velocity(self, T, seed, n):
        out = np.zeros((n, 3))
        self.L.ref_mb_velocity(c_double(T), c_int(seed), c_int(n), _d(out))
        return out

    def work_at(self, pts):
        pts = _f64(pts).reshape(-1, 3)
        w = np.zeros(pts.shape[0])
        sec = np.zeros(pts.shape[0], dtype=np.int32)
        self.L.ref_work_at(c_int(pts.shape[0]), _d(pts), _d(w), _i(sec))
        return w, sec


# ---------------------------------------------------------------------------
# One simulation of the CPU restatement (rd_oracle_emit.c)
class Record(ctypes.Structure):
    """Per-step record; same layout as rd_step_record (include/rumdeed.h)."""

    _fields_ = [("step", c_int32), ("n", c_int32), ("emitted", c_int32), ("n_round", c_int32),
                ("absorbed", c_int32), ("absorbed_top", c_int32), ("absorbed_bot", c_int32),
                ("dropped", c_int32), ("start_rounds", c_int32), ("supply_level", c_int32),
                ("supply_cells", c_int32), ("supply_fail", c_int32),
                ("n_sup", c_double), ("ramo_current", c_double), ("mh_std", c_double),
                ("a_rate", c_double), ("df_avg", c_double), ("f_avg", c_double * 3),
                ("ramo", c_double * 3), ("avg_elec_vel", c_double * 3), ("supply_err", c_double)]


def record_dict(r):
    d = {k: getattr(r, k) for k, _ in Record._fields_}
    for k in ("f_avg", "ramo", "avg_elec_vel"):
        d[k] = list(getattr(r, k))
    return d


EMIT_RECORD = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("emitter", "<i4"),
                        ("section", "<i4"), ("id", "<i4"), ("species", "<i4")])
ABSORB_RECORD = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("vx", "<f8"),
                          ("vy", "<f8"), ("vz", "<f8"), ("emitter", "<i4"), ("section", "<i4"),
                          ("id", "<i4"), ("where", "<i4")])
PLANE_RECORD = np.dtype([("x", "<f8"), ("y", "<f8"), ("vx", "<f8"), ("vy", "<f8"),
                         ("vz", "<f8"), ("emitter", "<i4"), ("section", "<i4"), ("id", "<i4"),
                         ("plane", "<i4")])


class Sim:
    """CPU restatement of one simulation (emission, Beeman step, removal)."""

    def __init__(self, cap, nthreads=None):
        L = lib()
        L.rdo_sim_create.restype = ctypes.c_void_p
        L.rdo_sim_create.argtypes = [c_int, c_int]
        for name in ("rdo_sim_destroy", "rdo_sim_set_geometry", "rdo_sim_set_time_step",
                     "rdo_sim_set_seed", "rdo_sim_set_emitter", "rdo_sim_add", "rdo_sim_put",
                     "rdo_sim_step", "rdo_sim_get", "rdo_sim_life", "rdo_sim_count"):
            getattr(L, name).restype = None
        L.rdo_sim_count.restype = c_int
        L.rdo_sim_set_time_step.argtypes = [ctypes.c_void_p, c_double]
        L.rdo_sim_set_seed.argtypes = [ctypes.c_void_p, c_uint64, c_uint64]
        L.rdo_sim_set_geometry.argtypes = [ctypes.c_void_p, ctypes.c_void_p, c_double]
        L.rdo_sim_set_emitter.argtypes = [ctypes.c_void_p, c_int, c_int, c_double, c_double,
                                          c_double, c_double, c_double, c_double, c_int, c_int,
                                          _dp, c_int, c_int]
        L.rdo_sim_set_quad_levels.restype = None
        L.rdo_sim_set_quad_levels.argtypes = [ctypes.c_void_p, c_int]
        L.rdo_sim_set_two_time_step.restype = None
        L.rdo_sim_set_two_time_step.argtypes = [ctypes.c_void_p, c_int]
        L.rdo_sim_step.argtypes = [ctypes.c_void_p, c_int, c_int, ctypes.c_void_p]
        L.rdo_sim_count.argtypes = [ctypes.c_void_p]
        self.L = L
        self.h = L.rdo_sim_create(int(cap), int(nthreads or default_threads()))
        self.cap = cap
        self.step_done = 0

    def __del__(self):
        try:
            if self.h:
                self.L.rdo_sim_destroy(ctypes.c_void_p(self.h))
                self.h = None
        except Exception:
            pass

    def set_geometry(self, g: Geom, box_top=0.0):
        self.L.rdo_sim_set_geometry(ctypes.c_void_p(self.h), ctypes.byref(g), c_double(box_top))

    def set_time_step(self, dt):
        self.L.rdo_sim_set_time_step(ctypes.c_void_p(self.h), dt)

    def set_seed(self, seed, stream=0):
        self.L.rdo_sim_set_seed(ctypes.c_void_p(self.h), int(seed), int(stream))

    def set_serial(self, on=True):
        """Sample emission sites with the reference's default serial
        samplers (each chain sees the electrons of the chains before it)
        instead of the lockstep (batch) semantics the device runs."""
        self.L.rdo_sim_set_serial.restype = None
        self.L.rdo_sim_set_serial.argtypes = [ctypes.c_void_p, c_int]
        self.L.rdo_sim_set_serial(ctypes.c_void_p(self.h), int(bool(on)))

    def set_emitter(self, mode, delay=0, x0=0.0, y0=0.0, lx=0.0, ly=0.0, T=300.0, w_tip=4.7,
                    wf=None, quad=0, ndim=0, quad_levels=0):
        """quad_levels: refinement doublings of the planar supply rule (-1
        none, 0 up to four, k up to k; never beyond 1024 points per side)."""
        wf = np.array([[4.7]]) if wf is None else np.atleast_2d(np.asarray(wf, dtype=np.float64))
        ny, nx = wf.shape
        tab = np.ascontiguousarray(wf.ravel())
        self._wf = tab
        self.L.rdo_sim_set_emitter(ctypes.c_void_p(self.h), mode, delay, x0, y0, lx, ly, T, w_tip,
                                   nx, ny, _d(tab), quad, ndim)
        self.L.rdo_sim_set_quad_levels(ctypes.c_void_p(self.h), c_int(quad_levels))

    def set_two_time_step(self, atom_time_interval):
        """two_time_step with atom_time_interval K (0: off)."""
        self.L.rdo_sim_set_two_time_step(ctypes.c_void_p(self.h), c_int(atom_time_interval))

    def put(self, pos, q, m, vel=None, species=None):
        pos = _f64(pos).reshape(-1, 3)
        n = pos.shape[0]
        sp = None if species is None else np.ascontiguousarray(species, np.int32)
        v = None if vel is None else _f64(vel).reshape(n, 3)
        self.L.rdo_sim_put(ctypes.c_void_p(self.h), c_int(n), _d(pos),
                           None if v is None else _d(v), _d(_f64(q)), _d(_f64(m)),
                           None if sp is None else _i(sp))

    def add(self, pos, vel=None, species=None, step=1, emit=1, sec=None):
        pos = _f64(pos).reshape(-1, 3)
        n = pos.shape[0]
        v = None if vel is None else _f64(vel).reshape(n, 3)
        sp = None if species is None else np.ascontiguousarray(species, np.int32)
        sc = None if sec is None else np.ascontiguousarray(sec, np.int32)
        self.L.rdo_sim_add(ctypes.c_void_p(self.h), c_int(n), _d(pos),
                           None if v is None else _d(v), None if sp is None else _i(sp),
                           c_int(step), c_int(emit), None if sc is None else _i(sc))

    def step(self, step=None, phases=7):
        if step is None:
            step = self.step_done + 1
        rec = Record()
        self.L.rdo_sim_step(ctypes.c_void_p(self.h), c_int(step), c_int(phases), ctypes.byref(rec))
        if phases & 4 or phases == 7:
            self.step_done = step
        return rec

    @property
    def n(self):
        return self.L.rdo_sim_count(ctypes.c_void_p(self.h))

    def step_logs(self):
        """The last step's emission and absorption records (Add_Particle's
        and Mark_Particles_Remove's density writes), as structured arrays
        with the fields of rd_emit_record / rd_absorb_record."""
        L = self.L
        for name in ("rdo_sim_emit_log", "rdo_sim_absorb_log"):
            getattr(L, name).restype = c_int
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        h = ctypes.c_void_p(self.h)
        em = np.zeros(L.rdo_sim_emit_log(h, None), dtype=EMIT_RECORD)
        ab = np.zeros(L.rdo_sim_absorb_log(h, None), dtype=ABSORB_RECORD)
        L.rdo_sim_emit_log(h, em.ctypes.data_as(ctypes.c_void_p))
        L.rdo_sim_absorb_log(h, ab.ctypes.data_as(ctypes.c_void_p))
        return em, ab

    def set_planes(self, z_m):
        """Recording planes (Check_Planes), heights in m."""
        z = np.ascontiguousarray(np.asarray(z_m, dtype=np.float64).ravel())
        self.L.rdo_sim_set_planes.restype = None
        self.L.rdo_sim_set_planes.argtypes = [ctypes.c_void_p, c_int, ctypes.c_void_p]
        self.L.rdo_sim_set_planes(ctypes.c_void_p(self.h), int(z.size),
                                  z.ctypes.data_as(ctypes.c_void_p))
        self._planes = z

    def plane_logs(self):
        """The last step's plane crossings, plane 1 first."""
        L = self.L
        L.rdo_sim_plane_log.restype = c_int
        L.rdo_sim_plane_log.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        h = ctypes.c_void_p(self.h)
        rec = np.zeros(L.rdo_sim_plane_log(h, None), dtype=PLANE_RECORD)
        L.rdo_sim_plane_log(h, rec.ctypes.data_as(ctypes.c_void_p))
        return rec

    def state(self):
        n = self.n
        arrs = [np.zeros((n, 3)) for _ in range(6)]
        ints = np.zeros((n, 5), dtype=np.int32)
        q = np.zeros(n)
        self.L.rdo_sim_get(ctypes.c_void_p(self.h), *[_d(a) for a in arrs], _i(ints), _d(q))
        keys = ["pos", "prev_pos", "vel", "acc", "acc_prev", "acc_prev2"]
        st = dict(zip(keys, arrs))
        st.update(species=ints[:, 0], step=ints[:, 1], emitter=ints[:, 2], section=ints[:, 3],
                  id=ints[:, 4], q=q)
        return st

    def life_time(self):
        h = np.zeros(2000, dtype=np.int32)
        self.L.rdo_sim_life(ctypes.c_void_p(self.h), _i(h))
        return h.reshape(2, 1000).T


def gtf(F, T, w):
    L = lib()
    L.rdo_gtf.restype = c_double
    L.rdo_gtf.argtypes = [c_double, c_double, c_double]
    return L.rdo_gtf(F, T, w)


def fn_functions(ic, F, w):
    """(v_y, t_y, Escape_Prob_log, Elec_Supply_log) of the planar FN model."""
    L = lib()
    for name in ("rdo_fn_v", "rdo_fn_t", "rdo_escape_log", "rdo_supply_log"):
        getattr(L, name).restype = c_double
        getattr(L, name).argtypes = [c_int, c_double, c_double]
    return (L.rdo_fn_v(ic, F, w), L.rdo_fn_t(ic, F, w), L.rdo_escape_log(ic, F, w),
            L.rdo_supply_log(ic, F, w))


def work_checker(table, x0, y0, lx, ly, pts):
    L = lib()
    L.rdo_work_checker.restype = c_double
    L.rdo_work_checker.argtypes = [_dp, c_int, c_int, c_double, c_double, c_double, c_double,
                                   c_double, c_double, POINTER(c_int)]
    tab = np.ascontiguousarray(np.atleast_2d(table), dtype=np.float64)
    ny, nx = tab.shape
    pts = _f64(pts).reshape(-1, 3)
    w = np.zeros(len(pts))
    sec = np.zeros(len(pts), dtype=np.int32)
    for k, p in enumerate(pts):
        s = c_int()
        w[k] = L.rdo_work_checker(_d(tab.ravel()), nx, ny, x0, y0, lx, ly, p[0], p[1],
                                  ctypes.byref(s))
        sec[k] = s.value
    return w, sec


def tip_surface(g: Geom):
    """The oracle's tip surface functions for geometry g: (xyz(xi, eta, phi),
    normal(p), area(xi1, xi2, phi1, phi2), eta(p), field_normal(p, E))."""
    L = lib()
    L.rdo_tip_xyz.argtypes = [ctypes.c_void_p, c_double, c_double, c_double, _dp]
    L.rdo_tip_normal.argtypes = [ctypes.c_void_p, _dp, _dp]
    L.rdo_tip_area.argtypes = [ctypes.c_void_p, c_double, c_double, c_double, c_double]
    L.rdo_tip_area.restype = c_double
    L.rdo_tip_eta.argtypes = [ctypes.c_void_p, _dp]
    L.rdo_tip_eta.restype = c_double
    L.rdo_field_normal.argtypes = [ctypes.c_void_p, _dp, _dp]
    L.rdo_field_normal.restype = c_double
    gp = ctypes.byref(g)

    def xyz(xi, eta, phi):
        o = np.zeros(3)
        L.rdo_tip_xyz(gp, xi, eta, phi, _d(o))
        return o

    def normal(p):
        o = np.zeros(3)
        L.rdo_tip_normal(gp, _d(_f64(p)), _d(o))
        return o

    def area(xi1, xi2, phi1, phi2):
        return L.rdo_tip_area(gp, xi1, xi2, phi1, phi2)

    def eta(p):
        return L.rdo_tip_eta(gp, _d(_f64(p)))

    def field_normal(p, E):
        return L.rdo_field_normal(gp, _d(_f64(p)), _d(_f64(E)))

    return xyz, normal, area, eta, field_normal


def mb_velocity(seed, stream, step, T, n):
    """Emission velocities (n,3) of candidates 0..n-1 of a thermal-field step."""
    L = lib()
    L.rdo_mb_velocity.argtypes = [c_uint64, c_uint64, c_int, c_double, c_int, _dp]
    L.rdo_mb_velocity.restype = None
    out = np.zeros((n, 3))
    L.rdo_mb_velocity(int(seed), int(stream), int(step), float(T), int(n), _d(out))
    return out


def sincos(x):
    L = lib()
    s = c_double()
    c = c_double()
    L.rdo_sincos.restype = None
    L.rdo_sincos.argtypes = [c_double, POINTER(c_double), POINTER(c_double)]
    L.rdo_sincos(float(x), ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value
