"""Host-side mirror of the RUMDEED hot-path interface over librumdeed.so.

The method names follow the reference routines they stand in for
(mod_verlet / mod_pair / the emission modules); every call goes through the
C ABI of the HIP library.  Arrays are numpy (n, 3) float64 in SI units, which
is the memory layout of the Fortran (1:3, 1:n) particle arrays.
"""
from __future__ import annotations

import ctypes
from ctypes import c_double, c_int32, c_void_p

import numpy as np

from . import _lib as L
from ._lib import RD_GEOM_PLANAR, RD_GEOM_TIP, RD_PREC_FP64, RD_PREC_MIXED, check, dptr, f64, iptr

PRECISIONS = {"fp64": RD_PREC_FP64, "mixed": RD_PREC_MIXED}
# planes_z default of the input namelist, nm (src/mod_global.F90:340-341)
PLANES_Z_NM = (5.0, 10.0, 25.0, 50.0, 75.0, 100.0, 125.0, 250.0, 500.0, 750.0)


def _prec(p):
    if isinstance(p, str):
        return PRECISIONS[p]
    return int(p)


class Diode:
    """One simulation context: particle state resident in HBM plus geometry."""

    def __init__(self, max_particles: int, device: int = 0):
        self._lib = L.lib()
        h = c_void_p()
        check(self._lib.rd_ctx_create(device, int(max_particles), ctypes.byref(h)))
        self._h = h
        self.max_particles = int(max_particles)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.rd_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- set-up (Read_Input_Variables, Init_Field_Emission_v2, Init_Emission_Tip)
    def set_planar(self, V: float, d: float, image_charge: bool = True, n_ic_max: int = 0,
                   box_top: float = 0.0):
        g = L.Geometry(kind=RD_GEOM_PLANAR, image_charge=int(bool(image_charge)),
                       n_ic_max=int(n_ic_max), V=float(V), d=float(d), box_top=float(box_top))
        check(self._lib.rd_set_geometry(self._h, ctypes.byref(g)))

    def set_tip(self, V: float, d_tip: float, R_base: float, h_tip: float,
                image_charge: bool = True, box_top: float = 0.0):
        g = L.Geometry(kind=RD_GEOM_TIP, image_charge=int(bool(image_charge)), V=float(V),
                       d_tip=float(d_tip), R_base=float(R_base), h_tip=float(h_tip),
                       box_top=float(box_top))
        check(self._lib.rd_set_geometry(self._h, ctypes.byref(g)))

    def set_time_step(self, dt: float):
        check(self._lib.rd_set_time_step(self._h, c_double(dt)))

    def set_two_time_step(self, atom_time_interval: int):
        """two_time_step with atom_time_interval K (0: off): ions move every
        K steps over K time steps (rd_set_two_time_step)."""
        check(self._lib.rd_set_two_time_step(self._h, int(atom_time_interval)))

    def set_logs(self, on: bool = True):
        """Collect each step's emission and absorption particle records."""
        check(self._lib.rd_set_logs(self._h, int(bool(on))))

    def step_logs(self):
        """The last step's (emission, absorption) records of the selected
        simulation, as numpy structured arrays (L.EMIT_RECORD,
        L.ABSORB_RECORD)."""
        ne, na = ctypes.c_int32(0), ctypes.c_int32(0)
        rc = self._lib.rd_get_logs(self._h, 0, None, ctypes.byref(ne), 0, None, ctypes.byref(na))
        if rc not in (0, L.RD_ERR_CAPACITY):
            check(rc)
        emit = np.zeros(ne.value, dtype=L.EMIT_RECORD)
        absorb = np.zeros(na.value, dtype=L.ABSORB_RECORD)
        check(self._lib.rd_get_logs(self._h, ne.value, emit.ctypes.data_as(ctypes.c_void_p),
                                    ctypes.byref(ne), na.value,
                                    absorb.ctypes.data_as(ctypes.c_void_p), ctypes.byref(na)))
        return emit, absorb

    def set_planes(self, z_m):
        """Recording planes (heights in m; <= 0 keeps the number, records
        nothing).  RUMDEED's default is PLANES_Z_NM scaled by 1e-9."""
        z = np.ascontiguousarray(np.asarray(z_m, dtype=np.float64).ravel())
        check(self._lib.rd_set_planes(self._h, int(z.size),
                                      z.ctypes.data_as(ctypes.c_void_p) if z.size else None))

    def plane_logs(self):
        """The last step's plane crossings (L.PLANE_RECORD), plane 1 first."""
        n = ctypes.c_int32(0)
        rc = self._lib.rd_get_plane_logs(self._h, 0, None, ctypes.byref(n))
        if rc not in (0, L.RD_ERR_CAPACITY):
            check(rc)
        rec = np.zeros(n.value, dtype=L.PLANE_RECORD)
        check(self._lib.rd_get_plane_logs(self._h, n.value, rec.ctypes.data_as(ctypes.c_void_p),
                                          ctypes.byref(n)))
        return rec

    def set_graphs(self, on: bool = True):
        """Launch each step's Metropolis-Hastings jump loop as one HIP graph."""
        check(self._lib.rd_set_graphs(self._h, int(bool(on))))

    def set_force_options(self, no_symmetric: bool = False, no_far_expansion: bool = False):
        """Forms of the mixed planar accelerations (rd_set_force_options):
        by default the pair-symmetric form from 16384 particles and the far
        image partners (n >= 2) as one FP64 expansion."""
        flags = (1 if no_symmetric else 0) | (2 if no_far_expansion else 0)
        check(self._lib.rd_set_force_options(self._h, flags))

    def set_partial_budget(self, nbytes: int = 0):
        """Bound on one pair launch's partial sums (rd_set_partial_budget;
        0 = 512 MiB); larger launches run in bands with identical results."""
        check(self._lib.rd_set_partial_budget(self._h, int(nbytes)))

    def work_bytes(self) -> int:
        """Device bytes held by the context's work buffers (rd_work_bytes)."""
        v = ctypes.c_int64()
        check(self._lib.rd_work_bytes(self._h, ctypes.byref(v)))
        return v.value

    def set_jump_pairing(self, min_workgroups: int):
        """Pool launches of at least min_workgroups jump workgroups give each
        workgroup four of their partial sums (results unchanged; 0: always)."""
        check(self._lib.rd_set_jump_pairing(self._h, int(min_workgroups)))

    def set_solo_jumps(self, on: bool):
        """Mixed precision: small simulations run all their jumps in one
        workgroup (default on; rd_set_solo_jumps)."""
        check(self._lib.rd_set_solo_jumps(self._h, int(bool(on))))

    def set_seed(self, seed: int, stream: int = 0):
        check(self._lib.rd_set_seed(self._h, int(seed), int(stream)))

    def set_emitter(self, mode: int, delay: int = 0, x0: float = 0.0, y0: float = 0.0,
                    lx: float = 0.0, ly: float = 0.0, T: float = 300.0, w_tip: float = 4.7,
                    wf=None, quad: int = 0, ndim: int = 0, quad_levels: int = 0):
        """Emitter + work function (`work` file: rows top first).  quad:
        supply points per side (0 = 32); quad_levels: doublings of the planar
        supply rule until two levels agree within the reference's Divonne
        tolerances (-1 none, 0 up to four, k up to k)."""
        wf = np.array([[4.7]]) if wf is None else np.atleast_2d(np.asarray(wf, dtype=np.float64))
        ny, nx = wf.shape
        self._wf = np.ascontiguousarray(wf.ravel())
        em = L.Emitter(mode=int(mode), delay=int(delay), x0=x0, y0=y0, lx=lx, ly=ly, T=T,
                       w_tip=w_tip, wf_nx=nx, wf_ny=ny, wf_table=dptr(self._wf),
                       quad_n=int(quad), ndim=int(ndim), quad_levels=int(quad_levels))
        check(self._lib.rd_set_emitter(self._h, ctypes.byref(em)))

    # ---- the time step (main.F90:175-219)
    def add_particles(self, pos, vel=None, species=None, step=1, emit=1, sec=None):
        """Add_Particle for every row of pos (n,3), in order."""
        pos = f64(pos).reshape(-1, 3)
        n = pos.shape[0]
        vel = None if vel is None else f64(vel).reshape(n, 3)
        sp = None if species is None else np.ascontiguousarray(species, dtype=np.int32)
        sc = None if sec is None else np.ascontiguousarray(sec, dtype=np.int32)
        check(self._lib.rd_add_particles(self._h, n, dptr(pos), dptr(vel), iptr(sp), int(step),
                                         int(emit), iptr(sc)))

    def do_emission(self, step: int, precision="fp64"):
        rec = L.StepRecord()
        check(self._lib.rd_do_emission(self._h, int(step), _prec(precision), ctypes.byref(rec)))
        return rec

    def update_position(self, step: int, precision="fp64"):
        rec = L.StepRecord()
        check(self._lib.rd_update_position(self._h, int(step), _prec(precision),
                                           ctypes.byref(rec)))
        return rec

    def remove_particles(self, step: int):
        rec = L.StepRecord()
        check(self._lib.rd_remove_particles(self._h, int(step), ctypes.byref(rec)))
        return rec

    def run(self, nsteps: int, precision="fp64"):
        recs = (L.StepRecord * max(int(nsteps), 1))()
        check(self._lib.rd_run(self._h, int(nsteps), _prec(precision), recs))
        return list(recs)[:nsteps]

    def get_state(self):
        n = self.n
        arrs = [np.zeros((n, 3)) for _ in range(6)]
        ints = np.zeros((n, 5), dtype=np.int32)
        check(self._lib.rd_get_state(self._h, *[dptr(a) for a in arrs], iptr(ints)))
        keys = ["pos", "prev_pos", "vel", "acc", "acc_prev", "acc_prev2"]
        st = dict(zip(keys, arrs))
        st.update(species=ints[:, 0], step=ints[:, 1], emitter=ints[:, 2], section=ints[:, 3],
                  id=ints[:, 4])
        st["q"] = self.get_particles()["q"]
        return st

    def life_time(self):
        h = np.zeros(2000, dtype=np.int32)
        check(self._lib.rd_get_life_time(self._h, iptr(h)))
        return h.reshape(2, 1000).T

    # ---- particle store (mod_pair)
    def put_particles(self, pos, q, m, vel=None, species=None):
        pos = f64(pos).reshape(-1, 3)
        n = pos.shape[0]
        q = f64(q).reshape(n)
        m = f64(m).reshape(n)
        vel = None if vel is None else f64(vel).reshape(n, 3)
        sp = None if species is None else np.ascontiguousarray(species, dtype=np.int32)
        check(self._lib.rd_put_particles(self._h, n, dptr(pos), dptr(vel), dptr(q), dptr(m),
                                         iptr(sp)))

    @property
    def n(self) -> int:
        v = c_int32()
        check(self._lib.rd_particle_count(self._h, ctypes.byref(v)))
        return v.value

    def get_particles(self):
        n = self.n
        pos = np.zeros((n, 3))
        vel = np.zeros((n, 3))
        q = np.zeros(n)
        m = np.zeros(n)
        sp = np.zeros(n, dtype=np.int32)
        check(self._lib.rd_get_particles(self._h, dptr(pos), dptr(vel), dptr(q), dptr(m),
                                         iptr(sp)))
        return dict(pos=pos, vel=vel, q=q, m=m, species=sp)

    # ---- forces and fields (mod_verlet)
    def calculate_acceleration_particles(self, precision="fp64"):
        """Calculate_Acceleration_Particles: accelerations (n,3)."""
        check(self._lib.rd_accel(self._h, _prec(precision)))
        out = np.zeros((self.n, 3))
        check(self._lib.rd_get_accel(self._h, dptr(out)))
        return out

    def calc_field_at_batch(self, pts, precision="fp64"):
        """Calc_Field_at_Batch: field (M,3) at points (M,3)."""
        pts = f64(pts).reshape(-1, 3)
        out = np.zeros_like(pts)
        check(self._lib.rd_field_batch(self._h, pts.shape[0], dptr(pts), dptr(out),
                                       _prec(precision)))
        return out

    def calc_field_at_batch_static(self, pts, precision="fp64"):
        """Calc_Field_at_Batch through rd_field_batch_static, which allocates
        no device memory (its work memory is a static array per device)."""
        pts = f64(pts).reshape(-1, 3)
        out = np.zeros_like(pts)
        check(self._lib.rd_field_batch_static(self._h, pts.shape[0], dptr(pts), dptr(out),
                                              _prec(precision)))
        return out

    def calc_field_at(self, pos, precision="fp64"):
        """Calc_Field_at: field (3,) at one point."""
        return self.calc_field_at_batch(np.asarray(pos, dtype=np.float64).reshape(1, 3),
                                        precision)[0]

    def field_grid(self, x0, y0, lx, ly, nx, ny, z=0.0, precision="fp64", copy=True):
        out = np.zeros((nx * ny, 3)) if copy else None
        check(self._lib.rd_field_grid(self._h, x0, y0, lx, ly, nx, ny, z, _prec(precision),
                                      dptr(out)))
        return out

    # ---- measurement
    def time_accel(self, precision="fp64", reps=10):
        tot = c_double()
        ker = c_double()
        check(self._lib.rd_time_accel(self._h, _prec(precision), reps, ctypes.byref(tot),
                                      ctypes.byref(ker)))
        return tot.value, ker.value

    def time_field_grid(self, x0, y0, lx, ly, nx, ny, z=0.0, precision="fp64", reps=10):
        tot = c_double()
        ker = c_double()
        check(self._lib.rd_time_field_grid(self._h, x0, y0, lx, ly, nx, ny, z, _prec(precision),
                                           reps, ctypes.byref(tot), ctypes.byref(ker)))
        return tot.value, ker.value


    def time_copy(self, nbytes: int, reps: int = 3) -> float:
        """Average time (ms) of a plain device copy of `nbytes` bytes."""
        ker = c_double()
        check(self._lib.rd_time_copy(self._h, int(nbytes), int(reps), ctypes.byref(ker)))
        return ker.value


def device_count() -> int:
    v = c_int32()
    code = L.lib().rd_device_count(ctypes.byref(v))
    return v.value if code == 0 else 0


class Pool(Diode):
    """Many independent simulations sharing one particle store on one device.

    Every `step()` advances all open simulations with one launch per kernel.
    The single-simulation methods inherited from Diode act on the selected
    slot (`select`)."""

    def __init__(self, total_particles: int, max_slots: int, device: int = 0):
        self._lib = L.lib()
        h = c_void_p()
        check(self._lib.rd_pool_create(device, int(total_particles), int(max_slots),
                                       ctypes.byref(h)))
        self._h = h
        self.max_particles = int(total_particles)
        self.max_slots = int(max_slots)

    @staticmethod
    def footprint(capacity: int) -> int:
        v = ctypes.c_int64()
        check(L.lib().rd_slot_footprint(int(capacity), ctypes.byref(v)))
        return v.value

    def open(self, capacity: int) -> int:
        s = c_int32()
        check(self._lib.rd_slot_open(self._h, int(capacity), ctypes.byref(s)))
        return s.value

    def try_open(self, capacity: int):
        s = c_int32()
        code = self._lib.rd_slot_open(self._h, int(capacity), ctypes.byref(s))
        if code == L.RD_OK:
            return s.value
        if code == 3:  # RD_ERR_CAPACITY
            return None
        check(code)

    def close_slot(self, slot: int):
        check(self._lib.rd_slot_close(self._h, int(slot)))

    def select(self, slot: int):
        check(self._lib.rd_slot_select(self._h, int(slot)))

    def slot_step(self, slot: int) -> int:
        v = c_int32()
        check(self._lib.rd_slot_step(self._h, int(slot), ctypes.byref(v)))
        return v.value

    def step(self, precision="fp64"):
        """One main-loop step of every open simulation; records indexed by slot."""
        recs = (L.StepRecord * self.max_slots)()
        check(self._lib.rd_pool_step(self._h, _prec(precision), recs))
        return recs
