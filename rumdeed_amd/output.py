"""RUMDEED's per-step output files, written from the device step records.

The reference's analysis scripts read a run's ``out/`` directory.  This module
writes the files they read for the emission modes built here, from the
records every ``rd_run`` / ``rd_pool_step`` returns (and the lifetime
histogram), with the reference's Fortran edit descriptors:

- ``emitted.dt``: Do_Field_Emission (src/mod_field_emission_v2.F90:202-204),
  Do_Field_Thermo_Emission (src/mod_field_thermo_emission.F90:130-132),
  Do_Emission_Tip (src/mod_emission_tip.f90:174-175);
- ``field.dt``: src/mod_field_emission_v2.F90:389-390 (planar FE) and
  src/mod_field_thermo_emission.F90:189-190 (GTF); the tip FE writes none;
- ``absorbed.dt``, ``absorbed_top.dt``, ``absorbed_bot.dt``,
  ``absorbed_recom.dt``: Write_Absorbed (src/mod_pair.F90:790-805);
- ``ramo_current.dt``: Write_Ramo_Current (src/mod_pair.F90:812-837), with
  the mobility of Average_Velocities (src/mod_verlet.F90:425-446);
- ``lifetime.dt``: Write_Life_Time (src/mod_pair.F90:774-784);
- ``integration.dt``: Do_Surface_Integration_FE
  (src/mod_field_emission_v2.F90:654-655), one line per planar FE emission
  step: emitter, regions, integrand evaluations, fail flag, integral, error
  estimate, probability.  Here the supply is the refined midpoint rule, so
  the regions are the cells of the level used, the evaluations the points of
  every level computed, the error the difference of the last two levels
  (0 for the fixed rule), the fail flag set when the last allowed level still
  misses the tolerance, and the probability (Divonne's chi-square) 0; the
  file is opened for every mode (src/main.F90:647) and stays empty for the
  others;
- ``volt.dt``: Set_Voltage (src/mod_verlet.F90:2050-2057), step 0 at
  start-up (src/main.F90:172) and then every step; its format has three
  descriptors for four items, so the zero RF voltage goes on a record of
  its own (format reversion);
- ``init.bin``: Write_Initial_Variables (src/main.F90:928-935);
- ``density_emit.bin``, ``density_emit_elec.bin`` (Add_Particle,
  src/mod_pair.F90:86-89 and 119-123), ``density_absorb_top.bin``,
  ``density_absorb_bot.bin`` (Mark_Particles_Remove, src/mod_pair.F90:240-256):
  Fortran stream files (src/main.F90:665-722), packed records with no
  markers, from the per-step particle records (Diode.step_logs()).  The
  ion, atom and recombination density files are created empty: these
  emission modes make none;
- ``planes-<k>.bin``: Check_Planes (src/mod_verlet.F90:343-367), one stream
  file per plane with height > 0 (src/main.F90:726-740), from the plane
  records (Diode.plane_logs());
- ``position.bin``: Write_Position (src/mod_pair.F90:841-867), written
  between the position update and the removal (src/main.F90:190-191), so
  it lists the particles marked for removal too; from Diode.get_state()
  taken between update_position and remove_particles.

Counts follow the reference's moments of writing: ``emitted.dt`` and
``ramo_current.dt`` see the particles after emission and before removal
(n at the end of the step plus the absorbed ones).  Only electrons are
emitted by these modes, so ion columns are zero and the particle and
electron columns agree.
"""
from __future__ import annotations

import math
import os

import numpy as np

# src/mod_global.F90
EPSILON_R = 1.0
M_EEFF = 1.0
M_IEFF = 1.0
LENGTH_SCALE = 1.0e-9
TIME_SCALE = 1.0e-12
VEL_SCALE = LENGTH_SCALE / TIME_SCALE
CUR_SCALE = 1.0
MAX_EMITTERS = 1
MAX_SECTIONS = 96 * 96
MAX_LIFE_TIME = 1000

EMIT_TIP_FE, EMIT_PLANAR_GTF, EMIT_PLANAR_FE = 3, 9, 10
SPECIES_ELEC = 1
WHERE_TOP, WHERE_BOT = 1, 2

# Stream-file record layouts (packed, little-endian as the reference's
# unformatted writes on x86 and AMD hosts).
DENSITY_EMIT = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("emitter", "<i4"),
                         ("section", "<i4"), ("id", "<i4"), ("species", "<i4")])
DENSITY_EMIT_SPECIES = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("emitter", "<i4"),
                                 ("id", "<i4")])
DENSITY_ABSORB_TOP = np.dtype([("x", "<f8"), ("y", "<f8"), ("vx", "<f8"), ("vy", "<f8"),
                               ("vz", "<f8"), ("emitter", "<i4"), ("section", "<i4"),
                               ("id", "<i4")])
DENSITY_ABSORB_BOT = np.dtype([("x", "<f8"), ("y", "<f8"), ("emitter", "<i4"),
                               ("section", "<i4"), ("id", "<i4")])
POSITION_PARTICLE = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("emitter", "<i4"),
                              ("section", "<i4"), ("id", "<i4")])
DENSITY_FILES = ("density_emit", "density_ion", "density_emit_ion", "density_emit_elec",
                 "density_emit_atom", "density_absorb_top", "density_absorb_bot",
                 "density_absorb_recom")


# ---- Fortran edit descriptors ---------------------------------------------
def fortran_I(v: int, w: int) -> str:
    """Iw: right-justified in w columns, asterisks when it does not fit."""
    s = str(int(v))
    return s.rjust(w) if len(s) <= w else "*" * w


def _exponent(e: int) -> str:
    """The exponent field of Ew.d / ESw.d: E+dd, or +ddd without the E when
    |e| > 99."""
    if abs(e) <= 99:
        return f"E{'-' if e < 0 else '+'}{abs(e):02d}"
    return f"{'-' if e < 0 else '+'}{abs(e):03d}"


def fortran_E(x: float, w: int, d: int) -> str:
    """Ew.d: 0.ddd...E+ee with d significant digits, right-justified."""
    if not math.isfinite(x):
        s = "NaN" if math.isnan(x) else ("-Infinity" if x < 0 else "Infinity")
        return s.rjust(w) if len(s) <= w else "*" * w
    if x == 0.0:
        s = ("-" if math.copysign(1.0, x) < 0 else "") + "0." + "0" * d + "E+00"
    else:
        m, e = f"{abs(x):.{d - 1}e}".split("e")
        digits = m.replace(".", "")
        s = ("-" if x < 0 else "") + "0." + digits + _exponent(int(e) + 1)
    return s.rjust(w) if len(s) <= w else "*" * w


def fortran_ES(x: float, w: int, d: int) -> str:
    """ESw.d: d.ddd...E+ee (scientific), right-justified."""
    if not math.isfinite(x):
        s = "NaN" if math.isnan(x) else ("-Infinity" if x < 0 else "Infinity")
        return s.rjust(w) if len(s) <= w else "*" * w
    m, e = f"{abs(x):.{d}e}".split("e")
    s = ("-" if math.copysign(1.0, x) < 0 else "") + m + _exponent(int(e))
    return s.rjust(w) if len(s) <= w else "*" * w


def supply_evaluations(cells: int, level: int) -> int:
    """Integrand evaluations of a planar supply whose level `level` has
    `cells` cells: every level 0..level was computed, each with a quarter of
    the next one's cells (level 0 alone for the fixed rule)."""
    return sum(cells >> (2 * k) for k in range(level + 1))


def integration_line(cells: int, level: int, fail: int, integral: float, error: float,
                     emitter: int = 1) -> str:
    """One integration.dt record, (i3, tr2, i8, tr2, i8, tr2, i4, tr2, ES12.4,
    tr2, ES12.4, tr2, ES12.4): emitter, regions, evaluations, fail, integral,
    error, probability (src/mod_field_emission_v2.F90:654-655)."""
    return (fortran_I(emitter, 3) + "  " + fortran_I(cells, 8) + "  " +
            fortran_I(supply_evaluations(cells, level), 8) + "  " + fortran_I(fail, 4) + "  " +
            fortran_ES(integral, 12, 4) + "  " + fortran_ES(error, 12, 4) + "  " +
            fortran_ES(0.0, 12, 4))


def volt_lines(cur_time: float, step: int, V_d: float, V_rf: float) -> list[str]:
    """The records of one (ES12.4, tr2, i8, tr2, ES18.8) write of cur_time,
    step, V_d, V_rf: the fourth item reverts to the start of the format on a
    new record."""
    return [fortran_ES(cur_time, 12, 4) + "  " + fortran_I(step, 8) + "  " +
            fortran_ES(V_d, 18, 8), fortran_ES(V_rf, 12, 4)]


# ---- one run's files -----------------------------------------------------
class OutputFiles:
    """The out/ files of one simulation.

    ``mode``: the rd_emitter mode (RD_EMIT_*); ``dt``: time step (s);
    ``V``: applied voltage (V); ``d``: gap (planar) or box_dim(3) (tip), m.
    ``densities``: also write the density_*.bin particle files (from
    ``write_particles``).  ``planes``: the recording plane heights (m) whose
    planes-<k>.bin files ``write_planes`` fills.  ``position_steps``: write
    position.bin (write_position_file), whose header holds this total step
    count.
    """

    def __init__(self, directory: str, mode: int, dt: float, V: float, d: float,
                 max_particles: int = 0, densities: bool = False, planes=None,
                 position_steps: int | None = None):
        self.dir = directory
        self.mode = mode
        self.dt = dt
        self.E_z = -V / d if d else 0.0
        self.V = V
        os.makedirs(directory, exist_ok=True)
        names = ["emitted", "absorbed", "absorbed_top", "absorbed_bot", "absorbed_recom",
                 "ramo_current", "volt", "integration"]
        if mode in (EMIT_PLANAR_FE, EMIT_PLANAR_GTF):
            names.append("field")
        self.f = {k: open(os.path.join(directory, k + ".dt"), "w") for k in names}
        self.bin = ({k: open(os.path.join(directory, k + ".bin"), "wb") for k in DENSITY_FILES}
                    if densities else {})
        self.planes = {}
        for k, z in enumerate([] if planes is None else list(planes), start=1):
            if z > 0.0:
                self.planes[k] = open(os.path.join(directory, f"planes-{k}.bin"), "wb")
                self.bin[f"planes-{k}"] = self.planes[k]
        self.position_steps = position_steps
        if position_steps is not None:
            self.bin["position"] = open(os.path.join(directory, "position.bin"), "wb")
        self.write_init_bin(max_particles)
        self.f["volt"].write("\n".join(volt_lines(0.0, 0, V, 0.0)) + "\n")

    def close(self, life_time=None):
        """Close the files; with the lifetime histogram (Diode.life_time(),
        1000 rows of electron and ion counts), write lifetime.dt too."""
        for f in (*self.f.values(), *self.bin.values()):
            f.close()
        self.f, self.bin, self.planes = {}, {}, {}
        if life_time is not None:
            self.write_life_time(life_time)

    def write_init_bin(self, max_particles: int):
        rec = np.zeros(1, dtype=np.dtype([
            ("epsilon_r", "<f8"), ("m_eeff", "<f8"), ("m_ieff", "<f8"),
            ("length_scale", "<f8"), ("time_scale", "<f8"), ("vel_scale", "<f8"),
            ("cur_scale", "<f8"), ("MAX_PARTICLES", "<i4"), ("MAX_EMITTERS", "<i4"),
            ("MAX_SECTIONS", "<i4"), ("MAX_LIFE_TIME", "<i4")]))
        rec[0] = (EPSILON_R, M_EEFF, M_IEFF, LENGTH_SCALE, TIME_SCALE, VEL_SCALE, CUR_SCALE,
                  max_particles, MAX_EMITTERS, MAX_SECTIONS, MAX_LIFE_TIME)
        rec.tofile(os.path.join(self.dir, "init.bin"))

    def write_life_time(self, hist):
        hist = np.asarray(hist)
        hist = hist.T if hist.shape == (MAX_LIFE_TIME, 2) else hist.reshape(2, MAX_LIFE_TIME)
        with open(os.path.join(self.dir, "lifetime.dt"), "w") as f:
            for i in range(1, MAX_LIFE_TIME + 1):
                time = i * self.dt / TIME_SCALE
                f.write(fortran_ES(time, 12, 4) + "  " + fortran_I(i, 6) + "  " +
                        fortran_I(hist[0, i - 1], 6) + "  " + fortran_I(hist[1, i - 1], 6) +
                        "\n")

    def write_step(self, rec):
        """One step's lines from its record (rd_step_record or its as_dict())."""
        r = rec if isinstance(rec, dict) else rec.as_dict()
        step = r["step"]
        cur_time = self.dt * step / TIME_SCALE
        n_mid = r["n"] + r["absorbed"]  # after emission, before removal
        # volt.dt (Set_Voltage, at the start of the position update)
        self.f["volt"].write("\n".join(volt_lines(cur_time, step, self.V, 0.0)) + "\n")
        # emitted.dt: (E14.6, *(tr8, i6))
        ints = [step, r["emitted"], n_mid]
        if self.mode != EMIT_TIP_FE:
            ints.append(r["emitted"])  # nrEmitted_emitters(1)
        self.f["emitted"].write(fortran_E(cur_time, 14, 6) +
                                "".join(" " * 8 + fortran_I(v, 6) for v in ints) + "\n")
        # field.dt: (i8, 7(tr2, E16.8))
        if "field" in self.f:
            df_avg = r["df_avg"] if self.mode == EMIT_PLANAR_FE else 0.0
            vals = [*r["f_avg"], r["n_sup"], df_avg, r["a_rate"], r["mh_std"]]
            self.f["field"].write(fortran_I(step, 8) +
                                  "".join("  " + fortran_E(v, 16, 8) for v in vals) + "\n")
        # integration.dt: (i3, tr2, i8, tr2, i8, tr2, i4, tr2, ES12.4, tr2, ES12.4, tr2, ES12.4)
        cells = r.get("supply_cells", 0)
        if self.mode == EMIT_PLANAR_FE and cells > 0:
            self.f["integration"].write(integration_line(cells, r["supply_level"],
                                                         r["supply_fail"], r["n_sup"],
                                                         r["supply_err"]) + "\n")
        # absorbed*.dt: (ES12.4, *(tr2, i8))
        def absorbed(name, ints):
            self.f[name].write(fortran_ES(cur_time, 12, 4) +
                               "".join("  " + fortran_I(v, 8) for v in [step, *ints]) + "\n")

        absorbed("absorbed", [r["absorbed"], r["absorbed"], 0])
        absorbed("absorbed_top", [r["absorbed_top"], r["absorbed_top"], 0, r["absorbed_top"]])
        absorbed("absorbed_bot", [r["absorbed_bot"], r["absorbed_bot"], 0])
        absorbed("absorbed_recom", [0, 0, 0])
        # ramo_current.dt
        v = r["avg_elec_vel"]
        speed = math.sqrt(v[0] ** 2 + v[1] ** 2 + v[2] ** 2)
        mob = speed / (-self.E_z) if self.E_z != 0.0 else 0.0
        ramo = r["ramo"]
        line = (fortran_ES(cur_time, 12, 4) + "  " + fortran_I(step, 8) + "  " +
                fortran_ES(sum(ramo) / CUR_SCALE, 12, 4) + "  " + fortran_ES(self.V, 12, 4) +
                "  " + fortran_I(n_mid, 6) + "  " + fortran_I(n_mid, 6) + "  " +
                fortran_I(0, 6) + "  " + fortran_ES(mob, 12, 4) + "  " +
                fortran_ES(speed, 12, 4) + "  " + fortran_ES(speed, 12, 4) + "  " +
                fortran_ES(0.0, 12, 4) + "".join("  " + fortran_ES(x, 12, 4) for x in ramo))
        self.f["ramo_current"].write(line + "\n")

    def write_planes(self, records):
        """One step's plane crossings (Diode.plane_logs(): plane 1 first,
        each plane in particle index order) appended to planes-<k>.bin."""
        records = np.asarray(records)
        for k, f in self.planes.items():
            r = records[records["plane"] == k]
            out = np.empty(r.shape[0], DENSITY_ABSORB_TOP)
            out["x"], out["y"] = r["x"] / LENGTH_SCALE, r["y"] / LENGTH_SCALE
            for name in ("vx", "vy", "vz", "emitter", "section", "id"):
                out[name] = r[name]
            f.write(out.tobytes())

    def write_position(self, step: int, pos, emitter, section, ids):
        """Write_Position's block for one step: (step, nrPart) and each
        particle's position (m), emitter, section and id; the header
        (steps, N_steps = 1) goes before step 1."""
        f = self.bin["position"]
        pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
        if step == 1:
            f.write(np.array([self.position_steps, 1], "<i4").tobytes())
        f.write(np.array([step, pos.shape[0]], "<i4").tobytes())
        rec = np.empty(pos.shape[0], POSITION_PARTICLE)
        rec["x"], rec["y"], rec["z"] = pos[:, 0], pos[:, 1], pos[:, 2]
        rec["emitter"], rec["section"], rec["id"] = emitter, section, ids
        f.write(rec.tobytes())

    def write_particles(self, emit, absorb):
        """One step's particle records (the structured arrays of
        Diode.step_logs(): emission in Add_Particle order, absorption in
        particle index order) appended to the density_*.bin files, positions
        divided by length_scale as the reference writes them."""
        if "density_emit" not in self.bin:
            raise ValueError("OutputFiles was opened without densities=True")
        emit = np.asarray(emit)
        absorb = np.asarray(absorb)
        rec = np.empty(emit.shape[0], DENSITY_EMIT)
        for k in ("x", "y", "z"):
            rec[k] = emit[k] / LENGTH_SCALE
        for k in ("emitter", "section", "id", "species"):
            rec[k] = emit[k]
        elec = rec[emit["species"] == SPECIES_ELEC]
        sp = np.empty(elec.shape[0], DENSITY_EMIT_SPECIES)
        for k in DENSITY_EMIT_SPECIES.names:
            sp[k] = elec[k]
        self.bin["density_emit"].write(rec.tobytes())
        self.bin["density_emit_elec"].write(sp.tobytes())
        top = absorb[absorb["where"] == WHERE_TOP]
        t = np.empty(top.shape[0], DENSITY_ABSORB_TOP)
        t["x"], t["y"] = top["x"] / LENGTH_SCALE, top["y"] / LENGTH_SCALE
        for k in ("vx", "vy", "vz", "emitter", "section", "id"):
            t[k] = top[k]
        bot = absorb[absorb["where"] == WHERE_BOT]
        b = np.empty(bot.shape[0], DENSITY_ABSORB_BOT)
        b["x"], b["y"] = bot["x"] / LENGTH_SCALE, bot["y"] / LENGTH_SCALE
        for k in ("emitter", "section", "id"):
            b[k] = bot[k]
        self.bin["density_absorb_top"].write(t.tobytes())
        self.bin["density_absorb_bot"].write(b.tobytes())
