"""The out/ file writer (rumdeed_amd/output.py): Fortran edit descriptors
against a Fortran compiler's own output, and the files of a run against the
records they come from."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from rumdeed_amd import output as O

HERE = os.path.dirname(os.path.abspath(__file__))
FLANG = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"

VALUES = [1.25e-15, 0.0, -3.5e-3, 7.123456789e5, 9.99999951e-1, 1.0, -2.0e-300,
          6.02214076e23, 1.0e100, 4.9999999e-5, 123.456, -0.0]


def python_lines():
    out = []
    for i, x in enumerate(VALUES, start=1):
        out.append(O.fortran_E(x, 14, 6) + "".join(" " * 8 + O.fortran_I(v, 6)
                                                   for v in (i, 10 * i, -i)))
        out.append(O.fortran_I(i, 8) + "".join("  " + O.fortran_E(v, 16, 8)
                                               for v in (x, -x, x * 3.0)))
        out.append(O.fortran_ES(x, 12, 4) + "".join("  " + O.fortran_I(v, 8)
                                                    for v in (i, 1000 * i, 0)))
        out.append(O.fortran_ES(x, 12, 4) + "  " + O.fortran_I(i, 6) + "  " + O.fortran_I(7 * i, 6))
        out.extend(O.volt_lines(x, i, 2.0 * x + 1.0, 0.0))
        out.append(O.fortran_I(i, 3) + "  " + O.fortran_I(1024 * i, 8) + "  " +
                   O.fortran_I(1280 * i, 8) + "  " + O.fortran_I(i % 2, 4) + "  " +
                   O.fortran_ES(x, 12, 4) + "  " + O.fortran_ES(-x, 12, 4) + "  " +
                   O.fortran_ES(0.0, 12, 4))
    return out


@pytest.mark.skipif(not os.path.exists(FLANG), reason="no Fortran compiler in this image")
def test_edit_descriptors_match_fortran(tmp_path):
    """E14.6, E16.8, ES12.4, i3, i4, i6, i8 and tr fields as a Fortran compiler
    writes them (zero, negative zero, rounding carries, three-digit
    exponents)."""
    exe = tmp_path / "ff"
    subprocess.run([FLANG, "-O0", os.path.join(HERE, "golden", "fortran_formats.f90"), "-o",
                    str(exe)], check=True, capture_output=True, cwd=tmp_path)
    got = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    assert got.splitlines() == python_lines()


def test_known_lines():
    """Values worked by hand from the descriptors."""
    assert O.fortran_E(1.25e-15, 14, 6) == "  0.125000E-14"
    assert O.fortran_E(-3.5e-3, 16, 8) == " -0.35000000E-02"
    assert O.fortran_ES(2.5e-13, 12, 4) == "  2.5000E-13"
    assert O.fortran_ES(9.99996e-1, 12, 4) == "  1.0000E+00"  # the carry
    assert O.fortran_I(123456789, 8) == "********"


def _record(step, **kw):
    r = dict(step=step, n=100 + step, emitted=7, n_round=20, absorbed=3, absorbed_top=2,
             absorbed_bot=1, dropped=0, start_rounds=1, supply_level=1, supply_cells=4096,
             supply_fail=0, n_sup=12.5, ramo_current=1e-6, mh_std=0.05, a_rate=0.3,
             df_avg=0.01, f_avg=[1.0, -2.0, -3.0e9], ramo=[1e-6, 0.0, 0.0],
             avg_elec_vel=[1.0, 2.0, 2.0e5], supply_err=0.25)
    r.update(kw)
    return r


def test_run_files(tmp_path):
    """A planar FE run's files hold one line per step with the record's
    values in the reference's columns; init.bin and lifetime.dt have the
    reference's layouts."""
    dt, V, d = 0.25e-15, 2000.0, 1000e-9
    f = O.OutputFiles(str(tmp_path), O.EMIT_PLANAR_FE, dt, V, d, max_particles=4096)
    recs = [_record(k) for k in range(1, 6)]
    for r in recs:
        f.write_step(r)
    life = np.zeros((1000, 2), np.int32)  # as Diode.life_time() returns it
    life[4, 0] = 9
    f.close(life)
    em = np.loadtxt(tmp_path / "emitted.dt")
    assert em.shape == (5, 5)
    assert np.allclose(em[:, 0], [dt * k / 1e-12 for k in range(1, 6)], rtol=1e-6)
    assert list(em[0, 1:]) == [1, 7, 101 + 3, 7]
    fl = np.loadtxt(tmp_path / "field.dt")
    assert fl.shape == (5, 8) and np.allclose(fl[0, 1:], [1.0, -2.0, -3.0e9, 12.5, 0.01, 0.3,
                                                          0.05], rtol=1e-8)
    ab = np.loadtxt(tmp_path / "absorbed_top.dt")
    assert list(ab[2, 1:]) == [3, 2, 2, 0, 2]
    ra = np.loadtxt(tmp_path / "ramo_current.dt")
    assert ra.shape == (5, 14)
    speed = np.sqrt(1 + 4 + 4e10)
    assert np.allclose(ra[0, [2, 3, 4, 5, 6]], [1e-6, V, 104, 104, 0], rtol=1e-4)
    assert np.allclose(ra[0, 7], speed / (V / d), rtol=1e-4)  # mobility
    assert np.allclose(ra[0, 8:11], [speed, speed, 0.0], rtol=1e-4)
    vt = (tmp_path / "volt.dt").read_text().splitlines()
    assert len(vt) == 2 * 6  # step 0 and steps 1-5, two records each
    assert vt[2] == O.fortran_ES(dt * 1 / 1e-12, 12, 4) + "  " + O.fortran_I(1, 8) + "  " + \
        O.fortran_ES(V, 18, 8) and vt[3] == "  0.0000E+00"
    it = (tmp_path / "integration.dt").read_text().splitlines()
    # 64 x 64 cells at level 1: the 32 x 32 level 0 was evaluated too
    assert it == ["  1      4096      5120     0    1.2500E+01    2.5000E-01    0.0000E+00"] * 5
    assert O.supply_evaluations(64 * 64, 3) == 32 * 32 + 16 * 16 + 8 * 8 + 64 * 64
    lt = np.loadtxt(tmp_path / "lifetime.dt")
    assert lt.shape == (1000, 4) and lt[4, 2] == 9 and lt[4, 1] == 5
    init = np.fromfile(tmp_path / "init.bin", dtype=np.uint8)
    assert init.size == 72  # 7 float64 + 4 int32
    head = np.frombuffer(init[:56].tobytes(), "<f8")
    tail = np.frombuffer(init[56:].tobytes(), "<i4")
    assert list(head) == [1.0, 1.0, 1.0, 1e-9, 1e-12, 1e-9 / 1e-12, 1.0]
    assert list(tail) == [4096, 1, 96 * 96, 1000]


def test_density_files(tmp_path):
    """density_*.bin: packed stream records in the reference's field order,
    positions divided by length_scale; top and bottom absorptions go to
    their own files, species files hold only their species."""
    import struct

    from rumdeed_amd import _lib as L

    em = np.zeros(3, L.EMIT_RECORD)
    em["x"], em["y"], em["z"] = [1e-9, 2e-9, -3e-9], [4e-9, 5e-9, 6e-9], 1e-9
    em["emitter"], em["section"], em["id"], em["species"] = 1, [1, 2, 3], [10, 11, 12], 1
    ab = np.zeros(3, L.ABSORB_RECORD)
    ab["x"], ab["y"] = [7e-9, 8e-9, 9e-9], [-1e-9, -2e-9, -3e-9]
    ab["vx"], ab["vy"], ab["vz"] = [1.0, 2.0, 3.0], [4.0, 5.0, 6.0], [7.0, 8.0, 9.0]
    ab["emitter"], ab["section"], ab["id"], ab["where"] = 1, [4, 5, 6], [0, 1, 2], [1, 2, 1]
    f = O.OutputFiles(str(tmp_path), O.EMIT_PLANAR_FE, 1e-15, 100.0, 1e-6, densities=True)
    f.write_particles(em, ab)
    f.write_particles(em[:1], ab[:0])
    f.close()
    b = (tmp_path / "density_emit.bin").read_bytes()
    assert len(b) == 4 * 40
    assert b[:40] == struct.pack("<3d4i", 1e-9 / 1e-9, 4e-9 / 1e-9, 1e-9 / 1e-9, 1, 1, 10, 1)
    assert b[120:] == b[:40]
    b = (tmp_path / "density_emit_elec.bin").read_bytes()
    assert len(b) == 4 * 32 and b[32:64] == struct.pack("<3d2i", 2.0, 5.0, 1.0, 1, 11)
    b = (tmp_path / "density_absorb_top.bin").read_bytes()
    assert b == (struct.pack("<5d3i", 7e-9 / 1e-9, -1.0, 1.0, 4.0, 7.0, 1, 4, 0) +
                 struct.pack("<5d3i", 9e-9 / 1e-9, -3e-9 / 1e-9, 3.0, 6.0, 9.0, 1, 6, 2))
    b = (tmp_path / "density_absorb_bot.bin").read_bytes()
    assert b == struct.pack("<2d3i", 8e-9 / 1e-9, -2e-9 / 1e-9, 1, 5, 1)
    for name in ("density_ion", "density_emit_ion", "density_emit_atom", "density_absorb_recom"):
        assert (tmp_path / (name + ".bin")).read_bytes() == b""


def test_plane_and_position_files(tmp_path):
    """planes-<k>.bin for each plane with height > 0 (numbered by position
    in the list), records of the density_absorb_top layout; position.bin
    with its (steps, 1) header and per-step (step, n) blocks of positions
    in m."""
    import struct

    from rumdeed_amd import _lib as L

    pl = np.zeros(3, L.PLANE_RECORD)
    pl["x"], pl["y"] = [1e-9, 2e-9, 3e-9], [4e-9, 5e-9, 6e-9]
    pl["vx"], pl["vy"], pl["vz"] = 1.5, 2.5, [3.5, 4.5, 5.5]
    pl["emitter"], pl["section"], pl["id"], pl["plane"] = 1, 7, [3, 4, 3], [1, 1, 3]
    f = O.OutputFiles(str(tmp_path), O.EMIT_PLANAR_FE, 1e-15, 100.0, 1e-6,
                      planes=[5e-9, -1.0, 25e-9], position_steps=20)
    f.write_planes(pl)
    f.write_position(1, [[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]], [1, 1], [2, 3], [8, 9])
    f.write_position(2, np.zeros((0, 3)), [], [], [])
    f.close()
    assert sorted(p.name for p in tmp_path.glob("planes-*")) == ["planes-1.bin", "planes-3.bin"]
    assert (tmp_path / "planes-1.bin").read_bytes() == (
        struct.pack("<5d3i", 1.0, 4.0, 1.5, 2.5, 3.5, 1, 7, 3) +
        struct.pack("<5d3i", 2e-9 / 1e-9, 5e-9 / 1e-9, 1.5, 2.5, 4.5, 1, 7, 4))
    assert (tmp_path / "planes-3.bin").read_bytes() == struct.pack(
        "<5d3i", 3e-9 / 1e-9, 6e-9 / 1e-9, 1.5, 2.5, 5.5, 1, 7, 3)
    assert (tmp_path / "position.bin").read_bytes() == (
        struct.pack("<2i", 20, 1) + struct.pack("<2i", 1, 2) +
        struct.pack("<3d3i", 1.0, 2.0, 3.0, 1, 2, 8) + struct.pack("<3d3i", 4.0, 5.0, 6.0, 1, 3, 9) +
        struct.pack("<2i", 2, 0))
