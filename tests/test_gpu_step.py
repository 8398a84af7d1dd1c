"""The device-resident time step (emission, Beeman, boundaries, Ramo,
compaction) against the oracle (bit for bit on the FP64 path), the
reference's own routines (golden steps.npz) and closed forms."""
import os

import numpy as np
import pytest

from conftest import golden, rel_err

pytestmark = pytest.mark.gpu

NM = 1e-9

REC_INT = ("n", "emitted", "n_round", "absorbed", "absorbed_top", "absorbed_bot", "dropped",
           "start_rounds", "supply_level", "supply_cells", "supply_fail")
REC_DBL = ("n_sup", "ramo_current", "mh_std", "a_rate", "df_avg", "supply_err")


def assert_records_equal(rg, ro, ctx=""):
    for k in REC_INT:
        assert getattr(rg, k) == getattr(ro, k), (ctx, k, getattr(rg, k), getattr(ro, k))
    for k in REC_DBL:
        assert getattr(rg, k) == getattr(ro, k), (ctx, k, getattr(rg, k), getattr(ro, k))
    assert list(rg.f_avg) == list(ro.f_avg), ctx
    assert list(rg.ramo) == list(ro.ramo), ctx
    assert list(rg.avg_elec_vel) == list(ro.avg_elec_vel), ctx


def assert_state_equal(sg, so, ctx=""):
    for k in ("pos", "prev_pos", "vel", "acc", "acc_prev", "acc_prev2"):
        assert np.array_equal(sg[k], so[k]), (ctx, k)
    for k in ("species", "step", "emitter", "section", "id"):
        assert np.array_equal(sg[k], so[k]), (ctx, k)


@pytest.fixture(scope="module")
def Diode():
    from rumdeed_amd.diode import Diode

    return Diode


def test_beeman_kinematics_gpu(Diode, oracle_mod):
    """Test_Beeman_Kinematics (src/mod_tests.F90:1403-1452) on the device."""
    c = oracle_mod.consts()
    d, V, dt, vx0 = 1000 * NM, 2.0e3, 0.25e-15, 1.0e3
    dd = Diode(16)
    dd.set_planar(V, d, False, 0)
    dd.set_time_step(dt)
    dd.add_particles(np.array([[0.0, 0.0, 500 * NM]]), np.array([[vx0, 0.0, 0.0]]))
    for k in range(1, 4):
        r = dd.update_position(k)
        dd.remove_particles(k)
    st = dd.get_state()
    a_z = c.q_0 * V / (c.m_0 * d)
    assert abs((st["pos"][0, 2] - 500 * NM) - 4.5 * a_z * dt ** 2) < 1e-9 * 4.5 * a_z * dt ** 2
    assert abs(st["vel"][0, 2] - 3 * a_z * dt) < 1e-9 * 3 * a_z * dt
    assert abs(st["pos"][0, 0] - 3 * vx0 * dt) < 1e-9 * 3 * vx0 * dt
    assert abs(r.ramo[0] - c.q_0 * 3 * a_z * dt / d) < 1e-9 * c.q_0 * 3 * a_z * dt / d


def test_steps_vs_reference_and_oracle(Diode, oracle_mod):
    """Update_Position + Remove_Particles: the device matches the reference's
    routines (golden) to 1e-12 with identical survivors/ids/counts, and the
    oracle bit for bit."""
    g = golden("steps.npz")
    dd = Diode(256)
    dd.set_planar(float(g["V"]), float(g["d"]), True, 1)
    dd.set_time_step(float(g["dt"]))
    so = oracle_mod.Sim(256, nthreads=2)
    so.set_geometry(oracle_mod.geom_planar(float(g["d"]), float(g["V"]), True, 1))
    so.set_time_step(float(g["dt"]))
    n = len(g["species0"])
    for i in range(n):
        kw = dict(species=g["species0"][i:i + 1], step=int(g["birth0"][i]), emit=1)
        dd.add_particles(g["pos0"][i:i + 1], g["vel0"][i:i + 1], **kw)
        so.add(g["pos0"][i:i + 1], g["vel0"][i:i + 1], **kw)
    for k in range(1, int(g["steps"]) + 1):
        rg = dd.update_position(k)
        ro = so.step(k, phases=2)
        ref = g[f"ramo_{k}"]
        assert abs(rg.ramo_current - ref[3]) <= 1e-12 * np.abs(ref[:3]).sum()
        rg2 = dd.remove_particles(k)
        so.step(k, phases=4)
        rem = g[f"removed_{k}"]
        assert rg2.absorbed == rem[0] and rg2.absorbed_top == rem[1] and rg2.absorbed_bot == rem[2]
        sg = dd.get_state()
        assert np.array_equal(sg["id"], g[f"id_{k}"])
        for key in ("pos", "vel", "prev_pos", "acc_prev", "acc_prev2"):
            assert rel_err(sg[key], g[f"{key}_{k}"], floor=1e-300).max() < 1e-12, (k, key)
        assert_state_equal(sg, so.state(), f"step {k}")
        assert list(rg.ramo) == list(ro.ramo)
    assert np.array_equal(dd.life_time(), g["life_time"][:, :2])


def _fe(target, oracle_mod, seed=5, V=1600.0, d=200 * NM, dt=0.5e-15, wf=((4.7,),), nic=1):
    if hasattr(target, "set_planar"):
        target.set_planar(V, d, True, nic)
        target.set_time_step(dt)
        target.set_seed(seed)
        target.set_emitter(10, 0, -50 * NM, -50 * NM, 100 * NM, 100 * NM, wf=np.array(wf))
    else:
        target.set_geometry(oracle_mod.geom_planar(d, V, True, nic))
        target.set_time_step(dt)
        target.set_seed(seed)
        target.set_emitter(10, 0, -50 * NM, -50 * NM, 100 * NM, 100 * NM, wf=np.array(wf))


def test_planar_fe_run_bit_exact(Diode, oracle_mod):
    """Planar field emission with space charge and absorption: every count,
    record and particle coordinate of the FP64 device run equals the oracle."""
    dd = Diode(4096)
    so = oracle_mod.Sim(4096, nthreads=8)
    _fe(dd, oracle_mod)
    _fe(so, oracle_mod)
    nsteps = 45
    recs = dd.run(nsteps, "fp64")
    absorbed = 0
    for k in range(1, nsteps + 1):
        ro = so.step(k)
        print(f"FE step {k}: n={recs[k-1].n} emitted={recs[k-1].emitted}", flush=True)
        assert_records_equal(recs[k - 1], ro, f"step {k}")
        absorbed += ro.absorbed
    assert absorbed > 0 and dd.n > 100
    assert_state_equal(dd.get_state(), so.state(), "end")
    assert np.array_equal(dd.life_time(), so.life_time())


def test_planar_fe_many_chains_bit_exact(Diode, oracle_mod):
    """About 2500 chains per step: the accept step of every jump spans three
    workgroups, whose batch counts set the shared width.  Records and state
    equal the oracle's."""
    def setup(t):
        if hasattr(t, "set_planar"):
            t.set_planar(8000.0, 1000 * NM, True, 0)
        else:
            t.set_geometry(oracle_mod.geom_planar(1000 * NM, 8000.0, True, 0))
        t.set_time_step(0.25e-15)
        t.set_seed(31)
        t.set_emitter(10, 0, -150 * NM, -150 * NM, 300 * NM, 300 * NM, wf=np.array([[4.7]]))

    dd = Diode(8192)
    so = oracle_mod.Sim(8192, nthreads=8)
    setup(dd)
    setup(so)
    recs = dd.run(10, "fp64")
    for k in range(1, 11):
        ro = so.step(k)
        print(f"many chains step {k}: n_round={recs[k-1].n_round} mh_std={recs[k-1].mh_std}",
              flush=True)
        assert_records_equal(recs[k - 1], ro, f"step {k}")
    assert recs[0].n_round > 2048
    assert_state_equal(dd.get_state(), so.state(), "end")


def test_output_files_equal_oracle(Diode, oracle_mod, tmp_path):
    """The out/ files written from the device's records are byte-identical to
    those written from the oracle's (planar FE; the tip's set too)."""
    from rumdeed_amd import output as O

    dd = Diode(4096)
    so = oracle_mod.Sim(4096, nthreads=8)
    _fe(dd, oracle_mod)
    _fe(so, oracle_mod)
    fg = O.OutputFiles(str(tmp_path / "gpu"), O.EMIT_PLANAR_FE, 0.5e-15, 1600.0, 200 * NM, 4096)
    fo = O.OutputFiles(str(tmp_path / "oracle"), O.EMIT_PLANAR_FE, 0.5e-15, 1600.0, 200 * NM,
                       4096)
    for k, r in enumerate(dd.run(25, "fp64"), start=1):
        fg.write_step(r)
        fo.write_step(oracle_mod.record_dict(so.step(k)))
    fg.close(dd.life_time())
    fo.close(so.life_time())
    names = sorted(os.listdir(tmp_path / "gpu"))
    assert names == sorted(os.listdir(tmp_path / "oracle")) and "ramo_current.dt" in names
    # one supply line per emission step
    assert len((tmp_path / "gpu" / "integration.dt").read_text().splitlines()) == 25
    for name in names:
        assert (tmp_path / "gpu" / name).read_bytes() == (tmp_path / "oracle" / name).read_bytes(), name


def test_particle_records_equal_oracle(Diode, oracle_mod, tmp_path):
    """Each step's emission and absorption particle records (the
    density_*.bin writes) equal the oracle's bit for bit, in the same order,
    and so do the density files written from them (planar FE with top
    absorption, then the tip)."""
    from rumdeed_amd import output as O

    from rumdeed_amd.diode import PLANES_Z_NM

    dd = Diode(4096)
    so = oracle_mod.Sim(4096, nthreads=8)
    _fe(dd, oracle_mod)
    _fe(so, oracle_mod)
    planes = [z * NM for z in PLANES_Z_NM]  # the input namelist's default planes
    planes[2] = 199.0 * NM  # one just below the anode (200 nm)
    dd.set_planes(planes)
    so.set_planes(planes)
    dd.set_logs(True)
    nsteps = 45
    kw = dict(densities=True, planes=planes, position_steps=nsteps)
    fg = O.OutputFiles(str(tmp_path / "gpu"), O.EMIT_PLANAR_FE, 0.5e-15, 1600.0, 200 * NM,
                       4096, **kw)
    fo = O.OutputFiles(str(tmp_path / "oracle"), O.EMIT_PLANAR_FE, 0.5e-15, 1600.0, 200 * NM,
                       4096, **kw)
    absorbed = crossings = 0
    for k in range(1, nsteps + 1):
        # the reference's step: emission, move, Write_Position, removal
        dd.do_emission(k, "fp64")
        dd.update_position(k, "fp64")
        sg = dd.get_state()
        rg = dd.remove_particles(k)
        so.step(k, phases=1)
        so.step(k, phases=2)
        s_o = so.state()
        ro = so.step(k, phases=4)
        assert_records_equal(rg, ro, f"step {k}")
        fg.write_position(k, sg["pos"], sg["emitter"], sg["section"], sg["id"])
        fo.write_position(k, s_o["pos"], s_o["emitter"], s_o["section"], s_o["id"])
        eg, ag = dd.step_logs()
        eo, ao = so.step_logs()
        pg, po = dd.plane_logs(), so.plane_logs()
        print(f"records step {k}: emitted={len(eg)} absorbed={len(ag)} crossings={len(pg)}",
              flush=True)
        assert eg.tobytes() == eo.tobytes(), f"emission records, step {k}"
        assert ag.tobytes() == ao.tobytes(), f"absorption records, step {k}"
        assert pg.tobytes() == po.tobytes(), f"plane records, step {k}"
        absorbed += len(ag)
        crossings += len(pg)
        fg.write_particles(eg, ag)
        fo.write_particles(eo, ao)
        fg.write_planes(pg)
        fo.write_planes(po)
    fg.close()
    fo.close()
    assert absorbed > 0 and crossings > 100
    names = sorted(os.listdir(tmp_path / "gpu"))
    assert names == sorted(os.listdir(tmp_path / "oracle")) and "planes-3.bin" in names
    for name in names:
        a = (tmp_path / "gpu" / name).read_bytes()
        assert a == (tmp_path / "oracle" / name).read_bytes(), name

    V, dtip, rb, ht = 1000.0, 1000 * NM, 250 * NM, 500 * NM
    dt_ = Diode(4096)
    dt_.set_tip(V, dtip, rb, ht, True, box_top=1000 * NM)
    dt_.set_time_step(0.25e-16)
    dt_.set_seed(9)
    dt_.set_emitter(3, 0, w_tip=4.7)
    dt_.set_logs(True)
    st = oracle_mod.Sim(4096, nthreads=8)
    st.set_geometry(oracle_mod.geom_tip(V, dtip, rb, ht, True), 1000 * NM)
    st.set_time_step(0.25e-16)
    st.set_seed(9)
    st.set_emitter(3, 0, w_tip=4.7)
    for k in range(1, 7):
        assert_records_equal(dt_.run(1, "fp64")[0], st.step(k), f"tip step {k}")
        eg, ag = dt_.step_logs()
        eo, ao = st.step_logs()
        assert eg.tobytes() == eo.tobytes() and ag.tobytes() == ao.tobytes(), f"tip step {k}"


def test_chunked_compaction_equals_oracle(Diode, oracle_mod):
    """A simulation of more than 8192 particles is compacted by chunks
    (k_compact_count / _move / _back, a workgroup per 1024 particles): 12 000
    electrons spread over a 200 nm gap with fast vertical velocities, so that
    over four 1 fs steps particles leave through both plates from every
    chunk.  Records, absorption records (in index order), lifetimes and the
    particle state equal the oracle's bit for bit."""
    rng = np.random.default_rng(23)
    n = 12000
    pos = np.column_stack([rng.uniform(-100, 100, n), rng.uniform(-100, 100, n),
                           rng.uniform(1, 199, n)]) * NM
    vel = np.column_stack([rng.normal(0, 2e5, n), rng.normal(0, 2e5, n),
                           rng.normal(0, 3e6, n)])
    dd = Diode(16384)
    so = oracle_mod.Sim(16384, nthreads=16)
    for t in (dd, so):
        if hasattr(t, "set_planar"):
            t.set_planar(1600.0, 200 * NM, True, 0)
        else:
            t.set_geometry(oracle_mod.geom_planar(200 * NM, 1600.0, True, 0))
        t.set_time_step(1e-15)
    dd.add_particles(pos, vel)
    so.add(pos, vel)
    dd.set_logs(True)
    removed = 0
    for k in range(1, 5):
        rg = dd.run(1, "fp64")[0]
        ro = so.step(k)
        print(f"chunked compaction step {k}: n={rg.n} absorbed={rg.absorbed}", flush=True)
        assert_records_equal(rg, ro, f"step {k}")
        _, ag = dd.step_logs()
        _, ao = so.step_logs()
        assert ag.tobytes() == ao.tobytes(), f"absorption records, step {k}"
        removed += rg.absorbed
    assert removed > 300 and rg.n > 8192  # every step compacted by chunks
    assert_state_equal(dd.get_state(), so.state(), "chunked compaction")
    assert np.array_equal(np.asarray(dd.life_time()).reshape(-1), np.asarray(so.life_time()).reshape(-1))


def test_particle_records_mixed(Diode):
    """Mixed precision: the records agree with the step's counts, ids run
    on, and each absorbed particle has left the particle list."""
    dd = Diode(4096)
    dd.set_planar(1600.0, 200 * NM, True, 1)
    dd.set_time_step(0.5e-15)
    dd.set_seed(5)
    dd.set_emitter(10, 0, -50 * NM, -50 * NM, 100 * NM, 100 * NM, wf=np.array([[4.7]]))
    dd.set_logs(True)
    next_id, absorbed = 0, 0
    for k in range(1, 46):
        ids_before = set(dd.get_state()["id"])
        r = dd.run(1, "mixed")[0]
        em, ab = dd.step_logs()
        ids_after = set(dd.get_state()["id"])
        assert len(em) == r.emitted and len(ab) == r.absorbed
        assert np.sum(ab["where"] == 1) == r.absorbed_top
        assert list(em["id"]) == list(range(next_id, next_id + len(em)))
        next_id += len(em)
        assert set(ab["id"]) == (ids_before | set(em["id"])) - ids_after
        absorbed += len(ab)
    assert absorbed > 0
    dd.set_logs(False)
    dd.run(1, "mixed")


def test_planar_gtf_run_bit_exact(Diode, oracle_mod):
    """Thermal-field emission, checkerboard work function, Poisson supply,
    Maxwell-Boltzmann velocities."""
    def setup(t):
        wf = np.array([[4.0, 4.5], [4.5, 4.0]])
        if hasattr(t, "set_planar"):
            t.set_planar(6000.0, 1000 * NM, True, 0)
            t.set_time_step(0.25e-15)
            t.set_seed(7)
            t.set_emitter(9, 0, -100 * NM, -100 * NM, 200 * NM, 200 * NM, T=1500.0, wf=wf)
        else:
            t.set_geometry(oracle_mod.geom_planar(1000 * NM, 6000.0, True, 0))
            t.set_time_step(0.25e-15)
            t.set_seed(7)
            t.set_emitter(9, 0, -100 * NM, -100 * NM, 200 * NM, 200 * NM, T=1500.0, wf=wf)

    dd = Diode(4096)
    so = oracle_mod.Sim(4096, nthreads=8)
    setup(dd)
    setup(so)
    recs = dd.run(20, "fp64")
    for k in range(1, 21):
        assert_records_equal(recs[k - 1], so.step(k), f"step {k}")
    st = dd.get_state()
    assert_state_equal(st, so.state(), "end")
    assert set(np.unique(st["section"])) <= {1, 2, 3, 4} and dd.n > 100


def test_tip_fe_run_bit_exact(Diode, oracle_mod):
    """Hyperboloid tip field emission (100 x 100 surface supply, MH in
    (xi, phi), emission along the surface normal, tip boundary)."""
    V, dtip, rb, ht = 1000.0, 1000 * NM, 250 * NM, 500 * NM
    dd = Diode(4096)
    dd.set_tip(V, dtip, rb, ht, True, box_top=1000 * NM)
    dd.set_time_step(0.25e-16)
    dd.set_seed(9)
    dd.set_emitter(3, 0, w_tip=4.7)
    so = oracle_mod.Sim(4096, nthreads=8)
    so.set_geometry(oracle_mod.geom_tip(V, dtip, rb, ht, True), 1000 * NM)
    so.set_time_step(0.25e-16)
    so.set_seed(9)
    so.set_emitter(3, 0, w_tip=4.7)
    recs = dd.run(6, "fp64")
    for k in range(1, 7):
        assert_records_equal(recs[k - 1], so.step(k), f"step {k}")
    assert_state_equal(dd.get_state(), so.state(), "end")


def _supply_case(target, oracle_mod, quad, levels, wf, pos=None, seed=11):
    """Thermal-field emission at 6 kV over 1000 nm, 200 x 200 nm emitter."""
    if hasattr(target, "set_planar"):
        target.set_planar(6000.0, 1000 * NM, True, 0)
    else:
        target.set_geometry(oracle_mod.geom_planar(1000 * NM, 6000.0, True, 0))
    target.set_time_step(0.25e-15)
    target.set_seed(seed)
    target.set_emitter(9, 0, -100 * NM, -100 * NM, 200 * NM, 200 * NM, T=1500.0,
                       wf=np.array(wf), quad=quad, quad_levels=levels)
    if pos is not None:
        n = len(pos)
        q, m = np.full(n, -1.602176634e-19), np.full(n, 9.1093837015e-31)
        if hasattr(target, "set_planar"):
            target.put_particles(pos, q, m)
        else:
            target.put(pos, q, m)


def _space_charge(n=30, seed=1):
    rng = np.random.default_rng(seed)
    return np.stack([rng.uniform(-100, 100, n), rng.uniform(-100, 100, n),
                     rng.uniform(2.0, 20.0, n)], 1) * NM


def test_supply_refinement_bit_exact(Diode, oracle_mod):
    """The refined supply rule (from 4 x 4 points, under 30 electrons 2-20 nm
    above the cathode, then under its own emission): the level reached, N_sup,
    the chains and every particle equal the oracle's on the FP64 path."""
    dd = Diode(4096)
    so = oracle_mod.Sim(4096, nthreads=8)
    pos = _space_charge()
    _supply_case(dd, oracle_mod, 4, 0, [[4.0]], pos)
    _supply_case(so, oracle_mod, 4, 0, [[4.0]], pos)
    recs = dd.run(8, "fp64")
    levels = []
    for k in range(1, 9):
        ro = so.step(k)
        assert_records_equal(recs[k - 1], ro, f"step {k}")
        levels.append(ro.supply_level)
    assert max(levels) >= 2 and max(levels) <= 4, levels
    assert_state_equal(dd.get_state(), so.state(), "end")


def test_supply_levels_in_a_pool(oracle_mod):
    """Three simulations whose supply stops at different levels (refined from
    4 x 4, the fixed 32 x 32 rule, a 7 x 7 checkerboard on 35 x 35 cells)
    advanced together in one pool equal the oracle running each alone."""
    from rumdeed_amd.diode import Pool

    wf7 = [[4.0 if (i + j) % 2 == 0 else 4.5 for i in range(7)] for j in range(7)]
    cases = [dict(quad=4, levels=0, wf=[[4.0]], pos=_space_charge(), seed=11),
             dict(quad=32, levels=-1, wf=[[4.2]], pos=_space_charge(20, 2), seed=12),
             dict(quad=32, levels=0, wf=wf7, pos=None, seed=13)]
    pool = Pool(3 * 2048, 8)
    slots, oracles = [], []
    for cs in cases:
        s = pool.open(2048)
        _supply_case(pool, oracle_mod, cs["quad"], cs["levels"], cs["wf"], cs["pos"], cs["seed"])
        slots.append(s)
        so = oracle_mod.Sim(2048, nthreads=8)
        _supply_case(so, oracle_mod, cs["quad"], cs["levels"], cs["wf"], cs["pos"], cs["seed"])
        oracles.append(so)
    seen = set()
    for k in range(1, 7):
        recs = pool.step("fp64")
        for s, so in zip(slots, oracles):
            ro = so.step(k)
            assert_records_equal(recs[s], ro, f"slot {s} step {k}")
            seen.add((s, ro.supply_level))
    assert (slots[1], 0) in seen and any(s == slots[0] and l >= 2 for s, l in seen), seen
    for s, so in zip(slots, oracles):
        pool.select(s)
        assert_state_equal(pool.get_state(), so.state(), f"slot {s}")


def test_pool_alone_equals_in_the_mix(oracle_mod):
    """Three different diodes advanced together in one pool give exactly the
    results each gives alone (the oracle runs each alone)."""
    from rumdeed_amd.diode import Pool

    cfgs = [dict(seed=5, V=1600.0, d=200 * NM), dict(seed=6, V=1500.0, d=180 * NM),
            dict(seed=7, V=1700.0, d=220 * NM, wf=((4.6, 4.8),))]
    pool = Pool(3 * 2048, 8)
    slots = []
    oracles = []
    for cfg in cfgs:
        s = pool.open(2048)
        _fe(pool, oracle_mod, **cfg)
        slots.append(s)
        so = oracle_mod.Sim(2048, nthreads=8)
        _fe(so, oracle_mod, **cfg)
        oracles.append(so)
    for k in range(1, 16):
        recs = pool.step("fp64")
        for s, so in zip(slots, oracles):
            assert_records_equal(recs[s], so.step(k), f"slot {s} step {k}")
    for s, so in zip(slots, oracles):
        pool.select(s)
        assert_state_equal(pool.get_state(), so.state(), f"slot {s}")


def test_mixed_precision_run_statistics(Diode, oracle_mod):
    """The mixed-precision step follows the FP64 step statistically: same
    supply within 1e-5, emitted totals within counting noise."""
    d64 = Diode(4096)
    d32 = Diode(4096)
    _fe(d64, oracle_mod, seed=11)
    _fe(d32, oracle_mod, seed=11)
    r64 = d64.run(30, "fp64")
    r32 = d32.run(30, "mixed")
    e64 = sum(r.emitted for r in r64)
    e32 = sum(r.emitted for r in r32)
    assert abs(r64[0].n_sup - r32[0].n_sup) < 1e-5 * r64[0].n_sup
    assert abs(e64 - e32) < 5 * np.sqrt(e64 + 1)


def test_mixed_tip_run_statistics(Diode):
    """Tip field emission in mixed precision follows the FP64 run: the same
    supply on the first step (empty diode) and emitted totals within counting
    noise."""
    def setup(dd):
        dd.set_tip(1000.0, 1000 * NM, 250 * NM, 500 * NM, True, box_top=1500 * NM)
        dd.set_time_step(0.25e-15)
        dd.set_seed(13)
        dd.set_emitter(3, 0, w_tip=4.5)

    d64 = Diode(1 << 14)
    d32 = Diode(1 << 14)
    setup(d64)
    setup(d32)
    r64 = d64.run(12, "fp64")
    r32 = d32.run(12, "mixed")
    assert r32[0].n_round == r64[0].n_round
    assert abs(r32[0].n_sup - r64[0].n_sup) <= 1e-12 * r64[0].n_sup
    e64 = sum(r.emitted for r in r64)
    e32 = sum(r.emitted for r in r32)
    print("tip emitted fp64", e64, "mixed", e32, flush=True)
    assert e64 > 10
    assert abs(e64 - e32) < 5 * np.sqrt(e64)


def test_mixed_precision_many_chains_statistics(Diode):
    """A 1000 x 1000 nm emitter at 8 GV/m starts hundreds of chains per step,
    so the jumps take the multi-workgroup path in both precisions; the mixed
    run follows the FP64 run within counting noise."""
    def setup(dd):
        dd.set_planar(8000.0, 1000 * NM, True, 0)
        dd.set_time_step(0.25e-15)
        dd.set_seed(21)
        dd.set_emitter(10, 0, -500 * NM, -500 * NM, 1000 * NM, 1000 * NM, wf=np.array([[4.7]]))

    d64 = Diode(1 << 15)
    d32 = Diode(1 << 15)
    setup(d64)
    setup(d32)
    r64 = d64.run(8, "fp64")
    r32 = d32.run(8, "mixed")
    print("chains per step", [r.n_round for r in r64], "emitted", [r.emitted for r in r64],
          [r.emitted for r in r32], flush=True)
    assert r64[0].n_round > 64
    assert r32[0].n_round == r64[0].n_round
    e64 = sum(r.emitted for r in r64)
    e32 = sum(r.emitted for r in r32)
    assert e64 > 100
    assert abs(e64 - e32) < 5 * np.sqrt(e64)
    # field-driven quantities agree closely while the trajectories have not diverged
    assert abs(r64[1].n_sup - r32[1].n_sup) < 1e-3 * r64[1].n_sup


def test_jump_graph_equals_separate_launches(oracle_mod):
    """Each step's Metropolis-Hastings jump loop launched as one HIP graph
    (captured when its launch parameters change, replayed while they do not,
    with the rest of the step launched eagerly in between): FP64 records equal
    the oracle's at every step, and mixed records and state equal those of
    separate launches."""
    from rumdeed_amd.diode import Pool

    cfgs = [dict(seed=5, V=1600.0, d=200 * NM), dict(seed=6, V=1500.0, d=180 * NM),
            dict(seed=7, V=1700.0, d=220 * NM, wf=((4.6, 4.8),))]

    def make(graphs):
        pool = Pool(3 * 2048, 4)
        pool.set_graphs(graphs)
        slots = []
        for cfg in cfgs:
            slots.append(pool.open(2048))
            _fe(pool, oracle_mod, **cfg)
        return pool, slots

    pool, slots = make(True)
    oracles = []
    for cfg in cfgs:
        so = oracle_mod.Sim(2048, nthreads=8)
        _fe(so, oracle_mod, **cfg)
        oracles.append(so)
    for k in range(1, 13):
        recs = pool.step("fp64")
        print(f"graph fp64 step {k}: n_round {[recs[s].n_round for s in slots]}", flush=True)
        for s, so in zip(slots, oracles):
            assert_records_equal(recs[s], so.step(k), f"slot {s} step {k}")
    for s, so in zip(slots, oracles):
        pool.select(s)
        assert_state_equal(pool.get_state(), so.state(), f"slot {s}")
    pool.close()

    runs = []
    for graphs in (True, False):
        pool, slots = make(graphs)
        recs = [[pool.step("mixed")[s].as_dict() for s in slots] for _ in range(12)]
        states = []
        for s in slots:
            pool.select(s)
            states.append(pool.get_state())
        runs.append((recs, states))
        pool.close()
    assert runs[0][0] == runs[1][0]
    for sg, se in zip(runs[0][1], runs[1][1]):
        assert_state_equal(sg, se, "mixed graph vs separate launches")


def test_mixed_pool_alone_equals_in_the_mix():
    """Mixed precision in a pool: a many-chain diode (jumps split over chain
    groups and source segments) and two small ones advance together; each
    gives exactly the records it gives alone, because every sum is split by
    the simulation's own size."""
    from rumdeed_amd.diode import Pool

    def cfg(pool, k):
        if k == 0:
            pool.set_planar(8000.0, 1000 * NM, True, 0)
            pool.set_time_step(0.25e-15)
            pool.set_seed(31)
            pool.set_emitter(10, 0, -400 * NM, -400 * NM, 800 * NM, 800 * NM,
                             wf=np.array([[4.7]]))
        else:
            pool.set_planar(1600.0 + 100.0 * k, 200 * NM, True, 1)
            pool.set_time_step(0.5e-15)
            pool.set_seed(31 + k)
            pool.set_emitter(10, 0, -50 * NM, -50 * NM, 100 * NM, 100 * NM,
                             wf=np.array([[4.7]]))

    caps = (8192, 1024, 1024)
    mix = Pool(sum(caps) + 1024, 4)
    slots = []
    for k, cap in enumerate(caps):
        s = mix.open(cap)
        cfg(mix, k)
        slots.append(s)
    steps = 6
    rec_mix = [mix.step("mixed") for _ in range(steps)]
    print("chains", [rec_mix[0][s].n_round for s in slots], flush=True)
    assert rec_mix[0][slots[0]].n_round > 64
    for k, cap in enumerate(caps):
        alone = Pool(cap + 1024, 2)
        s = alone.open(cap)
        cfg(alone, k)
        for t in range(steps):
            ra = alone.step("mixed")[s].as_dict()
            rm = rec_mix[t][slots[k]].as_dict()
            assert ra == rm, (k, t)
        alone.close()
    mix.close()


def test_mixed_jump_pairing_leaves_results_unchanged():
    """A pool's mixed jump launches may give each workgroup two partial sums
    (rd_set_jump_pairing); the partials keep their size and order, so every
    record and particle is the same with pairing always on and never on.
    The large diode starts with 3000 electrons, so its chain fields span
    several partials."""
    from rumdeed_amd.diode import Pool

    rng = np.random.default_rng(5)
    seed_pos = np.column_stack([rng.uniform(-400, 400, 3000), rng.uniform(-400, 400, 3000),
                                rng.uniform(20, 900, 3000)]) * NM

    def run(threshold):
        pool = Pool(8192 + 2 * 1024 + 1024, 4)
        pool.set_jump_pairing(threshold)
        slots = []
        for k, cap in enumerate((8192, 1024, 1024)):
            s = pool.open(cap)
            if k == 0:
                pool.set_planar(8000.0, 1000 * NM, True, 0)
                pool.set_time_step(0.25e-15)
                pool.set_seed(41)
                pool.set_emitter(10, 0, -400 * NM, -400 * NM, 800 * NM, 800 * NM,
                                 wf=np.array([[4.7]]))
                pool.add_particles(seed_pos)
            else:
                pool.set_planar(1600.0 + 100.0 * k, 200 * NM, True, 1)
                pool.set_time_step(0.5e-15)
                pool.set_seed(41 + k)
                pool.set_emitter(10, 0, -50 * NM, -50 * NM, 100 * NM, 100 * NM,
                                 wf=np.array([[4.7]]))
            slots.append(s)
        recs = [[r.as_dict() for r in pool.step("mixed")] for _ in range(3)]
        states = []
        for s in slots:
            pool.select(s)
            states.append(pool.get_state())
        pool.close()
        return slots, recs, states

    slots, rec_on, st_on = run(0)
    _, rec_off, st_off = run(1 << 40)
    print("chains", [rec_on[0][s]["n_round"] for s in slots],
          "electrons", [rec_on[0][s]["n"] for s in slots], flush=True)
    assert rec_on[0][slots[0]]["n_round"] > 64 and rec_on[0][slots[0]]["n"] > 2048
    assert rec_on == rec_off
    for a, b in zip(st_on, st_off):
        assert_state_equal(a, b, "jump pairing on vs off")


def test_mixed_serving_alone_equals_in_the_mix():
    """Mixed precision: every sum is split by the simulation's own size only,
    so a request gives identical records alone and in the mix.  The fourth
    request (8 kV over 1000 nm, 1000 x 1000 nm emitter) has its capacity
    given, so it fits both the mixed pool and its own."""
    from rumdeed_amd.serve import Request, Server

    reqs = [Request(rid=i, V=v, gap=g * NM, side=s * NM, phi=p, steps=st, seed=i, capacity=c)
            for i, (v, g, s, p, st, c) in enumerate([(1600.0, 200, 100, 4.7, 30, 0),
                                                     (1500.0, 180, 60, 4.5, 20, 0),
                                                     (1700.0, 220, 80, 4.6, 25, 0),
                                                     (8000.0, 1000, 1000, 4.7, 6, 1 << 16)])]
    srv = Server(total_particles=1 << 17, max_slots=8, precision="mixed")
    mix, stats = srv.run(reqs, realtime=False, keep_records=True)
    srv.close()
    assert stats["refused"] == 0 and stats["max_active"] == 4
    for r in reqs:
        s1 = Server(total_particles=1 << 17, max_slots=2, precision="mixed")
        alone, _ = s1.run([r], realtime=False, keep_records=True)
        s1.close()
        a, b = alone[0], mix[r.rid]
        print("request", r.rid, "emitted", a.emitted, flush=True)
        assert a.valid and b.valid
        assert a.emitted > 0
        assert a.records == b.records, r.rid


def test_serving_alone_equals_in_the_mix():
    """Continuous batching: requests served together (overlapping in the
    pool) give the same per-step records as each request served alone."""
    from rumdeed_amd.serve import Request, Server

    reqs = [Request(rid=i, V=v, gap=g * NM, side=s * NM, phi=p, steps=st, seed=i)
            for i, (v, g, s, p, st) in enumerate([(1600.0, 200, 100, 4.7, 30),
                                                  (1500.0, 180, 60, 4.5, 20),
                                                  (1700.0, 220, 80, 4.6, 25)])]
    srv = Server(total_particles=1 << 16, max_slots=8, precision="fp64")
    mix, stats = srv.run(reqs, realtime=False, keep_records=True)
    srv.close()
    assert stats["max_active"] == 3
    for r in reqs:
        s1 = Server(total_particles=1 << 15, max_slots=2, precision="fp64")
        alone, _ = s1.run([r], realtime=False, keep_records=True)
        s1.close()
        assert alone[0].records == mix[r.rid].records, r.rid
        assert alone[0].emitted == mix[r.rid].emitted > 0
