"""The oracle's time step (Beeman, boundaries, Ramo, removal) and emission
physics, pinned against the reference's own routines (golden vectors),
closed forms and statistics.  CPU only."""
import numpy as np
import pytest

from conftest import golden, rel_err

NM = 1e-9


def _sim_from_golden(oracle_mod, g):
    c = oracle_mod.consts()
    s = oracle_mod.Sim(256, nthreads=2)
    s.set_geometry(oracle_mod.geom_planar(float(g["d"]), float(g["V"]), True, 1))
    s.set_time_step(float(g["dt"]))
    n = len(g["species0"])
    for i in range(n):
        s.add(g["pos0"][i:i + 1], g["vel0"][i:i + 1], species=g["species0"][i:i + 1],
              step=int(g["birth0"][i]), emit=1)
    assert s.n == n
    return s, c


def test_steps_vs_reference_update_and_removal(oracle_mod):
    """Update_Position + Remove_Particles of the reference (golden steps.npz):
    same survivors in the same order, same ids, positions/velocities to 1e-12,
    same removal counts, same lifetime histogram."""
    g = golden("steps.npz")
    s, c = _sim_from_golden(oracle_mod, g)
    for k in range(1, int(g["steps"]) + 1):
        s.step(k, phases=2)
        rec_ramo = None
        r = s.step(k, phases=4)
        st = s.state()
        rem = g[f"removed_{k}"]
        assert r.absorbed == rem[0] and r.absorbed_top == rem[1] and r.absorbed_bot == rem[2]
        assert np.array_equal(st["id"], g[f"id_{k}"])
        assert np.array_equal(st["species"], g[f"species_{k}"])
        assert np.array_equal(st["step"], g[f"step_{k}"])
        for key in ("pos", "vel", "prev_pos", "acc_prev", "acc_prev2"):
            assert rel_err(st[key], g[f"{key}_{k}"], floor=1e-300).max() < 1e-12, (k, key)
        del rec_ramo
    life = s.life_time()
    assert np.array_equal(life, g["life_time"][:, :2])


def test_ramo_current_vs_reference(oracle_mod):
    g = golden("steps.npz")
    s, c = _sim_from_golden(oracle_mod, g)
    for k in range(1, int(g["steps"]) + 1):
        r = s.step(k, phases=2)
        ref = g[f"ramo_{k}"]
        assert abs(r.ramo_current - ref[3]) <= 1e-12 * np.abs(ref[:3]).sum()
        assert np.allclose(list(r.ramo), ref[:3], rtol=1e-12, atol=1e-12 * np.abs(ref[:3]).sum())
        s.step(k, phases=4)


def test_beeman_kinematics(oracle_mod):
    """Test_Beeman_Kinematics (src/mod_tests.F90:1403-1452): one electron in a
    uniform field with the seeded history; after 3 steps z - z0 = 9/2 a dt^2,
    v_z = 3 a dt, x - x0 = 3 v_x dt, I = q_0 v_z / d."""
    c = oracle_mod.consts()
    d, V, dt, vx0 = 1000 * NM, 2.0e3, 0.25e-15, 1.0e3
    s = oracle_mod.Sim(8, nthreads=1)
    s.set_geometry(oracle_mod.geom_planar(d, V, False, 0))
    s.set_time_step(dt)
    s.add(np.array([[0.0, 0.0, 500 * NM]]), np.array([[vx0, 0.0, 0.0]]))
    for k in range(1, 4):
        r = s.step(k, phases=6)
    st = s.state()
    a_z = c.q_0 * V / (c.m_0 * d)
    assert abs((st["pos"][0, 2] - 500 * NM) - 4.5 * a_z * dt ** 2) < 1e-9 * 4.5 * a_z * dt ** 2
    assert abs(st["vel"][0, 2] - 3 * a_z * dt) < 1e-9 * 3 * a_z * dt
    assert abs(st["pos"][0, 0] - 3 * vx0 * dt) < 1e-9 * 3 * vx0 * dt
    assert abs(r.ramo[0] - c.q_0 * 3 * a_z * dt / d) < 1e-9 * c.q_0 * 3 * a_z * dt / d


def test_fn_functions_known_answers(oracle_mod):
    """Test_FN_Functions (src/mod_tests.F90:1688-1736), phi = 4.7 eV."""
    v, t, lnD, _ = oracle_mod.fn_functions(1, -2.0e9, 4.7)
    assert abs(v - 0.8253581935658024) < 1e-13
    assert abs(t - 1.0292422630703880) < 1e-13
    assert abs(lnD - (-28.723444978828507)) < 1e-12
    v, t, lnD, _ = oracle_mod.fn_functions(1, -4.0e9, 4.7)
    assert abs(v - 0.6808388366998485) < 1e-13
    assert abs(t - 1.0484437096180323) < 1e-13
    assert abs(lnD - (-11.846999895226958)) < 1e-12
    v, t, _, _ = oracle_mod.fn_functions(1, -1.0e12, 4.7)
    assert v == 0.0  # l clamps to 1
    v, t, _, _ = oracle_mod.fn_functions(0, -2.0e9, 4.7)
    assert v == 1.0 and t == 1.0
    # supply log = 2 ln(-F) - 2 ln t - ln w
    _, t, _, sl = oracle_mod.fn_functions(1, -3.0e9, 4.5)
    assert abs(sl - (2 * np.log(3.0e9) - 2 * np.log(t) - np.log(4.5))) < 1e-12


def test_gtf_vs_reference_table(oracle_mod):
    g = golden("gtf.npz")
    for i, F in enumerate(g["F"]):
        for j, T in enumerate(g["T"]):
            for k, w in enumerate(g["w"]):
                ref = g["J"][i, j, k]
                val = oracle_mod.gtf(-F, T, w)
                assert abs(val - ref) <= 1e-11 * abs(ref) + 1e-300, (F, T, w, val, ref)


def test_checkerboard_vs_reference(oracle_mod):
    g = golden("work_checker.npz")
    w, sec = oracle_mod.work_checker(g["table"], float(g["x0"]), float(g["y0"]), float(g["lx"]),
                                     float(g["ly"]), g["pts"])
    assert np.array_equal(w, g["w"])
    assert np.array_equal(sec, g["sec"])


def test_sincos_within_one_ulp(oracle_mod):
    rng = np.random.default_rng(4)
    xs = np.concatenate([rng.uniform(-7, 7, 4000), np.linspace(0, 2 * np.pi, 257)])
    for x in xs:
        s, c = oracle_mod.sincos(x)
        assert abs(s - np.sin(x)) <= max(np.spacing(abs(np.sin(x))), 2.3e-16)
        assert abs(c - np.cos(x)) <= max(np.spacing(abs(np.cos(x))), 2.3e-16)


def _fe_sim(oracle_mod, V=20e3, d=2500 * NM, wf=((4.7,),), seed=1, side=100.0, dt=0.1e-15,
            ic=True, nic=1, quad=32):
    s = oracle_mod.Sim(20000, nthreads=4)
    s.set_geometry(oracle_mod.geom_planar(d, V, ic, nic))
    s.set_time_step(dt)
    s.set_seed(seed)
    s.set_emitter(10, 0, -side / 2 * NM, -side / 2 * NM, side * NM, side * NM, wf=np.array(wf),
                  quad=quad)
    return s


def test_supply_integral_uniform_field(oracle_mod):
    """Empty diode: the field is uniform, so the midpoint supply integral is
    exactly A * Elec_Supply_V2(F)."""
    c = oracle_mod.consts()
    V, d, dt = 20e3, 2500 * NM, 0.1e-15
    s = _fe_sim(oracle_mod, V=V, d=d, dt=dt)
    r = s.step(1, phases=1)
    F = -V / d
    _, t, _, _ = oracle_mod.fn_functions(1, F, 4.7)
    sup = dt / c.q_0 * c.a_FN / (t * t * 4.7) * F * F
    A = (100 * NM) ** 2
    assert abs(r.n_sup - A * sup) < 1e-12 * A * sup
    assert r.n_round == int(np.floor(A * sup + 0.5))


def test_emission_site_statistics(oracle_mod):
    """Two-section work function under a uniform field: emitted electrons
    split between the sections as supply * escape probability (chi-square),
    and the emitted count matches the expected sum of D over the supply."""
    V, d, dt = 20e3, 2500 * NM, 0.1e-15
    wf = ((4.3, 4.7),)
    # one emission step from an empty diode: every chain sees the uniform field
    s = _fe_sim(oracle_mod, V=V, d=d, dt=dt, wf=wf, seed=3, ic=True, side=1000.0)
    F = -V / d
    r = s.step(1, phases=1)
    st = s.state()
    counts = np.array([np.sum(st["section"] == 1), np.sum(st["section"] == 2)], dtype=float)
    assert r.emitted == counts.sum()
    w = np.array([4.3, 4.7])
    S = np.array([oracle_mod.fn_functions(1, F, wi)[1] for wi in w])
    sup = 1.0 / (S * S * w)  # supply per area, common factors dropped
    D = np.exp([oracle_mod.fn_functions(1, F, wi)[2] for wi in w])
    p = sup * D / np.sum(sup * D)
    N = counts.sum()
    assert N > 500
    chi2 = np.sum((counts - N * p) ** 2 / (N * p))
    assert chi2 < 10.8  # 1 dof, p = 0.001


def test_poisson_counts(oracle_mod):
    """Thermal-field emission draws the number of chains from Rand_Poisson:
    mean and variance over steps agree with the supply."""
    ns, sups = [], []
    for seed in range(150):
        # first step of an empty diode: the same supply for every seed
        s = oracle_mod.Sim(4000, nthreads=1)
        s.set_geometry(oracle_mod.geom_planar(1000 * NM, 6000.0, True, 0))
        s.set_time_step(0.25e-15)
        s.set_seed(seed)
        s.set_emitter(9, 0, -50 * NM, -50 * NM, 100 * NM, 100 * NM, T=1500.0,
                      wf=np.array([[2.5]]), quad=8, ndim=1)
        r = s.step(1, phases=1)
        ns.append(r.n_round)
        sups.append(r.n_sup)
    lam = sups[0]
    assert np.all(np.array(sups) == lam) and lam > 100
    z = (np.array(ns) - lam) / np.sqrt(lam)
    assert abs(z.mean()) < 4 / np.sqrt(len(z))
    assert 0.7 < z.var() < 1.4


def test_particle_records(oracle_mod):
    """The per-step particle records (the density_emit / density_absorb
    writes): counts equal the step record's, ids run on from step to step,
    planar FE sites sit 1 nm above the cathode, and every absorbed particle
    is one that vanished, listed in index order with its position after the
    Beeman update and the velocity it had before this step's velocity
    update."""
    s = oracle_mod.Sim(4096, nthreads=4)
    s.set_geometry(oracle_mod.geom_planar(200 * NM, 1600.0, True, 1))
    s.set_time_step(0.5e-15)
    s.set_seed(5)
    s.set_emitter(10, 0, -50 * NM, -50 * NM, 100 * NM, 100 * NM, wf=np.array([[4.7]]))
    planes = np.array([5.0, 10.0, -1.0, 50.0, 75.0, 100.0, 125.0, 150.0, 199.0, 250.0]) * NM
    s.set_planes(planes)
    dt = 0.5e-15
    next_id = 0
    top = bot = crossings = 0
    for k in range(1, 46):
        before = s.state()
        r = s.step(k)
        em, ab = s.step_logs()
        after = s.state()
        # plane crossings from the states on either side of the step:
        # z_prev < z_p < z for each particle in id (= index) order
        zold = dict(zip(before["id"].tolist(), before["pos"][:, 2]))
        zold.update(zip(em["id"].tolist(), em["z"]))
        vold = dict(zip(before["id"].tolist(), map(tuple, before["vel"])))
        vold.update({int(i): (0.0, 0.0, 0.0) for i in em["id"]})
        znew = dict(zip(after["id"].tolist(), after["pos"][:, 2]))
        znew.update(zip(ab["id"].tolist(), ab["z"]))
        want = [(p + 1, i) for p, zp in enumerate(planes) if zp > 0
                for i in sorted(zold) if zold[i] < zp < znew[i]]
        pl = s.plane_logs()
        assert [(int(a), int(b)) for a, b in zip(pl["plane"], pl["id"])] == want
        for rec in pl:
            assert (rec["vx"], rec["vy"], rec["vz"]) == vold[int(rec["id"])]
        crossings += len(pl)
        assert len(em) == r.emitted and len(ab) == r.absorbed
        assert np.sum(ab["where"] == 1) == r.absorbed_top
        assert np.sum(ab["where"] == 2) == r.absorbed_bot
        assert list(em["id"]) == list(range(next_id, next_id + len(em)))
        next_id += len(em)
        assert np.all(em["z"] == 1.0 * NM) and np.all(em["species"] == 1)
        assert np.all(em["emitter"] == 1) and np.all(em["section"] == 1)
        ids = list(before["id"]) + list(em["id"])
        gone = [i for i in ids if i not in set(after["id"])]
        assert list(ab["id"]) == gone
        idx = {int(i): j for j, i in enumerate(before["id"])}
        for a in ab:
            j = idx.get(int(a["id"]))
            if j is None:  # emitted this step with zero velocity
                assert a["vx"] == a["vy"] == a["vz"] == 0.0
                continue
            assert [a["vx"], a["vy"], a["vz"]] == list(before["vel"][j])
            x = (before["pos"][j] + before["vel"][j] * dt +
                 (4 * before["acc"][j] - before["acc_prev"][j]) / 6 * dt * dt)
            assert abs(a["x"] - x[0]) <= 1e-15 * 200 * NM and abs(a["y"] - x[1]) <= 1e-15 * 200 * NM
            assert (x[2] > 200 * NM) == (a["where"] == 1)
        top += r.absorbed_top
        bot += r.absorbed_bot
    assert top > 0 and crossings > 100
