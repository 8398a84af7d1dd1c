"""The two-time-step integrator of the oracle (two_time_step with
atom_time_interval K: Update_Particle_Position / _Acceleration / _Velocity,
src/mod_verlet.F90:182-193, 396-419, 886-897, and the species routines
:234-290, 511-580, 899-1151), pinned against the reference's own routines
(golden tts.npz, written by the flang build of the reference) and against
the reference's equivalence test (src/mod_tests.F90:1455-1541).  CPU only."""
import numpy as np

from conftest import golden, rel_err

NM = 1e-9


def _sim(oracle_mod, g, K):
    s = oracle_mod.Sim(256, nthreads=2)
    s.set_geometry(oracle_mod.geom_planar(float(g["d"]), float(g["V"]), True, 1))
    s.set_time_step(float(g["dt"]))
    s.set_two_time_step(K)
    n = len(g["species0"])
    for i in range(n):
        s.add(g["pos0"][i:i + 1], g["vel0"][i:i + 1], species=g["species0"][i:i + 1],
              step=int(g["birth0"][i]), emit=1)
    assert s.n == n
    return s


def test_two_time_step_vs_reference():
    """K = 3, electrons and ions (an ion first in index order), ten steps with
    removal: the same survivors, ids and removal counts as the reference;
    positions, velocities and all three accelerations within 1e-12; the
    Ramo current per species within 1e-12 of its scale."""
    import oracle as O

    g = golden("tts.npz")
    s = _sim(O, g, int(g["K"]))
    for k in range(1, int(g["steps"]) + 1):
        r = s.step(k, phases=2)
        ref = g[f"ramo_{k}"]
        assert np.allclose(list(r.ramo), ref[:3], rtol=1e-12, atol=1e-12 * np.abs(ref[:3]).sum())
        r = s.step(k, phases=4)
        st = s.state()
        rem = g[f"removed_{k}"]
        assert r.absorbed == rem[0] and r.absorbed_top == rem[1] and r.absorbed_bot == rem[2], k
        assert np.array_equal(st["id"], g[f"id_{k}"])
        assert np.array_equal(st["species"], g[f"species_{k}"])
        for key in ("pos", "vel", "prev_pos", "acc", "acc_prev", "acc_prev2"):
            assert rel_err(st[key], g[f"{key}_{k}"], floor=1e-300).max() < 1e-12, (k, key)
    assert np.array_equal(s.life_time(), g["life_time"][:, :2])


def test_ions_rest_between_their_steps():
    """Between multiples of K an ion keeps its position, velocity and
    acceleration history; its current acceleration only grows by the
    electrons' reactions."""
    import oracle as O

    g = golden("tts.npz")
    s = _sim(O, g, 3)
    s.step(1, phases=2)
    a = s.state()
    s.step(2, phases=2)
    b = s.state()
    ion = a["species"] == 2
    for key in ("pos", "vel", "prev_pos", "acc_prev", "acc_prev2"):
        assert np.array_equal(a[key][ion], b[key][ion]), key
    assert not np.array_equal(a["acc"][ion], b["acc"][ion])
    assert not np.array_equal(a["pos"][~ion], b["pos"][~ion])


def test_interval_one_equals_the_standard_integrator():
    """Test_TTS_Equivalence (src/mod_tests.F90:1455-1541): two electrons and
    an ion a few nm apart, three steps; with atom_time_interval = 1 and the
    electrons first in index order the two-time-step integrator reproduces
    the standard one (here bit for bit: the same terms in the same order)."""
    import oracle as O

    R0 = np.array([[1.0, -2.0, 500.0], [3.0, 4.0, 503.0], [-2.0, 1.0, 505.0]]) * NM
    V0 = np.array([[1.0e3, 0.0, 0.0], [0.0, -1.0e3, 0.0], [0.0, 0.0, 0.0]])
    spc = np.array([1, 1, 2], dtype=np.int32)
    out = []
    for K in (0, 1):
        s = O.Sim(16, nthreads=1)
        s.set_geometry(O.geom_planar(1000 * NM, 2.0e3, True, 0))
        s.set_time_step(0.25e-15)
        s.set_two_time_step(K)
        for i in range(3):
            s.add(R0[i:i + 1], V0[i:i + 1], species=spc[i:i + 1], step=1, emit=1)
        recs = [s.step(k, phases=2) for k in (1, 2, 3)]
        out.append((s.state(), [list(r.ramo) for r in recs]))
    (a, ra), (b, rb) = out
    for key in ("pos", "vel", "acc", "acc_prev", "acc_prev2"):
        assert np.array_equal(a[key], b[key]), key
    assert ra == rb


def test_electron_ion_role_follows_species():
    """Between an electron and an ion the electron takes the first role of the
    pair in the two-time-step accelerations (its image series is evaluated at
    the electron), whatever their index order; the standard accelerations
    give it to the lower index.  With N_ic_max = 1 the same-charge partners
    of the series are not mirror-symmetric (the reference's lower-index
    convention), so an ion placed first in index order gives different
    accelerations in the two modes, and equal ones placed last."""
    import oracle as O

    g = O.geom_planar(1000 * NM, 2.0e3, True, 1)
    pos = np.array([[0.0, 0.0, 300.0], [2.0, 1.0, 310.0], [-3.0, 2.0, 290.0]]) * NM
    q = np.array([1.0, -1.0, -1.0]) * 1.602176634e-19
    m = np.array([4.6e-26, 9.1093837015e-31, 9.1093837015e-31])
    sp = np.array([2, 1, 1], dtype=np.int32)
    std = O.accel(g, pos, q, m)
    tts = O.accel_tts(g, pos, q, m, sp, 1, 1)
    assert not np.array_equal(std[1:], tts[1:])
    order = [1, 2, 0]  # electrons first: the roles agree
    std2 = O.accel(g, pos[order], q[order], m[order])
    tts2 = O.accel_tts(g, pos[order], q[order], m[order], sp[order], 1, 1)
    assert np.array_equal(std2, tts2)
