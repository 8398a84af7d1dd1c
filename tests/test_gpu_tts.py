"""The two-time-step integrator on the device (rd_set_two_time_step:
two_time_step with atom_time_interval K, src/mod_verlet.F90:182-193,
234-290, 396-419, 511-580, 886-1151) against the oracle, which is pinned to
the reference's own routines (tests/test_oracle_tts.py)."""
import numpy as np
import pytest

from conftest import golden, rel_err
from test_gpu_step import assert_state_equal

pytestmark = pytest.mark.gpu

NM = 1e-9


@pytest.fixture(scope="module")
def Diode():
    from rumdeed_amd.diode import Diode

    return Diode


def _fill(t, g, K, oracle_mod=None):
    if oracle_mod is None:
        t.set_planar(float(g["V"]), float(g["d"]), True, 1)
    else:
        t.set_geometry(oracle_mod.geom_planar(float(g["d"]), float(g["V"]), True, 1))
    t.set_time_step(float(g["dt"]))
    t.set_two_time_step(K)
    n = len(g["species0"])
    for i in range(n):
        kw = dict(species=g["species0"][i:i + 1], step=int(g["birth0"][i]), emit=1)
        if oracle_mod is None:
            t.add_particles(g["pos0"][i:i + 1], g["vel0"][i:i + 1], **kw)
        else:
            t.add(g["pos0"][i:i + 1], g["vel0"][i:i + 1], **kw)


def test_two_time_step_fp64_bit_exact_and_vs_reference(Diode, oracle_mod):
    """The golden scenario (K = 3, 48 particles, every sixth an ion, ten steps
    with removal): FP64 device state equals the oracle bit for bit after every
    step, and the reference's within 1e-12."""
    g = golden("tts.npz")
    dd = Diode(256)
    _fill(dd, g, 3)
    so = oracle_mod.Sim(256, nthreads=2)
    _fill(so, g, 3, oracle_mod)
    for k in range(1, int(g["steps"]) + 1):
        print(f"tts step {k}", flush=True)
        rg = dd.update_position(k, "fp64")
        ro = so.step(k, phases=2)
        assert list(rg.ramo) == list(ro.ramo), k
        dd.remove_particles(k)
        so.step(k, phases=4)
        st = dd.get_state()
        assert_state_equal(st, so.state(), f"step {k}")
        for key in ("pos", "vel", "acc", "acc_prev", "acc_prev2"):
            assert rel_err(st[key], g[f"{key}_{k}"], floor=1e-300).max() < 1e-12, (k, key)
    dd.close()


def test_two_time_step_tip_fp64_bit_exact(Diode, oracle_mod):
    """The hyperboloid tip with K = 2: electrons and ions (ions first and
    last in index order), eight steps, FP64 device equals the oracle."""
    V, dtip, rb, ht = 1000.0, 1000 * NM, 250 * NM, 500 * NM
    rng = np.random.Generator(np.random.Philox(41))
    n = 40
    pos = np.stack([rng.uniform(-80, 80, n), rng.uniform(-80, 80, n),
                    rng.uniform(560, 950, n)], 1) * NM
    vel = rng.normal(0.0, 1.0e5, (n, 3))
    sp = np.where(np.arange(n) % 5 == 0, 2, 1).astype(np.int32)
    sp[-1] = 2
    dd = Diode(256)
    dd.set_tip(V, dtip, rb, ht, True, box_top=1000 * NM)
    dd.set_time_step(0.25e-15)
    dd.set_two_time_step(2)
    dd.add_particles(pos, vel, species=sp, step=1, emit=1)
    so = oracle_mod.Sim(256, nthreads=2)
    so.set_geometry(oracle_mod.geom_tip(V, dtip, rb, ht, True), 1000 * NM)
    so.set_time_step(0.25e-15)
    so.set_two_time_step(2)
    so.add(pos, vel, species=sp, step=1, emit=1)
    for k in range(1, 9):
        dd.update_position(k, "fp64")
        so.step(k, phases=2)
        dd.remove_particles(k)
        so.step(k, phases=4)
        assert_state_equal(dd.get_state(), so.state(), f"tip step {k}")
    with pytest.raises(RuntimeError):
        dd.update_position(9, "mixed")  # mixed two-time-step runs on the planar geometry only
    dd.close()


def test_two_time_step_mixed_close_to_fp64(Diode):
    """Mixed precision (the gather form with the two-time-step roles) against
    FP64 from the same state, one step at a time: accelerations within 1e-5
    relative for every particle (particles marked for removal have charge 0
    and zero acceleration in both)."""
    g = golden("tts.npz")
    out = {}
    for prec in ("fp64", "mixed"):
        dd = Diode(256)
        _fill(dd, g, 3)
        accs = []
        for k in range(1, 7):
            dd.update_position(k, prec)
            accs.append(dd.get_state()["acc"])
        out[prec] = accs
        dd.close()
    for k in range(6):
        a64, am = out["fp64"][k], out["mixed"][k]
        assert a64.shape == am.shape
        assert rel_err(am, a64, floor=1e-300).max() < 1e-5, k


def test_interval_one_equals_standard_on_device(Diode):
    """Test_TTS_Equivalence (src/mod_tests.F90:1455-1541) on the device:
    K = 1 with the electrons first reproduces the standard integrator bit for
    bit, in FP64 and in mixed precision."""
    R0 = np.array([[1.0, -2.0, 500.0], [3.0, 4.0, 503.0], [-2.0, 1.0, 505.0]]) * NM
    V0 = np.array([[1.0e3, 0.0, 0.0], [0.0, -1.0e3, 0.0], [0.0, 0.0, 0.0]])
    spc = np.array([1, 1, 2], dtype=np.int32)
    for prec in ("fp64", "mixed"):
        st = []
        for K in (0, 1):
            dd = Diode(16)
            dd.set_planar(2.0e3, 1000 * NM, True, 0)
            dd.set_time_step(0.25e-15)
            dd.set_two_time_step(K)
            dd.add_particles(R0, V0, species=spc, step=1, emit=1)
            ramo = [list(dd.update_position(k, prec).ramo) for k in (1, 2, 3)]
            st.append((dd.get_state(), ramo))
            dd.close()
        (a, ra), (b, rb) = st
        for key in ("pos", "vel", "acc", "acc_prev", "acc_prev2"):
            assert np.array_equal(a[key], b[key]), (prec, key)
        assert ra == rb, prec
