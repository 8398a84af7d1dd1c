"""The large space-charge-limited thermo-field emission diode at 100000
electrons: checkerboard cathode of 10 x 10 squares of 100 nm (4.0 / 4.5 eV)
under a 1000 x 1000 nm emitter, T = 1500 K, gap 1000 nm, 6 kV (at the
stated 20 V the cathode field is 2e7 V/m and nothing is emitted), dt =
0.25 fs, seed 7, image charge with N_ic_max = 0 (the reference's emission
loop, src/mod_field_thermo_emission.F90:136-193; one main-loop step,
src/main.F90:175-219).

The electrons (synthetic: uniform over the emitter, heights z = 1000 u^2 nm,
denser near the cathode as under space charge) are put on the device and in
the oracle; then
  - the FP64 step (emission with the refined supply rule over the 40 x 40
    section-aligned grid, Metropolis-Hastings sites, Beeman update with the
    all-pairs force in chunk bands, removal) equals the oracle bit for bit:
    every count, N_sup, the supply level and every particle;
  - the mixed accelerations (pair-symmetric form) meet the conditioning-aware
    bars against the long-double oracle on 1024 sampled targets."""
import numpy as np
import pytest

from conftest import rel_err
from test_gpu_config2 import _assert_bars, _report, _stats, _targets
from test_gpu_step import assert_records_equal, assert_state_equal

pytestmark = pytest.mark.gpu

NM = 1e-9
N = 100000


def _state(n=N, seed=7):
    rng = np.random.default_rng(seed)
    z = np.maximum(1000.0 * rng.uniform(0, 1, n) ** 2, 0.5)
    pos = np.stack([rng.uniform(-500, 500, n), rng.uniform(-500, 500, n), z], 1) * NM
    return pos, np.full(n, -1.602176634e-19), np.full(n, 9.1093837015e-31)


WF = np.where((np.arange(10)[:, None] + np.arange(10)[None, :]) % 2 == 0, 4.0, 4.5)


def _setup(t, oracle_mod):
    if hasattr(t, "set_planar"):
        t.set_planar(6000.0, 1000 * NM, True, 0)
    else:
        t.set_geometry(oracle_mod.geom_planar(1000 * NM, 6000.0, True, 0))
    t.set_time_step(0.25e-15)
    t.set_seed(7)
    t.set_emitter(9, 0, -500 * NM, -500 * NM, 1000 * NM, 1000 * NM, T=1500.0, wf=WF)


def test_config5_step_fp64_bit_exact(oracle_mod):
    from rumdeed_amd.diode import Diode

    pos, q, m = _state()
    cap = N + 8192
    dd = Diode(cap)
    _setup(dd, oracle_mod)
    dd.put_particles(pos, q, m)
    print(f"config-5 FP64 step, N={N}", flush=True)
    rec = dd.run(1, "fp64")[0]
    so = oracle_mod.Sim(cap)
    _setup(so, oracle_mod)
    so.put(pos, q, m)
    ro = so.step(1)
    print(f"config-5: emitted={ro.emitted} n_round={ro.n_round} N_sup={ro.n_sup:.4f} "
          f"level={ro.supply_level} absorbed={ro.absorbed}", flush=True)
    assert_records_equal(rec, ro, "config-5 step 1")
    assert_state_equal(dd.get_state(), so.state(), "config-5 step 1")
    assert ro.emitted > 0 and rec.n >= N
    dd.close()


def test_config5_mixed_forces_vs_oracle(oracle_mod):
    from rumdeed_amd.diode import Diode

    pos, q, m = _state()
    dd = Diode(N + 16)
    _setup(dd, oracle_mod)
    dd.put_particles(pos, q, m)
    am = dd.calculate_acceleration_particles("mixed")
    a64 = dd.calculate_acceleration_particles("fp64")
    dd.close()
    g = oracle_mod.geom_planar(1000 * NM, 6000.0, True, 0)
    t = _targets(N)
    ref64 = oracle_mod.accel_subset(g, pos, q, m, t)
    assert np.array_equal(a64[t], ref64)
    refld = oracle_mod.accel_subset(g, pos, q, m, t, long_double=True)
    scale = oracle_mod.accel_scale(g, pos, q, m, t)
    s_ld = _stats(am[t], refld, scale)
    s_64 = _stats(am, a64)
    _report({"case": "config5_force", "n": N, "n_ic_max": 0, "form": "pair-symmetric",
             "targets": int(t.size), "mixed_vs_long_double": s_ld, "mixed_vs_fp64_all": s_64})
    _assert_bars(s_ld)
    assert s_64["normwise"] < 1e-6
