"""HIP force/field kernels against the oracle and the reference's golden
vectors.  FP64 kernels must be bit-identical to the oracle (same summation
order); mixed-precision kernels must stay within 1e-5 relative error of the
long-double oracle."""
import numpy as np
import pytest

from conftest import golden, rel_err

pytestmark = pytest.mark.gpu

NM = 1e-9
MIXED_TOL = 1e-5


@pytest.fixture(scope="module")
def Diode():
    from rumdeed_amd.diode import Diode

    return Diode


def _diode_planar(Diode, pos, q, m, d, V, ic, nic, species=None):
    dd = Diode(max(len(pos), 1) + 16)
    dd.set_planar(V, d, ic, nic)
    dd.put_particles(pos, q, m, species=species)
    return dd


@pytest.mark.parametrize("n", [1, 2, 255, 256, 257, 1000, 3000])
@pytest.mark.parametrize("ic,nic", [(0, 0), (1, 0), (1, 1), (1, 10)])
def test_planar_accel_fp64_bit_exact(Diode, oracle_mod, n, ic, nic):
    from rumdeed_amd import workloads as W

    pos, q, m = W.snapshot(n, seed=11 + n)
    print(f"planar fp64 n={n} ic={ic} nic={nic}", flush=True)
    dd = _diode_planar(Diode, pos, q, m, W.SNAPSHOT_D, W.SNAPSHOT_V, ic, nic)
    a_gpu = dd.calculate_acceleration_particles("fp64")
    g = oracle_mod.geom_planar(W.SNAPSHOT_D, W.SNAPSHOT_V, bool(ic), nic)
    a_cpu = oracle_mod.accel(g, pos, q, m)
    assert np.array_equal(a_gpu, a_cpu)


@pytest.mark.parametrize("nic", [0, 1, 10])
def test_planar_golden_snapshot(Diode, nic):
    s = golden("snapshot512.npz")
    dd = _diode_planar(Diode, s["pos"], s["q"], s["m"], float(s["d"]), float(s["V"]), 1, nic)
    a = dd.calculate_acceleration_particles("fp64")
    assert rel_err(a, s[f"acc_n{nic}"]).max() < 1e-12
    f = dd.calc_field_at_batch(s["grid"], "fp64")
    assert rel_err(f, s[f"field_n{nic}"]).max() < 1e-12
    am = dd.calculate_acceleration_particles("mixed")
    assert rel_err(am, s[f"acc_n{nic}"]).max() < MIXED_TOL
    fm = dd.calc_field_at_batch(s["grid"], "mixed")
    assert rel_err(fm, s[f"field_n{nic}"]).max() < MIXED_TOL


def test_forty_and_three_body_golden(Diode):
    f = golden("forty.npz")
    for ic, nic in [(0, 0), (1, 0), (1, 2), (1, 10)]:
        dd = _diode_planar(Diode, f["pos"], f["q"], f["m"], float(f["d"]), float(f["V"]), ic, nic,
                           species=f["species"])
        a = dd.calculate_acceleration_particles("fp64")
        assert rel_err(a, f[f"acc_gather_ic{ic}_n{nic}"]).max() < 1e-12
        am = dd.calculate_acceleration_particles("mixed")
        assert rel_err(am, f[f"acc_gather_ic{ic}_n{nic}"]).max() < MIXED_TOL
        fl = dd.calc_field_at_batch(f["pts"], "fp64")
        assert rel_err(fl, f[f"field_ic{ic}_n{nic}"]).max() < 1e-12
    t = golden("three_body.npz")
    dd = _diode_planar(Diode, t["pos"], t["q"], t["m"], float(t["d"]), float(t["V"]), 0, 0,
                       species=t["species"])
    assert rel_err(dd.calculate_acceleration_particles("fp64"), t["acc"]).max() < 1e-12
    assert rel_err(dd.calc_field_at(t["probe"]), t["field"]).max() < 1e-12


def test_tip_fp64_bit_exact_and_golden(Diode, oracle_mod):
    t = golden("tip64.npz")
    dd = Diode(128)
    dd.set_tip(float(t["V"]), float(t["d_tip"]), float(t["R_base"]), float(t["h_tip"]))
    dd.put_particles(t["pos"], t["q"], t["m"])
    a = dd.calculate_acceleration_particles("fp64")
    g = oracle_mod.geom_tip(float(t["V"]), float(t["d_tip"]), float(t["R_base"]), float(t["h_tip"]))
    assert np.array_equal(a, oracle_mod.accel(g, t["pos"], t["q"], t["m"]))
    assert rel_err(a, t["acc_gather"]).max() < 1e-12
    f = dd.calc_field_at_batch(t["pts"], "fp64")
    assert np.array_equal(f, oracle_mod.field(g, t["pos"], t["q"], t["pts"]))
    assert rel_err(f, t["field"]).max() < 1e-12


@pytest.mark.parametrize("n", [0, 1, 300, 2500])
@pytest.mark.parametrize("nic", [0, 10])
def test_field_fp64_bit_exact(Diode, oracle_mod, n, nic):
    from rumdeed_amd import workloads as W

    pos, q, m = W.snapshot(max(n, 1), seed=5)
    pos, q, m = pos[:n], q[:n], m[:n]
    dd = _diode_planar(Diode, pos, q, m, W.SNAPSHOT_D, W.SNAPSHOT_V, 1, nic)
    pts = W.grid_points(24, 20)
    f = dd.calc_field_at_batch(pts, "fp64")
    g = oracle_mod.geom_planar(W.SNAPSHOT_D, W.SNAPSHOT_V, True, nic)
    assert np.array_equal(f, oracle_mod.field(g, pos, q, pts))
    if n == 0:
        assert np.all(f[:, :2] == 0) and np.all(f[:, 2] == -W.SNAPSHOT_V / W.SNAPSHOT_D)
    grid = dd.field_grid(*W.SNAPSHOT_BOX, 24, 20, 0.0, "fp64")
    assert np.array_equal(grid, f)


@pytest.mark.parametrize("n", [4096])
@pytest.mark.parametrize("nic", [0, 1, 10])
def test_mixed_error_vs_long_double(Diode, oracle_mod, n, nic):
    """Mixed precision against the long-double oracle on the force/field
    snapshot: per-particle relative error <= 1e-5."""
    from rumdeed_amd import workloads as W

    pos, q, m = W.snapshot(n, seed=2)
    dd = _diode_planar(Diode, pos, q, m, W.SNAPSHOT_D, W.SNAPSHOT_V, 1, nic)
    g = oracle_mod.geom_planar(W.SNAPSHOT_D, W.SNAPSHOT_V, True, nic)
    ref = oracle_mod.accel(g, pos, q, m, long_double=True)
    am = dd.calculate_acceleration_particles("mixed")
    e = rel_err(am, ref)
    print(f"mixed accel n={n} nic={nic}: max {e.max():.2e} p99 {np.percentile(e, 99):.2e}")
    assert e.max() < MIXED_TOL
    a64 = dd.calculate_acceleration_particles("fp64")
    assert rel_err(a64, ref).max() < 1e-12
    pts = W.grid_points(32, 32)
    fref = oracle_mod.field(g, pos, q, pts, long_double=True)
    fm = dd.calc_field_at_batch(pts, "mixed")
    ef = rel_err(fm, fref)
    print(f"mixed field n={n} nic={nic}: max {ef.max():.2e}")
    assert ef.max() < MIXED_TOL


@pytest.mark.parametrize("n,ic,nic", [(16384, 1, 10), (16500, 0, 0), (16500, 1, 1), (40000, 1, 10)])
def test_mixed_pair_symmetric_form(Diode, n, ic, nic):
    """From 16384 particles the mixed planar force evaluates each unordered
    pair once (blocks of 1024, the last one partial here for n != 16384).
    Against the FP64 kernel, which the other tests pin bit-exactly to the
    oracle: per-particle relative error <= 1e-5; repeated calls are
    bit-identical."""
    from rumdeed_amd import workloads as W

    pos, q, m = W.snapshot(n, seed=21)
    q = q.copy()
    q[7::97] = 0.0  # absorbed particles: no force on them, none from them
    dd = _diode_planar(Diode, pos, q, m, W.SNAPSHOT_D, W.SNAPSHOT_V, ic, nic)
    a64 = dd.calculate_acceleration_particles("fp64")
    am = dd.calculate_acceleration_particles("mixed")
    keep = q != 0.0
    e = rel_err(am[keep], a64[keep])
    print(f"symmetric mixed n={n} ic={ic} nic={nic}: max {e.max():.2e} p99 {np.percentile(e, 99):.2e}")
    assert e.max() < MIXED_TOL
    assert np.all(am[~keep] == 0.0)
    assert np.array_equal(dd.calculate_acceleration_particles("mixed"), am)


@pytest.mark.parametrize("nic", [0, 1, 10])
def test_mixed_field_on_and_off_the_cathode(Diode, oracle_mod, nic):
    """Points on the cathode plane take the paired-image form of the series
    (lateral image terms cancel there); a wavefront holding any point off the
    plane takes the general form.  Both within 1e-5 of the long-double
    oracle, and the lateral field on the plane is zero."""
    from rumdeed_amd import workloads as W

    pos, q, m = W.snapshot(2000, seed=9)
    dd = _diode_planar(Diode, pos, q, m, W.SNAPSHOT_D, W.SNAPSHOT_V, 1, nic)
    g = oracle_mod.geom_planar(W.SNAPSHOT_D, W.SNAPSHOT_V, True, nic)
    on = W.grid_points(16, 16)  # 256 points, z = 0
    mixed = on.copy()
    mixed[5::64, 2] = 1e-12  # one point off the plane in every wavefront
    above = on.copy()
    above[:, 2] = np.linspace(1e-9, 900e-9, len(above))
    for pts in (on, mixed, above):
        ref = oracle_mod.field(g, pos, q, pts, long_double=True)
        f = dd.calc_field_at_batch(pts, "mixed")
        e = rel_err(f, ref)
        print(f"nic={nic} z in [{pts[:, 2].min():.1e}, {pts[:, 2].max():.1e}]: max {e.max():.2e}")
        assert e.max() < MIXED_TOL
    f = dd.calc_field_at_batch(on, "mixed")
    assert np.all(f[:, :2] == 0.0)


@pytest.mark.parametrize("d_nm", [200, 1000, 5000])
@pytest.mark.parametrize("nic", [1, 10])
def test_mixed_cathode_field_far_series_and_fallback(Diode, oracle_mod, nic, d_nm):
    """On the cathode the image partners k != 0 of a source are summed as a
    power series in t = s / d^2 (s the squared lateral distance) while
    t <= 0.1, and one by one beyond it.  Sources spread five gaps wide put
    the pairs of every wavefront on both sides of that bound.  Gaps of 200 nm
    to 5 um (the series' coefficients scale as d^-2 in t and stay in FP32's
    range; in s they would fall as d^-(2m+2)), at a vacuum field of 2 GV/m:
    per-point error <= 1e-5 against the long-double oracle where the field is
    not a cancellation (kappa <= 50), scaled error <= 1e-6 everywhere."""
    from rumdeed_amd import workloads as W
    from test_gpu_config2 import _assert_bars, _stats

    d = d_nm * 1e-9
    V = W.SNAPSHOT_V * d / W.SNAPSHOT_D
    pos, q, m = W.snapshot(1500, seed=13)
    pos[:, :2] *= 5.0 * d / 200e-9  # lateral +- 2.5 d
    pos[:, 2] *= d / W.SNAPSHOT_D  # z in (0, d]
    dd = _diode_planar(Diode, pos, q, m, d, V, 1, nic)
    g = oracle_mod.geom_planar(d, V, True, nic)
    on = W.grid_points(16, 16, box=(-2.5 * d, -2.5 * d, 5 * d, 5 * d))
    rho2 = ((on[:, None, :2] - pos[None, :, :2]) ** 2).sum(-1)
    near = (rho2 <= 0.1 * d * d).mean()
    assert 0.001 < near < 0.5, near
    ref = oracle_mod.field(g, pos, q, on, long_double=True)
    scale = oracle_mod.field_scale(g, pos, q, on)
    f = dd.calc_field_at_batch(on, "mixed")
    st = _stats(f, ref, scale)
    print(f"cathode field d={d_nm} nm nic={nic}: {near:.3f} of the pairs in the series range, "
          f"{st}", flush=True)
    _assert_bars(st)
    if d_nm == 200:
        assert st["max"] < MIXED_TOL
    assert np.all(f[:, :2] == 0.0)


def vacuum_accel(oracle_mod, g, p, q, m):
    return q / m * oracle_mod.vacuum_field(g, p)


@pytest.mark.parametrize("n", [1, 300, 3000])
@pytest.mark.parametrize("ic", [0, 1])
def test_mixed_tip_error_vs_long_double(Diode, oracle_mod, n, ic):
    """Mixed precision on the hyperboloid tip (vacuum field in FP64, pair and
    sphere-image terms in FP32): per-particle relative error <= 1e-5 against
    the long-double oracle, for accelerations and for the field at points."""
    V, dtip, rb, ht = 1000.0, 1000e-9, 250e-9, 500e-9
    rng = np.random.Generator(np.random.Philox(17))
    pos = np.column_stack([rng.uniform(-150e-9, 150e-9, n), rng.uniform(-150e-9, 150e-9, n),
                           rng.uniform(ht + 5e-9, ht + dtip - 5e-9, n)])
    q = np.full(n, -1.602176634e-19)
    m = np.full(n, 9.1093837015e-31)
    dd = Diode(n + 16)
    dd.set_tip(V, dtip, rb, ht, bool(ic))
    dd.put_particles(pos, q, m)
    g = oracle_mod.geom_tip(V, dtip, rb, ht, bool(ic))
    ref = oracle_mod.accel(g, pos, q, m, long_double=True)
    am = dd.calculate_acceleration_particles("mixed")
    e = rel_err(am, ref)
    print(f"mixed tip accel n={n} ic={ic}: max {e.max():.2e}")
    assert e.max() < MIXED_TOL
    if n > 1:
        # the particle part alone (the vacuum field dominates the total):
        # relative to its own size, floored at 1% of its mean
        vac = np.array([vacuum_accel(oracle_mod, g, p, qq, mm) for p, qq, mm in zip(pos, q, m)])
        pr, pm = ref - vac, am - vac
        floor = 0.01 * np.linalg.norm(pr, axis=1).mean()
        ep = rel_err(pm, pr, floor=floor)
        print(f"mixed tip particle part n={n} ic={ic}: max {ep.max():.2e}")
        assert ep.max() < MIXED_TOL
    pts = np.column_stack([rng.uniform(-100e-9, 100e-9, 64), rng.uniform(-100e-9, 100e-9, 64),
                           rng.uniform(ht + 1e-9, ht + dtip - 1e-9, 64)])
    fref = oracle_mod.field(g, pos, q, pts, long_double=True)
    fm = dd.calc_field_at_batch(pts, "mixed")
    ef = rel_err(fm, fref)
    print(f"mixed tip field n={n} ic={ic}: max {ef.max():.2e}")
    assert ef.max() < MIXED_TOL


def test_absorbed_particles_have_no_effect(Diode, oracle_mod):
    """Particles with charge 0 (marked for removal, src/mod_pair.F90:206-207)
    contribute nothing but still receive their vacuum-field acceleration."""
    from rumdeed_amd import workloads as W

    pos, q, m = W.snapshot(600, seed=3)
    q = q.copy()
    q[::3] = 0.0
    dd = _diode_planar(Diode, pos, q, m, W.SNAPSHOT_D, W.SNAPSHOT_V, 1, 1)
    a = dd.calculate_acceleration_particles("fp64")
    g = oracle_mod.geom_planar(W.SNAPSHOT_D, W.SNAPSHOT_V, True, 1)
    assert np.array_equal(a, oracle_mod.accel(g, pos, q, m))
    assert np.all(a[::3] == 0.0)  # q_i = 0 -> no force at all
    keep = np.ones(600, bool)
    keep[::3] = False
    am = dd.calculate_acceleration_particles("mixed")
    ref = oracle_mod.accel(g, pos[keep], q[keep], m[keep], long_double=True)
    assert rel_err(am[keep], ref).max() < MIXED_TOL


@pytest.mark.parametrize("n,form", [(4096, "gather"), (16384, "symmetric"), (20000, "symmetric")])
def test_mixed_far_partner_expansion(Diode, oracle_mod, n, form):
    """With N_ic_max >= 2 the mixed accelerations take the image partners
    n >= 2 from one FP64 expansion of their field (the pair kernels keep
    n <= 1).  Against the form that sums every partner pair by pair: within
    1e-6 relative of the total (the far part is about 1e-3 of it); against
    the long-double oracle on sampled targets: within 1e-5; absorbed
    particles (charge 0) feel and exert nothing."""
    from rumdeed_amd import workloads as W

    pos, q, m = W.snapshot(n, seed=23)
    q = q.copy()
    q[11::101] = 0.0
    dd = _diode_planar(Diode, pos, q, m, W.SNAPSHOT_D, W.SNAPSHOT_V, 1, 10)
    if form == "gather":
        dd.set_force_options(no_symmetric=True)
    a_exp = dd.calculate_acceleration_particles("mixed")
    dd.set_force_options(no_symmetric=form == "gather", no_far_expansion=True)
    a_dir = dd.calculate_acceleration_particles("mixed")
    dd.close()
    keep = q != 0.0
    e = rel_err(a_exp[keep], a_dir[keep])
    print(f"far expansion n={n} {form}: vs direct max {e.max():.2e} p99 {np.percentile(e, 99):.2e}")
    assert e.max() < 1e-6
    assert np.all(a_exp[~keep] == 0.0)
    g = oracle_mod.geom_planar(W.SNAPSHOT_D, W.SNAPSHOT_V, True, 10)
    t = np.flatnonzero(keep)[::max(1, n // 512)]
    ref = oracle_mod.accel_subset(g, pos, q, m, t, long_double=True)
    el = rel_err(a_exp[t], ref)
    print(f"far expansion n={n}: vs long double max {el.max():.2e}")
    assert el.max() < MIXED_TOL


def test_far_partner_expansion_falls_back_when_it_would_not_converge(Diode):
    """Particles spread laterally over several gaps: the expansion would not
    converge, so the pair kernels keep every partner -- the accelerations
    equal the pair-by-pair form bit for bit."""
    from rumdeed_amd import workloads as W

    pos, q, m = W.snapshot(3000, seed=4)
    pos[:, :2] *= 20.0  # lateral +-2000 nm, gap 1000 nm
    dd = _diode_planar(Diode, pos, q, m, W.SNAPSHOT_D, W.SNAPSHOT_V, 1, 10)
    a = dd.calculate_acceleration_particles("mixed")
    dd.set_force_options(no_far_expansion=True)
    b = dd.calculate_acceleration_particles("mixed")
    dd.close()
    assert np.array_equal(a, b)
