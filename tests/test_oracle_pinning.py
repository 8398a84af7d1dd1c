"""Pins for the oracle functions of the tip surface, the tip supply sum and
the Maxwell-Boltzmann emission velocity: each is checked against the
reference's own routine (the flang build in oracle/_ref) and against what
the geometry or the distribution fixes independently of both.  CPU only."""
import math

import numpy as np
import pytest

NM = 1e-9
TIP = (1000.0, 1000 * NM, 250 * NM, 500 * NM)  # V, d_tip, R_base, h_tip


@pytest.fixture(scope="module")
def ref(oracle_mod):
    if not oracle_mod.Ref.available():
        pytest.skip("reference harness not built (oracle/_ref)")
    r = oracle_mod.Ref.get()
    return r


@pytest.fixture(scope="module")
def tip(oracle_mod, ref):
    V, dtip, rb, ht = TIP
    ref.setup_tip(V, dtip, rb, ht)
    g = oracle_mod.geom_tip(V, dtip, rb, ht, True)
    return g, oracle_mod.tip_surface(g)


def _grid(g, n=23):
    xis = 1.0 + (g.max_xi - 1.0) * (np.arange(n) + 0.37) / n
    phis = 2 * np.pi * (np.arange(n) + 0.11) / n
    return [(x, p) for x in xis for p in phis]


def test_tip_surface_vs_reference(oracle_mod, ref, tip):
    """xyz_corr, the (xi, eta, phi) round trip, surface_normal, Tip_Area and
    Field_normal (src/mod_hyperboloid_tip.f90:25-163): oracle against the
    reference's own functions at 529 surface points."""
    g, (xyz, normal, area, eta, fnorm) = tip
    scale = g.a_foci
    for xi, phi in _grid(g):
        p_ref = ref.tip_xyz(xi, g.eta_1, phi)
        p = xyz(xi, g.eta_1, phi)
        assert np.abs(p - p_ref).max() <= 4e-15 * scale
        c = ref.tip_coords(p_ref)
        assert abs(c[0] - xi) < 1e-12 and abs(c[1] - g.eta_1) < 1e-12
        assert abs(eta(p_ref) - c[1]) < 1e-15
        n_ref = ref.tip_normal(p_ref)
        assert np.abs(normal(p_ref) - n_ref).max() < 4e-16
        E = ref.tip_vacuum(p_ref)
        assert abs(fnorm(p_ref, E) - ref.field_normal(p_ref, E)) <= 4e-15 * np.linalg.norm(E)
        d_xi, d_phi = (g.max_xi - 1.0) / 100, 2 * np.pi / 100
        a_ref = ref.tip_area(xi, xi + d_xi, phi, phi + d_phi)
        assert abs(area(xi, xi + d_xi, phi, phi + d_phi) - a_ref) <= 1e-13 * a_ref


def _surface_element(ref, g, xi, phi, h=1e-6):
    """|dr/dxi x dr/dphi| from central differences of the reference's own
    xyz_corr: the area element fixed by the embedding alone."""
    e = g.eta_1
    rx = (ref.tip_xyz(xi + h, e, phi) - ref.tip_xyz(xi - h, e, phi)) / (2 * h)
    rp = (ref.tip_xyz(xi, e, phi + h) - ref.tip_xyz(xi, e, phi - h)) / (2 * h)
    return np.linalg.norm(np.cross(rx, rp)), rx, rp


def test_tip_area_is_the_integral_of_the_surface_element(oracle_mod, ref, tip):
    """Tip_Area of a cell equals the integral of the area element of the
    surface r(xi, phi) = xyz_corr(xi, eta_1, phi) over the cell (8 x 8
    Gauss-Legendre on finite-difference tangents).  A wrong factor, bound or
    sign in the closed form fails this."""
    g, (xyz, normal, area, eta, fnorm) = tip
    xg, wg = np.polynomial.legendre.leggauss(8)
    d_xi = (g.max_xi - 1.0) / 7.0
    for i0 in range(7):
        xi1, xi2 = 1.0 + i0 * d_xi + 1e-5, 1.0 + (i0 + 1) * d_xi
        phi1, phi2 = 0.3, 1.1
        tot = 0.0
        for a, wa in zip(xg, wg):
            xi = 0.5 * (xi1 + xi2) + 0.5 * (xi2 - xi1) * a
            for b, wb in zip(xg, wg):
                phi = 0.5 * (phi1 + phi2) + 0.5 * (phi2 - phi1) * b
                tot += wa * wb * _surface_element(ref, g, xi, phi)[0]
        tot *= 0.25 * (xi2 - xi1) * (phi2 - phi1)
        assert abs(area(xi1, xi2, phi1, phi2) - tot) < 1e-6 * tot, (i0, tot)
    # the whole surface: the cells of the supply sum tile it exactly
    whole = area(1.0, g.max_xi, 0.0, 2 * np.pi)
    nr = 100
    lx, lp = (g.max_xi - 1.0) / nr, 2 * np.pi / nr
    cells = sum(area(1.0 + (i - 1) * lx, 1.0 + i * lx, (j - 1) * lp, j * lp)
                for i in range(1, nr + 1) for j in (1, 37, 100)) * nr / 3
    assert abs(cells - whole) < 1e-12 * whole


def test_tip_normal_is_the_unit_outward_normal(oracle_mod, ref, tip):
    """surface_normal is a unit vector orthogonal to both tangents of the
    surface and points into the vacuum gap (+z at the apex)."""
    g, (xyz, normal, area, eta, fnorm) = tip
    for xi, phi in _grid(g, 9):
        p = ref.tip_xyz(xi, g.eta_1, phi)
        n = normal(p)
        _, rx, rp = _surface_element(ref, g, xi, phi)
        assert abs(np.linalg.norm(n) - 1.0) < 1e-15
        assert abs(n @ rx) < 1e-7 * np.linalg.norm(rx)
        assert abs(n @ rp) < 1e-7 * np.linalg.norm(rp) + 1e-20
        assert n[2] > 0.0
    apex = xyz(1.0, g.eta_1, 0.0)
    assert np.allclose(normal(apex), [0.0, 0.0, 1.0], atol=1e-15)


def test_tip_supply_on_an_empty_diode(oracle_mod, ref, tip):
    """With no electrons the surface field is the vacuum field alone, so the
    tip supply (Do_Field_Emission_Tip_OLDCODE, src/mod_emission_tip.f90:
    417-487) is the sum over the 100 x 100 (xi, phi) cells of the
    reference's Tip_Area times the electron supply at the cell centre's
    normal field (reference field_E_Hyperboloid and Field_normal)."""
    g, _ = tip
    c = oracle_mod.consts()
    V, dtip, rb, ht = TIP
    dt, w = 0.25e-15, 4.7
    s = oracle_mod.Sim(64, nthreads=2)
    s.set_geometry(g, 1000 * NM)
    s.set_time_step(dt)
    s.set_seed(5)
    s.set_emitter(3, 0, w_tip=w)
    r = s.step(1, phases=1)
    nr = 100
    lx, lp = (g.max_xi - 1.0) / nr, 2 * np.pi / nr
    tot = 0.0
    F_sum = 0.0
    for i in range(1, nr + 1):
        xc = 1.0 + (i - 0.5) * lx
        for j in range(1, nr + 1):
            pc = (j - 0.5) * lp
            p = ref.tip_xyz(xc, g.eta_1, pc)
            F = ref.field_normal(p, ref.tip_vacuum(p))
            F_sum += F
            if F >= 0.0:
                continue
            A = ref.tip_area(1.0 + (i - 1) * lx, 1.0 + i * lx, (j - 1) * lp, j * lp)
            t = oracle_mod.fn_functions(1, F, w)[1]
            tot += A * c.a_FN * F * F * dt / (c.q_0 * w * t * t)
    assert tot > 1.0
    assert abs(r.n_sup - tot) < 1e-12 * tot
    assert r.n_round == math.floor(tot + 0.5)
    assert abs(r.f_avg[2] - F_sum / nr ** 2) < 1e-12 * abs(F_sum / nr ** 2)


def _mb_moments_ok(v, sd, tol_sigma=5.0):
    n = len(v)
    # x, y ~ N(0, sd^2); z = |N(0, sd^2)|
    for k in (0, 1):
        m, var = v[:, k].mean(), v[:, k].var()
        assert abs(m) < tol_sigma * sd / math.sqrt(n)
        assert abs(var / sd ** 2 - 1.0) < tol_sigma * math.sqrt(2.0 / n)
        kurt = np.mean(v[:, k] ** 4) / var ** 2
        assert abs(kurt - 3.0) < tol_sigma * math.sqrt(24.0 / n)
    mz = math.sqrt(2.0 / math.pi) * sd
    vz = (1.0 - 2.0 / math.pi) * sd ** 2
    assert np.all(v[:, 2] >= 0.0)
    assert abs(v[:, 2].mean() - mz) < tol_sigma * math.sqrt(vz / n)
    assert abs(v[:, 2].var() / vz - 1.0) < tol_sigma * 0.02 * math.sqrt(2e5 / n) + 0.01
    # the three components are uncorrelated
    cc = np.corrcoef(v.T)
    assert np.abs(cc - np.eye(3)).max() < tol_sigma / math.sqrt(n)


def test_mb_velocity_moments_and_reference(oracle_mod, ref):
    """Get_MB_Velocity (src/mod_velocity.f90:53-68): x and y normal with
    standard deviation sqrt(k_b T / m_0), z the absolute value of one.  The
    oracle's draws and the reference's own (intrinsic generator) both match
    those moments, and the two samples agree by a two-sample
    Kolmogorov-Smirnov test per component."""
    from scipy import stats

    c = oracle_mod.consts()
    T = 1500.0
    sd = math.sqrt(c.k_b * T / c.m_0)
    n = 200_000
    vo = np.concatenate([oracle_mod.mb_velocity(7, 0, step, T, n // 4) for step in range(1, 5)])
    vr = ref.mb_velocity(T, 12345, n)
    _mb_moments_ok(vo, sd)
    _mb_moments_ok(vr, sd)
    for k in range(3):
        assert stats.ks_2samp(vo[:, k], vr[:, k]).pvalue > 1e-4, k
