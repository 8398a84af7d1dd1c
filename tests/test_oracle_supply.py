"""The planar supply integral of the oracle (Divonne in the reference,
src/mod_field_emission_v2.F90 and src/mod_field_thermo_emission.F90 through
src/mod_cuba_integration.F90): the midpoint rule refined by doubling the
points per side until two levels agree within the reference's Divonne
tolerances (src/mod_global.F90:89-90).  Pinned against the fixed rule on
1024 x 1024 points (the limit the refinement approaches) and the closed form
of the empty diode."""
import numpy as np
import pytest

NM = 1e-9
Q0 = 1.602176634e-19
M0 = 9.1093837015e-31


def _sim(O, mode, quad, levels, V, wf, T=300.0, dt=0.25e-15, lx=200 * NM, pos=None):
    t = O.Sim(1000, nthreads=8)
    t.set_geometry(O.geom_planar(1000 * NM, V, True, 0))
    t.set_time_step(dt)
    t.set_seed(7)
    t.set_emitter(mode, 0, -lx / 2, -lx / 2, lx, lx, T=T, wf=wf, quad=quad, quad_levels=levels)
    if pos is not None and len(pos):
        t.put(pos, np.full(len(pos), -Q0), np.full(len(pos), M0))
    return t.step(1, phases=1)


def _tol(v):
    return max(0.5, 1e-3 * abs(v))


def test_quad_level_limits(oracle_mod):
    L = oracle_mod.lib()
    assert [L.rdo_quad_lmax(32, k) for k in (-1, 0, 1, 3, 8)] == [0, 4, 1, 3, 5]
    assert L.rdo_quad_lmax(512, 0) == 1 and L.rdo_quad_lmax(1024, 0) == 0
    assert L.rdo_quad_lmax(100, 8) == 3  # 800 <= 1024 < 1600


def test_uniform_field_stops_at_the_first_refinement(oracle_mod):
    """Without particles the integrand is constant: level 1 agrees with
    level 0, and both equal the closed form A * supply(F)."""
    r = _sim(oracle_mod, 10, 32, 0, 2000.0, np.array([[2.5]]))
    assert r.supply_level == 1
    c = oracle_mod.consts()
    F = -2000.0 / (1000 * NM)
    _, t, _, _ = oracle_mod.fn_functions(1, F, 2.5)
    exact = (200 * NM) ** 2 * 0.25e-15 / c.q_0 * c.a_FN / (t * t * 2.5) * F * F
    assert abs(r.n_sup - exact) < 1e-12 * exact


def test_checkerboard_cells_follow_the_sections(oracle_mod):
    """Thermal-field emission of the empty diode over a 7 x 7 checkerboard of
    work functions: the grid takes 7 * ceil(32 / 7) = 35 points per side, so
    every cell lies in one square and the rule is exact at every level: the
    closed form is the mean of the two current densities times the area."""
    wf = np.array([[4.0 if (i + j) % 2 == 0 else 4.5 for i in range(7)] for j in range(7)])
    c = oracle_mod.consts()
    J = [oracle_mod.gtf(-6e9, 1500.0, w) for w in (4.0, 4.5)]
    exact = (25 * J[0] + 24 * J[1]) / 49 * (200 * NM) ** 2 * 25e-15 / c.q_0
    for quad, levels in ((32, -1), (32, 0), (100, -1)):
        r = _sim(oracle_mod, 9, quad, levels, 6000.0, wf, T=1500.0, dt=25e-15)
        assert abs(r.n_sup - exact) < 1e-12 * exact, (quad, levels, r.n_sup, exact)
    assert oracle_mod.lib().rdo_quad_side(32, 7) == 35


def _space_charge(n=30, seed=1):
    rng = np.random.default_rng(seed)
    return np.stack([rng.uniform(-100, 100, n), rng.uniform(-100, 100, n),
                     rng.uniform(2.0, 20.0, n)], 1) * NM


GTF = dict(T=1500.0, dt=25e-15)


def test_refinement_converges_to_the_fine_rule(oracle_mod):
    """Thermal-field emission under 30 electrons 2-20 nm above the cathode,
    starting from a 4 x 4 rule: the fixed rule is off by several tolerances;
    the refinement doubles until two levels agree and lands within the
    tolerance of the 1024 x 1024 rule."""
    pos = _space_charge()
    wf = np.array([[4.0]])
    fixed = _sim(oracle_mod, 9, 4, -1, 6000.0, wf, pos=pos, **GTF)
    refined = _sim(oracle_mod, 9, 4, 0, 6000.0, wf, pos=pos, **GTF)
    fine = _sim(oracle_mod, 9, 1024, -1, 6000.0, wf, pos=pos, **GTF)
    assert fixed.supply_level == 0 and fine.supply_level == 0
    assert refined.supply_level >= 2
    tol = _tol(fine.n_sup)
    assert abs(fixed.n_sup - fine.n_sup) > 3 * tol
    assert abs(refined.n_sup - fine.n_sup) <= tol
    # the level it stopped at is the fixed rule on that many points
    G = 4 << refined.supply_level
    same = _sim(oracle_mod, 9, G, -1, 6000.0, wf, pos=pos, **GTF)
    assert same.n_sup == refined.n_sup
    # and the level before it is within the tolerance, the one before that not
    prev = _sim(oracle_mod, 9, G // 2, -1, 6000.0, wf, pos=pos, **GTF)
    assert abs(same.n_sup - prev.n_sup) <= _tol(same.n_sup)
    prev2 = _sim(oracle_mod, 9, G // 4, -1, 6000.0, wf, pos=pos, **GTF)
    assert abs(prev.n_sup - prev2.n_sup) > _tol(prev.n_sup)


def test_space_charge_refinement_within_tolerance(oracle_mod):
    """Planar field emission under 20 electrons 0.5-3 nm above the cathode:
    the refined rule from the default 32 x 32 stays within the tolerance of
    the fine rule."""
    rng = np.random.default_rng(1)
    n = 20
    pos = np.stack([rng.uniform(-90, 90, n), rng.uniform(-90, 90, n),
                    rng.uniform(0.5, 3.0, n)], 1) * NM
    wf = np.array([[2.5]])
    refined = _sim(oracle_mod, 10, 32, 0, 5000.0, wf, pos=pos)
    fine = _sim(oracle_mod, 10, 1024, -1, 5000.0, wf, pos=pos)
    assert refined.supply_level >= 1
    assert abs(refined.n_sup - fine.n_sup) <= _tol(fine.n_sup)


def test_level_limit(oracle_mod):
    """quad_levels = -1 is the midpoint rule on quad x quad points alone;
    quad_levels = 1 stops after one doubling, at the rule on 2 quad."""
    pos = _space_charge()
    wf = np.array([[4.0]])
    a = _sim(oracle_mod, 9, 4, -1, 6000.0, wf, pos=pos, **GTF)
    b = _sim(oracle_mod, 9, 8, -1, 6000.0, wf, pos=pos, **GTF)
    c = _sim(oracle_mod, 9, 4, 1, 6000.0, wf, pos=pos, **GTF)
    assert a.supply_level == 0 and c.supply_level == 1
    assert c.n_sup == b.n_sup and a.n_sup != b.n_sup


def test_supply_diagnostics(oracle_mod):
    """The record's integration.dt fields: cells of the level used, the
    difference of the last two levels, and the fail flag, which is set only
    when the last allowed level still misses the tolerance."""
    pos = _space_charge()
    wf = np.array([[4.0]])
    runs = {lv: _sim(oracle_mod, 9, 4, lv, 6000.0, wf, pos=pos, **GTF) for lv in (-1, 1, 0)}
    rules = {G: _sim(oracle_mod, 9, G, -1, 6000.0, wf, pos=pos, **GTF) for G in (4, 8)}
    fixed, one, refined = runs[-1], runs[1], runs[0]
    assert (fixed.supply_cells, fixed.supply_err, fixed.supply_fail) == (16, 0.0, 0)
    # one doubling allowed: the 8 x 8 rule against the 4 x 4 rule, which are
    # several tolerances apart (test_refinement_converges_to_the_fine_rule)
    d = abs(rules[8].n_sup - rules[4].n_sup)
    assert one.supply_cells == 64 and one.supply_err == d and d > _tol(one.n_sup)
    assert one.supply_fail == 1
    # refined until converged: the last difference within the tolerance
    G = 4 << refined.supply_level
    assert refined.supply_cells == G * G and refined.supply_fail == 0
    assert 0.0 < refined.supply_err <= _tol(refined.n_sup)
