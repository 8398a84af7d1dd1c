"""The CPU restatement (oracle) of the force and field path, pinned against
the reference's golden vectors, closed forms and invariants.  CPU only."""
import numpy as np
import pytest

from conftest import golden, rel_err

NM = 1e-9


def test_constants_match_reference_values(oracle_mod):
    c = oracle_mod.consts()
    # CODATA values as src/mod_global.F90:26-43 states them
    assert c.q_0 == 1.602176634e-19 and c.m_0 == 9.1093837015e-31
    assert c.epsilon_0 == 1.0 / (1.25663706212e-6 * 299792458.0 ** 2)
    assert c.div_fac_c == 1.0 / (4.0 * np.pi * c.epsilon_0)
    assert c.eps_soft == 1e-9 * 1e-9
    # FN constants: a_FN = q^2/(16 pi^2 hbar), b_FN = -(4/(3 hbar)) sqrt(2 m q)
    hbar = 6.62607015e-34 / (2 * np.pi)
    assert np.isclose(c.a_FN, c.q_0 ** 2 / (16 * np.pi ** 2 * hbar), rtol=1e-15)
    assert np.isclose(c.b_FN, -4.0 / (3.0 * hbar) * np.sqrt(2 * c.m_0 * c.q_0), rtol=1e-15)


def test_three_body_closed_form(oracle_mod):
    """Test_Acceleration_Without_Image_Charge (src/mod_tests.F90:405-519):
    accelerations from Coulomb's law written out per pair."""
    g3 = golden("three_body.npz")
    c = oracle_mod.consts()
    g = oracle_mod.geom_planar(float(g3["d"]), float(g3["V"]), False, 0)
    acc = oracle_mod.accel(g, g3["pos"], g3["q"], g3["m"])
    # independent closed form
    pos, q, m = g3["pos"], g3["q"], g3["m"]
    k = 1.0 / (4 * np.pi * c.epsilon_0)
    E = np.array([0, 0, -float(g3["V"]) / float(g3["d"])])
    exp = np.zeros((3, 3))
    for i in range(3):
        exp[i] = q[i] * E / m[i]
        for j in range(3):
            if i != j:
                r = pos[i] - pos[j]
                exp[i] += k * q[i] * q[j] / m[i] * r / np.linalg.norm(r) ** 3
    # the reference softens r by 1e-18 m: relative effect 3e-18/r_min ~ 1e-10
    assert rel_err(acc, exp).max() < 1e-9
    assert rel_err(acc, g3["acc"]).max() < 1e-14
    # field probe against the reference value and the reference test's Python value
    f = oracle_mod.field(g, pos, q, g3["probe"])
    assert rel_err(f, g3["field"]).max() < 1e-14
    assert rel_err(f, g3["field_python"]).max() < 0.02  # the reference test's tolerance


@pytest.mark.parametrize("ic,nic", [(0, 0), (1, 0), (1, 2), (1, 10)])
def test_forty_particles_vs_reference(oracle_mod, ic, nic):
    """40-particle generator of Test_Planar_Specialized_Acceleration
    (src/mod_tests.F90:1641-1651) against the reference's own gather and
    symmetric pair loops."""
    f = golden("forty.npz")
    g = oracle_mod.geom_planar(float(f["d"]), float(f["V"]), bool(ic), nic)
    acc = oracle_mod.accel(g, f["pos"], f["q"], f["m"])
    assert rel_err(acc, f[f"acc_gather_ic{ic}_n{nic}"]).max() < 1e-14
    assert rel_err(acc, f[f"acc_planar_ic{ic}_n{nic}"]).max() < 1e-13
    fld = oracle_mod.field(g, f["pos"], f["q"], f["pts"])
    assert rel_err(fld, f[f"field_ic{ic}_n{nic}"]).max() < 1e-14
    acc_ld = oracle_mod.accel(g, f["pos"], f["q"], f["m"], long_double=True)
    assert rel_err(acc, acc_ld).max() < 1e-13


@pytest.mark.parametrize("nic", [0, 1, 10])
def test_snapshot512_vs_reference(oracle_mod, nic):
    s = golden("snapshot512.npz")
    g = oracle_mod.geom_planar(float(s["d"]), float(s["V"]), True, nic)
    acc = oracle_mod.accel(g, s["pos"], s["q"], s["m"])
    assert rel_err(acc, s[f"acc_n{nic}"]).max() < 1e-13
    fld = oracle_mod.field(g, s["pos"], s["q"], s["grid"])
    assert rel_err(fld, s[f"field_n{nic}"]).max() < 1e-13


def test_tip_vs_reference(oracle_mod):
    t = golden("tip64.npz")
    g = oracle_mod.geom_tip(float(t["V"]), float(t["d_tip"]), float(t["R_base"]), float(t["h_tip"]))
    params = np.array([g.a_foci, g.shift_z, g.pre_fac_E_tip, g.pre_fac_E_tip_unit_voltage,
                       g.r_tip, g.eta_1, g.h_tip, g.theta_tip])
    assert np.array_equal(params, t["tip_params"])
    acc = oracle_mod.accel(g, t["pos"], t["q"], t["m"])
    assert rel_err(acc, t["acc_gather"]).max() < 1e-13
    assert rel_err(acc, t["acc_generic"]).max() < 1e-12
    fld = oracle_mod.field(g, t["pos"], t["q"], t["pts"])
    assert rel_err(fld, t["field"]).max() < 1e-13


def test_tip_geometry_closed_form(oracle_mod):
    """The apex (0, 0, h_tip) lies on the tip surface eta = eta_1, and the
    apex radius of curvature of the hyperboloid, a*sin(theta)*tan(theta),
    is 50 nm for d_tip = 1000, R_base = 250, h_tip = 500 nm."""
    g = oracle_mod.geom_tip(100.0, 1000 * NM, 250 * NM, 500 * NM)
    assert abs(g.r_tip - 50 * NM) < 1e-20
    z = 500 * NM
    r_p = np.sqrt((z + g.a_foci - g.shift_z) ** 2)
    r_m = np.sqrt((z - g.a_foci - g.shift_z) ** 2)
    eta = (r_p - r_m) / (2 * g.a_foci)
    assert abs(eta - g.eta_1) < 1e-12
    # the vacuum field at the apex points along z
    e = oracle_mod.vacuum_field(g, np.array([0.0, 0.0, z]))
    assert abs(e[0]) == 0 and abs(e[1]) == 0 and e[2] != 0


def test_image_charge_closed_forms(oracle_mod):
    """Test_Image_Charge (src/mod_tests.F90:523-695): the field of one charge
    and its image partners written out with explicit image positions and
    charges; n = 0 and N_ic_max = 1 (five partners)."""
    c = oracle_mod.consts()
    d = 100 * NM
    k = 1.0 / (4 * np.pi * c.epsilon_0)
    src = np.array([[6.0, -24.0, 98.53]]) * NM
    qs = np.array([c.q_0])
    p = np.array([[3.0, -10.0, 2.0]]) * NM
    for nic, images in [(0, [(-1, -src[0, 2])]),
                        (1, [(-1, -src[0, 2]), (-1, -2 * d - src[0, 2]), (1, -2 * d + src[0, 2]),
                             (-1, 2 * d - src[0, 2]), (1, 2 * d + src[0, 2])])]:
        g = oracle_mod.geom_planar(d, 0.0, True, nic)
        f = oracle_mod.field(g, src, qs, p)[0]
        exp = k * qs[0] * (p[0] - src[0]) / np.linalg.norm(p[0] - src[0]) ** 3
        for sgn, zi in images:
            r = p[0] - np.array([src[0, 0], src[0, 1], zi])
            exp = exp + k * sgn * qs[0] * r / np.linalg.norm(r) ** 3
        # closed form without the 1e-18 m softening of r: 3e-18/r ~ 3e-11
        assert np.linalg.norm(f - exp) / np.linalg.norm(exp) < 1e-9


def test_newton_third_law(oracle_mod):
    """Without image charges and vacuum field the pair forces cancel:
    sum_i m_i a_i = 0 (a sign or index slip in the pair term breaks this)."""
    rng = np.random.default_rng(0)
    n = 300
    pos = rng.uniform(0, 100, (n, 3)) * NM
    c = oracle_mod.consts()
    q = np.where(rng.random(n) < 0.8, -c.q_0, c.q_0)
    m = np.where(q < 0, c.m_0, c.m_N2p)
    g = oracle_mod.geom_planar(1000 * NM, 0.0, False, 0)
    acc = oracle_mod.accel(g, pos, q, m)
    F = (m[:, None] * acc)
    assert np.linalg.norm(F.sum(0)) < 1e-12 * np.abs(F).sum()


def test_lower_index_image_semantics(oracle_mod):
    """With images on, the force on the higher-index particle of a pair is the
    image term of the lower-index one mirrored in x and y, not the textbook
    image force (src/mod_verlet.F90:736-742, :1316-1330)."""
    c = oracle_mod.consts()
    d = 1000 * NM
    pos = np.array([[0.0, 0.0, 100.0], [30.0, 10.0, 400.0]]) * NM
    q = np.array([-c.q_0, -c.q_0])
    m = np.array([c.m_0, c.m_0])
    g = oracle_mod.geom_planar(d, 0.0, True, 1)
    a = oracle_mod.accel(g, pos, q, m)
    a_sw = oracle_mod.accel(g, pos[::-1].copy(), q, m)[::-1]
    # swapping the index order changes the result (order dependence is real)
    assert np.linalg.norm(a - a_sw) > 1e-6 * np.linalg.norm(a)


def test_long_double_agrees(oracle_mod):
    s = golden("snapshot512.npz")
    g = oracle_mod.geom_planar(float(s["d"]), float(s["V"]), True, 10)
    a = oracle_mod.accel(g, s["pos"], s["q"], s["m"])
    b = oracle_mod.accel(g, s["pos"], s["q"], s["m"], long_double=True)
    assert rel_err(a, b).max() < 1e-13
