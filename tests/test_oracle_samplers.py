"""Metropolis-Hastings emission-site samplers of the oracle, statistically.

The device runs the reference's lockstep (mh_batch) semantics: all chains of
a step see the particles of the start of the step.  The reference's default
is the serial sampler: each chain runs to completion in turn, sees the
electrons emitted by the chains before it, and updates the width once per
chain.  The oracle implements both (Sim.set_serial).  Here:

  - serial and lockstep samples agree (two-sample tests on independent seed
    sets) for planar FE under space charge, thermal-field (GTF) on a
    checkerboard and the hyperboloid tip, at the reference's chain lengths;
  - on a frozen field (empty diode) the emitted sites follow the analytic
    target: the section split of the GTF current J(F, T, phi) and the tip's
    distribution over xi of supply x escape probability x area element.

CPU only."""
import math

import numpy as np
import pytest
from scipy import stats

NM = 1e-9
P_MIN = 1e-3  # reject equality / goodness of fit below this p-value


def _sim(oracle_mod, mode, seed, serial, pre=None, ndim=0, w_tip=4.0, wf=None):
    s = oracle_mod.Sim(20000, nthreads=oracle_mod.default_threads())
    if mode == "fe":
        s.set_geometry(oracle_mod.geom_planar(2500 * NM, 20e3, True, 1))
        s.set_time_step(0.1e-15)
        s.set_emitter(10, 0, -100 * NM, -100 * NM, 200 * NM, 200 * NM,
                      wf=np.array([[4.3, 4.7]]), ndim=ndim)
    elif mode == "gtf":
        s.set_geometry(oracle_mod.geom_planar(1000 * NM, 6e3, True, 0))
        s.set_time_step(0.25e-15)
        s.set_emitter(9, 0, -500 * NM, -500 * NM, 1000 * NM, 1000 * NM, T=1500.0,
                      wf=np.array(wf if wf is not None else [[4.0, 4.5], [4.5, 4.0]]), ndim=ndim)
    else:
        s.set_geometry(oracle_mod.geom_tip(1000.0, 1000 * NM, 250 * NM, 500 * NM, True),
                       1000 * NM)
        s.set_time_step(0.25e-15)
        s.set_emitter(3, 0, w_tip=w_tip, ndim=ndim)
    s.set_seed(seed)
    if pre is not None:
        s.add(pre)
    s.set_serial(serial)
    r = s.step(1, phases=1)
    em, _ = s.step_logs()
    return r, em


def _space_charge(n=150, side=200.0, seed=1):
    rng = np.random.default_rng(seed)
    return np.column_stack([rng.uniform(-side / 2, side / 2, n), rng.uniform(-side / 2, side / 2, n),
                            rng.uniform(1.0, 20.0, n)]) * NM


def _both(oracle_mod, mode, nseeds, **kw):
    out = {}
    for serial in (False, True):
        runs = [_sim(oracle_mod, mode, sd + 1000 * serial, serial, **kw) for sd in range(nseeds)]
        out[serial] = ([r.emitted for r, _ in runs], np.concatenate([e for _, e in runs]),
                       runs[0][0].n_round)
    return out


def _contingency_p(a, b, k):
    t = np.array([np.bincount(a, minlength=k + 1)[1:], np.bincount(b, minlength=k + 1)[1:]])
    return stats.chi2_contingency(t)[1]


def _counts_z(ca, cb):
    ca, cb = np.asarray(ca, float), np.asarray(cb, float)
    se = math.sqrt(ca.var(ddof=1) / len(ca) + cb.var(ddof=1) / len(cb))
    return abs(ca.mean() - cb.mean()) / se


def test_fe_serial_and_lockstep_agree_under_space_charge(oracle_mod):
    """Planar FE, two work-function sections, 150 electrons within 20 nm of
    the cathode: emitted counts, section split and emission positions of the
    serial and lockstep samplers agree."""
    out = _both(oracle_mod, "fe", 30, pre=_space_charge())
    (cb, eb, nr), (cs, es, _) = out[False], out[True]
    print(f"FE chains/step {nr}: emitted lockstep {np.mean(cb):.2f} serial {np.mean(cs):.2f}")
    assert nr > 300 and len(eb) > 400 and len(es) > 400
    assert _counts_z(cb, cs) < 4.0
    assert _contingency_p(eb["section"], es["section"], 2) > P_MIN
    for k in ("x", "y"):
        assert stats.ks_2samp(eb[k], es[k]).pvalue > P_MIN, k


def test_gtf_serial_and_lockstep_agree(oracle_mod):
    """Thermal-field emission on a 2 x 2 checkerboard of 4.0 / 4.5 eV: the
    sections and positions of the emitted electrons agree between the
    serial and lockstep samplers."""
    out = _both(oracle_mod, "gtf", 30)
    (cb, eb, nr), (cs, es, _) = out[False], out[True]
    print(f"GTF chains/step {nr}: sites lockstep {len(eb)} serial {len(es)}")
    assert len(eb) > 5000 and len(es) > 5000
    assert _contingency_p(eb["section"], es["section"], 4) > P_MIN
    for k in ("x", "y"):
        assert stats.ks_2samp(eb[k], es[k]).pvalue > P_MIN, k


def _xi(g, p):
    a, sz = g.a_foci, g.shift_z
    rp = np.sqrt(p[:, 0] ** 2 + p[:, 1] ** 2 + (p[:, 2] + a - sz) ** 2)
    rm = np.sqrt(p[:, 0] ** 2 + p[:, 1] ** 2 + (p[:, 2] - a - sz) ** 2)
    return (rp + rm) / (2 * a)


def _sites(em):
    return np.column_stack([em["x"], em["y"], em["z"]])


def test_tip_serial_and_lockstep_agree(oracle_mod):
    """Hyperboloid tip, 4.0 eV: emitted counts and the distribution of the
    emission sites over xi and phi agree between the samplers."""
    g = oracle_mod.geom_tip(1000.0, 1000 * NM, 250 * NM, 500 * NM, True)
    out = _both(oracle_mod, "tip", 24)
    (cb, eb, nr), (cs, es, _) = out[False], out[True]
    print(f"tip chains/step {nr}: emitted lockstep {np.mean(cb):.1f} serial {np.mean(cs):.1f}")
    assert len(eb) > 800 and len(es) > 800
    assert _counts_z(cb, cs) < 4.0
    assert stats.ks_2samp(_xi(g, _sites(eb)), _xi(g, _sites(es))).pvalue > P_MIN
    phb = np.arctan2(eb["y"], eb["x"])
    phs = np.arctan2(es["y"], es["x"])
    assert stats.ks_2samp(phb, phs).pvalue > P_MIN


@pytest.mark.parametrize("wf,ndim", [([[4.4, 4.5], [4.5, 4.4]], 0),
                                     ([[4.0, 4.5], [4.5, 4.0]], 200)])
def test_gtf_sampler_follows_the_current_density(oracle_mod, wf, ndim):
    """Frozen field (empty diode, uniform 6 GV/m, 1500 K): the emitted
    electrons split over the checkerboard sections as J_GTF(F, T, phi) times
    the section area, for both samplers.  The reference's 25-jump chains
    (no warm-up) are close enough to equilibrium at low contrast (4.4 / 4.5
    eV); at high contrast (4.0 / 4.5 eV) chains of 200 jumps are."""
    F = -6e9
    J = {w: oracle_mod.gtf(F, 1500.0, w) for w in (4.0, 4.4, 4.5)}
    w_sec = [wf[1][0], wf[1][1], wf[0][0], wf[0][1]]  # sections 1..4: bottom row first
    p = np.array([J[w] for w in w_sec])
    p /= p.sum()
    for serial in (False, True):
        secs = np.concatenate([_sim(oracle_mod, "gtf", sd + 77 * serial, serial, wf=wf,
                                    ndim=ndim)[1]["section"] for sd in range(24)])
        n = np.bincount(secs, minlength=5)[1:]
        pv = stats.chisquare(n, n.sum() * p).pvalue
        print(f"GTF {wf[0]} ndim={ndim or 25} serial={serial}: {n / n.sum()} vs {p}, p={pv:.3g}")
        assert n.sum() > 2000 and pv > P_MIN


def test_tip_sampler_follows_supply_times_escape(oracle_mod):
    """Frozen field (empty diode; 56 electrons a step are far too few to
    change it): the tip's emission sites are distributed over xi as
    S(F) D(F) sqrt(xi^2 - eta_1^2), F the normal vacuum field at xi (the
    chain target times the escape probability; Tip_fe_target_log,
    src/mod_emission_tip.f90:1213-1219).  Chi-square over 8 bins of equal
    expected count, both samplers."""
    g = oracle_mod.geom_tip(1000.0, 1000 * NM, 250 * NM, 500 * NM, True)
    xyz, normal, area, eta, fnorm = oracle_mod.tip_surface(g)
    w = 4.0
    c = oracle_mod.consts()

    def dens(xi):
        p = xyz(xi, g.eta_1, 0.0)
        F = fnorm(p, oracle_mod.vacuum_field(g, p))
        if F >= 0.0:
            return 0.0
        v, t, lnD, _ = oracle_mod.fn_functions(1, F, w)
        sw = math.sqrt(w)
        D = math.exp(c.b_FN * sw ** 3 * v / abs(F))
        return F * F / (t * t) * D * math.sqrt(xi * xi - g.eta_1 ** 2)

    xs = np.linspace(1.0, g.max_xi, 4001)
    d = np.array([dens(x) for x in xs])
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (d[1:] + d[:-1]) * np.diff(xs))])
    cdf /= cdf[-1]
    edges = np.interp(np.linspace(0, 1, 9), cdf, xs)
    edges[0], edges[-1] = 1.0, g.max_xi
    for serial in (False, True):
        sites = np.concatenate([_sites(_sim(oracle_mod, "tip", sd + 500 * serial, serial,
                                            ndim=200)[1]) for sd in range(30)])
        # the site is 1 nm along the normal from the surface point; step back
        surf = sites - np.array([normal(p) for p in sites]) * NM
        xi = _xi(g, surf)
        n, _ = np.histogram(xi, edges)
        pv = stats.chisquare(n).pvalue
        print(f"tip serial={serial}: {len(xi)} sites, bins {n}, p={pv:.3g}")
        assert len(xi) > 1000 and pv > P_MIN
