"""One-workgroup jump loops of small planar simulations (k_mh_solo32)
against the two-launch jump path (k_mh_field32 + k_mh_settle), mixed
precision.

Both paths propose from the same random numbers, reflect and accept with the
same arithmetic and form the same FP32 pair terms; only the grouping of their
sums differs: FP64 additions in another order (about 1e-16 of the particle
sum), and on the cathode with N_ic_max = 0 other FP32 blocks of at most 16
terms (within 1e-6 of the particle sum, which is itself a small part of the
field while the diode holds few electrons).  An accept test decides the same
way unless its margin lies within that difference, so the chains take
identical sites here: the integer records and the particle state agree
exactly, and the record values that carry a chain field agree to 1e-7
(measured: 1.1e-9 with N_ic_max = 0).  Each case runs the same diode with the
solo path on and off (rd_set_solo_jumps); each case's chain count (<= 256)
and chains x electrons x the term's cost stay within the solo rule
(solo_jumps), as the printed counts show."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NM = 1e-9
REC_INT = ("n", "emitted", "n_round", "absorbed", "absorbed_top", "absorbed_bot", "dropped",
           "start_rounds", "supply_level", "supply_cells", "supply_fail")
REC_DBL = ("n_sup", "ramo_current", "mh_std", "a_rate", "df_avg", "supply_err")
STATE = ("pos", "prev_pos", "vel", "acc", "acc_prev", "acc_prev2", "species", "step", "emitter",
         "section", "id")


def _planar(dd, V, d, ic, nic, side, seed, dt=0.5e-15, seed_pos=None):
    dd.set_planar(V, d, ic, nic)
    dd.set_time_step(dt)
    dd.set_seed(seed)
    dd.set_emitter(10, 0, -side / 2, -side / 2, side, side, wf=np.array([[4.7]]))
    if seed_pos is not None:
        dd.add_particles(seed_pos)


def _run(setup, steps, solo, cap=1 << 14):
    from rumdeed_amd.diode import Diode

    dd = Diode(cap)
    dd.set_solo_jumps(solo)
    setup(dd)
    recs = dd.run(steps, "mixed")
    st = dd.get_state()
    dd.close()
    return recs, st


def _compare(setup, steps, name, cap=1 << 14):
    ra, sa = _run(setup, steps, True, cap)
    rb, sb = _run(setup, steps, False, cap)
    print(name, "chains", [r.n_round for r in ra], "electrons", [r.n for r in ra], flush=True)
    for k, (a, b) in enumerate(zip(ra, rb)):
        for f in REC_INT:
            assert getattr(a, f) == getattr(b, f), (name, k + 1, f)
        for f in REC_DBL:
            x, y = getattr(a, f), getattr(b, f)
            assert abs(x - y) <= 1e-7 * max(abs(x), abs(y), 1e-300), (name, k + 1, f, x, y)
        for f in ("f_avg", "ramo", "avg_elec_vel"):
            x, y = np.array(getattr(a, f)), np.array(getattr(b, f))
            assert np.all(np.abs(x - y) <= 1e-7 * np.maximum(np.abs(y), 1e-300)), (name, k + 1, f)
    for f in STATE:
        assert np.array_equal(sa[f], sb[f]), (name, f)
    return ra


def test_solo_planar_fe_image_series():
    """Planar FE, 1600 V over 200 nm, image series N_ic_max = 1 (chain points
    on the cathode: the surface form), 40 x 40 nm emitter: about 90 chains."""
    recs = _compare(lambda dd: _planar(dd, 1600.0, 200 * NM, True, 1, 40 * NM, 5), 24,
                    "fe nic=1")
    assert all(r.n_round <= 256 for r in recs) and sum(r.emitted for r in recs) > 20


def test_solo_planar_fe_single_image():
    """Planar FE, 8 kV over 1000 nm, N_ic_max = 0 (FP32 blocks of the lane's
    own terms), 50 x 50 nm emitter: about 70 chains."""
    recs = _compare(lambda dd: _planar(dd, 8000.0, 1000 * NM, True, 0, 50 * NM, 3,
                                       dt=0.25e-15), 24, "fe nic=0")
    assert all(r.n_round <= 256 for r in recs) and sum(r.emitted for r in recs) > 20


def test_solo_planar_fe_without_images():
    """Planar FE without image charges (the general pair term, FP64 sums):
    256 chains over 200 electrons placed in the gap."""
    rng = np.random.default_rng(3)
    pos = np.column_stack([rng.uniform(-100, 100, 200), rng.uniform(-100, 100, 200),
                           rng.uniform(20, 180, 200)]) * NM
    recs = _compare(lambda dd: _planar(dd, 1000.0, 200 * NM, False, 0, 100 * NM, 9, seed_pos=pos),
                    4, "fe no images")
    assert 64 < recs[0].n_round <= 256 and recs[0].n == 200


def test_solo_planar_staged_sources():
    """More electrons than the workgroup holds in LDS (3000 > 2048): the
    sources are staged per jump (about 22 chains on a 20 x 20 nm emitter)."""
    rng = np.random.default_rng(17)
    pos = np.column_stack([rng.uniform(-300, 300, 3000), rng.uniform(-300, 300, 3000),
                           rng.uniform(30, 180, 3000)]) * NM
    recs = _compare(lambda dd: _planar(dd, 1600.0, 200 * NM, True, 1, 20 * NM, 7, seed_pos=pos),
                    3, "fe staged")
    assert recs[0].n_round >= 1 and recs[0].n + recs[0].absorbed >= 3000


def test_solo_many_chains():
    """About 230 chains (the most the solo path takes is 256: four source
    rows per chain), 8 kV over 1000 nm on a 90 x 90 nm emitter."""
    recs = _compare(lambda dd: _planar(dd, 8000.0, 1000 * NM, True, 0, 90 * NM, 21,
                                       dt=0.25e-15), 8, "many chains")
    assert all(200 <= r.n_round <= 256 for r in recs)
    assert sum(r.emitted for r in recs) > 10
