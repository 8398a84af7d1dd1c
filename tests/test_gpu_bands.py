"""Banded pair launches (rd_set_partial_budget): a launch whose partial sums
exceed the budget runs in bands whose partials are added, in the order of
one unbanded launch, to running sums.  Every result is bit-identical for
every budget, and the memory the sums take is bounded by the budget."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NM = 1e-9


@pytest.fixture(scope="module")
def Diode():
    from rumdeed_amd.diode import Diode

    return Diode


def _cloud(n, seed=4):
    rng = np.random.default_rng(seed)
    pos = np.stack([rng.uniform(-300, 300, n), rng.uniform(-300, 300, n),
                    rng.uniform(0.5, 900, n)], 1) * NM
    return pos, np.full(n, -1.602176634e-19), np.full(n, 9.1093837015e-31)


@pytest.mark.parametrize("nic", [0, 10])
def test_planar_forces_identical_for_every_budget(Diode, nic):
    """FP64 (k_accel64 in up to 79 chunk bands) and the pair-symmetric mixed
    form (up to 20 row bands) at N = 20000 give the same bits under a 1 MiB,
    a 7 MiB and the default budget."""
    n = 20000
    pos, q, m = _cloud(n)
    dd = Diode(n + 16)
    dd.set_planar(3000.0, 1000 * NM, True, nic)
    dd.put_particles(pos, q, m)
    ref = {p: dd.calculate_acceleration_particles(p) for p in ("fp64", "mixed")}
    for budget in (1 << 20, 7 << 20):
        print(f"bands: nic={nic} budget={budget}", flush=True)
        dd.set_partial_budget(budget)
        for p in ("fp64", "mixed"):
            a = dd.calculate_acceleration_particles(p)
            assert np.array_equal(a, ref[p]), (p, budget)
    dd.set_partial_budget(0)
    assert np.array_equal(dd.calculate_acceleration_particles("mixed"), ref["mixed"])
    dd.close()


def test_tip_forces_and_fields_identical_for_every_budget(Diode):
    n = 6000
    rng = np.random.default_rng(5)
    pos = np.stack([rng.uniform(-200, 200, n), rng.uniform(-200, 200, n),
                    rng.uniform(520, 980, n)], 1) * NM
    q, m = np.full(n, -1.602176634e-19), np.full(n, 9.1093837015e-31)
    dd = Diode(n + 16)
    dd.set_tip(1000.0, 1000 * NM, 250 * NM, 500 * NM, True, box_top=1000 * NM)
    dd.put_particles(pos, q, m)
    a = dd.calculate_acceleration_particles("fp64")
    f = dd.field_grid(-100 * NM, -100 * NM, 200 * NM, 200 * NM, 64, 64, 700 * NM, "fp64")
    dd.set_partial_budget(1 << 18)
    assert np.array_equal(dd.calculate_acceleration_particles("fp64"), a)
    assert np.array_equal(
        dd.field_grid(-100 * NM, -100 * NM, 200 * NM, 200 * NM, 64, 64, 700 * NM, "fp64"), f)
    dd.close()


def test_cathode_grid_identical_for_every_budget(Diode):
    """The FP64 field on the 256 x 256 cathode grid (k_field64 in chunk
    bands) and the refined supply of an emission step keep their bits."""
    n = 12000
    pos, q, m = _cloud(n, 6)
    dd = Diode(n + 16)
    dd.set_planar(3000.0, 1000 * NM, True, 10)
    dd.put_particles(pos, q, m)
    args = (-100 * NM, -100 * NM, 200 * NM, 200 * NM, 256, 256, 0.0, "fp64")
    f = dd.field_grid(*args)
    dd.set_partial_budget(3 << 20)
    assert np.array_equal(dd.field_grid(*args), f)
    dd.close()


def test_pool_identical_for_every_budget():
    """Two simulations of a pool, one in the symmetric form, one in the
    gather form, under a budget that bands the first."""
    from rumdeed_amd.diode import Pool

    out = []
    for budget in (0, 1 << 20):
        pool = Pool(40000, 4)
        pool.set_partial_budget(budget)
        res = []
        for n, seed in ((20000, 1), (5000, 2)):
            pool.open(n + 16)
            pos, q, m = _cloud(n, seed)
            pool.set_planar(3000.0, 1000 * NM, True, 1)
            pool.put_particles(pos, q, m)
        for s in (0, 1):
            pool.select(s)
            res.append({p: pool.calculate_acceleration_particles(p) for p in ("fp64", "mixed")})
        out.append(res)
        pool.close()
    for s in (0, 1):
        for p in ("fp64", "mixed"):
            assert np.array_equal(out[0][s][p], out[1][s][p]), (s, p)


def test_partial_memory_is_bounded(Diode):
    """N = 200000: unbanded FP64 partials would take 3.75 GB (3 x 782 chunks
    x N doubles) and the symmetric ones 0.94 GB; under a 64 MiB budget the
    context's work buffers stay below 256 MiB (rd_work_bytes), and the forces
    equal the default budget's."""
    n = 200000
    pos, q, m = _cloud(n, 7)
    dd = Diode(n + 16)
    dd.set_planar(3000.0, 1000 * NM, True, 0)
    dd.put_particles(pos, q, m)
    dd.set_partial_budget(64 << 20)
    print(f"N={n}: FP64 force under a 64 MiB budget", flush=True)
    a64 = dd.calculate_acceleration_particles("fp64")
    print(f"N={n}: mixed force under a 64 MiB budget", flush=True)
    am = dd.calculate_acceleration_particles("mixed")
    held = dd.work_bytes()
    print(f"N={n}: work buffers {held / 2**20:.1f} MiB", flush=True)
    assert held < 256 << 20, held
    dd.set_partial_budget(0)
    assert np.array_equal(dd.calculate_acceleration_particles("mixed"), am)
    assert np.array_equal(dd.calculate_acceleration_particles("fp64"), a64)
    print(f"N={n}: default budget, work buffers {dd.work_bytes() / 2**20:.1f} MiB", flush=True)
    dd.close()
