"""rd_field_batch_static, the allocation-free variant of rd_field_batch:
through the C ABI, its fields equal rd_field_batch's bit for bit, in both
precisions and both geometries, on two calls in a row with the second one
smaller, and with another context using the same device's static work array
in between.  Sizes beyond the array are refused with RD_ERR_CAPACITY."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NM = 1e-9


def _planar(n, nic, seed):
    from rumdeed_amd import workloads as W
    from rumdeed_amd.diode import Diode

    pos, q, m = W.snapshot(n, seed=seed)
    dd = Diode(n + 16)
    dd.set_planar(W.SNAPSHOT_V, W.SNAPSHOT_D, True, nic)
    dd.put_particles(pos, q, m)
    return dd


def _tip(n, seed):
    from rumdeed_amd.diode import Diode

    rng = np.random.Generator(np.random.Philox(seed))
    ht, dtip = 500 * NM, 1000 * NM
    pos = np.column_stack([rng.uniform(-150, 150, n) * NM, rng.uniform(-150, 150, n) * NM,
                           rng.uniform(ht + 5 * NM, ht + dtip - 5 * NM, n)])
    dd = Diode(n + 16)
    dd.set_tip(1000.0, dtip, 250 * NM, ht, True)
    dd.put_particles(pos, np.full(n, -1.602176634e-19), np.full(n, 9.1093837015e-31))
    return dd


def _points(k, seed, z_top):
    from rumdeed_amd import workloads as W

    g = int(np.sqrt(k))
    on = W.grid_points(g, g)  # on the cathode: the paired-image form
    rng = np.random.default_rng(seed)
    above = np.column_stack([rng.uniform(-100, 100, k) * NM, rng.uniform(-100, 100, k) * NM,
                             rng.uniform(1.0, z_top, k) * NM])
    return np.concatenate([on, above])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("nic", [0, 10])
def test_static_field_equals_field_batch_planar(precision, nic):
    a = _planar(3000, nic, seed=5)
    b = _planar(700, 1, seed=6)
    for k in (1600, 300):  # the second call smaller than the first
        print(f"static field planar nic={nic} {precision} points={2 * k}", flush=True)
        pts = _points(k, k, 990.0)
        ref = a.calc_field_at_batch(pts, precision)
        got = a.calc_field_at_batch_static(pts, precision)
        assert np.array_equal(got, ref)
        # another context uses the same device's static array in between
        pb = _points(50, 9, 500.0)
        assert np.array_equal(b.calc_field_at_batch_static(pb, precision),
                              b.calc_field_at_batch(pb, precision))
    a.close()
    b.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_static_field_equals_field_batch_tip(precision):
    t = _tip(2000, seed=3)
    for k in (900, 100):
        print(f"static field tip {precision} points={k}", flush=True)
        rng = np.random.default_rng(k)
        pts = np.column_stack([rng.uniform(-100, 100, k) * NM, rng.uniform(-100, 100, k) * NM,
                               rng.uniform(505, 1495, k) * NM])
        assert np.array_equal(t.calc_field_at_batch_static(pts, precision),
                              t.calc_field_at_batch(pts, precision))
    t.close()


def test_static_field_refuses_sizes_beyond_the_array():
    from rumdeed_amd import _lib as L

    a = _planar(4096, 0, seed=5)
    pts = np.zeros((4_000_000, 3))
    with pytest.raises(L.RumdeedError) as err:
        a.calc_field_at_batch_static(pts, "fp64")
    assert "static work array" in str(err.value)
    # the entry point still works after the refusal
    p = _points(100, 1, 900.0)
    assert np.array_equal(a.calc_field_at_batch_static(p, "fp64"), a.calc_field_at_batch(p, "fp64"))
    a.close()
