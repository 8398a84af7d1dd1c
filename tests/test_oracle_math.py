"""The oracle's own elementary functions and generator against numpy/libm."""
import numpy as np


def _ulp_err(a, b):
    return abs(a - b) / np.spacing(abs(b)) if b != 0 else abs(a)


def test_log_within_one_ulp(oracle_mod):
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(1e-6, 4.0, 3000), np.exp(rng.uniform(-700, 700, 3000)),
                         [5e-324, 1e-310, 0.5, 1.0, 2.0, np.sqrt(2.0), 1 / np.sqrt(2.0)]])
    worst = max(_ulp_err(oracle_mod.log(x), np.log(x)) for x in xs)
    assert worst <= 1.0
    assert oracle_mod.log(1.0) == 0.0
    assert np.isnan(oracle_mod.log(-1.0)) and oracle_mod.log(0.0) == -np.inf


def test_exp_within_one_ulp(oracle_mod):
    rng = np.random.default_rng(2)
    xs = np.concatenate([rng.uniform(-5, 5, 3000), rng.uniform(-700, 700, 3000),
                         [0.0, 1.0, -1.0, 709.7, -700.0]])
    worst = max(_ulp_err(oracle_mod.exp(x), np.exp(x)) for x in xs)
    assert worst <= 1.0
    assert oracle_mod.exp(0.0) == 1.0
    assert oracle_mod.exp(800.0) == np.inf and oracle_mod.exp(-800.0) == 0.0


def test_philox4x64_known_answers(oracle_mod):
    # Random123 known-answer vector for philox4x64-10 (zero counter, zero key)
    assert oracle_mod.philox4x64([0, 0, 0, 0], [0, 0]) == [
        0x16554d9eca36314c, 0xdb20fe9d672d0fdc, 0xd7e772cee186176b, 0x7e68b68aec7ba23b]
    # numpy's Philox (4x64-10) increments the counter before each block
    key = [0x123456789abcdef0, 0x0fedcba987654321]
    bg = np.random.Philox(key=np.array(key, dtype=np.uint64),
                          counter=np.array([41, 7, 3, 9], dtype=np.uint64))
    assert [int(v) for v in bg.random_raw(4)] == oracle_mod.philox4x64([42, 7, 3, 9], key)
