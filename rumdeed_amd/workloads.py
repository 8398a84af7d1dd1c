"""Synthetic inputs of the force/field snapshot workload.

Snapshot: N electrons at positions uniform in a 200 x 200 x 1000 nm box above
a planar cathode (x, y in [-100, 100) nm, z in (0, 1000] nm), zero velocity,
charge -q_0, mass m_0; gap d = 1000 nm; the surface field is evaluated on the
256 x 256 cell centres of the 200 x 200 nm patch at z = 0.  Positions come
from numpy's Philox4x64-10 bit generator with the stated seed.
"""
from __future__ import annotations

import numpy as np

Q_0 = 1.602176634e-19
M_0 = 9.1093837015e-31
NM = 1.0e-9

SNAPSHOT_D = 1000.0 * NM
SNAPSHOT_V = 2000.0          # the vacuum term only adds q E_z / m
SNAPSHOT_BOX = (-100.0 * NM, -100.0 * NM, 200.0 * NM, 200.0 * NM)  # x0, y0, lx, ly
SNAPSHOT_GRID = 256


def snapshot(n: int, seed: int = 2):
    """Positions (n,3) [m], charges, masses of the force/field snapshot."""
    rng = np.random.Generator(np.random.Philox(seed))
    x = rng.uniform(-100.0, 100.0, n)
    y = rng.uniform(-100.0, 100.0, n)
    z = 1000.0 - rng.uniform(0.0, 1000.0, n)   # (0, 1000]
    pos = np.stack([x, y, z], axis=1) * NM
    q = np.full(n, -Q_0)
    m = np.full(n, M_0)
    return pos, q, m


def grid_points(nx=SNAPSHOT_GRID, ny=SNAPSHOT_GRID, box=SNAPSHOT_BOX, z=0.0):
    """Cell centres of the surface grid, x fastest (same order as rd_field_grid)."""
    x0, y0, lx, ly = box
    ix = np.arange(nx)
    iy = np.arange(ny)
    X = x0 + (ix + 0.5) * (lx / nx)
    Y = y0 + (iy + 0.5) * (ly / ny)
    gx, gy = np.meshgrid(X, Y)  # gy rows -> y slowest
    return np.stack([gx.ravel(), gy.ravel(), np.full(nx * ny, z)], axis=1)


SYM_MIN = 16384  # particles from which the mixed planar force is pair-symmetric
SYM_REACTION_FLOPS = 9  # reaction on the source: z terms 2, charge products 4, sums 3
FAR_RHO_MAX = 0.42  # convergence ratio up to which the far partners are an expansion


def far_expansion_used(pos, d: float, n_ic_max: int) -> bool:
    """Whether the mixed planar accelerations take the image partners n >= 2
    from the FP64 expansion (the library's rule): N_ic_max >= 2 and
    2 R / (3 d) <= FAR_RHO_MAX, R the half-diagonal of the particles' lateral
    bounding box extended by d / 2 vertically."""
    if n_ic_max < 2 or len(pos) == 0:
        return False
    hx = 0.5 * (pos[:, 0].max() - pos[:, 0].min())
    hy = 0.5 * (pos[:, 1].max() - pos[:, 1].min())
    R = (hx * hx + hy * hy + 0.25 * d * d) ** 0.5
    return 2.0 * R / (3.0 * d) <= FAR_RHO_MAX


def force_flops(n: int, n_ic_max: int, precision: str = "mixed", far_expansion=False) -> int:
    """Algorithmic flops of the pair kernel of one planar acceleration
    evaluation of n particles: the gather form evaluates n (n - 1) ordered
    pairs; the pair-symmetric form (mixed, n >= SYM_MIN) evaluates
    n (n - 1) / 2 pairs, each with the reaction on the source added.  With
    the far expansion the pair kernel takes the partners n <= 1 only (the
    far ones are O(n) work in FP64, outside this count)."""
    nic = min(n_ic_max, 1) if (precision == "mixed" and far_expansion) else n_ic_max
    if precision == "mixed" and n >= SYM_MIN:
        return n * (n - 1) // 2 * (pair_flops(nic) + SYM_REACTION_FLOPS)
    return n * (n - 1) * pair_flops(nic)


def pair_flops(n_ic_max: int) -> int:
    """Algorithmic flops of one planar pair interaction (FMA = 2, sqrt = 1,
    divide = 1), counted from the reference op sequence
    (src/mod_verlet.F90:1292-1330 and src/acc_ic_planar_series.inc):
    Coulomb 21, each image partner 15, image accumulation 6."""
    P = 1 + 4 * n_ic_max
    return 21 + 15 * P + 6
