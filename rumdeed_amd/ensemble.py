"""Ensembles of one diode over many seeds, on one device.

Every seed is a simulation of one pool (rumdeed_amd.diode.Pool), so all
seeds advance together, one launch per kernel per step.  Each simulation
draws its random numbers from its own seed, so its records are those it
gives alone.  Used for the ensemble-mean current (the statistical parity
criterion against the reference, whose RNG stream is its runtime's) and for
the reference's whole-system test scenarios (src/mod_tests.F90:2013-2219).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .diode import Pool

NM = 1e-9
RD_EMIT_TIP_FE = 3
RD_EMIT_PLANAR_GTF = 9
RD_EMIT_PLANAR_FE = 10


@dataclass
class DiodeSpec:
    """One diode: planar (gap d, voltage V, square emitter of side `side`
    centred on the axis, work-function table `wf`, rows top first) or the
    hyperboloid tip (d_tip, R_base, h_tip, work function phi)."""
    kind: str = "planar"
    mode: int = RD_EMIT_PLANAR_FE
    V: float = 0.0
    d: float = 0.0
    side: float = 0.0
    wf: list = field(default_factory=lambda: [[4.7]])
    T: float = 300.0
    d_tip: float = 0.0
    R_base: float = 0.0
    h_tip: float = 0.0
    phi: float = 4.7
    box_top: float = 0.0
    dt: float = 0.25e-15
    image_charge: bool = True
    nic: int = 0

    def configure(self, pool: Pool, seed: int):
        if self.kind == "planar":
            pool.set_planar(self.V, self.d, self.image_charge, self.nic, box_top=self.box_top)
        else:
            pool.set_tip(self.V, self.d_tip, self.R_base, self.h_tip, self.image_charge,
                         box_top=self.box_top)
        pool.set_time_step(self.dt)
        pool.set_seed(seed)
        if self.mode == RD_EMIT_TIP_FE:
            pool.set_emitter(RD_EMIT_TIP_FE, 0, w_tip=self.phi)
        else:
            s = self.side
            pool.set_emitter(self.mode, 0, -s / 2, -s / 2, s, s, T=self.T, wf=np.array(self.wf))


def run_ensemble(spec: DiodeSpec, seeds, steps: int, precision: str = "mixed",
                 capacity: int = 1 << 16, state_every: int = 0, progress=None):
    """Run `steps` main-loop steps of `spec` for every seed in one pool.
    Returns, per seed, the list of step records (dicts) and the particle
    states taken every `state_every` steps (0: the final state only)."""
    seeds = list(seeds)
    pool = Pool(capacity * len(seeds) + 1024, len(seeds) + 1)
    slots = []
    for sd in seeds:
        s = pool.open(capacity)
        spec.configure(pool, sd)
        slots.append(s)
    recs = [[] for _ in seeds]
    states = [[] for _ in seeds]
    for k in range(1, steps + 1):
        rr = pool.step(precision)
        for i, s in enumerate(slots):
            recs[i].append(rr[s].as_dict())
        take = (state_every and k % state_every == 0) or k == steps
        if take:
            for i, s in enumerate(slots):
                pool.select(s)
                states[i].append((k, pool.get_state()))
        if progress and (k % 200 == 0 or k == steps):
            progress(k)
    pool.close()
    return recs, states


def steady_current(recs, skip_frac: float = 0.5) -> float:
    """Mean Ramo current (A) over the last (1 - skip_frac) of a run's steps,
    the second half by default (ramo_current.dt, column 3)."""
    r = recs[int(len(recs) * skip_frac):]
    return float(np.mean([x["ramo_current"] for x in r]))
