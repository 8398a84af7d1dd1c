"""Continuous-batching request loop for many independent diode simulations
on one GPU.

Requests (planar field-emission diodes of different voltage, gap, emitter
size and work function) arrive over time.  All admitted simulations live in
one device pool (rumdeed_amd.diode.Pool) and advance together: every loop
iteration is one main-loop step of every open simulation, executed with one
launch per kernel.  A request is admitted as soon as it has arrived, a
simulation slot is free and its HBM footprint (particle capacity from the
space-charge limit of its surface field) fits in the budget; it leaves the
pool when its last step is done, freeing its slot and memory for the next.
Each simulation draws its random numbers from streams keyed by its own seed
and step, so its result does not depend on which other simulations share
the device.
"""
from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from ._lib import RD_EMIT_PLANAR_FE
from .diode import Pool

NM = 1e-9
EPS0 = 1.0 / (1.25663706212e-6 * 299792458.0 ** 2)
Q0 = 1.602176634e-19


@dataclass
class Request:
    rid: int
    V: float           # applied voltage (V)
    gap: float         # d (m)
    side: float        # square emitter side (m)
    phi: float         # work function (eV)
    steps: int = 2000
    dt: float = 0.25e-15
    arrival: float = 0.0  # seconds after the stream starts
    seed: int = 0
    n_ic_max: int = 0
    capacity: int = 0     # particle capacity; 0 = estimate


@dataclass
class Result:
    rid: int
    arrival: float
    admitted: float
    finished: float
    steps: int
    emitted: int = 0
    absorbed: int = 0
    dropped: int = 0
    n_final: int = 0
    ramo_second_half: float = 0.0
    capacity: int = 0
    error: str = ""       # set when the request was refused (never admitted)
    records: list = field(default_factory=list)

    @property
    def latency(self):
        return self.finished - self.arrival

    @property
    def valid(self) -> bool:
        """Served to the end with every emitted electron kept: a refused
        request, or one whose slot overflowed (electrons dropped as
        Add_Particle drops them beyond MAX_PARTICLES), is not a valid result."""
        return not self.error and self.dropped == 0


def capacity_estimate(req: Request) -> int:
    """Upper bound of electrons in flight: three times the surface charge
    that cancels the vacuum field over the emitter (space-charge limit)."""
    if req.capacity > 0:
        return req.capacity
    F = abs(req.V) / req.gap
    n = 3.0 * EPS0 * F * req.side * req.side / Q0
    return int(math.ceil(max(n, 256.0) / 256.0) * 256)


def stream_requests(n: int = 256, rate: float = 4.0, seed: int = 2024, steps: int = 2000,
                    regime: str = "intended"):
    """The batched serving stream: Poisson arrivals at `rate` per second,
    V ~ U(1,10), gap ~ U(200,2000) nm, emitter side ~ U(20,200) nm,
    work function ~ U(4,5) eV, seed = request id.

    regime "as-written" applies V volts; "intended" applies a vacuum field of
    V GV/m (V * gap[nm] volts), the field-emission regime in which the
    stated electron counts (tens to ~20k) occur."""
    rng = np.random.Generator(np.random.Philox(seed))
    gaps = rng.uniform(200.0, 2000.0, n) * NM
    Vs = rng.uniform(1.0, 10.0, n)
    sides = rng.uniform(20.0, 200.0, n) * NM
    phis = rng.uniform(4.0, 5.0, n)
    arr = np.cumsum(rng.exponential(1.0 / rate, n))
    reqs = []
    for i in range(n):
        V = Vs[i] if regime == "as-written" else Vs[i] * gaps[i] / NM
        reqs.append(Request(rid=i, V=float(V), gap=float(gaps[i]), side=float(sides[i]),
                            phi=float(phis[i]), steps=steps, arrival=float(arr[i]), seed=i))
    return reqs


MH_JUMPS = 200  # Metropolis-Hastings jumps per chain of the planar FE sampler


class Server:
    def __init__(self, total_particles: int = 1 << 22, max_slots: int = 512,
                 hbm_budget: int = 64 << 30, precision: str = "mixed", device: int = 0,
                 graphs: bool = False):
        self.pool = Pool(total_particles, max_slots, device)
        self.pool.set_graphs(graphs)  # jump loop of each step as one HIP graph
        self.total_particles = int(total_particles)
        self.budget = int(hbm_budget)
        self.precision = precision
        self.max_slots = max_slots

    def close(self):
        self.pool.close()

    def _configure(self, slot: int, r: Request):
        p = self.pool
        p.select(slot)
        p.set_planar(r.V, r.gap, True, r.n_ic_max)
        p.set_time_step(r.dt)
        p.set_seed(r.seed)
        p.set_emitter(RD_EMIT_PLANAR_FE, 0, -r.side / 2, -r.side / 2, r.side, r.side,
                      wf=np.array([[r.phi]]))

    def run(self, requests, realtime: bool = True, keep_records: bool = False,
            progress=None, on_step=None):
        """Serve `requests`; arrival times are honoured against the wall clock
        when `realtime` (otherwise every request is available at once).
        progress(result, n_done) runs as each request completes, on_step(
        pool_steps, n_active) after every pool step.

        A request that can never be admitted -- its particle capacity is
        larger than the pool's store, its footprint larger than the HBM
        budget, or it does not fit even when no other simulation is open --
        is refused at once: it comes back as a Result with `error` set and
        is not waited for, so it never blocks the requests behind it."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        active = {}   # slot -> (request, result, footprint, steps done)
        results = []
        used = 0
        t0 = time.perf_counter()
        loop_steps = 0
        occupancy = []
        mh_pairs = 0
        force_pairs = 0
        refused = []
        while pending or active:
            now = time.perf_counter() - t0
            while pending and (not realtime or pending[0].arrival <= now):
                r = pending[0]
                cap = capacity_estimate(r)
                fp = Pool.footprint(cap)
                why = ""
                if cap > self.total_particles:
                    why = f"capacity {cap} exceeds the pool's {self.total_particles} particles"
                elif fp > self.budget:
                    why = f"footprint {fp} B exceeds the HBM budget of {self.budget} B"
                if why:
                    pending.popleft()
                    refused.append(self._refused(r, cap, now, realtime, why))
                    continue
                if used + fp > self.budget or len(active) >= self.max_slots:
                    break
                slot = self.pool.try_open(cap)
                if slot is None:
                    if not active:
                        # the store is empty and still has no contiguous range
                        pending.popleft()
                        refused.append(self._refused(
                            r, cap, now, realtime,
                            f"capacity {cap} does not fit the empty pool store"))
                        continue
                    break
                self._configure(slot, r)
                pending.popleft()
                used += fp
                # without real-time arrivals every request is available at t = 0
                res = Result(rid=r.rid, arrival=r.arrival if realtime else 0.0, admitted=now,
                             finished=0.0,
                             steps=r.steps, capacity=cap)
                active[slot] = [r, res, fp, 0, 0]  # ..., steps done, electrons in flight
            if not active:
                if pending:
                    time.sleep(max(0.0, min(0.05, pending[0].arrival - now)))
                continue
            recs = self.pool.step(self.precision)
            loop_steps += 1
            occupancy.append(len(active))
            if on_step:
                on_step(loop_steps, len(active))
            now = time.perf_counter() - t0
            for slot in list(active):
                r, res, fp, done, n_prev = active[slot]
                rec = recs[slot]
                done += 1
                active[slot][3] = done
                active[slot][4] = rec.n
                # work of the step: every chain's field on each of the 200
                # jumps sees the electrons of the start of the step; the
                # forces see those plus the new ones
                mh_pairs += rec.n_round * n_prev * MH_JUMPS
                n_mid = n_prev + rec.emitted
                force_pairs += n_mid * max(n_mid - 1, 0)
                res.emitted += rec.emitted
                res.absorbed += rec.absorbed
                res.dropped = rec.dropped
                res.n_final = rec.n
                if done > r.steps // 2:
                    res.ramo_second_half += rec.ramo_current
                if keep_records:
                    res.records.append(rec.as_dict())
                if done >= r.steps:
                    res.ramo_second_half /= (r.steps - r.steps // 2)
                    res.finished = now
                    results.append(res)
                    self.pool.close_slot(slot)
                    used -= fp
                    del active[slot]
                    if progress:
                        progress(res, len(results))
        makespan = time.perf_counter() - t0
        lat = np.array([r.latency for r in results]) if results else np.zeros(1)
        stats = {
            "requests": len(results),
            "refused": len(refused),
            "invalid": sum(1 for r in results if not r.valid),
            "makespan_s": makespan,
            "sims_per_hour": len(results) / makespan * 3600.0 if makespan > 0 else 0.0,
            "latency_p50_s": float(np.percentile(lat, 50)),
            "latency_p99_s": float(np.percentile(lat, 99)),
            "latency_max_s": float(lat.max()),
            "pool_steps": loop_steps,
            "mean_active": float(np.mean(occupancy)) if occupancy else 0.0,
            "max_active": int(max(occupancy)) if occupancy else 0,
            "mh_field_pair_evals": mh_pairs,
            "force_pair_evals": force_pairs,
            "pair_evals_per_s": (mh_pairs + force_pairs) / makespan if makespan > 0 else 0.0,
        }
        return sorted(results + refused, key=lambda r: r.rid), stats

    @staticmethod
    def _refused(r: Request, cap: int, now: float, realtime: bool, why: str) -> Result:
        arrival = r.arrival if realtime else 0.0
        return Result(rid=r.rid, arrival=arrival, admitted=now, finished=now, steps=0,
                      capacity=cap, error=why)
