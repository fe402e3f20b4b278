"""Serving loop: independent torus runs admitted by HBM budget and advanced
together with continuous batching.

The reference runs one torus per process, its time loop stepping until the
end time (src/fluid/main.cpp:169-187).  Here many such runs share one
device engine: a request arrives at a time of its own, is admitted as soon as
its blocks fit in the engine's block budget (the HBM the engine reserved),
gets its initial data on the device, and from then on advances in every
shared step together with whatever else is resident -- each run with its own
CFL dt.  A run that has taken its steps leaves at once and frees its slots
for the next arrivals.  Every per-run quantity (dt, con2prim, ghost fills,
reductions) is computed from that run's blocks only, so a run's result is
the same alone and in any mix.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

import numpy as np

from . import sph


@dataclass
class Request:
    rid: int
    arrival: float          # seconds after the start of the loop
    grid: dict
    physics: dict
    torus: dict
    seed: int
    nsteps: int = 100
    amr: bool = False
    # filled by the server
    blocks: Optional[np.ndarray] = None
    admitted: float = -1.0
    finished: float = -1.0
    steps: int = 0
    t: float = 0.0
    result: Optional[dict] = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def make_requests(n: int, rate: float, seed: int = 2026, nsteps: int = 100) -> List[Request]:
    """``n`` requests with Poisson arrivals of ``rate`` per second: grids
    from 64^3 to 192 x 96 x 96 zones (uniform, or two levels refined on the
    torus), spin a ~ U(0, 0.99), beta in {30, 100, 300}, PCG64 seeds of
    their own."""
    rng = np.random.default_rng(seed)
    sizes = [(64, 64, 64), (96, 64, 64), (128, 64, 64), (128, 96, 96), (192, 96, 96)]
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        nr, nth, nph = sizes[rng.integers(len(sizes))]
        amr = bool(rng.integers(2))
        a = float(rng.uniform(0.0, 0.99))
        beta = float(rng.choice([30.0, 100.0, 300.0]))
        if amr:
            grid = dict(nr=nr // 2 if (nr // 2) % 16 == 0 else nr, nth=nth // 2 if (nth // 2) % 16 == 0 else nth,
                        nph=nph // 2 if (nph // 2) % 16 == 0 else nph, nlevels=2, r_in=sph.R_IN, r_out=sph.R_OUT)
        else:
            grid = dict(nr=nr, nth=nth, nph=nph, nlevels=1, r_in=sph.R_IN, r_out=sph.R_OUT)
        phys = dict(sph.PHYSICS, a=a)
        tor = dict(sph.TORUS, beta=beta)
        out.append(Request(i, t, grid, phys, tor, seed=int(rng.integers(1 << 31)), nsteps=nsteps, amr=amr))
    return out


def request_blocks(req: Request) -> np.ndarray:
    if req.blocks is None:
        g = req.grid
        req.blocks = sph.amr_blocks(g, g["nlevels"], req.physics, req.torus) if g["nlevels"] > 1 else \
            sph.uniform_blocks(g, 16)
    return req.blocks


class Server:
    """Continuous batching over one ``sph.Engine`` of ``budget_blocks``
    block slots (bz = 16); ``mixed`` is the engine's precision (False FP64,
    True mixed, 2 mixed with the hardware FP32 reciprocal and square root)."""

    def __init__(self, budget_blocks: int, mixed=True, schedule: int = 0,
                 clock: Callable[[], float] = time.perf_counter, keep_state: bool = False):
        self.eng = sph.Engine(16, budget_blocks)
        self.eng.set_precision(mixed)
        self.eng.set_schedule(schedule)
        self.budget = budget_blocks
        self.used = 0
        self.clock = clock
        self.keep_state = keep_state
        self.log: List[tuple] = []   # (time, batch runs, batch blocks)

    def run(self, requests: List[Request], realtime: bool = True) -> Dict[str, float]:
        """Serve ``requests`` (sorted by arrival).  With ``realtime`` the
        arrivals follow the wall clock; otherwise every request is present
        from the start (a closed batch).  Returns the latency and throughput
        summary."""
        for r in requests:
            request_blocks(r)
            if len(r.blocks) > self.budget:
                raise ValueError(f"request {r.rid} needs {len(r.blocks)} blocks, budget {self.budget}")
        pending = sorted(requests, key=lambda r: r.arrival)
        active: List[tuple] = []   # (request, run)
        t0 = self.clock()
        zone_updates = 0
        now = 0.0
        while pending or active:
            now = self.clock() - t0
            # admission: arrivals in order while their blocks fit
            while pending and (not realtime or pending[0].arrival <= now) and \
                    self.used + len(pending[0].blocks) <= self.budget:
                req = pending.pop(0)
                run = self.eng.add_run(req.physics, req.grid, req.blocks)
                run.init_torus(req.torus, req.seed)
                self.used += len(req.blocks)
                req.admitted = self.clock() - t0
                active.append((req, run))
            if not active:
                if pending:
                    wait = pending[0].arrival - (self.clock() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.01))
                continue
            runs = [x[1] for x in active]
            dts = self.eng.step(runs, 1)[0]
            nb = 0
            for (req, run), dt in zip(active, dts):
                req.steps += 1
                req.t += dt
                nb += run.nb
            zone_updates += nb * 16 * 16 * 16
            self.log.append((self.clock() - t0, len(active), nb))
            done = [(req, run) for (req, run) in active if req.steps >= req.nsteps]
            if done:
                tnow = self.clock() - t0
                for req, run in done:
                    req.finished = tnow
                    d = run.diagnostics()
                    req.result = dict(t=req.t, steps=req.steps, **d)
                    if self.keep_state:
                        req.result["state"] = run.state()
                    self.eng.remove_run(run)
                    self.used -= len(req.blocks)
                active = [x for x in active if x[0].steps < x[0].nsteps]
        wall = self.clock() - t0
        lat = np.array([r.latency for r in requests])
        return dict(requests=len(requests), wall_s=wall, zone_updates=zone_updates,
                    zone_updates_per_s=zone_updates / wall if wall > 0 else float("nan"),
                    p50_latency_s=float(np.percentile(lat, 50)), p99_latency_s=float(np.percentile(lat, 99)),
                    max_batch_runs=max((x[1] for x in self.log), default=0),
                    mean_batch_blocks=float(np.mean([x[2] for x in self.log])) if self.log else 0.0)
