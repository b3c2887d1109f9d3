"""Served cases: a seeded stream of independent heated-body cases and the
request loop that runs them on one GPU.

Each case is a reference case directory (the pre-processor's output that
``solver_exec`` reads, ``run_solver.sh:41``), written by
:func:`leswht_amd.case.write_case`. The stream draws, per case:
- cells n1m in {64, 96, ..., 256}, n2m in {64, 96, ..., 192}, n3m in {8, 16, 32};
- Re log-uniform in [40, 400], Pr in {0.71, 7.0};
- a circle (radius 0.5), a square (half side 0.5, angle uniform in [0, pi)) or
  an ellipse (semi-axes 0.5 and 0.5/aspect, aspect uniform in [1, 3], angle
  uniform in [0, pi)) at the origin of [-15, 30] x [-15, 15] x [0, pi];
- a share of large LES cases (384 x 256 x 64, Re 3900, dynamic SGS);
- a step budget, and an arrival time from a Poisson process.
The request loop itself is native (``lw_server_*`` in lw_serve.cpp): this
module only prepares directories and submits them at their arrival times.
"""
from __future__ import annotations

import dataclasses
import math
import os
import time
from typing import List, Optional

import numpy as np

from . import case as C
from .solver import CASE_DONE, Server


@dataclasses.dataclass
class ServedCase:
    name: str
    spec: C.CaseSpec
    nsteps: int
    arrival: float  # seconds after the stream starts
    large: bool = False

    @property
    def cells(self) -> int:
        return self.spec.n1m * self.spec.n2m * self.spec.n3m

    @property
    def body(self):
        return C.body_params(self.spec.body, self.spec.body_args)


def _nlev(n1m: int) -> int:
    """Multigrid levels: x-semicoarsening halves n1m per level; up to 4
    levels while n1m stays divisible."""
    lev = 0
    while lev < 4 and n1m % (2 ** (lev + 1)) == 0 and n1m // (2 ** (lev + 1)) >= 4:
        lev += 1
    return max(lev, 1)


def draw_cases(n: int, seed: int = 20261203, rate: float = 10.0, steps=(200, 2000),
               large_share: float = 0.1, large_steps=(200, 2000), noise: float = 1e-3) -> List[ServedCase]:
    """n seeded cases with Poisson arrivals at ``rate`` per second."""
    rng = np.random.default_rng(seed)
    out = []
    t = 0.0
    for q in range(n):
        t += float(rng.exponential(1.0 / rate))
        large = bool(rng.uniform() < large_share)
        if large:
            s = C.Settings(re=3900.0, pr=0.71, iles=1, idvmon=1, idtopt=1, cflfac=0.5, nlev=5, resid1=1e-6)
            spec = C.CaseSpec(384, 256, 64, noise=1e-2, seed=seed + q, settings=s)
            nsteps = int(rng.integers(large_steps[0], large_steps[1] + 1))
        else:
            n1m = int(rng.choice(np.arange(64, 257, 32)))
            n2m = int(rng.choice(np.arange(64, 193, 32)))
            n3m = int(rng.choice([8, 16, 32]))
            re = float(math.exp(rng.uniform(math.log(40.0), math.log(400.0))))
            pr = float(rng.choice([0.71, 7.0]))
            kind = int(rng.integers(0, 3))
            angle = float(rng.uniform(0.0, math.pi))
            if kind == 0:
                body, args = "cylinder", (0.5,)
            elif kind == 1:
                body, args = "square", (0.5, angle)
            else:
                aspect = float(rng.uniform(1.0, 3.0))
                body, args = "ellipse", (0.5, 0.5 / aspect, angle)
            s = C.Settings(re=re, pr=pr, iles=0, idtopt=1, cflfac=0.5, nlev=_nlev(n1m), resid1=1e-6)
            spec = C.CaseSpec(n1m, n2m, n3m, body=body, body_args=args, noise=noise, seed=seed + q, settings=s)
            nsteps = int(rng.integers(steps[0], steps[1] + 1))
        out.append(ServedCase("case%03d" % q, spec, nsteps, t, large))
    return out


def prepare(cases: List[ServedCase], root: str) -> List[str]:
    """Writes each case's directory under ``root``; returns the solver
    working directories."""
    dirs = []
    for c in cases:
        dirs.append(C.write_case(C.build_case(c.spec), os.path.join(root, c.name)))
    return dirs


def fld_path(solver_dir: str) -> str:
    return os.path.join(solver_dir, "..", "output", "field", "fld_final")


def run_stream(cases: List[ServedCase], dirs: List[str], budget_bytes: int, slots: int,
               honour_arrivals: bool = True, write_fields: bool = True, batch_max_cells: int = 0) -> dict:
    """Submits every case at its arrival time (or all at once), waits for all
    and returns per-case results and the server's counters. Time-to-solution
    is t_done - t_submit of each case. ``batch_max_cells`` > 0 batches the
    cases of at most that many cells (continuous batching)."""
    srv = Server(budget_bytes, slots, batch_max_cells)
    tickets = []
    t0 = time.perf_counter()
    try:
        for c, d in zip(cases, dirs):
            if honour_arrivals:
                wait = c.arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(wait)
            tickets.append(srv.submit(d, c.nsteps, fld_path(d) if write_fields else None, body=c.body))
        res = [srv.wait(t, raise_on_error=False) for t in tickets]
        info = srv.stats()
    finally:
        srv.close()
    wall = time.perf_counter() - t0
    # failed cases are counted and reported, and left out of the latency and
    # throughput figures (their partial steps are not a solution)
    done = [r for r in res if r["status"] == CASE_DONE]
    failed = [dict(ticket=r["ticket"], status=r["status"], error=r.get("error", "")) for r in res
              if r["status"] != CASE_DONE]
    if not done:
        return dict(results=res, info=info, wall=wall, p50=float("nan"), p99=float("nan"), cell_updates=0,
                    throughput=0.0, failed=failed)
    lat = np.array([r["t_done"] - r["t_submit"] for r in done])
    updates = sum(r["cells"] * r["steps_done"] for r in done)
    first = min(r["t_submit"] for r in done)
    last = max(r["t_done"] for r in done)
    return dict(results=res, info=info, wall=wall, p50=float(np.percentile(lat, 50)),
                p99=float(np.percentile(lat, 99)), cell_updates=updates,
                throughput=updates / (last - first), failed=failed)


def run_alone(case: ServedCase, solver_dir: str, budget_bytes: Optional[int] = None) -> dict:
    """One case through a one-slot server, nothing else on the device."""
    srv = Server(budget_bytes or (64 << 30), 1)
    try:
        t = srv.submit(solver_dir, case.nsteps, fld_path(solver_dir), body=case.body)
        return srv.wait(t)
    finally:
        srv.close()
