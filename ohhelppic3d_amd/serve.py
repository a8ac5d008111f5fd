"""Continuous-batching serving loop for independent PIC runs.

Requests (a grid, particles per cell, a step count and the plasma parameters
of one run) arrive over time.  The loop keeps every admitted run's particles
and fields resident on the device and advances all of them together in
``Group`` launch sequences of a few steps, so runs that arrive while others
are on the GPU join at the next sequence and their work overlaps on the
device; new runs are created by worker threads while the group steps.  Each
launch sequence is a replayed hipGraph, captured again when the membership
changes; a capture waits for no creator (when one is inside a library call
then, that sequence is enqueued launch by launch).  Admission is FIFO
against a device-memory budget (the library's footprint estimate per run) and
a slot limit; a run that finishes frees its slot and memory at once.  A run
whose footprint exceeds the whole budget is rejected; a run that fails
(creation or a step) is dropped with its error and the others continue.

Requests are plain dicts / JSON lines (``Request.from_dict``)::

    {"rid": 7, "arrival": 1.25, "n": [64, 64, 64], "ppc": [40, 22], "steps": 900,
     "vth_e": 1.3, "omega_ce": 0.4, "seed": 1007, "dt": 0.1}

Physics per run as in the reference's two-species setup: electrons q/m -1,
ions q/m +1/100, macro charges -1/ppc_e and +1/ppc_i (omega_pe = 1),
v_th,i = v_th,e / 10, uniform B0z = omega_ce, seeded uniform Maxwellian.
"""
from __future__ import annotations

import json
import math
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor
from dataclasses import asdict, dataclass
from typing import Callable, Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib as L
from .ensemble import InstanceSpec
from .pic import Group, OhpicError, SimConfig, Simulation, estimate_footprint

GRIDS = ((32, 32, 32), (64, 64, 64), (96, 96, 96), (128, 128, 128))
GRID_WEIGHTS = (4, 3, 2, 1)
HBM_BUDGET = 240 * 10 ** 9


@dataclass
class Request:
    rid: int
    arrival: float               # seconds (wall clock) or steps (step clock)
    n: Tuple[int, int, int]
    ppc: Tuple[int, int]
    steps: int
    vth_e: float
    omega_ce: float
    seed: int
    dt: float = 0.1

    @staticmethod
    def from_dict(d: dict) -> "Request":
        try:
            r = Request(rid=int(d["rid"]), arrival=float(d.get("arrival", 0.0)),
                        n=tuple(int(v) for v in d["n"]), ppc=tuple(int(v) for v in d["ppc"]),
                        steps=int(d["steps"]), vth_e=float(d["vth_e"]),
                        omega_ce=float(d.get("omega_ce", 0.0)), seed=int(d.get("seed", 42)),
                        dt=float(d.get("dt", 0.1)))
        except (KeyError, TypeError, ValueError) as e:
            raise ValueError(f"bad request {d!r}: {e}") from None
        if len(r.n) != 3 or min(r.n) < 1:
            raise ValueError(f"request {r.rid}: grid must be 3 positive lengths")
        if len(r.ppc) != 2 or min(r.ppc) < 0:
            raise ValueError(f"request {r.rid}: ppc must be 2 non-negative counts")
        if r.steps < 0 or not r.dt > 0 or not r.vth_e >= 0 or not math.isfinite(r.vth_e) \
                or not math.isfinite(r.omega_ce) or not math.isfinite(r.arrival):
            raise ValueError(f"request {r.rid}: steps, dt and vth_e must be finite and "
                             "non-negative (dt positive)")
        return r

    def to_dict(self) -> dict:
        d = asdict(self)
        d["n"] = list(self.n)
        d["ppc"] = list(self.ppc)
        return d

    @property
    def spec(self) -> InstanceSpec:
        return InstanceSpec(n=self.n, ppc=self.ppc, vth_e=self.vth_e, omega_ce=self.omega_ce,
                            seed=self.seed, dt=self.dt)

    @property
    def particles(self) -> int:
        return self.spec.particles

    @property
    def particle_steps(self) -> int:
        return self.particles * self.steps


def read_requests(lines) -> List[Request]:
    """Requests from JSON lines (blank lines and '#' comments skipped)."""
    out = []
    for ln in lines:
        ln = ln.strip()
        if ln and not ln.startswith("#"):
            out.append(Request.from_dict(json.loads(ln)))
    return out


def sample_requests(count: int, rate: float, seed: int = 7, grids=GRIDS, weights=GRID_WEIGHTS,
                    ppc_range=(16, 64), steps_range=(200, 2000)) -> List[Request]:
    """Requests with Poisson arrivals of the given rate (per second; rate <= 0
    or inf puts every arrival at 0).  Grid drawn with the given weights;
    per request particles per cell uniform in ppc_range per species, step
    count uniform in steps_range, v_th,e uniform in [0.5, 2], omega_ce
    uniform in [0, 1]."""
    rng = np.random.default_rng(seed)
    p = np.asarray(weights, np.float64)
    p /= p.sum()
    t = 0.0
    out = []
    for i in range(count):
        if rate > 0 and math.isfinite(rate):
            t += float(rng.exponential(1.0 / rate))
        n = tuple(grids[int(rng.choice(len(grids), p=p))])
        ppc = tuple(int(v) for v in rng.integers(ppc_range[0], ppc_range[1] + 1, size=2))
        steps = int(rng.integers(steps_range[0], steps_range[1] + 1))
        out.append(Request(rid=i, arrival=t if rate > 0 and math.isfinite(rate) else 0.0,
                           n=n, ppc=ppc, steps=steps, vth_e=float(rng.uniform(0.5, 2.0)),
                           omega_ce=float(rng.uniform(0.0, 1.0)), seed=10_000 + i))
    return out


class Scheduler:
    """FIFO admission against a memory budget and a slot limit (host logic
    only: ``footprint(request)`` gives the bytes a run will hold)."""

    def __init__(self, budget_bytes: int, max_active: int, footprint: Callable[[Request], int]):
        self.budget = int(budget_bytes)
        self.max_active = int(max_active)
        self.footprint = footprint
        self.waiting: deque = deque()
        self.active: Dict[int, int] = {}
        self.used = 0
        self.peak_used = 0

    def arrive(self, req: Request) -> bool:
        """Queue a request; False (rejected) when it can never fit."""
        fp = int(self.footprint(req))
        if fp > self.budget:
            return False
        self.waiting.append((req, fp))
        return True

    def admit(self) -> List[Request]:
        out = []
        while self.waiting and len(self.active) < self.max_active:
            req, fp = self.waiting[0]
            if self.used + fp > self.budget:
                break  # the head waits for memory; no overtaking, so no starvation
            self.waiting.popleft()
            self.active[req.rid] = fp
            self.used += fp
            self.peak_used = max(self.peak_used, self.used)
            out.append(req)
        return out

    def complete(self, req: Request) -> None:
        self.used -= self.active.pop(req.rid)

    def idle(self) -> bool:
        return not self.waiting and not self.active


@dataclass
class Outcome:
    rid: int
    status: str                   # "done", "rejected" or "failed"
    arrival: float
    admitted: Optional[float] = None
    finished: Optional[float] = None
    particles: int = 0
    steps: int = 0
    error: str = ""

    @property
    def time_to_completion(self) -> Optional[float]:
        return None if self.finished is None else self.finished - self.arrival


@dataclass
class ServeReport:
    outcomes: List[Outcome]
    makespan: float               # first arrival to last completion
    group_steps: int
    mean_active: float
    peak_bytes: int
    clock: str
    graph_sequences: int = 0      # launch sequences replayed from a captured graph
    direct_sequences: int = 0     # launch sequences enqueued launch by launch

    def summary(self) -> dict:
        done = [o for o in self.outcomes if o.status == "done"]
        ttc = np.asarray([o.time_to_completion for o in done], np.float64)
        psteps = float(sum(o.particles * o.steps for o in done))
        return {
            "requests": len(self.outcomes), "done": len(done),
            "rejected": sum(o.status == "rejected" for o in self.outcomes),
            "failed": sum(o.status == "failed" for o in self.outcomes),
            "ttc_p50": float(np.percentile(ttc, 50)) if len(ttc) else None,
            "ttc_p99": float(np.percentile(ttc, 99)) if len(ttc) else None,
            "ttc_unit": "s" if self.clock == "wall" else "steps",
            "particle_steps": psteps,
            "particle_steps_per_s": psteps / self.makespan if self.makespan > 0 else None,
            "makespan": self.makespan, "group_steps": self.group_steps,
            "mean_active": self.mean_active, "peak_admitted_bytes": self.peak_bytes,
            "graph_sequences": self.graph_sequences, "direct_sequences": self.direct_sequences,
        }


class ServingLoop:
    """Admit arriving runs into one device-resident group and advance the
    group until every run has finished: each group call takes up to `batch`
    steps (one launch sequence, one host synchronisation), never past the end
    of an active run, and with the "steps" clock never past the next arrival;
    with the wall clock an arrival waits at most one such call.  With the wall
    clock, admitted runs are created (allocated, seeded, binned on their own
    streams) by `creators` worker threads while the group keeps stepping, and
    join the group at the first step after they are ready.  The report counts
    the launch sequences replayed from graphs and those enqueued directly."""

    def __init__(self, budget_bytes: int = HBM_BUDGET, max_active: int = 256,
                 tile=(8, 8, 8), chunks: int = 1, precision: int = L.MIXED,
                 use_graph: bool = True, track_ids: bool = False, device: int = 0,
                 batch: int = 8, creators: int = 2,
                 weight_order: int = L.WEIGHTS_STANDARD):
        self.budget = int(budget_bytes)
        self.max_active = int(max_active)
        self.tile = tuple(tile)
        self.chunks = int(chunks)
        self.precision = precision
        # one CIC weight order for every run: a group's members share it
        self.weight_order = int(weight_order)
        self.use_graph = use_graph
        self.track_ids = track_ids
        self.device = device
        self.batch = int(batch)
        self.creators = int(creators)

    def _create(self, req: Request):
        """(Simulation, None), or (None, error message) if creation failed."""
        sim = None
        try:
            sim = req.spec.create(self.tile, self.chunks, self.precision, self.track_ids,
                                  self.device, weight_order=self.weight_order)
            return sim, None
        except OhpicError as e:
            if sim is not None:
                sim.close()
            return None, str(e)

    def config(self, req: Request) -> SimConfig:
        return req.spec.config(self.tile, self.chunks, self.precision, self.track_ids,
                               self.device, weight_order=self.weight_order)

    def footprint(self, req: Request) -> int:
        return estimate_footprint(self.config(req))

    def run(self, requests: Sequence[Request], clock: str = "wall",
            on_complete: Optional[Callable[[Request, Simulation], None]] = None,
            progress: Optional[Callable[[str], None]] = None) -> ServeReport:
        """clock "wall": arrivals in seconds after the start; "steps":
        arrivals counted in group steps (a reproducible mix for tests)."""
        if clock not in ("wall", "steps"):
            raise ValueError("clock must be 'wall' or 'steps'")
        say = progress or (lambda s: None)
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        outcomes = {r.rid: Outcome(r.rid, "waiting", r.arrival, particles=r.particles,
                                   steps=r.steps) for r in pending}
        if len(outcomes) != len(pending):
            raise ValueError("request ids must be unique")
        sched = Scheduler(self.budget, self.max_active, self.footprint)
        active: Dict[int, list] = {}
        # wall clock: runs are created by worker threads while the group steps
        # (the library's calls release the interpreter lock); steps clock:
        # created in line, so the mix is reproducible
        pool = ThreadPoolExecutor(self.creators) if clock == "wall" and self.creators > 0 else None
        # the group captures a launch sequence only while no creator thread is
        # inside a library call (their HIP calls could invalidate the
        # capture), and otherwise enqueues that sequence directly
        grp = Group(self.device, self.use_graph, batch=self.batch)
        creating: Dict[int, tuple] = {}
        t0 = time.perf_counter()
        last_say = t0
        nsteps = 0
        active_sum = 0
        pi = 0

        def now() -> float:
            return time.perf_counter() - t0 if clock == "wall" else float(nsteps)

        def join(r: Request, made) -> None:
            # made: (Simulation, None) or (None, error message)
            sim, err = made
            o = outcomes[r.rid]
            if sim is not None:
                try:
                    grp.attach(sim)
                except OhpicError as e:
                    sim.close()
                    sim, err = None, str(e)
            if sim is None:
                sched.complete(r)
                o.status = "failed"
                o.error = err
                return
            active[r.rid] = [r, sim, r.steps]
            if r.steps == 0:
                finish(r.rid, "done")

        def finish(rid: int, status: str, err: str = "") -> None:
            req, sim, _ = active.pop(rid)
            o = outcomes[rid]
            o.status = status
            o.error = err
            o.finished = now() if status == "done" else None
            try:
                if grp.contains(sim):
                    grp.detach(sim)
                if status == "done" and on_complete is not None:
                    on_complete(req, sim)
            finally:
                sim.close()
                sched.complete(req)

        try:
            while pi < len(pending) or not sched.idle():
                t = now()
                while pi < len(pending) and pending[pi].arrival <= t:
                    r = pending[pi]
                    pi += 1
                    if not sched.arrive(r):
                        o = outcomes[r.rid]
                        o.status = "rejected"
                        o.error = (f"footprint {self.footprint(r)} B exceeds the budget "
                                   f"{self.budget} B")
                for r in sched.admit():
                    outcomes[r.rid].admitted = now()
                    if pool is not None:
                        creating[r.rid] = (r, pool.submit(self._create, r))
                    else:
                        join(r, self._create(r))
                for rid in [rid for rid, (_, fut) in creating.items() if fut.done()]:
                    r, fut = creating.pop(rid)
                    join(r, fut.result())
                if not active and creating:
                    # nothing to step: wait for the first run being created
                    rid = next(iter(creating))
                    r, fut = creating.pop(rid)
                    join(r, fut.result())
                    continue
                if not active:
                    if pi < len(pending):
                        if clock == "wall":
                            time.sleep(max(0.0, pending[pi].arrival - now()))
                        else:
                            nsteps = max(nsteps, int(math.ceil(pending[pi].arrival)))
                        continue
                    if sched.waiting:  # cannot happen: the head fits an empty budget
                        raise RuntimeError("serving loop stalled with waiting requests")
                    continue
                k = min(self.batch, min(entry[2] for entry in active.values()))
                if clock == "steps" and pi < len(pending):
                    k = min(k, max(1, int(math.ceil(pending[pi].arrival)) - nsteps))
                err = None
                before = grp.stats()["steps"]
                try:
                    grp.step(k)
                except OhpicError as e:
                    err = str(e)
                taken = grp.stats()["steps"] - before
                if err is not None and taken == 0 and all(grp.contains(e[1])
                                                          for e in active.values()):
                    # the group failed before any step and left every member
                    # attached: the same call would fail again, so the active
                    # runs fail with its error
                    say(f"[serve] group step failed: {err}")
                    for rid in list(active):
                        finish(rid, "failed", err)
                    continue
                nsteps += taken
                active_sum += len(active) * taken
                for rid in list(active):
                    entry = active[rid]
                    if err is not None and not grp.contains(entry[1]):
                        finish(rid, "failed", err)
                        continue
                    entry[2] -= taken
                    if entry[2] <= 0:
                        finish(rid, "done")
                if progress and (nsteps // 200 != (nsteps - taken) // 200 or
                                 time.perf_counter() - last_say >= 10.0):
                    last_say = time.perf_counter()
                    say(f"[serve] step {nsteps}: {len(active)} active, "
                        f"{len(sched.waiting)} waiting, {sched.used / 1e9:.1f} GB admitted")
        finally:
            for rid, (r, fut) in list(creating.items()):
                sim, _ = fut.result()
                if sim is not None:
                    sim.close()
                sched.complete(r)
                outcomes[rid].status = "failed"
                outcomes[rid].error = "serving loop stopped"
            creating.clear()
            if pool is not None:
                pool.shutdown(wait=True)
            for rid in list(active):
                finish(rid, "failed", "serving loop stopped")
            if grp.last_error():
                say(f"[serve] group's last error: {grp.last_error()}")
            gstats = grp.stats()
            grp.close()
        out = [outcomes[r.rid] for r in pending]
        first = min((o.arrival for o in out), default=0.0)
        last = max((o.finished for o in out if o.finished is not None), default=first)
        return ServeReport(outcomes=out, makespan=last - first, group_steps=nsteps,
                           mean_active=active_sum / max(nsteps, 1), peak_bytes=sched.peak_used,
                           clock=clock, graph_sequences=gstats["replayed"],
                           direct_sequences=gstats["direct"])
