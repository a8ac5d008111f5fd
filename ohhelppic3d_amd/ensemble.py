"""Batched ensemble of independent periodic PIC instances (one group).

Each instance is the reference's two-species plasma (electrons q/m -1, ions
q/m +1/100, macro charges -1/ppc and +1/ppc so omega_pe = 1 and the box is
neutral, uniform B0 along z with omega_ce = B0, dt 0.1) on its own non-cubic
periodic grid.  All instances are advanced together by one ``Group``: per
step one launch of each field kernel over a table of per-instance FFT jobs
and one launch of the particle kernel over the work items of every instance,
replayed from a captured hipGraph.  Tile shape and chunks per tile come from
``autotune``.

The mixed-precision form uses the standard CIC weight order by default
(``InstanceSpec.config(weight_order=...)``): with the reference's mirrored
weights (linear_interpolator.f90:54-57) a particle's deposit jumps by two
nodes when it crosses a cell face, so any 32-bit position that lands on the
other side of a face than the FP64 one moves a whole macro charge, and rho
after a push differs from FP64 by ~5e-4 relative L2.  With standard weights
the deposit is continuous and the mixed form stays within 1e-7 of FP64.
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib as L
from .pic import Group, SimConfig, Simulation, autotune

AXES = (16, 24, 32, 48, 64)


@dataclass
class InstanceSpec:
    n: Tuple[int, int, int]
    ppc: Tuple[int, int]        # particles per cell, electrons and ions
    vth_e: float                # ion thermal speed is vth_e / 10
    omega_ce: float             # electron cyclotron frequency = B0z
    seed: int                   # PCG32 initstate of the seeded loader
    dt: float = 0.1

    @property
    def cells(self) -> int:
        return self.n[0] * self.n[1] * self.n[2]

    @property
    def counts(self) -> Tuple[int, int]:
        return (self.ppc[0] * self.cells, self.ppc[1] * self.cells)

    @property
    def particles(self) -> int:
        return sum(self.counts)

    def config(self, tile=(0, 0, 0), chunks: int = 0, precision: int = L.MIXED,
               track_ids: bool = False, device: int = 0, slack: float = 0.3,
               weight_order: int = L.WEIGHTS_STANDARD) -> SimConfig:
        # bins 30% above their count: small instances have few particles per
        # tile and larger relative fluctuations, and a bin that overflows
        # costs that step a list-mode retry
        return SimConfig(n=self.n, qm=(-1.0, 0.01),
                         q_macro=(-1.0 / self.ppc[0], 1.0 / self.ppc[1]),
                         capacity=self.counts, dt=self.dt, b0=(0.0, 0.0, self.omega_ce),
                         precision=precision, weight_order=weight_order, tile=tuple(tile),
                         chunks=chunks, track_ids=track_ids, device=device, slack=slack)

    def create(self, tile=(0, 0, 0), chunks: int = 0, precision: int = L.MIXED,
               track_ids: bool = False, device: int = 0,
               weight_order: int = L.WEIGHTS_STANDARD) -> Simulation:
        sim = Simulation(self.config(tile, chunks, precision, track_ids, device,
                                     weight_order=weight_order))
        try:
            sim.seed_uniform_maxwell(list(self.counts), [self.vth_e, self.vth_e / 10.0],
                                     seed=(self.seed, 52))
        except Exception:
            sim.close()
            raise
        return sim


def sample_ensemble(count: int = 128, seed: int = 2024, axes: Sequence[int] = AXES,
                    ppc_range=(8, 64)) -> List[InstanceSpec]:
    """Instances drawn from one seed: each axis length independently from
    ``axes``, particles per cell uniform in ppc_range per species, v_th,e
    uniform in [0.5, 2], omega_ce uniform in [0, 1]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = tuple(int(v) for v in rng.choice(np.asarray(axes), 3))
        ppc = tuple(int(v) for v in rng.integers(ppc_range[0], ppc_range[1] + 1, size=2))
        vth_e = float(rng.uniform(0.5, 2.0))
        omega_ce = float(rng.uniform(0.0, 1.0))
        out.append(InstanceSpec(n=n, ppc=ppc, vth_e=vth_e, omega_ce=omega_ce, seed=1000 + i))
    return out


def run_ensemble(specs: Sequence[InstanceSpec], steps: int, warmup: int = 2,
                 tune: Optional[dict] = None, tune_sample: int = 16, use_graph: bool = True,
                 precision: int = L.MIXED, batch: int = 8,
                 progress: Optional[Callable[[str], None]] = None,
                 weight_order: int = L.WEIGHTS_STANDARD) -> dict:
    """Advance every instance ``steps`` steps as one group, in launch
    sequences of ``batch`` steps, and time each sequence (host wall clock
    around the synchronous call); a step's time is its sequence's over its
    length."""
    say = progress or (lambda s: None)
    t0 = time.perf_counter()
    if tune is None:
        cfgs = [s.config(precision=precision, weight_order=weight_order) for s in specs]
        tune = autotune(cfgs, sample=tune_sample, steps=3)
        say(f"[ensemble] autotune: tile {tune['tile']}, chunks {tune['chunks']} "
            f"({tune['candidates']} candidates, {time.perf_counter() - t0:.1f} s)")
    t1 = time.perf_counter()
    sims = []
    grp = Group(use_graph=use_graph, batch=batch)
    try:
        for s in specs:
            sims.append(s.create(tune["tile"], tune["chunks"], precision,
                                 weight_order=weight_order))
            grp.attach(sims[-1])
        say(f"[ensemble] {len(sims)} instances created and seeded in "
            f"{time.perf_counter() - t1:.1f} s")
        if warmup > 0:
            grp.step(warmup)
        per_step = []
        done = 0
        while done < steps:
            k = min(batch, steps - done)
            ts = time.perf_counter()
            grp.step(k)
            dt = time.perf_counter() - ts
            per_step.extend([dt / k] * k)
            if (done + k) // max(1, steps // 5) != done // max(1, steps // 5):
                say(f"[ensemble] step {done + k}/{steps}: {dt / k * 1e3:.2f} ms")
            done += k
        stats = grp.stats()
        energy = [sim.diagnostics()["total_energy"] for sim in sims]
    finally:
        grp.close()
        for sim in sims:
            sim.close()
    total = float(np.sum(per_step))
    npart = sum(s.particles for s in specs)
    a = np.asarray(per_step) * 1e3
    return {
        "instances": len(specs), "particles": npart, "steps": steps, "warmup": warmup,
        "value": npart * steps / total, "unit": "particle-steps/s",
        "step_ms_p50": float(np.percentile(a, 50)), "step_ms_p99": float(np.percentile(a, 99)),
        "step_ms_max": float(a.max()), "completion_s": total,
        "tile": list(tune["tile"]), "chunks": tune["chunks"],
        "autotune_particle_steps_per_s": tune.get("particle_steps_per_s"),
        "weight_order": weight_order, "graph": use_graph, "batch": batch,
        "graph_captures": stats["captures"],
        "work_items": stats["items"],
        "energy_finite": bool(np.all(np.isfinite(energy))),
    }


def run_sequential(specs: Sequence[InstanceSpec], steps: int, tune: dict,
                   precision: int = L.MIXED, weight_order: int = L.WEIGHTS_STANDARD) -> dict:
    """The same instances stepped one after another with Simulation.step
    (no group): the batching gain is value(group) / value(sequential)."""
    sims = [s.create(tune["tile"], tune["chunks"], precision, weight_order=weight_order)
            for s in specs]
    try:
        for sim in sims:
            sim.step(1)
            sim.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            for sim in sims:
                sim.step(1)
        for sim in sims:
            sim.synchronize()
        total = time.perf_counter() - t
    finally:
        for sim in sims:
            sim.close()
    npart = sum(s.particles for s in specs)
    return {"value": npart * steps / total, "steps": steps,
            "step_ms": total / steps * 1e3}


FIELD_NAMES = ("rho", "phi", "Ex", "Ey", "Ez")


def record_fields(sim: Simulation, steps: int) -> List[List[np.ndarray]]:
    """rho, phi, Ex, Ey, Ez after each of ``steps`` single steps (the first
    holds the loaded state's fields: a step deposits and solves before it
    pushes)."""
    out = []
    for _ in range(steps):
        sim.step(1)
        out.append([sim.grid(w) for w in (L.FIELD_RHO, L.FIELD_PHI, L.FIELD_EX, L.FIELD_EY,
                                          L.FIELD_EZ)])
    return out


def field_errors(got: Sequence[Sequence[np.ndarray]],
                 ref: Sequence[Sequence[np.ndarray]]) -> dict:
    """Relative L2 of each field per step: {"step1": {"rho": ..}, ..} plus the
    largest rho/E error over the steps after the first push."""
    out = {}
    worst = 0.0
    for k, (g, r) in enumerate(zip(got, ref)):
        e = {}
        for name, a, b in zip(FIELD_NAMES, g, r):
            den = float(np.linalg.norm(b))
            e[name] = float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))
        out[f"step{k + 1}"] = e
        if k > 0:
            worst = max(worst, e["rho"], e["Ex"], e["Ey"], e["Ez"])
    out["max_rho_E_after_push"] = worst if len(got) > 1 else None
    return out


def error_vs_fp64(spec: InstanceSpec, steps: int = 4, tile=(0, 0, 0), chunks: int = 0,
                  weight_order: int = L.WEIGHTS_STANDARD) -> dict:
    """The mixed-precision run of one instance against its FP64 run from the
    same seeded state, field by field over its first ``steps`` steps."""
    res = {}
    for prec in (L.FP64, L.MIXED):
        with spec.create(tile, chunks, prec, weight_order=weight_order) as sim:
            res[prec] = record_fields(sim, steps)
    out = field_errors(res[L.MIXED], res[L.FP64])
    out["instance"] = {"n": list(spec.n), "ppc": list(spec.ppc), "seed": spec.seed,
                       "weight_order": weight_order}
    return out
