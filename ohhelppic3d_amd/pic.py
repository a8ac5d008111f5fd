"""Host-side mirror of ohhelppic3d's per-timestep interfaces over libohpic.so.

The reference drives its hot path through polymorphic Fortran objects, one call
per particle (src/ohhelppic3d.f90:311-358).  Here each object keeps the
reference's name and role but acts on every particle of a device-resident
simulation at once:

    reference (file:line)                                   here
    t_LinearScatter%scatter   linear_scatter.f90:22-55      LinearScatter.scatter_all
    t_PoissonFieldSolver%solve poisson_field_solver.f90:90   PoissonFieldSolver.solve
    t_LinearInterpolator%interp linear_interpolator.f90:22   (fused into move_all)
    t_ParticleMoverBoris%move particle_mover_boris.f90:21   ParticleMoverBoris.move_all
    t_ParticleBoundaries%apply particle_boundaries.f90:34    ParticleBoundaries.apply_all
    mainstep                  ohhelppic3d.f90:286-381       mainstep / Simulation.step

Errors raise ``OhpicError`` (the reference aborts with error stop/MPI_Abort).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _lib as L
from ._lib import OhpicError, check


def charge_per_macro_particle(plasma_frequency, charge_to_mass_ratio, ppc):
    """q_macro = wp^2 / (q/m) / ppc  (parameters.f90:140-149)."""
    wp = np.asarray(plasma_frequency, np.float64)
    qm = np.asarray(charge_to_mass_ratio, np.float64)
    return wp * wp / qm / np.asarray(ppc, np.float64)


@dataclass
class SimConfig:
    n: Sequence[int] = (32, 32, 32)
    qm: Sequence[float] = (-1.0, 0.01)
    q_macro: Sequence[float] = (-1.0 / 16, 1.0 / 16)
    capacity: Sequence[int] = (16 * 32 ** 3, 16 * 32 ** 3)
    dt: float = 0.1
    b0: Sequence[float] = (0.0, 0.0, 0.5)
    precision: int = L.FP64
    weight_order: int = L.WEIGHTS_REFERENCE
    tile: Sequence[int] = (0, 0, 0)
    slack: float = 0.15
    chunks: int = 0
    track_ids: bool = False
    device: int = 0
    charge_bits: int = 0

    def to_c(self) -> L.Config:
        c = L.Config()
        L.lib().ohpic_config_default(C.byref(c))
        c.nx, c.ny, c.nz = (int(v) for v in self.n)
        ns = len(self.qm)
        if not (len(self.q_macro) == ns == len(self.capacity)):
            raise OhpicError(L.EINVAL, "qm, q_macro and capacity must have one entry per species")
        c.nspecies = ns
        for i in range(ns):
            c.qm[i] = self.qm[i]
            c.q_macro[i] = self.q_macro[i]
            c.capacity[i] = int(self.capacity[i])
        c.dt = self.dt
        for i in range(3):
            c.b0[i] = self.b0[i]
            c.tile[i] = int(self.tile[i])
        c.precision = self.precision
        c.weight_order = self.weight_order
        c.slack = self.slack
        c.chunks = self.chunks
        c.track_ids = 1 if self.track_ids else 0
        c.device = self.device
        c.charge_bits = self.charge_bits
        return c


class Simulation:
    """One device-resident PIC instance (an ``ohpic_ctx``)."""

    def __init__(self, cfg: SimConfig):
        self.cfg = cfg
        self._lib = L.lib()
        self._ctx = C.c_void_p()
        c = cfg.to_c()
        check(self._lib.ohpic_create(C.byref(c), C.byref(self._ctx)))
        got = L.Config()
        self._chk(self._lib.ohpic_get_config(self._ctx, C.byref(got)))
        self.tile = (got.tile[0], got.tile[1], got.tile[2])
        self.charge_bits = got.charge_bits
        self.n = tuple(int(v) for v in cfg.n)
        self.nspecies = len(cfg.qm)

    # -- lifetime ---------------------------------------------------------
    def close(self) -> None:
        if self._ctx:
            self._lib.ohpic_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        check(rc, self._ctx)

    # -- particles --------------------------------------------------------
    def load_particles(self, ispec: int, x, y, z, vx, vy, vz, pid=None) -> None:
        arrs = [np.ascontiguousarray(a, np.float64) for a in (x, y, z, vx, vy, vz)]
        n = len(arrs[0])
        if any(len(a) != n for a in arrs):
            raise OhpicError(L.EINVAL, "particle arrays differ in length")
        p = None if pid is None else L.i64ptr(np.ascontiguousarray(pid, np.int64))
        self._chk(self._lib.ohpic_load_particles(self._ctx, ispec, n,
                                                 *[L.dptr(a) for a in arrs], p))

    def particle_count(self, ispec: int) -> int:
        n = C.c_int64()
        self._chk(self._lib.ohpic_particle_count(self._ctx, ispec, C.byref(n)))
        return n.value

    def store_particles(self, ispec: int, with_ids: bool = False):
        n = self.particle_count(ispec)
        out = [np.empty(n, np.float64) for _ in range(6)]
        pid = np.empty(n, np.int64) if with_ids else None
        got = C.c_int64()
        self._chk(self._lib.ohpic_store_particles(
            self._ctx, ispec, n, C.byref(got), *[L.dptr(a) for a in out],
            L.i64ptr(pid) if with_ids else None))
        res = dict(zip(("x", "y", "z", "vx", "vy", "vz"), [a[: got.value] for a in out]))
        if with_ids:
            res["pid"] = pid[: got.value]
        return res

    def load_oh_particles(self, records: np.ndarray) -> None:
        rec = np.ascontiguousarray(records, L.OH_PARTICLE_DTYPE)
        self._chk(self._lib.ohpic_load_oh_particles(self._ctx, len(rec), rec.ctypes.data))

    def store_oh_particles(self) -> np.ndarray:
        tot = sum(self.particle_count(s) for s in range(self.nspecies))
        rec = np.zeros(tot, L.OH_PARTICLE_DTYPE)
        got = C.c_int64()
        self._chk(self._lib.ohpic_store_oh_particles(self._ctx, tot, C.byref(got),
                                                     rec.ctypes.data))
        return rec[: got.value]

    def seed_uniform_maxwell(self, np_per_species, vth, seed=(42, 52)) -> None:
        npa = np.ascontiguousarray(np_per_species, np.int64)
        vt = np.ascontiguousarray(vth, np.float64)
        self._chk(self._lib.ohpic_seed_uniform_maxwell(self._ctx, seed[0], seed[1],
                                                       L.i64ptr(npa), L.dptr(vt)))

    def seed_hump_maxwell(self, np_per_species, vth, amp, sigma, seed=(42, 52),
                          budget=2052) -> None:
        npa = np.ascontiguousarray(np_per_species, np.int64)
        vt = np.ascontiguousarray(vth, np.float64)
        self._chk(self._lib.ohpic_seed_hump_maxwell(self._ctx, seed[0], seed[1], L.i64ptr(npa),
                                                    L.dptr(vt), amp, sigma, budget))

    # -- operators --------------------------------------------------------
    def deposit(self) -> None:
        self._chk(self._lib.ohpic_deposit(self._ctx))

    def field_solve(self) -> None:
        self._chk(self._lib.ohpic_field_solve(self._ctx))

    def gather_push(self) -> None:
        self._chk(self._lib.ohpic_gather_push(self._ctx))

    def boundary(self) -> None:
        self._chk(self._lib.ohpic_boundary(self._ctx))

    def step(self, nsteps: int = 1) -> None:
        self._chk(self._lib.ohpic_step(self._ctx, int(nsteps)))

    def synchronize(self) -> None:
        self._chk(self._lib.ohpic_synchronize(self._ctx))

    # -- fields -----------------------------------------------------------
    def get_field(self, which: int, lb=(0, 0, 0), ub=None, nelem: Optional[int] = None,
                  comp: int = 0) -> np.ndarray:
        """values(nelem, lb:ub) in Fortran order (returned as an F-ordered array)."""
        if ub is None:
            ub = tuple(m - 1 for m in self.n)
        if nelem is None:
            nelem = 6 if which == L.FIELD_EB else 1
        shape = (nelem,) + tuple(u - l + 1 for l, u in zip(lb, ub))
        out = np.zeros(shape, np.float64, order="F")
        lb_a = np.asarray(lb, np.int32)
        ub_a = np.asarray(ub, np.int32)
        buf = np.zeros(int(np.prod(shape)), np.float64)
        self._chk(self._lib.ohpic_get_field(self._ctx, which, L.dptr(buf), nelem,
                                            L.i32ptr(lb_a), L.i32ptr(ub_a), comp))
        out[...] = buf.reshape(shape, order="F")
        return out

    def grid(self, which: int) -> np.ndarray:
        """The periodic grid as a flat array, x fastest (nx*ny*nz)."""
        v = self.get_field(which)
        return np.asarray(v[0]).reshape(-1, order="F")

    def put_grid(self, which: int, values) -> None:
        v = np.ascontiguousarray(values, np.float64).reshape(-1)
        lb = np.zeros(3, np.int32)
        ub = np.asarray([m - 1 for m in self.n], np.int32)
        self._chk(self._lib.ohpic_put_field(self._ctx, which, L.dptr(v), 1, L.i32ptr(lb),
                                            L.i32ptr(ub), 0))

    # -- diagnostics ------------------------------------------------------
    def diagnostics(self) -> dict:
        d = L.Diag()
        self._chk(self._lib.ohpic_diagnostics(self._ctx, C.byref(d)))
        ns = self.nspecies
        return dict(step=d.step, ke_old=np.array(d.ke_old[:ns]), ke_new=np.array(d.ke_new[:ns]),
                    field_energy=d.field_energy, field_energy_e2=d.field_energy_e2,
                    net_charge=d.net_charge, total_energy=d.total_energy, repacks=d.repacks,
                    fast_particles=d.fast_particles, rebalances=d.rebalances,
                    compactions=d.compactions)

    def energy(self) -> np.ndarray:
        """[kinetic, field, total, net charge] of the last step (ohpic_energy)."""
        out = np.zeros(4)
        self._chk(self._lib.ohpic_energy(self._ctx, L.dptr(out)))
        return out

    def cell_counts(self, ispec: int = -1) -> np.ndarray:
        c = np.zeros(self.n[0] * self.n[1] * self.n[2], np.int32)
        self._chk(self._lib.ohpic_cell_counts(self._ctx, ispec, L.i32ptr(c)))
        return c

    def enable_timing(self, on: bool = True) -> None:
        self._chk(self._lib.ohpic_enable_timing(self._ctx, 1 if on else 0))

    def kernel_times(self) -> dict:
        names = (C.c_char * 32 * 16)()
        ms = np.zeros(16)
        cnt = np.zeros(16, np.int64)
        k = C.c_int()
        self._chk(self._lib.ohpic_kernel_times(self._ctx, 16, C.cast(names, C.c_void_p),
                                               L.dptr(ms), L.i64ptr(cnt), C.byref(k)))
        return {names[i].value.decode(): (ms[i], int(cnt[i])) for i in range(k.value)}

    def memory_footprint(self) -> int:
        b = C.c_int64()
        self._chk(self._lib.ohpic_memory_footprint(self._ctx, C.byref(b)))
        return b.value


class Group:
    """Instances advanced together (``ohpic_group``): one launch per kernel and
    step over every member's work items and FFT jobs, replayed from a captured
    hipGraph.  Members must share device, precision, weight order and tile;
    each member's result equals stepping it alone with ``Simulation.step``."""

    def __init__(self, device: int = 0, use_graph: bool = True, persistent: bool = True,
                 batch: Optional[int] = None):
        self._lib = L.lib()
        self._g = C.c_void_p()
        check(self._lib.ohpic_group_create(int(device), C.byref(self._g)))
        self.members: list = []
        if not use_graph:
            self.use_graph(False)
        if not persistent:
            self.use_persistent(False)
        if batch is not None:
            self.set_batch(batch)

    def _chk(self, rc):
        L.check_group(rc, self._g)

    def close(self) -> None:
        if self._g:
            self._lib.ohpic_group_destroy(self._g)
            self._g = C.c_void_p()
            self.members = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def attach(self, sim: Simulation) -> None:
        self._chk(self._lib.ohpic_group_attach(self._g, sim._ctx))
        self.members.append(sim)

    def detach(self, sim: Simulation) -> None:
        self._chk(self._lib.ohpic_group_detach(self._g, sim._ctx))
        if sim in self.members:
            self.members.remove(sim)

    def __len__(self) -> int:
        n = C.c_int()
        self._chk(self._lib.ohpic_group_size(self._g, C.byref(n)))
        return n.value

    def contains(self, sim: Simulation) -> bool:
        if not sim._ctx:
            return False
        y = C.c_int()
        self._chk(self._lib.ohpic_group_contains(self._g, sim._ctx, C.byref(y)))
        return bool(y.value)

    def step(self, nsteps: int = 1) -> None:
        """Advance every member by nsteps, in launch sequences of `batch` steps
        (one host synchronisation each) while that many remain; on a member
        failure the others complete the sequence, failed members are detached,
        and OhpicError is raised (``stats()["steps"]`` counts the steps taken)."""
        rc = self._lib.ohpic_group_step(self._g, int(nsteps))
        if rc != L.OK:
            self.members = [m for m in self.members if self.contains(m)]
        self._chk(rc)

    def use_graph(self, on: bool) -> None:
        self._chk(self._lib.ohpic_group_use_graph(self._g, 1 if on else 0))

    def use_persistent(self, on: bool) -> None:
        self._chk(self._lib.ohpic_group_use_persistent(self._g, 1 if on else 0))

    def set_batch(self, steps: int) -> None:
        """Steps per launch sequence (1 to 16, default 8)."""
        self._chk(self._lib.ohpic_group_set_batch(self._g, int(steps)))

    def last_error(self) -> str:
        """The group's last error message (also of a failed graph capture that
        it recovered from by launching that sequence directly); "" if none."""
        m = self._lib.ohpic_group_last_error(self._g)
        return m.decode() if isinstance(m, bytes) else (m or "")

    def stats(self) -> dict:
        s, c, i = C.c_int64(), C.c_int64(), C.c_int64()
        self._chk(self._lib.ohpic_group_stats(self._g, C.byref(s), C.byref(c), C.byref(i)))
        r, d = C.c_int64(), C.c_int64()
        self._chk(self._lib.ohpic_group_sequences(self._g, C.byref(r), C.byref(d)))
        return {"steps": s.value, "captures": c.value, "items": i.value,
                "replayed": r.value, "direct": d.value}


def autotune(cfgs: Sequence[SimConfig], sample: int = 16, steps: int = 3) -> dict:
    """Fastest (tile, chunks) for a group of these configurations, timed on up
    to ``sample`` of them (``ohpic_autotune``)."""
    arr = (L.Config * len(cfgs))(*[c.to_c() for c in cfgs])
    r = L.TuneResult()
    check(L.lib().ohpic_autotune(arr, len(cfgs), int(sample), int(steps), C.byref(r)))
    return {"tile": tuple(r.tile), "chunks": r.chunks,
            "particle_steps_per_s": r.particle_steps_per_s, "candidates": r.candidates}


def estimate_footprint(cfg: SimConfig) -> int:
    b = C.c_int64()
    c = cfg.to_c()
    check(L.lib().ohpic_estimate_footprint(C.byref(c), C.byref(b)))
    return b.value


def fft3d_r2r(a: np.ndarray, n, backward: bool = False, device: int = 0) -> np.ndarray:
    """t_MPIFFTSolver3d forward (R2HC per axis) / backward (HC2R per axis, / N)."""
    src = np.ascontiguousarray(a, np.float64).reshape(-1)
    out = np.empty_like(src)
    check(L.lib().ohpic_fft3d_r2r(device, int(n[0]), int(n[1]), int(n[2]), 1 if backward else 0,
                                  L.dptr(src), L.dptr(out)))
    return out


# --------------------------------------------------------------------------
# Objects named after the reference's polymorphic types (whole-array methods)
# --------------------------------------------------------------------------
class LinearScatter:
    """t_LinearScatter (linear_scatter.f90): CIC charge deposit of all particles."""

    def scatter_all(self, sim: Simulation) -> None:
        sim.deposit()


class PoissonFieldSolver:
    """t_PoissonFieldSolver (poisson_field_solver.f90:90-175), periodic."""

    def solve(self, sim: Simulation) -> None:
        sim.field_solve()


class LinearInterpolator:
    """t_LinearInterpolator: the gather runs inside ParticleMoverBoris.move_all."""


class ParticleMoverBoris:
    """t_ParticleMoverBoris (particle_mover_boris.f90:21-64) with the CIC gather."""

    def move_all(self, sim: Simulation) -> None:
        sim.gather_push()


class ParticleBoundaries:
    """t_ParticleBoundaries (particle_boundaries.f90:23-43): periodic on 3 axes."""

    def __init__(self, nx: int, ny: int, nz: int):
        self.n = (nx, ny, nz)

    def apply_all(self, sim: Simulation) -> None:
        if tuple(sim.n) != self.n:
            raise OhpicError(L.EINVAL, "boundary extents differ from the simulation grid")
        sim.boundary()


def create_particle_mover(name: str) -> ParticleMoverBoris:
    """particle_mover_factory.f90:20-31 ('boris' only; unknown names raise)."""
    if name == "boris":
        return ParticleMoverBoris()
    raise OhpicError(L.EINVAL, f"unknown particle mover {name!r}")


def mainstep(sim: Simulation, scatter: LinearScatter, solver: PoissonFieldSolver,
             mover: ParticleMoverBoris, boundaries: ParticleBoundaries) -> None:
    """One mainstep (ohhelppic3d.f90:286-381) through the separate operators:
    rho := 0 and deposit; solve; gather + Boris; boundary."""
    scatter.scatter_all(sim)
    solver.solve(sim)
    mover.move_all(sim)
    boundaries.apply_all(sim)
