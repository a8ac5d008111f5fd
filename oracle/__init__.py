"""CPU FP64 oracle for the ohhelppic3d per-timestep loop (ctypes over liboracle.so).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this package, and only as the
checker or the timed CPU baseline.  The product package ``ohhelppic3d_amd``
never imports it.

Each wrapper names the reference function it restates; the C source
``oracle/ohpic_oracle.c`` carries the file:line citations.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_REF_PCG = os.path.join(_HERE, "_ref", "libpcgref.so")

REF_PI = 3.14159265358979  # src/utils/science_constants.f90:4

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class Pcg32State(C.Structure):
    _fields_ = [("state", C.c_uint64), ("inc", C.c_uint64)]


class OrcSim(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("nspecies", C.c_int),
        ("dt", C.c_double),
        ("b0", C.c_double * 3),
        ("weight_order", C.c_int),
        ("qm", C.c_double * 8),
        ("q", C.c_double * 8),
        ("np", C.c_int64 * 8),
        ("x", C.c_void_p * 8), ("y", C.c_void_p * 8), ("z", C.c_void_p * 8),
        ("vx", C.c_void_p * 8), ("vy", C.c_void_p * 8), ("vz", C.c_void_p * 8),
        ("rho", C.c_void_p), ("phi", C.c_void_p),
        ("ex", C.c_void_p), ("ey", C.c_void_p), ("ez", C.c_void_p),
        ("ke_old", C.c_double * 8),
        ("ke_new", C.c_double * 8),
        ("field_energy", C.c_double),
        ("net_charge", C.c_double),
    ]


_lib = None


def build() -> None:
    """Compile liboracle.so (and oracle/_ref when the reference is mounted)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.POINTER(Pcg32State)
        L.orc_pcg32_srandom.argtypes = [P, C.c_uint64, C.c_uint64]
        L.orc_pcg32_random.argtypes = [P]
        L.orc_pcg32_random.restype = C.c_uint32
        L.orc_pcg32_boundedrand.argtypes = [P, C.c_uint32]
        L.orc_pcg32_boundedrand.restype = C.c_uint32
        L.orc_pcg32_double.argtypes = [P]
        L.orc_pcg32_double.restype = C.c_double
        L.orc_pcg32_advance.argtypes = [P, C.c_uint64]
        L.orc_load_uniform_maxwell.argtypes = [P, C.c_int64, C.c_int, C.c_int, C.c_int,
                                               C.c_double] + [_dp] * 6
        L.orc_load_uniform_maxwell.restype = C.c_int64
        L.orc_load_hump_maxwell.argtypes = [P, C.c_uint64, C.c_uint64, C.c_int64, C.c_int,
                                            C.c_int, C.c_int, C.c_double, C.c_double,
                                            C.c_double] + [_dp] * 6
        L.orc_load_hump_maxwell.restype = C.c_int
        L.orc_apply_periodic.argtypes = [C.c_int, C.c_double]
        L.orc_apply_periodic.restype = C.c_double
        dref = C.POINTER(C.c_double)
        L.orc_boris_move.argtypes = [dref] * 6 + [C.c_double, _dp, C.c_double]
        L.orc_interp.argtypes = [_dp, C.c_int, _i32p, _i32p, _dp, C.c_int, _dp]
        L.orc_scatter.argtypes = [_dp, C.c_int, _i32p, _i32p, _dp, _dp, C.c_int]
        L.orc_deposit_periodic.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int64, _dp, _dp,
                                           _dp, C.c_double, C.c_int, _dp]
        L.orc_gather_periodic.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp,
                                          C.c_double, C.c_double, C.c_double, C.c_int, _dp]
        L.orc_r2r_axis.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp]
        L.orc_r2hc_1d.argtypes = [C.c_int, _dp, _dp]
        L.orc_hc2r_1d.argtypes = [C.c_int, _dp, _dp]
        L.orc_wave_number_sq.argtypes = [C.c_int, C.c_int]
        L.orc_wave_number_sq.restype = C.c_double
        L.orc_poisson_periodic.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.orc_efield.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        L.orc_step.argtypes = [C.POINTER(OrcSim)]
        L.orc_step_threads.argtypes = [C.POINTER(OrcSim), C.c_int]
        L.orc_step_threads.restype = C.c_int
        L.orc_kinetic_energy.argtypes = [C.c_int64, _dp, _dp, _dp, C.c_double]
        L.orc_kinetic_energy.restype = C.c_double
        L.orc_cell_counts.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int64, _dp, _dp, _dp,
                                      _i32p]
        _lib = L
    return _lib


# --------------------------------------------------------------------------- PCG
class Pcg32:
    """PCG32 (pcg_setseq_64 / XSH-RR), restating lib/pcg-c."""

    def __init__(self, initstate: int = 42, initseq: int = 52):
        self.st = Pcg32State()
        lib().orc_pcg32_srandom(C.byref(self.st), initstate & (2**64 - 1), initseq & (2**64 - 1))

    def random(self) -> int:
        return lib().orc_pcg32_random(C.byref(self.st))

    def boundedrand(self, bound: int) -> int:
        return lib().orc_pcg32_boundedrand(C.byref(self.st), bound)

    def double(self) -> float:
        return lib().orc_pcg32_double(C.byref(self.st))

    def advance(self, delta: int) -> None:
        lib().orc_pcg32_advance(C.byref(self.st), delta & (2**64 - 1))

    def copy(self) -> "Pcg32":
        o = Pcg32.__new__(Pcg32)
        o.st = Pcg32State(self.st.state, self.st.inc)
        return o


def ref_pcg_lib():
    """The reference's own PCG-C built from its sources (None if not built)."""
    if not os.path.exists(_REF_PCG):
        return None
    L = C.CDLL(_REF_PCG)
    P = C.POINTER(Pcg32State)
    L.pcg_setseq_64_srandom_r.argtypes = [P, C.c_uint64, C.c_uint64]
    L.pcg_setseq_64_xsh_rr_32_random_r.argtypes = [P]
    L.pcg_setseq_64_xsh_rr_32_random_r.restype = C.c_uint32
    L.pcg_setseq_64_advance_r.argtypes = [P, C.c_uint64]
    L.pcg32_random_double_r.argtypes = [P]
    L.pcg32_random_double_r.restype = C.c_double
    return L


# ----------------------------------------------------------------------- loaders
def load_uniform_maxwell(rng: Pcg32, np_: int, n: tuple, vth: float):
    x, y, z, vx, vy, vz = (np.empty(np_, np.float64) for _ in range(6))
    lib().orc_load_uniform_maxwell(C.byref(rng.st), np_, n[0], n[1], n[2], vth,
                                   x, y, z, vx, vy, vz)
    return x, y, z, vx, vy, vz


def load_hump_maxwell(rng0: Pcg32, base: int, budget: int, np_: int, n: tuple, amp: float,
                      sigma: float, vth: float):
    x, y, z, vx, vy, vz = (np.empty(np_, np.float64) for _ in range(6))
    rc = lib().orc_load_hump_maxwell(C.byref(rng0.st), base, budget, np_, n[0], n[1], n[2],
                                     amp, sigma, vth, x, y, z, vx, vy, vz)
    if rc != 0:
        raise RuntimeError("rejection sampling ran out of draws")
    return x, y, z, vx, vy, vz


# --------------------------------------------------------- per-particle operators
def apply_periodic(n: int, p: float) -> float:
    return lib().orc_apply_periodic(n, p)


def boris_move(pos, vel, qm, eb, dt):
    """particle_mover_boris.f90:21-64 on one particle; returns (pos, vel)."""
    cs = [C.c_double(float(v)) for v in list(pos) + list(vel)]
    lib().orc_boris_move(*[C.byref(c) for c in cs], qm, np.ascontiguousarray(eb, np.float64),
                         dt)
    v = [c.value for c in cs]
    return np.array(v[:3]), np.array(v[3:])


def interp(values, lb, ub, pos, weight_order=0):
    """linear_interpolator.f90:22-69 over values(nelem, lb:ub) in Fortran order."""
    nelem = values.shape[0]
    flat = np.ascontiguousarray(values.reshape(-1, order="F"))
    ret = np.zeros(nelem)
    lib().orc_interp(flat, nelem, np.asarray(lb, np.int32), np.asarray(ub, np.int32),
                     np.asarray(pos, np.float64), weight_order, ret)
    return ret


def deposit_periodic(n, x, y, z, q, weight_order=0, rho=None):
    nx, ny, nz = n
    if rho is None:
        rho = np.zeros(nx * ny * nz)
    lib().orc_deposit_periodic(nx, ny, nz, len(x), np.ascontiguousarray(x),
                               np.ascontiguousarray(y), np.ascontiguousarray(z), q,
                               weight_order, rho)
    return rho


def gather_periodic(n, e, pos, weight_order=0):
    out = np.zeros(3)
    lib().orc_gather_periodic(n[0], n[1], n[2], e[0], e[1], e[2], pos[0], pos[1], pos[2],
                              weight_order, out)
    return out


def r2hc_1d(a):
    a = np.ascontiguousarray(a, np.float64)
    out = np.empty_like(a)
    lib().orc_r2hc_1d(len(a), a, out)
    return out


def hc2r_1d(a):
    a = np.ascontiguousarray(a, np.float64)
    out = np.empty_like(a)
    lib().orc_hc2r_1d(len(a), a, out)
    return out


def r2r_axis(n, arr, axis, kind):
    a = np.ascontiguousarray(arr, np.float64).copy()
    lib().orc_r2r_axis(n[0], n[1], n[2], axis, kind, a)
    return a


def wave_number_sq(k, n):
    return lib().orc_wave_number_sq(k, n)


def poisson(n, rho):
    phi = np.empty(n[0] * n[1] * n[2])
    lib().orc_poisson_periodic(n[0], n[1], n[2], np.ascontiguousarray(rho, np.float64), phi)
    return phi


def efield(n, phi):
    g = n[0] * n[1] * n[2]
    ex, ey, ez = np.empty(g), np.empty(g), np.empty(g)
    lib().orc_efield(n[0], n[1], n[2], np.ascontiguousarray(phi, np.float64), ex, ey, ez)
    return ex, ey, ez


def cell_counts(n, x, y, z):
    c = np.zeros(n[0] * n[1] * n[2], np.int32)
    lib().orc_cell_counts(n[0], n[1], n[2], len(x), np.ascontiguousarray(x),
                          np.ascontiguousarray(y), np.ascontiguousarray(z), c)
    return c


# -------------------------------------------------------------------- simulation
class OracleSim:
    """Time-stepper over mainstep's order (ohhelppic3d.f90:286-381), FP64."""

    def __init__(self, n, species, dt, b0, weight_order=0):
        """species: list of dicts {qm, q, x, y, z, vx, vy, vz} (arrays copied)."""
        self.n = tuple(n)
        g = n[0] * n[1] * n[2]
        self.rho, self.phi = np.zeros(g), np.zeros(g)
        self.ex, self.ey, self.ez = np.zeros(g), np.zeros(g), np.zeros(g)
        self.sp = []
        s = OrcSim()
        s.nx, s.ny, s.nz = n
        s.nspecies = len(species)
        s.dt = dt
        for i in range(3):
            s.b0[i] = b0[i]
        s.weight_order = weight_order
        for i, d in enumerate(species):
            arrs = {k: np.ascontiguousarray(d[k], np.float64).copy()
                    for k in ("x", "y", "z", "vx", "vy", "vz")}
            self.sp.append(arrs)
            s.qm[i], s.q[i], s.np[i] = d["qm"], d["q"], len(arrs["x"])
            for k in ("x", "y", "z", "vx", "vy", "vz"):
                getattr(s, k)[i] = arrs[k].ctypes.data
        for k in ("rho", "phi", "ex", "ey", "ez"):
            setattr(s, k, getattr(self, k).ctypes.data)
        self.s = s

    def step(self, nsteps: int = 1):
        for _ in range(nsteps):
            lib().orc_step(C.byref(self.s))

    def step_threads(self, nsteps: int = 1, threads: int = 0) -> int:
        """orc_step_threads: the same step over OpenMP threads (0: default
        count); returns the thread count used."""
        used = 1
        for _ in range(nsteps):
            used = lib().orc_step_threads(C.byref(self.s), threads)
            if used < 0:
                raise MemoryError("orc_step_threads: no memory for per-thread charge arrays")
        return used

    @property
    def ke_old(self):
        return np.array(self.s.ke_old[: self.s.nspecies])

    @property
    def ke_new(self):
        return np.array(self.s.ke_new[: self.s.nspecies])

    @property
    def field_energy(self):
        return self.s.field_energy

    @property
    def net_charge(self):
        return self.s.net_charge
