"""ctypes bindings of libohpic.so (the C ABI declared in include/ohpic.h).

There is no fallback: if the library is missing or fails to load, every
entry point raises ``OhpicError``.  Build it with ``python -m ohhelppic3d_amd.build``
(or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("OHPIC_LIB", os.path.join(PKG, "libohpic.so"))

MAX_SPECIES = 8

OK, EINVAL, EHIP, ENOMEM, ESTATE, ENODEV, EIO, ENOLIB = 0, -1, -2, -3, -4, -6, -7, -8
FP64, MIXED = 0, 1
WEIGHTS_REFERENCE, WEIGHTS_STANDARD = 0, 1
FIELD_RHO, FIELD_PHI, FIELD_EX, FIELD_EY, FIELD_EZ, FIELD_EB = range(6)

# every symbol include/ohpic.h declares
EXPORTS = (
    "ohpic_config_default", "ohpic_create", "ohpic_destroy", "ohpic_last_error",
    "ohpic_get_config", "ohpic_load_particles", "ohpic_particle_count",
    "ohpic_store_particles", "ohpic_load_oh_particles", "ohpic_store_oh_particles",
    "ohpic_seed_uniform_maxwell", "ohpic_seed_hump_maxwell", "ohpic_deposit",
    "ohpic_field_solve", "ohpic_gather_push", "ohpic_boundary", "ohpic_step",
    "ohpic_synchronize", "ohpic_get_field", "ohpic_put_field", "ohpic_fft3d_r2r",
    "ohpic_fft3d_r2r_noalloc",
    "ohpic_diagnostics", "ohpic_energy", "ohpic_cell_counts", "ohpic_enable_timing", "ohpic_kernel_times",
    "ohpic_memory_footprint", "ohpic_estimate_footprint", "ohpic_boris_batch",
    "ohpic_interp_batch", "ohpic_scatter_batch", "ohpic_wrap_batch",
    "ohpic_group_create", "ohpic_group_destroy", "ohpic_group_attach", "ohpic_group_detach",
    "ohpic_group_size", "ohpic_group_contains", "ohpic_group_step", "ohpic_group_use_graph",
    "ohpic_group_use_persistent", "ohpic_group_set_batch", "ohpic_group_stats",
    "ohpic_group_sequences",
    "ohpic_group_last_error", "ohpic_autotune",
    "ohpic_snapshot_open", "ohpic_snapshot_write", "ohpic_snapshot_write_host",
    "ohpic_snapshot_close", "ohpic_snapshot_last_error", "ohpic_snapshot_read",
)


class OhpicError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ohpic error {code}: {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("nspecies", C.c_int32),
        ("qm", C.c_double * MAX_SPECIES),
        ("q_macro", C.c_double * MAX_SPECIES),
        ("capacity", C.c_int64 * MAX_SPECIES),
        ("dt", C.c_double),
        ("b0", C.c_double * 3),
        ("precision", C.c_int32),
        ("weight_order", C.c_int32),
        ("tile", C.c_int32 * 3),
        ("slack", C.c_double),
        ("chunks", C.c_int32),
        ("track_ids", C.c_int32),
        ("device", C.c_int32),
        ("charge_bits", C.c_int32),
    ]


class Diag(C.Structure):
    _fields_ = [
        ("step", C.c_int64),
        ("ke_old", C.c_double * MAX_SPECIES),
        ("ke_new", C.c_double * MAX_SPECIES),
        ("field_energy", C.c_double),
        ("field_energy_e2", C.c_double),
        ("net_charge", C.c_double),
        ("total_energy", C.c_double),
        ("repacks", C.c_int64),
        ("fast_particles", C.c_int64),
        ("rebalances", C.c_int64),
        ("compactions", C.c_int64),
    ]


class TuneResult(C.Structure):
    _fields_ = [
        ("tile", C.c_int32 * 3),
        ("chunks", C.c_int32),
        ("particle_steps_per_s", C.c_double),
        ("candidates", C.c_int32),
    ]


class OhParticle(C.Structure):
    """oh_particle record: x, y, z, vx, vy, vz (f64), pid (i64), nid, spec (i32)."""
    _fields_ = [
        ("x", C.c_double), ("y", C.c_double), ("z", C.c_double),
        ("vx", C.c_double), ("vy", C.c_double), ("vz", C.c_double),
        ("pid", C.c_int64), ("nid", C.c_int32), ("spec", C.c_int32),
    ]


OH_PARTICLE_DTYPE = np.dtype([
    ("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("vx", "<f8"), ("vy", "<f8"), ("vz", "<f8"),
    ("pid", "<i8"), ("nid", "<i4"), ("spec", "<i4")])

_lib = None


def available() -> bool:
    return os.path.exists(LIB_PATH)


def lib():
    """Load libohpic.so (raises OhpicError when it is absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OhpicError(ENODEV, f"{LIB_PATH} not built: run python -m ohhelppic3d_amd.build")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    dp = C.POINTER(C.c_double)
    i64p = C.POINTER(C.c_int64)
    i32p = C.POINTER(C.c_int32)
    sig = {
        "ohpic_config_default": [C.POINTER(Config)],
        "ohpic_create": [C.POINTER(Config), C.POINTER(vp)],
        "ohpic_destroy": [vp],
        "ohpic_get_config": [vp, C.POINTER(Config)],
        "ohpic_load_particles": [vp, C.c_int, i64, dp, dp, dp, dp, dp, dp, i64p],
        "ohpic_particle_count": [vp, C.c_int, i64p],
        "ohpic_store_particles": [vp, C.c_int, i64, i64p, dp, dp, dp, dp, dp, dp, i64p],
        "ohpic_load_oh_particles": [vp, i64, vp],
        "ohpic_store_oh_particles": [vp, i64, i64p, vp],
        "ohpic_seed_uniform_maxwell": [vp, C.c_uint64, C.c_uint64, i64p, dp],
        "ohpic_seed_hump_maxwell": [vp, C.c_uint64, C.c_uint64, i64p, dp, dbl, dbl,
                                    C.c_uint64],
        "ohpic_deposit": [vp],
        "ohpic_field_solve": [vp],
        "ohpic_gather_push": [vp],
        "ohpic_boundary": [vp],
        "ohpic_step": [vp, C.c_int],
        "ohpic_synchronize": [vp],
        "ohpic_get_field": [vp, C.c_int, dp, C.c_int, i32p, i32p, C.c_int],
        "ohpic_put_field": [vp, C.c_int, dp, C.c_int, i32p, i32p, C.c_int],
        "ohpic_fft3d_r2r": [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp],
        "ohpic_fft3d_r2r_noalloc": [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp],
        "ohpic_diagnostics": [vp, C.POINTER(Diag)],
        "ohpic_energy": [vp, dp],
        "ohpic_cell_counts": [vp, C.c_int, i32p],
        "ohpic_enable_timing": [vp, C.c_int],
        "ohpic_kernel_times": [vp, C.c_int, vp, dp, i64p, C.POINTER(C.c_int)],
        "ohpic_memory_footprint": [vp, i64p],
        "ohpic_estimate_footprint": [C.POINTER(Config), i64p],
        "ohpic_boris_batch": [C.c_int, i64, dp, dp, dp, dp, dp, dp, dp, dbl, dbl],
        "ohpic_interp_batch": [C.c_int, i64, dp, dp, C.c_int, i32p, i32p, C.c_int, dp],
        "ohpic_scatter_batch": [C.c_int, i64, dp, dp, C.c_int, i32p, i32p, dp, C.c_int],
        "ohpic_wrap_batch": [C.c_int, i64, dp, C.c_int32, C.c_int],
        "ohpic_group_create": [C.c_int, C.POINTER(vp)],
        "ohpic_group_destroy": [vp],
        "ohpic_group_attach": [vp, vp],
        "ohpic_group_detach": [vp, vp],
        "ohpic_group_size": [vp, C.POINTER(C.c_int)],
        "ohpic_group_contains": [vp, vp, C.POINTER(C.c_int)],
        "ohpic_group_step": [vp, C.c_int],
        "ohpic_group_use_graph": [vp, C.c_int],
        "ohpic_group_use_persistent": [vp, C.c_int],
        "ohpic_group_set_batch": [vp, C.c_int],
        "ohpic_group_stats": [vp, i64p, i64p, i64p],
        "ohpic_group_sequences": [vp, i64p, i64p],
        "ohpic_autotune": [C.POINTER(Config), C.c_int, C.c_int, C.c_int, C.POINTER(TuneResult)],
        "ohpic_snapshot_open": [C.c_char_p, C.c_char_p, i64p, i64p, C.POINTER(vp)],
        "ohpic_snapshot_write": [vp, C.c_char_p, vp, C.c_int],
        "ohpic_snapshot_write_host": [vp, C.c_char_p, dp, i64p, i64p],
        "ohpic_snapshot_close": [vp],
        "ohpic_snapshot_read": [C.c_char_p, C.c_char_p, C.c_char_p, dp, i64, i64p],
    }
    for name in EXPORTS:
        f = getattr(L, name)
        if name in ("ohpic_last_error", "ohpic_group_last_error", "ohpic_snapshot_last_error"):
            f.argtypes = [vp]
            f.restype = C.c_char_p
            continue
        f.argtypes = sig[name]
        f.restype = C.c_int
    _lib = L
    return L


def check(rc: int, ctx=None) -> None:
    if rc != OK:
        msg = lib().ohpic_last_error(ctx)
        raise OhpicError(rc, msg.decode() if msg else "")


def check_group(rc: int, group) -> None:
    if rc != OK:
        msg = lib().ohpic_group_last_error(group)
        raise OhpicError(rc, msg.decode() if msg else "")


def check_snapshot(rc: int, snap) -> None:
    if rc != OK:
        msg = lib().ohpic_snapshot_last_error(snap)
        raise OhpicError(rc, msg.decode() if msg else "")


def dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def i64ptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def i32ptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))
