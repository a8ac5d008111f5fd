"""Snapshot files in the reference's HDF5 layout.

Mirrors ``t_Hdf5ForOhfield`` (src/hdf5/hdf5_for_ohfield.f90:11-89) over the
C ABI's ``ohpic_snapshot_*``: one file, one group, one FP64 dataset per write
of the global shape ``(nx+1, ny+1, nz+1)`` (Fortran order, x first), holding
the nodes 0..n of each axis with node n equal to node 0 (periodic).  The
reference's run writes phi after every step to ``phisp00_0000.h5``, group
``phisp``, datasets ``'1'``, ``'2'``, ... (ohhelppic3d.f90:79-99);
``run_with_snapshots`` is that loop.

Arrays passed in and returned here are indexed ``[ix, iy, iz]`` (the
reference's ``values(1, ix, iy, iz)`` without the component axis).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Callable, Optional, Sequence

import numpy as np

from . import _lib as L
from .pic import Simulation


class Hdf5ForOhfield:
    """new_Hdf5ForOhfield(filename, group_name, global_shape, global_offset)
    (hdf5_for_ohfield.f90:27-43): creates (truncates) the file and the group."""

    def __init__(self, filename: str, group_name: str, global_shape: Sequence[int],
                 global_offset: Sequence[int] = (0, 0, 0)):
        self._lib = L.lib()
        self.filename = os.fspath(filename)
        self.group_name = group_name
        self.global_shape = tuple(int(v) for v in global_shape)
        self.global_offset = tuple(int(v) for v in global_offset)
        gs = np.asarray(self.global_shape, np.int64)
        go = np.asarray(self.global_offset, np.int64)
        h = C.c_void_p()
        rc = self._lib.ohpic_snapshot_open(self.filename.encode(), group_name.encode(),
                                           L.i64ptr(gs), L.i64ptr(go), C.byref(h))
        L.check_snapshot(rc, None)
        self._h = h

    def write(self, dataset_name: str, sim: Simulation, which: int = L.FIELD_PHI) -> None:
        """hdf5ForOhfield_write(dataset_name, ohfield, ps) (:45-76): field
        ``which`` of the simulation (nodes 0..n per axis) as a new dataset."""
        self._open()
        rc = self._lib.ohpic_snapshot_write(self._h, str(dataset_name).encode(), sim._ctx, which)
        L.check_snapshot(rc, self._h)

    def write_block(self, dataset_name: str, values: np.ndarray,
                    offset: Sequence[int] = (0, 0, 0)) -> None:
        """write_dataset3d_double (hdf5.f90:140-185): a 3-D block indexed
        [ix, iy, iz] written at ``offset`` of a new dataset of the global shape."""
        self._open()
        v = np.asarray(values, np.float64)
        if v.ndim != 3:
            raise ValueError("write_block: values must be 3-D, indexed [ix, iy, iz]")
        flat = np.ascontiguousarray(v.reshape(-1, order="F"))
        ls = np.asarray(v.shape, np.int64)
        lo = np.asarray([int(o) for o in offset], np.int64)
        rc = self._lib.ohpic_snapshot_write_host(self._h, str(dataset_name).encode(),
                                                 L.dptr(flat), L.i64ptr(ls), L.i64ptr(lo))
        L.check_snapshot(rc, self._h)

    def close(self) -> None:
        """hdf5ForOhfield_close (:78-86)."""
        h, self._h = getattr(self, "_h", None), None
        if h:
            L.check_snapshot(self._lib.ohpic_snapshot_close(h), None)

    def _open(self):
        if not self._h:
            raise L.OhpicError(L.ESTATE, "snapshot file is closed")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dataset_shape(filename: str, group: str, dataset: str) -> tuple:
    """(nx, ny, nz) of a dataset, Fortran order."""
    shape = np.zeros(3, np.int64)
    rc = L.lib().ohpic_snapshot_read(os.fspath(filename).encode(), group.encode(),
                                     str(dataset).encode(), None, 0, L.i64ptr(shape))
    L.check_snapshot(rc, None)
    return tuple(int(v) for v in shape)


def read_dataset(filename: str, group: str, dataset: str) -> np.ndarray:
    """A dataset of a snapshot file as an array indexed [ix, iy, iz]."""
    shape = dataset_shape(filename, group, dataset)
    out = np.empty(int(np.prod(shape)), np.float64)
    sh = np.zeros(3, np.int64)
    rc = L.lib().ohpic_snapshot_read(os.fspath(filename).encode(), group.encode(),
                                     str(dataset).encode(), L.dptr(out), out.size, L.i64ptr(sh))
    L.check_snapshot(rc, None)
    return out.reshape(shape, order="F")


def run_with_snapshots(sim: Simulation, nstep: int, filename: str = "phisp00_0000.h5",
                       group: str = "phisp", every: int = 1, which: int = L.FIELD_PHI,
                       progress: Optional[Callable[[int], None]] = None) -> int:
    """The reference's time loop with its phi output (ohhelppic3d.f90:79-99):
    step, then write field ``which`` as dataset ``str(istep)``, for istep =
    1..nstep (every ``every`` steps).  Returns the number of datasets written."""
    if every < 1:
        raise ValueError("every must be >= 1")
    n = sim.n
    written = 0
    with Hdf5ForOhfield(filename, group, (n[0] + 1, n[1] + 1, n[2] + 1)) as out:
        for istep in range(1, nstep + 1):
            sim.step(1)
            if istep % every == 0:
                out.write(str(istep), sim, which)
                written += 1
            if progress is not None:
                progress(istep)
    return written
