"""The Fortran mirror driving the GPU: a mainstep through polymorphic
t_Scatter / t_FieldSolver / t_ParticleMover / t_ParticleBoundary objects,
the per-particle bindings, the FFT and the fused step, all against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import fortran_io as fio
from helpers import rel_l2, two_species

pytestmark = pytest.mark.gpu


def _have_hdf5():
    from ohhelppic3d_amd import _lib
    probe = np.zeros(3, np.int64)
    return _lib.lib().ohpic_snapshot_read(b"/nonexistent.h5", b"g", b"d", None, 0,
                                          _lib.i64ptr(probe)) != _lib.ENOLIB


def test_fortran_mainstep_and_bindings_match_oracle(orc, tmp_path):
    if not os.path.exists(fio.EXE):
        pytest.fail("ohpic_fortran_check not built (make -C ohhelppic3d_amd/fortran)")
    n = (16, 12, 8)
    nsteps, k = 2, 48
    dt, b0 = 0.1, (0.0, 0.0, 0.5)
    sp = two_species(orc, n, 4)
    recs = fio.write_input(str(tmp_path / "in.bin"), n, sp, dt, b0, nsteps, k)
    h5 = str(tmp_path / "phisp00_0000.h5") if _have_hdf5() else None
    r = subprocess.run([fio.EXE, str(tmp_path / "in.bin"), str(tmp_path / "out.bin")] +
                       ([h5] if h5 else []), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    out = fio.read_output(str(tmp_path / "out.bin"), n, k)

    # part 1: mainstep through the interfaces == oracle mainstep
    osim = orc.OracleSim(n, sp, dt, b0, 0)
    osim.step(nsteps)
    parts = np.sort(out["parts"], order="pid")
    ref = np.concatenate([np.stack([osim.sp[s][c] for c in ("x", "y", "z", "vx", "vy", "vz")], 1)
                          for s in range(2)])
    got = np.stack([parts[c] for c in ("x", "y", "z", "vx", "vy", "vz")], 1)
    assert len(parts) == len(ref)
    for c in range(6):
        assert rel_l2(got[:, c], ref[:, c]) <= 1e-12
    rho = out["rho"][: n[2], : n[1], : n[0]].reshape(-1)
    assert rel_l2(rho, osim.rho) <= 1e-12
    phi = out["phi"][1:-2, 1:-2, 1:-2].reshape(-1)
    assert rel_l2(phi, osim.phi) <= 1e-12
    ex = out["eb"][1:-2, 1:-2, 1:-2, 0].reshape(-1)
    assert rel_l2(ex, osim.ex) <= 1e-12
    assert np.all(out["eb"][..., 5] == 0.5)
    # halo nodes are periodic images
    assert np.array_equal(out["phi"][0], out["phi"][n[2]])

    # part 4: fused step == mainstep through the separate operators (bitwise)
    fused = np.sort(out["parts_fused"], order="pid")
    for c in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.array_equal(fused[c], parts[c])

    # part 2: per-particle bindings == oracle per-particle functions
    eb_f = np.transpose(out["eb"], (3, 2, 1, 0))  # values(6, x, y, z)
    lb, ub = [-1, -1, -1], [n[0] + 1, n[1] + 1, n[2] + 1]
    for i in range(k):
        p = recs[i]
        pos = [p["x"], p["y"], p["z"]]
        e_ref = orc.interp(eb_f, lb, ub, pos, 0)
        assert np.allclose(out["ebi"][i], e_ref, rtol=1e-13, atol=1e-15)
        qm = sp[p["spec"] - 1]["qm"]
        x1, v1 = orc.boris_move(pos, [p["vx"], p["vy"], p["vz"]], qm, e_ref, dt)
        for _ in range(2):
            x1 = [orc.apply_periodic(m, v) for m, v in zip(n, x1)]
        pr = out["probe"][i]
        assert np.allclose([pr["x"], pr["y"], pr["z"]], x1, rtol=1e-13, atol=1e-14)
        assert np.allclose([pr["vx"], pr["vy"], pr["vz"]], v1, rtol=1e-13, atol=1e-15)
    rho2 = np.zeros((1, n[0] + 1, n[1] + 1, n[2] + 1))
    for i in range(k):
        p = recs[i]
        orc.lib()  # the oracle's per-particle scatter on the same halo layout
        flat = np.ascontiguousarray(rho2.reshape(-1, order="F"))
        orc.lib().orc_scatter(flat, 1, np.array([0, 0, 0], np.int32),
                              np.array([n[0], n[1], n[2]], np.int32),
                              np.array([p["x"], p["y"], p["z"]]),
                              np.array([sp[p["spec"] - 1]["q"]]), 0)
        rho2 = flat.reshape(rho2.shape, order="F")
    got2 = np.transpose(out["rho2"], (2, 1, 0))
    assert np.allclose(got2, rho2[0], rtol=0, atol=1e-14)

    # part 3: the reference's FFT impulse KAT through t_MPIFFTSolver3d%forward
    kk = np.arange(63)
    exp = np.where(kk <= 31, np.cos(kk / 63 * 2 * np.pi), np.sin(kk / 63 * 2 * np.pi))
    assert np.sum(np.abs(out["spectrum"] - exp)) <= 1e-8

    # part 5: t_Hdf5ForOhfieldHip%write of the host phi (nodes 0..n) and
    # %write_device of the fused run's phi, in the reference's phisp layout
    if h5:
        from ohhelppic3d_amd import snapshot
        want = np.transpose(out["phi"][1:n[2] + 2, 1:n[1] + 2, 1:n[0] + 2], (2, 1, 0))
        assert np.array_equal(snapshot.read_dataset(h5, "phisp", "1"), want)
        assert rel_l2(snapshot.read_dataset(h5, "phisp", "2"), want) <= 1e-12
