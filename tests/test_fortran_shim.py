"""The Fortran mirror of the reference's interfaces (flang, iso_c_binding):
it builds, and without a GPU it stops with the library's error message."""
import os
import subprocess

import numpy as np
import pytest

import fortran_io as fio


def test_fortran_shim_builds_and_reports_missing_gpu(tmp_path):
    if not os.path.exists("/opt/rocm/lib/llvm/bin/flang"):
        pytest.skip("flang not available")
    from ohhelppic3d_amd import _lib
    if not _lib.available():
        from ohhelppic3d_amd import build
        build.build()
    fio.build()
    assert os.path.exists(fio.EXE)
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present: the GPU test covers the shim")
    sp = [dict(qm=-1.0, q=-0.5, x=np.array([1.5]), y=np.array([0.5]), z=np.array([2.5]),
               vx=np.zeros(1), vy=np.zeros(1), vz=np.zeros(1))]
    inp = tmp_path / "in.bin"
    fio.write_input(str(inp), (4, 4, 4), sp, 0.1, (0, 0, 0.5), 1, 1)
    r = subprocess.run([fio.EXE, str(inp), str(tmp_path / "out.bin")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "ohpic error -6" in r.stdout and "no HIP device" in r.stdout
