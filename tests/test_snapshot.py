"""CPU checks of the snapshot files (ohpic_snapshot_*, the reference's
t_Hdf5ForOhfield layout): host blocks written and read back, the file's
dimension order and values as HDF5's own h5dump reads them, and the error
paths.  No GPU is needed for host blocks."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump"
                                    if os.path.exists("/opt/conda/bin/h5dump") else None)


@pytest.fixture(scope="module")
def snap():
    from ohhelppic3d_amd import _lib, snapshot
    if not _lib.available():
        from ohhelppic3d_amd import build
        build.build()
    probe = np.zeros(3, np.int64)
    rc = _lib.lib().ohpic_snapshot_read(b"/nonexistent.h5", b"g", b"d", None, 0,
                                        _lib.i64ptr(probe))
    if rc == _lib.ENOLIB:
        pytest.skip("HDF5 library not available")
    return snapshot


def test_block_round_trip_and_fill(snap, tmp_path):
    f = tmp_path / "phisp00_0000.h5"
    rng = np.random.default_rng(1)
    full = rng.standard_normal((9, 7, 5))          # [ix, iy, iz], global shape
    part = rng.standard_normal((3, 4, 2))
    with snap.Hdf5ForOhfield(f, "phisp", full.shape) as out:
        out.write_block("1", full)
        out.write_block("2", part, offset=(5, 1, 3))
    assert snap.dataset_shape(f, "phisp", "1") == (9, 7, 5)
    assert np.array_equal(snap.read_dataset(f, "phisp", "1"), full)
    got = snap.read_dataset(f, "phisp", "2")
    want = np.zeros(full.shape)
    want[5:8, 1:5, 3:5] = part
    assert np.array_equal(got, want)


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not installed")
def test_layout_as_h5dump_reads_it(snap, tmp_path):
    # the reference hands Fortran-order shapes to the Fortran HDF5 API, so a
    # C reader sees dimensions (z, y, x); value (ix, iy, iz) sits at [iz][iy][ix]
    f = tmp_path / "layout.h5"
    g = np.arange(4 * 3 * 2, dtype=np.float64).reshape((4, 3, 2), order="F") * 0.5
    with snap.Hdf5ForOhfield(f, "phisp", g.shape) as out:
        out.write_block("7", g)
    head = subprocess.run([H5DUMP, "-H", str(f)], capture_output=True, text=True,
                          check=True).stdout
    assert 'GROUP "phisp"' in head and 'DATASET "7"' in head
    assert "H5T_IEEE_F64LE" in head
    assert "SIMPLE { ( 2, 3, 4 ) / ( 2, 3, 4 ) }" in head
    for ix, iy, iz in ((0, 0, 0), (3, 0, 0), (1, 2, 1), (3, 2, 1)):
        out = subprocess.run([H5DUMP, "-d", "/phisp/7", "-s", f"{iz},{iy},{ix}",
                              "-c", "1,1,1", "-y", "-w", "0", str(f)],
                             capture_output=True, text=True, check=True).stdout
        body = out[out.index("DATA {") + 6: out.index("}", out.index("DATA {"))]
        assert float(body.strip().split()[-1].rstrip(",")) == g[ix, iy, iz]


def test_snapshot_errors(snap, tmp_path):
    from ohhelppic3d_amd import _lib as L
    f = tmp_path / "e.h5"
    out = snap.Hdf5ForOhfield(f, "phisp", (4, 4, 4))
    out.write_block("1", np.ones((4, 4, 4)))
    with pytest.raises(L.OhpicError) as e:
        out.write_block("1", np.ones((4, 4, 4)))       # dataset exists
    assert e.value.code == L.EIO and "'1'" in str(e.value)
    with pytest.raises(L.OhpicError) as e:
        out.write_block("2", np.ones((3, 3, 3)), offset=(2, 0, 0))   # outside
    assert e.value.code == L.EINVAL
    with pytest.raises(L.OhpicError) as e:
        out.write_block("", np.ones((4, 4, 4)))
    assert e.value.code == L.EINVAL
    with pytest.raises(ValueError):
        out.write_block("3", np.ones(16))
    out.close()
    out.close()                                         # idempotent
    with pytest.raises(L.OhpicError) as e:
        out.write_block("4", np.ones((4, 4, 4)))
    assert e.value.code == L.ESTATE
    with pytest.raises(L.OhpicError) as e:
        snap.Hdf5ForOhfield(f, "phisp", (0, 4, 4))
    assert e.value.code == L.EINVAL
    with pytest.raises(L.OhpicError) as e:
        snap.Hdf5ForOhfield(tmp_path / "no_dir" / "x.h5", "phisp", (4, 4, 4))
    assert e.value.code == L.EIO
    with pytest.raises(L.OhpicError) as e:
        snap.read_dataset(f, "phisp", "9")
    assert e.value.code == L.EIO
    with pytest.raises(L.OhpicError) as e:
        snap.read_dataset(tmp_path / "missing.h5", "phisp", "1")
    assert e.value.code == L.EIO
    # the file holds what was written before the errors
    assert np.array_equal(snap.read_dataset(f, "phisp", "1"), np.ones((4, 4, 4)))


def test_missing_hdf5_is_reported(snap, tmp_path):
    # OHPIC_HDF5_LIB names a library that is not HDF5: open fails with ENOLIB
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from ohhelppic3d_amd import snapshot, _lib as L\n"
            "try:\n"
            "    snapshot.Hdf5ForOhfield(%r, 'phisp', (2, 2, 2))\n"
            "except L.OhpicError as e:\n"
            "    print(e.code, e)\n") % (ROOT, str(tmp_path / "x.h5"))
    env = dict(os.environ, OHPIC_HDF5_LIB="libm.so.6")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                         env=env, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.startswith("-8 ") and "lacks H5open" in out.stdout
