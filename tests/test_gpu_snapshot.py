"""GPU checks of the snapshot output (ohpic_snapshot_write): device fields
written in the reference's phisp layout read back bit for bit, with the
periodic node n equal to node 0, and the per-step phi datasets of a run from
the reference's parameters.toml equal the oracle's phi (1e-12 relative L2
after one step, 1e-11 after three)."""
import os

import numpy as np
import pytest

from helpers import rel_l2, two_species

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def ohp():
    import ohhelppic3d_amd as o
    from ohhelppic3d_amd import _lib
    if not o.available():
        pytest.fail("libohpic.so is not built: the GPU path has no fallback")
    probe = np.zeros(3, np.int64)
    if _lib.lib().ohpic_snapshot_read(b"/nonexistent.h5", b"g", b"d", None, 0,
                                      _lib.i64ptr(probe)) == _lib.ENOLIB:
        pytest.skip("HDF5 library not available")
    return o


def periodic_nodes(flat, n):
    """(n+1)^3 block indexed [ix, iy, iz] of a periodic grid given x fastest."""
    g = np.asarray(flat).reshape(n, order="F")
    return np.pad(g, [(0, 1)] * 3, mode="wrap")


def test_fields_round_trip_with_periodic_node(ohp, orc, tmp_path):
    from ohhelppic3d_amd import snapshot
    n = (16, 12, 8)
    cfg = ohp.SimConfig(n=n, qm=(-1.0, 0.01), q_macro=(-0.25, 0.25), capacity=(4 * 1536,) * 2,
                        dt=0.1, b0=(0.0, 0.0, 0.5))
    f = tmp_path / "fields.h5"
    with ohp.Simulation(cfg) as sim:
        sim.seed_uniform_maxwell([4 * 1536] * 2, [1.0, 0.1], seed=(42, 52))
        sim.step(1)
        which = {"rho": ohp.FIELD_RHO, "phi": ohp.FIELD_PHI, "ex": ohp.FIELD_EX,
                 "ey": ohp.FIELD_EY, "ez": ohp.FIELD_EZ}
        with snapshot.Hdf5ForOhfield(f, "fields", (17, 13, 9)) as out:
            for name, w in which.items():
                out.write(name, sim, w)
            with pytest.raises(ohp.OhpicError):
                out.write("eb", sim, ohp.FIELD_EB)   # one component per dataset
        want = {k: periodic_nodes(sim.grid(w), n) for k, w in which.items()}
    for k, w in want.items():
        got = snapshot.read_dataset(f, "fields", k)
        assert got.shape == (17, 13, 9)
        assert np.array_equal(got, w), k
        assert np.array_equal(got[16], got[0]) and np.array_equal(got[:, :, 8], got[:, :, 0])


def test_phisp_run_from_parameters_matches_oracle(ohp, orc, tmp_path):
    # the reference's run: config-0 parameters, phi written after every step
    # to phisp00_0000.h5 / phisp / '1', '2', ...
    from ohhelppic3d_amd import params, snapshot
    spec = params.load_parameters(os.path.join(os.path.dirname(__file__), "golden",
                                               "parameters_config0.toml"))
    n = spec.cfg.n
    osim = orc.OracleSim(n, two_species(orc, n, 16), spec.cfg.dt, spec.cfg.b0, 0)
    f = tmp_path / "phisp00_0000.h5"
    with spec.create() as sim:
        assert snapshot.run_with_snapshots(sim, 3, f) == 3
        assert sim.diagnostics()["step"] == 3
    for k in (1, 2, 3):
        osim.step(1)
        got = snapshot.read_dataset(f, "phisp", str(k))
        assert got.shape == (n[0] + 1, n[1] + 1, n[2] + 1)
        # one step from identical state: 1e-12; later steps start from
        # particles that already differ by rounding
        assert rel_l2(got, periodic_nodes(osim.phi, n)) <= (TOL if k == 1 else 1e-11), k
