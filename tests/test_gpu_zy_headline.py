"""The headline configuration (256^3 cells x 64 particles per cell per species,
2^31 particles) stepped on the GPU and checked through properties that hold at
any size: every particle is still binned, the plasma stays neutral, and the
kinetic energy is that of the loaded Maxwellians.  The FP64 context takes
about 238 GB of device memory, so this file is collected after the others."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ohp():
    import ohhelppic3d_amd as o
    if not o.available():
        pytest.fail("libohpic.so is not built: the GPU path has no fallback")
    return o


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_headline_configuration_step_properties(ohp, precision):
    n = (256, 256, 256)
    ppc = 64
    npcl = ppc * n[0] * n[1] * n[2]
    vth = (1.0, 0.1)
    qm = (-1.0, 0.01)
    q = (-1.0 / ppc, 1.0 / ppc)
    prec = ohp.FP64 if precision == "fp64" else ohp.MIXED
    cfg = ohp.SimConfig(n=n, qm=list(qm), q_macro=list(q), capacity=[npcl, npcl], dt=0.1,
                        b0=(0.0, 0.0, 0.5), precision=prec)
    with ohp.Simulation(cfg) as sim:
        sim.seed_uniform_maxwell([npcl, npcl], list(vth), seed=(42, 52))
        sim.step(2)
        d = sim.diagnostics()
        counts = sim.cell_counts()
    assert int(counts.sum(dtype=np.int64)) == 2 * npcl
    assert d["repacks"] == 0
    # 1/2 m <|v|^2> N with m = q / (q/m) and <|v|^2> = 3 v_th^2 per species
    ke_expected = sum(0.5 * (q[s] / qm[s]) * 3.0 * vth[s] ** 2 * npcl for s in range(2))
    ke = 0.5 * (np.sum(d["ke_old"]) + np.sum(d["ke_new"]))
    assert abs(ke / ke_expected - 1.0) < 1e-2
    # equal numbers of opposite charges: the sum of rho is zero up to the
    # rounding of the fixed-point deposit
    assert abs(d["net_charge"]) <= 1e-9 * 2 * npcl / ppc
    assert np.isfinite(d["total_energy"]) and d["field_energy"] > 0
