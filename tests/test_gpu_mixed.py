"""GPU checks of the mixed-precision form (tile-relative cell + 24-bit fixed-point
in-cell offset per axis, FP32 v, FP32 push/gather/deposit, FP64 charge grid,
FFT, E and diagnostics) against
the FP64 oracle.  Bars: rho and E within 1e-5 relative L2 of FP64 after a step
from the same state with the standard CIC weights (with the reference's
mirrored weights the bar holds for the loaded state, and after a push the
excess is pinned to the particles that cross a face); over 1000 steps the
total energy within 0.5% of the FP64 oracle's and the drift the mixed form
adds below 0.25%."""
import numpy as np
import pytest

from helpers import by_pid, energy_histories, rel_l2, sim_for, two_species

pytestmark = pytest.mark.gpu

TOL_MIXED = 1e-5


@pytest.fixture(scope="module")
def ohp():
    import ohhelppic3d_amd as o
    if not o.available():
        pytest.fail("libohpic.so is not built: the GPU path has no fallback")
    return o


def _mixed_and_oracle(ohp, orc, wo, steps, n=(32, 32, 32), ppc=16):
    """The mixed form and the FP64 oracle from the same loaded state after
    `steps` steps (step 1 solves the loaded state; each later step adds one
    push): fields, particles ordered by id, per-cell counts."""
    sp = two_species(orc, n, ppc)
    osim = orc.OracleSim(n, sp, 0.1, (0, 0, 0.5), wo)
    osim.step(steps)
    with sim_for(sp, n, precision=ohp.MIXED, weight_order=wo) as sim:
        sim.step(steps)
        got = {"rho": sim.grid(ohp.FIELD_RHO), "phi": sim.grid(ohp.FIELD_PHI),
               "Ex": sim.grid(ohp.FIELD_EX), "Ey": sim.grid(ohp.FIELD_EY),
               "Ez": sim.grid(ohp.FIELD_EZ),
               "sp": [by_pid(sim.store_particles(s, True)) for s in range(2)],
               "counts": sim.cell_counts()}
    return sp, osim, got


def _field_errors(got, osim):
    return {"rho": rel_l2(got["rho"], osim.rho), "phi": rel_l2(got["phi"], osim.phi),
            "Ex": rel_l2(got["Ex"], osim.ex), "Ey": rel_l2(got["Ey"], osim.ey),
            "Ez": rel_l2(got["Ez"], osim.ez)}


@pytest.mark.parametrize("steps", [1, 2, 3])
def test_mixed_vs_fp64_oracle_standard_weights(ohp, orc, steps):
    # step 1: the loaded state's fields (positions converted to the 24-bit
    # offsets); steps 2 and 3: fields deposited after one and two FP32 pushes.
    # Standard CIC weights: rho, phi, E and every particle within 1e-5.
    n = (32, 32, 32)
    _, osim, got = _mixed_and_oracle(ohp, orc, 1, steps, n)
    err = _field_errors(got, osim)
    for s in range(2):
        for k in ("x", "y", "z", "vx", "vy", "vz"):
            err[f"{k}{s}"] = rel_l2(got["sp"][s][k], osim.sp[s][k])
    print(f"mixed vs FP64 after step {steps} (standard weights):",
          {k: f"{v:.2e}" for k, v in err.items()})
    for k, v in err.items():
        assert v <= TOL_MIXED, k
    ref_counts = sum(orc.cell_counts(n, r["x"], r["y"], r["z"]) for r in osim.sp)
    # cell indices agree except for particles within float rounding of a face
    assert np.sum(got["counts"] != ref_counts) <= 16


def _with_face_grazers(orc, n, ppc, m=1024, seed=5):
    """The two-species plasma plus a third, ballistic species (q/m 1e-12, so
    the push leaves v unchanged to ~1e-14 and x^{n+1} = x + v dt) whose m
    particles land, in FP64, within 1e-10..3e-8 of a cell face along one
    axis: a 32-bit position ends on either side of that face."""
    sp = two_species(orc, n, ppc)
    rng = np.random.default_rng(seed)
    pos = [rng.uniform(0, n[a], m) for a in range(3)]
    vel = [rng.uniform(-2.0, 2.0, m) for _ in range(3)]
    axis = rng.integers(0, 3, m)
    eps = rng.uniform(1e-10, 3e-8, m) * rng.choice([-1.0, 1.0], m)
    for i in range(m):
        a = axis[i]
        face = np.floor(pos[a][i]) + (1.0 if rng.random() < 0.5 else 0.0)
        vel[a][i] = (face + eps[i] - pos[a][i]) / 0.1
    sp.append(dict(qm=1e-12, q=1.0 / ppc, x=pos[0], y=pos[1], z=pos[2], vx=vel[0], vy=vel[1],
                   vz=vel[2]))
    return sp


@pytest.mark.parametrize("wo", [0, 1])
def test_mixed_weights_error_comes_from_face_crossers(ohp, orc, wo):
    """The deposit after one push, mixed against FP64, with particles that
    land near cell faces.  With the reference's mirrored weights
    (linear_scatter.f90:41-52) the deposit jumps when a particle crosses a
    face: (a) the particles whose mixed cell differs from the oracle's all
    lie within the largest position difference of a face; (b) the charge of
    the particles whose cells agree is within 1e-5 of the oracle's; (c) the
    whole rho error exceeds 1e-5 and is bounded by that part, the deposit's
    own rounding and 2|q| per face-crosser.  With the standard weights the
    same particles give rho within 1e-5: the deposit is continuous at a
    face."""
    n = (32, 32, 32)
    ppc = 16
    sp = _with_face_grazers(orc, n, ppc)
    ns = len(sp)
    osim = orc.OracleSim(n, sp, 0.1, (0, 0, 0.5), wo)
    osim.step(1)
    with sim_for(sp, n, precision=ohp.MIXED, weight_order=wo) as sim:
        sim.step(1)
        got = {"sp": [by_pid(sim.store_particles(s, True)) for s in range(ns)]}
        # rho^1: the deposit of the positions after the push (the start of
        # the next step)
        sim.deposit()
        got["rho"] = sim.grid(ohp.FIELD_RHO)
    rho_o = np.zeros(n[0] * n[1] * n[2])
    for s in range(ns):
        o = osim.sp[s]
        orc.deposit_periodic(n, o["x"], o["y"], o["z"], sp[s]["q"], wo, rho_o)
    rho_m = np.zeros_like(rho_o)
    same_m = np.zeros_like(rho_o)
    same_o = np.zeros_like(rho_o)
    ncross = 0
    for s in range(ns):
        g, o, q = got["sp"][s], osim.sp[s], sp[s]["q"]
        orc.deposit_periodic(n, g["x"], g["y"], g["z"], q, wo, rho_m)
        diff = np.zeros(len(g["x"]), bool)
        dmax = 0.0
        for a, k in enumerate("xyz"):
            diff |= np.floor(g[k]) != np.floor(o[k])
            d = np.abs(g[k] - o[k])
            dmax = max(dmax, float(np.max(np.minimum(d, n[a] - d))))
        assert dmax <= 1e-6
        # (a) every crosser sits within dmax of a face
        near = np.zeros_like(diff)
        for k in ("x", "y", "z"):
            f = o[k] - np.floor(o[k])
            near |= np.minimum(f, 1.0 - f) <= dmax
        assert np.all(near[diff]), "a particle changed cell away from any face"
        ncross += int(diff.sum())
        keep = ~diff
        orc.deposit_periodic(n, g["x"][keep], g["y"][keep], g["z"][keep], q, wo, same_m)
        orc.deposit_periodic(n, o["x"][keep], o["y"][keep], o["z"][keep], q, wo, same_o)
    nrm = np.linalg.norm(rho_o)
    # the GPU's FP32 deposit of its own positions
    err_dep = np.linalg.norm(got["rho"] - rho_m) / nrm
    # (b) particles whose cells agree
    err_same = np.linalg.norm(same_m - same_o) / nrm
    err = rel_l2(got["rho"], rho_o)
    q = abs(sp[0]["q"])
    bound = err_same + err_dep + 2 * q * ncross / nrm
    print(f"weights {wo}, after one push: rho rel L2 {err:.2e}, agreeing cells "
          f"{err_same:.2e}, own deposit {err_dep:.2e}, {ncross} face-crossers, "
          f"bound {bound:.2e}")
    assert err_dep <= 1e-6
    assert err_same <= TOL_MIXED
    # the grazers make crossers certain: about half of them end on the other side
    assert ncross >= 100
    if wo == 0:
        assert TOL_MIXED < err <= bound
    else:
        assert err <= TOL_MIXED


def test_mixed_loaded_state_vs_fp64_oracle_reference_weights(ohp, orc):
    # the fields of the loaded state (before any push) with the mirrored
    # weights: converting the positions moves none across a face
    n = (32, 32, 32)
    _, osim, got = _mixed_and_oracle(ohp, orc, 0, 1, n)
    err = _field_errors(got, osim)
    for v in err.values():
        assert v <= TOL_MIXED
    ref_counts = sum(orc.cell_counts(n, r["x"], r["y"], r["z"]) for r in osim.sp)
    assert np.sum(got["counts"] != ref_counts) <= 16


def test_mixed_energy_drift_and_history(ohp, orc):
    # 1000 steps of the mixed form against the FP64 oracle: total energy
    # within 0.5% at every step, and the drift the mixed form adds to the
    # algorithm's own (FP64) drift below 0.25%
    tot_m, tot_o, fe_m, fe_o = energy_histories(ohp, orc, ohp.MIXED, 1000)
    drift_m = np.max(np.abs(tot_m - tot_m[0])) / tot_m[0]
    drift_o = np.max(np.abs(tot_o - tot_o[0])) / tot_o[0]
    print(f"energy drift over 1000 steps: mixed {drift_m:.3e}, FP64 oracle {drift_o:.3e}; "
          f"max |mixed - FP64| / FP64 = {np.max(np.abs(tot_m - tot_o) / tot_o):.2e}")
    assert np.max(np.abs(tot_m - tot_o) / tot_o) <= 0.005
    assert abs(drift_m - drift_o) <= 0.0025
    assert abs(fe_m[500:].mean() - fe_o[500:].mean()) <= 0.01 * fe_o[500:].mean()


def test_mixed_separate_operators_equal_fused(ohp, orc):
    n = (16, 16, 16)
    sp = two_species(orc, n, 8)
    with sim_for(sp, n, precision=ohp.MIXED) as a, sim_for(sp, n, precision=ohp.MIXED) as b:
        parts = (ohp.LinearScatter(), ohp.PoissonFieldSolver(),
                 ohp.create_particle_mover("boris"), ohp.ParticleBoundaries(*n))
        for _ in range(3):
            ohp.mainstep(a, *parts)
        b.step(3)
        assert np.array_equal(a.grid(ohp.FIELD_PHI), b.grid(ohp.FIELD_PHI))
        for s in range(2):
            pa, pb = by_pid(a.store_particles(s, True)), by_pid(b.store_particles(s, True))
            for k in pa:
                assert np.array_equal(pa[k], pb[k]), k


def test_mixed_is_deterministic(ohp, orc):
    n = (16, 16, 16)
    sp = two_species(orc, n, 8)
    out = []
    for chunks in (0, 3):
        with sim_for(sp, n, precision=ohp.MIXED, chunks=chunks) as sim:
            sim.step(6)
            out.append((sim.grid(ohp.FIELD_RHO), sim.grid(ohp.FIELD_PHI),
                        [by_pid(sim.store_particles(s, True)) for s in range(2)]))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    for s in range(2):
        for k in out[0][2][s]:
            assert np.array_equal(out[0][2][s][k], out[1][2][s][k])


def test_mixed_loader_and_roundtrip(ohp, orc):
    n = (16, 8, 8)
    npcl = [4096, 2048]
    cfg = ohp.SimConfig(n=n, qm=[-1.0, 0.01], q_macro=[-0.1, 0.1], capacity=npcl,
                        track_ids=True, precision=ohp.MIXED)
    with ohp.Simulation(cfg) as sim:
        sim.seed_uniform_maxwell(npcl, [1.0, 0.1], seed=(42, 52))
        g = [by_pid(sim.store_particles(s, True)) for s in range(2)]
    rng = orc.Pcg32(42, 52)
    for s, vt in enumerate((1.0, 0.1)):
        x, y, z, vx, vy, vz = orc.load_uniform_maxwell(rng, npcl[s], n, vt)
        assert np.max(np.abs(g[s]["x"] - x)) <= 2 ** -24
        assert np.array_equal(np.floor(g[s]["z"]), np.floor(z)) or \
            np.sum(np.floor(g[s]["z"]) != np.floor(z)) <= 2
        assert rel_l2(g[s]["vx"], vx) <= 1e-7


def test_mixed_wrap_and_repack(ohp, orc):
    n = (16, 8, 8)
    rng = np.random.default_rng(9)
    m = 2000
    x = rng.uniform(7.0, 8.0, m)
    y = rng.uniform(0, 8, m)
    z = rng.uniform(0, 8, m)
    vx = np.full(m, -80.25)  # crosses the periodic boundary, lands in another tile
    zero = np.zeros(m)
    sp = [dict(qm=-1.0, q=-1e-3, x=x, y=y, z=z, vx=vx, vy=zero, vz=zero),
          dict(qm=0.01, q=1e-3, x=x.copy(), y=y.copy(), z=z.copy(), vx=zero, vy=zero, vz=zero)]
    osim = orc.OracleSim(n, sp, 0.1, (0, 0, 0), 0)
    osim.step(1)
    with sim_for(sp, n, precision=ohp.MIXED, tile=(8, 8, 8), slack=0.01, b0=(0, 0, 0)) as sim:
        sim.step(1)
        d = sim.diagnostics()
        g = [by_pid(sim.store_particles(s, True)) for s in range(2)]
    assert d["repacks"] >= 1
    for s in range(2):
        assert len(g[s]["x"]) == m
        assert np.all((g[s]["x"] >= 0) & (g[s]["x"] < 16))
        assert np.max(np.abs(g[s]["x"] - osim.sp[s]["x"])) <= 1e-4


@pytest.mark.parametrize("n,ppc", [((64, 64, 64), 16), ((128, 128, 128), 4)])
def test_device_seeded_rho_matches_oracle(ohp, orc, n, ppc):
    """rho^0 of a device-seeded load, FP64 and mixed, against the oracle's
    deposit of the oracle loader's particles (same PCG32 stream)."""
    g = n[0] * n[1] * n[2]
    npcl = ppc * g
    rng = orc.Pcg32(42, 52)
    ref = np.zeros(g)
    for qm, vt in ((-1.0, 1.0), (0.01, 0.1)):
        x, y, z, *_ = orc.load_uniform_maxwell(rng, npcl, n, vt)
        orc.deposit_periodic(n, x, y, z, (1.0 if qm > 0 else -1.0) / ppc, 0, ref)
    errs = {}
    for prec in (ohp.FP64, ohp.MIXED):
        cfg = ohp.SimConfig(n=n, qm=[-1.0, 0.01], q_macro=[-1.0 / ppc, 1.0 / ppc],
                            capacity=[npcl, npcl], precision=prec, tile=(8, 8, 8))
        with ohp.Simulation(cfg) as sim:
            sim.seed_uniform_maxwell([npcl, npcl], [1.0, 0.1], seed=(42, 52))
            sim.deposit()
            rho = sim.grid(ohp.FIELD_RHO)
            counts = sim.cell_counts()
        errs[prec] = (rel_l2(rho, ref), int(counts.sum()))
    print(f"device-seeded rho^0 vs oracle at {n} x {ppc}: FP64 {errs[ohp.FP64][0]:.2e}, "
          f"mixed {errs[ohp.MIXED][0]:.2e}")
    assert errs[ohp.FP64][0] <= 1e-12
    assert errs[ohp.MIXED][0] <= TOL_MIXED
    assert errs[ohp.FP64][1] == errs[ohp.MIXED][1] == 2 * npcl


def test_mixed_from_hump_load_within_1e5(ohp, orc):
    # the Gaussian-hump density (cells holding ~1x to 10x the edge density):
    # two mixed steps against the FP64 oracle, standard weights
    n = (24, 24, 24)
    ppc = 8
    npcl = ppc * n[0] * n[1] * n[2]
    sp = []
    for s, (qm, vt) in enumerate(((-1.0, 1.0), (0.01, 0.1))):
        rng0 = orc.Pcg32(42, 52)
        x, y, z, vx, vy, vz = orc.load_hump_maxwell(rng0, s << 40, 2052, npcl, n, 9.0, 4.0, vt)
        sp.append(dict(qm=qm, q=(1.0 if qm > 0 else -1.0) / ppc, x=x, y=y, z=z, vx=vx, vy=vy,
                       vz=vz))
    osim = orc.OracleSim(n, sp, 0.1, (0, 0, 0.5), 1)
    osim.step(2)
    with sim_for(sp, n, precision=ohp.MIXED, weight_order=1) as sim:
        sim.step(2)
        rho = sim.grid(ohp.FIELD_RHO)
        e = [sim.grid(w) for w in (ohp.FIELD_EX, ohp.FIELD_EY, ohp.FIELD_EZ)]
    assert rel_l2(rho, osim.rho) <= TOL_MIXED
    for a, b in zip(e, (osim.ex, osim.ey, osim.ez)):
        assert rel_l2(a, b) <= TOL_MIXED
