"""GPU parity: libohpic.so (through its C ABI) against the FP64 oracle.

Bars: 1e-12 relative L2 for rho, phi, E, x, v after one step from identical
state; bit-exact particle cell indices and per-cell counts; bit-exact where
the arithmetic is exact (wrap edge cases, integer work).
"""
import numpy as np
import pytest

from helpers import by_pid, energy_histories, rel_l2, sim_for, two_species

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def ohp():
    import ohhelppic3d_amd as o
    if not o.available():
        pytest.fail("libohpic.so is not built: the GPU path has no fallback")
    return o


def _flat(n, a3):
    return np.asarray(a3).reshape(-1)


# ------------------------------------------------------------------------- FFT
def test_fft_impulse_response_reference_kat(ohp):
    # test/test_mpi_fftw3_solver.f90:53-97 (n = 1 x 1 x 63, impulse at z = 1)
    n = (1, 1, 63)
    a = np.zeros(63)
    a[1] = 1.0
    out = ohp.fft3d_r2r(a, n)
    k = np.arange(63)
    # R2HC of a unit impulse at index 1: X_k = exp(-2 pi i k / 63), so slot
    # k <= 31 holds Re X_k = cos(2 pi k / 63) and slot k > 31 holds
    # Im X_(63-k) = sin(2 pi k / 63)
    expected = np.where(k <= (63 + 1) // 2 - 1, np.cos(k / 63 * 2 * np.pi),
                        np.sin(k / 63 * 2 * np.pi))
    assert np.sum(np.abs(out - expected)) <= 1e-8
    assert np.max(np.abs(out - expected)) <= 1e-14


@pytest.mark.parametrize("n", [(8, 8, 8), (6, 10, 12), (16, 4, 9), (32, 32, 32), (24, 48, 16),
                               (5, 7, 3), (1, 1, 7), (64, 2, 3), (96, 1, 4),
                               # every radix up to 13 and the longest axes
                               (11, 13, 22), (26, 14, 1), (1024, 2, 3), (2, 1024, 2),
                               (3, 2, 1024), (256, 6, 5), (2, 640, 3), (1, 1, 1)])
def test_fft_r2hc_matches_oracle_and_roundtrips(ohp, orc, n):
    rng = np.random.default_rng(7)
    a = rng.standard_normal(n[0] * n[1] * n[2])
    ref = a.copy()
    for ax in range(3):
        ref = orc.r2r_axis(n, ref, ax, 0)
    got = ohp.fft3d_r2r(a, n)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(got - ref)) <= 1e-13 * scale
    back = ohp.fft3d_r2r(got, n, backward=True)
    assert np.max(np.abs(back - a)) <= 1e-13 * np.max(np.abs(a))


# --------------------------------------------------------------------- Poisson
@pytest.mark.parametrize("n", [(8, 8, 8), (16, 12, 10), (32, 32, 32), (24, 16, 48),
                               (11, 13, 14), (256, 4, 6), (5, 1024, 2), (2, 3, 1024),
                               # 256-point z lines (the 16 x 16 register pass): full
                               # groups, and an odd nx*ny with a partial last group
                               (8, 4, 256), (16, 16, 256), (5, 3, 256),
                               # an x row longer than one block's 2-point sweep (E)
                               (1024, 2, 3)])
def test_field_solve_matches_oracle(ohp, orc, n):
    rng = np.random.default_rng(11)
    rho = rng.standard_normal(n[0] * n[1] * n[2])
    cfg = ohp.SimConfig(n=n, qm=[-1.0], q_macro=[-1.0], capacity=[0])
    with ohp.Simulation(cfg) as sim:
        sim.put_grid(ohp.FIELD_RHO, rho)
        sim.field_solve()
        phi = sim.grid(ohp.FIELD_PHI)
        e = [sim.grid(w) for w in (ohp.FIELD_EX, ohp.FIELD_EY, ohp.FIELD_EZ)]
        d = sim.diagnostics()
    phi_ref = orc.poisson(n, rho)
    e_ref = orc.efield(n, phi_ref)
    assert rel_l2(phi, phi_ref) <= TOL
    for a, b in zip(e, e_ref):
        assert rel_l2(a, b) <= TOL
    assert abs(d["field_energy"] - 0.5 * rho @ phi_ref) <= 1e-12 * abs(0.5 * rho @ phi_ref)
    assert abs(d["net_charge"] - rho.sum()) <= 1e-10


def test_field_solve_256_cubed_properties(ohp):
    # the headline grid, through properties that hold at any size: the 7-point
    # Laplacian of phi gives back rho minus its mean, phi has zero mean, E is
    # the node-centred difference of phi, and the field energy is 1/2 rho.phi
    n = (256, 256, 256)
    rng = np.random.default_rng(5)
    rho = rng.standard_normal(n[0] * n[1] * n[2])
    cfg = ohp.SimConfig(n=n, qm=[-1.0], q_macro=[-1.0], capacity=[0])
    with ohp.Simulation(cfg) as sim:
        sim.put_grid(ohp.FIELD_RHO, rho)
        sim.field_solve()
        phi = sim.grid(ohp.FIELD_PHI)
        ex = sim.grid(ohp.FIELD_EX)
        ez = sim.grid(ohp.FIELD_EZ)
        d = sim.diagnostics()
    p = phi.reshape(n[2], n[1], n[0])
    lap = -6.0 * p
    for ax in range(3):
        lap += np.roll(p, 1, axis=ax) + np.roll(p, -1, axis=ax)
    scale = np.max(np.abs(phi))
    assert np.max(np.abs(-lap.reshape(-1) - (rho - rho.mean()))) <= 1e-12 * scale
    assert abs(phi.mean()) <= 1e-12 * scale
    ex_ref = 0.5 * ((np.roll(p, 1, axis=2) - p) + (p - np.roll(p, -1, axis=2)))
    ez_ref = 0.5 * ((np.roll(p, 1, axis=0) - p) + (p - np.roll(p, -1, axis=0)))
    assert np.max(np.abs(ex - ex_ref.reshape(-1))) <= 1e-15 * np.max(np.abs(ex_ref))
    assert np.max(np.abs(ez - ez_ref.reshape(-1))) <= 1e-15 * np.max(np.abs(ez_ref))
    we = 0.5 * rho @ phi
    assert abs(d["field_energy"] - we) <= 1e-12 * abs(we)


def test_field_get_with_halo_is_periodic(ohp):
    n = (6, 5, 4)
    rng = np.random.default_rng(3)
    rho = rng.standard_normal(120)
    cfg = ohp.SimConfig(n=n, qm=[-1.0], q_macro=[-1.0], capacity=[0])
    with ohp.Simulation(cfg) as sim:
        sim.put_grid(ohp.FIELD_RHO, rho)
        v = sim.get_field(ohp.FIELD_RHO, lb=(-1, -1, -1), ub=(7, 6, 5), nelem=2, comp=1)
        eb = sim.get_field(ohp.FIELD_EB, lb=(0, 0, 0), ub=(5, 4, 3))
    r3 = rho.reshape(4, 5, 6)  # [z][y][x]
    for ix in range(-1, 8):
        for iy in range(-1, 7):
            for iz in range(-1, 6):
                assert v[1, ix + 1, iy + 1, iz + 1] == r3[iz % 4, iy % 5, ix % 6]
    assert np.all(v[0] == 0.0)
    assert np.all(eb[5] == 0.5) and np.all(eb[3] == 0.0)


# --------------------------------------------------------------------- deposit
@pytest.mark.parametrize("wo", [0, 1])
def test_deposit_matches_oracle(ohp, orc, wo):
    n = (16, 16, 16)
    sp = two_species(orc, n, 4)
    with sim_for(sp, n, weight_order=wo) as sim:
        sim.deposit()
        rho = sim.grid(ohp.FIELD_RHO)
    ref = np.zeros(n[0] * n[1] * n[2])
    for d in sp:
        orc.deposit_periodic(n, d["x"], d["y"], d["z"], d["q"], wo, ref)
    assert rel_l2(rho, ref) <= TOL
    assert abs(rho.sum()) <= 1e-10


# ----------------------------------------------------------------- full step
def _compare_step(ohp, orc, n, ppc, nsteps, tol_x, tol_v, wo=0, tile=(0, 0, 0), b0=(0, 0, 0.5),
                  sp=None):
    if sp is None:
        sp = two_species(orc, n, ppc)
    osim = orc.OracleSim(n, sp, 0.1, b0, wo)
    with sim_for(sp, n, weight_order=wo, tile=tile, b0=b0) as sim:
        sim.step(nsteps)
        osim.step(nsteps)
        out = dict(rho=sim.grid(ohp.FIELD_RHO), phi=sim.grid(ohp.FIELD_PHI),
                   ex=sim.grid(ohp.FIELD_EX), ey=sim.grid(ohp.FIELD_EY),
                   ez=sim.grid(ohp.FIELD_EZ), diag=sim.diagnostics(),
                   counts=sim.cell_counts(),
                   sp=[by_pid(sim.store_particles(s, with_ids=True)) for s in range(2)])
    assert rel_l2(out["rho"], osim.rho) <= tol_x
    assert rel_l2(out["phi"], osim.phi) <= tol_x
    for k, ref in (("ex", osim.ex), ("ey", osim.ey), ("ez", osim.ez)):
        assert rel_l2(out[k], ref) <= tol_x
    ref_counts = np.zeros(n[0] * n[1] * n[2], np.int32)
    for s in range(2):
        g, r = out["sp"][s], osim.sp[s]
        assert len(g["x"]) == len(r["x"])
        assert np.array_equal(g["pid"], np.arange(len(r["x"])))
        for k in ("x", "y", "z"):
            assert rel_l2(g[k], r[k]) <= tol_x, k
            assert np.array_equal(g[k].astype(np.int64), r[k].astype(np.int64)), k
        for k in ("vx", "vy", "vz"):
            assert rel_l2(g[k], r[k]) <= tol_v, k
        ref_counts += orc.cell_counts(n, r["x"], r["y"], r["z"])
    assert np.array_equal(out["counts"], ref_counts)
    d = out["diag"]
    assert np.allclose(d["ke_old"], osim.ke_old, rtol=1e-12, atol=0)
    assert np.allclose(d["ke_new"], osim.ke_new, rtol=1e-12, atol=0)
    assert abs(d["field_energy"] - osim.field_energy) <= 1e-10 * abs(osim.field_energy)
    return out


def test_one_step_config0_matches_oracle(ohp, orc):
    # 32^3 periodic, 16 ppc per species (2^20 particles), dt 0.1, B0z 0.5
    _compare_step(ohp, orc, (32, 32, 32), 16, 1, TOL, TOL)


@pytest.mark.parametrize("n,tile", [((16, 16, 16), (4, 4, 4)), ((16, 16, 16), (8, 8, 4)),
                                    ((16, 16, 16), (16, 8, 4)), ((32, 32, 32), (8, 8, 8))])
def test_one_step_tiles_match_oracle(ohp, orc, n, tile):
    # (8, 8, 8) is the tile of the 256^3 headline instantiation
    _compare_step(ohp, orc, n, 8, 1, TOL, TOL, tile=tile)


def test_config0_hundred_steps_match_oracle_every_step(ohp, orc):
    """The CPU reference run's configuration (32^3, 16 ppc per species, dt
    0.1, B0z 0.5, 100 steps, rho, phi, E, x and v output every step) on the
    GPU beside the oracle, compared after every step.  Rounding-level
    differences grow with time in this (chaotic) system: two FP64 orderings
    of the oracle itself (serial and threaded deposit sums) start at 1e-15
    and reach 1.3e-12 in phi by step 100, so the bar is 1e-12 through step
    20 and 1e-9 to step 100; per-cell particle counts are bit-exact every
    step."""
    n = (32, 32, 32)
    sp = two_species(orc, n, 16)
    osim = orc.OracleSim(n, sp, 0.1, (0, 0, 0.5), 0)
    worst = []
    with sim_for(sp, n) as sim:
        for k in range(1, 101):
            sim.step(1)
            osim.step(1)
            e = {"rho": rel_l2(sim.grid(ohp.FIELD_RHO), osim.rho),
                 "phi": rel_l2(sim.grid(ohp.FIELD_PHI), osim.phi),
                 "Ex": rel_l2(sim.grid(ohp.FIELD_EX), osim.ex),
                 "Ey": rel_l2(sim.grid(ohp.FIELD_EY), osim.ey),
                 "Ez": rel_l2(sim.grid(ohp.FIELD_EZ), osim.ez)}
            ref_counts = np.zeros(n[0] * n[1] * n[2], np.int32)
            for s in range(2):
                g, r = by_pid(sim.store_particles(s, True)), osim.sp[s]
                for c in ("x", "y", "z", "vx", "vy", "vz"):
                    e[f"{c}{s}"] = rel_l2(g[c], r[c])
                ref_counts += orc.cell_counts(n, r["x"], r["y"], r["z"])
            assert np.array_equal(sim.cell_counts(), ref_counts), k
            w = max(e.values())
            worst.append(w)
            assert w <= (TOL if k <= 20 else 1e-9), (k, e)
    print("config 0, worst rel L2 per step at 1, 10, 20, 50, 100:",
          [f"{worst[i - 1]:.1e}" for i in (1, 10, 20, 50, 100)])


def test_one_step_standard_weights_noncubic(ohp, orc):
    _compare_step(ohp, orc, (24, 16, 8), 8, 1, TOL, TOL, wo=1)


def test_ten_steps_match_oracle(ohp, orc):
    _compare_step(ohp, orc, (16, 16, 16), 8, 10, 1e-11, 1e-11)


def test_separate_operators_equal_fused_step(ohp, orc):
    n = (16, 16, 16)
    sp = two_species(orc, n, 8)
    with sim_for(sp, n) as a, sim_for(sp, n) as b:
        parts = (ohp.LinearScatter(), ohp.PoissonFieldSolver(),
                 ohp.create_particle_mover("boris"), ohp.ParticleBoundaries(*n))
        for _ in range(3):
            ohp.mainstep(a, *parts)
        b.step(3)
        assert np.array_equal(a.grid(ohp.FIELD_PHI), b.grid(ohp.FIELD_PHI))
        for s in range(2):
            pa = by_pid(a.store_particles(s, True))
            pb = by_pid(b.store_particles(s, True))
            for k in pa:
                assert np.array_equal(pa[k], pb[k]), k


def test_step_is_deterministic(ohp, orc):
    n = (16, 16, 16)
    sp = two_species(orc, n, 8)
    res = []
    for _ in range(2):
        with sim_for(sp, n, chunks=3) as sim:
            sim.step(5)
            res.append((sim.grid(ohp.FIELD_PHI), sim.grid(ohp.FIELD_RHO),
                        [by_pid(sim.store_particles(s, True)) for s in range(2)]))
    assert np.array_equal(res[0][0], res[1][0])
    assert np.array_equal(res[0][1], res[1][1])
    for s in range(2):
        for k in res[0][2][s]:
            assert np.array_equal(res[0][2][s][k], res[1][2][s][k])


def test_energy_history_matches_oracle(ohp, orc):
    # 1000 steps: total energy within 0.5% of the oracle at every step; the
    # field energy (a sum over fluctuating modes of a chaotic system) within
    # 0.5% at every step until the two FP64 runs decorrelate (measured: first
    # 0.5% deviation near step 340-370), and its time mean over the last 500
    # steps within 1% afterwards
    tot_g, tot_o, fe_g, fe_o = energy_histories(ohp, orc, ohp.FP64, 1000)
    assert np.max(np.abs(tot_g - tot_o) / np.abs(tot_o)) <= 0.005
    assert np.max(np.abs(fe_g[:250] - fe_o[:250]) / np.abs(fe_o[:250])) <= 0.005
    assert abs(fe_g[500:].mean() - fe_o[500:].mean()) <= 0.01 * fe_o[500:].mean()
    # the GPU run's own drift equals the algorithm's (the oracle's) to 0.25%
    drift_g = np.max(np.abs(tot_g - tot_g[0])) / tot_g[0]
    drift_o = np.max(np.abs(tot_o - tot_o[0])) / tot_o[0]
    assert abs(drift_g - drift_o) <= 0.0025


# ------------------------------------------------------------------ boundaries
def test_wrap_edge_cases_bit_exact(ohp, orc):
    # q = 0 and B = 0 -> E = 0 exactly, v unchanged, x' = x + v dt exact (dt = 1/8)
    n = (16, 8, 8)
    dt = 0.125
    x = np.array([0.0, 15.5, 0.25, 0.5, 0.5, 3.0, 2.0, 0.5])
    v = np.array([-2.0 ** -57, 4.0, -4.0, -132.0, -200.0, 0.0, 8.0, 48.0])
    y = np.full_like(x, 1.5)
    z = np.full_like(x, 6.25)
    zero = np.zeros_like(x)
    sp = [dict(qm=-1.0, q=0.0, x=x, y=y, z=z, vx=v, vy=zero, vz=zero)]
    osim = orc.OracleSim(n, sp, dt, (0, 0, 0), 0)
    osim.step(1)
    with sim_for(sp, n, dt=dt, b0=(0, 0, 0), tile=(4, 4, 4)) as sim:
        sim.step(1)
        g = by_pid(sim.store_particles(0, True))
        d = sim.diagnostics()
    assert np.array_equal(g["x"], osim.sp[0]["x"])
    # -tiny -> n -> 0; n -> 0; -0.25 -> 15.75; -16 -> 0; -24.5 -> -8.5 -> 7.5 (two wraps)
    assert list(g["x"]) == [0.0, 0.0, 15.75, 0.0, 7.5, 3.0, 3.0, 6.5]
    # two particles end outside the deposit halo of their old tile
    assert d["fast_particles"] == 2


def test_repack_after_bin_overflow(ohp, orc):
    # every particle starts in tile 0 and moves one cell into tile 1, whose
    # bin capacity is the minimum: the overflow list and a repack absorb them
    n = (16, 8, 8)
    rng = np.random.default_rng(5)
    m = 3000
    x = rng.uniform(3.0, 4.0, m) * 1.0 + 4.0  # cells 7 (tile 0 with T=8 covers 0..7)
    y = rng.uniform(0, 8, m)
    z = rng.uniform(0, 8, m)
    vx = np.full(m, 10.0)
    zero = np.zeros(m)
    sp = [dict(qm=-1.0, q=-1e-3, x=x, y=y, z=z, vx=vx, vy=zero, vz=zero),
          dict(qm=0.01, q=1e-3, x=x.copy(), y=y.copy(), z=z.copy(), vx=zero, vy=zero,
               vz=zero)]
    osim = orc.OracleSim(n, sp, 0.1, (0, 0, 0.5), 0)
    osim.step(2)
    with sim_for(sp, n, tile=(8, 8, 8), slack=0.01) as sim:
        sim.step(2)
        d = sim.diagnostics()
        g = [by_pid(sim.store_particles(s, True)) for s in range(2)]
    assert d["repacks"] >= 1
    for s in range(2):
        assert len(g[s]["x"]) == m
        for k in ("x", "y", "z", "vx", "vy", "vz"):
            assert rel_l2(g[s][k], osim.sp[s][k]) <= TOL, (s, k)


# ------------------------------------------------------------------- particles
def test_particle_io_roundtrip_and_aos(ohp, orc):
    n = (8, 8, 8)
    sp = two_species(orc, n, 2)
    with sim_for(sp, n) as sim:
        for s in range(2):
            g = by_pid(sim.store_particles(s, True))
            for k in ("x", "y", "z", "vx", "vy", "vz"):
                assert np.array_equal(g[k], sp[s][k])
        rec = sim.store_oh_particles()
        assert len(rec) == 2 * 2 * 512
        assert set(np.unique(rec["spec"])) == {1, 2}
        rec2 = rec.copy()
        rec2["vx"] *= 2
        sim.load_oh_particles(rec2)
        g0 = by_pid(sim.store_particles(0, True))
        assert np.array_equal(g0["vx"], 2 * sp[0]["vx"])


def test_device_loader_matches_oracle(ohp, orc):
    n = (16, 8, 8)
    npcl = [4096, 2048]
    vth = [1.0, 0.1]
    cfg = ohp.SimConfig(n=n, qm=[-1.0, 0.01], q_macro=[-0.1, 0.1], capacity=npcl,
                        track_ids=True)
    with ohp.Simulation(cfg) as sim:
        sim.seed_uniform_maxwell(npcl, vth, seed=(42, 52))
        g = [by_pid(sim.store_particles(s, True)) for s in range(2)]
    rng = orc.Pcg32(42, 52)
    for s in range(2):
        x, y, z, vx, vy, vz = orc.load_uniform_maxwell(rng, npcl[s], n, vth[s])
        assert np.array_equal(g[s]["x"], x) and np.array_equal(g[s]["y"], y)
        assert np.array_equal(g[s]["z"], z)
        for a, b in ((g[s]["vx"], vx), (g[s]["vy"], vy), (g[s]["vz"], vz)):
            assert rel_l2(a, b) <= 1e-14


def test_hump_loader_matches_oracle(ohp, orc):
    n = (16, 16, 16)
    npcl = [3000, 3000]
    cfg = ohp.SimConfig(n=n, qm=[-1.0, 0.01], q_macro=[-0.1, 0.1], capacity=npcl,
                        track_ids=True)
    with ohp.Simulation(cfg) as sim:
        sim.seed_hump_maxwell(npcl, [1.0, 0.1], 9.0, 3.0, seed=(42, 52), budget=2052)
        g = [by_pid(sim.store_particles(s, True)) for s in range(2)]
    for s, vt in enumerate((1.0, 0.1)):
        rng0 = orc.Pcg32(42, 52)
        x, y, z, vx, vy, vz = orc.load_hump_maxwell(rng0, s << 40, 2052, npcl[s], n, 9.0, 3.0,
                                                    vt)
        assert np.array_equal(g[s]["x"], x) and np.array_equal(g[s]["z"], z)
        assert rel_l2(g[s]["vy"], vy) <= 1e-14


def test_energy_matches_diagnostics_and_oracle(ohp, orc):
    n = (16, 16, 16)
    sp = two_species(orc, n, 8)
    osim = orc.OracleSim(n, sp, 0.1, (0, 0, 0.5), 0)
    with sim_for(sp, n) as sim:
        sim.step(2)
        osim.step(2)
        e = sim.energy()
        d = sim.diagnostics()
    assert e[1] == d["field_energy"] and e[2] == d["total_energy"] and e[3] == d["net_charge"]
    ke = 0.5 * (osim.ke_old.sum() + osim.ke_new.sum())
    assert abs(e[0] - ke) <= 1e-12 * ke
    assert abs(e[1] - osim.field_energy) <= 1e-10 * abs(osim.field_energy)


def test_three_species_match_oracle(ohp, orc):
    # electrons, ions and a second, heavier ion species (per-species bins and
    # kinetic-energy slots beyond two)
    n = (16, 12, 8)
    ppc = 6
    g = n[0] * n[1] * n[2]
    rng = orc.Pcg32(42, 52)
    sp = []
    for qm, q, vt in ((-1.0, -2.0, 1.0), (0.01, 1.0, 0.1), (0.005, 1.0, 0.07)):
        x, y, z, vx, vy, vz = orc.load_uniform_maxwell(rng, ppc * g, n, vt)
        sp.append(dict(qm=qm, q=q / ppc, x=x, y=y, z=z, vx=vx, vy=vy, vz=vz))
    osim = orc.OracleSim(n, sp, 0.1, (0, 0, 0.5), 0)
    with sim_for(sp, n) as sim:
        sim.step(2)
        osim.step(2)
        rho, phi = sim.grid(ohp.FIELD_RHO), sim.grid(ohp.FIELD_PHI)
        d = sim.diagnostics()
        parts = [by_pid(sim.store_particles(s, True)) for s in range(3)]
        counts = sim.cell_counts()
    assert rel_l2(rho, osim.rho) <= TOL and rel_l2(phi, osim.phi) <= TOL
    ref_counts = np.zeros(g, np.int32)
    for s in range(3):
        for k in ("x", "y", "z", "vx", "vy", "vz"):
            assert rel_l2(parts[s][k], osim.sp[s][k]) <= TOL, (s, k)
        ref_counts += orc.cell_counts(n, osim.sp[s]["x"], osim.sp[s]["y"], osim.sp[s]["z"])
    assert np.array_equal(counts, ref_counts)
    assert np.allclose(d["ke_new"][:3], osim.ke_new[:3], rtol=1e-12, atol=0)


def test_two_steps_from_hump_load_match_oracle(ohp, orc):
    # the Gaussian-hump density n(r) = n0 (1 + 9 exp(-|r-c|^2 / (2 sigma^2)))
    # at 24^3 with sigma 4 (cells hold ~1x to 10x the edge density): the
    # imbalanced bins against the oracle, at the FP64 bars
    n = (24, 24, 24)
    ppc = 8
    npcl = ppc * n[0] * n[1] * n[2]
    sp = []
    for s, (qm, vt) in enumerate(((-1.0, 1.0), (0.01, 0.1))):
        rng0 = orc.Pcg32(42, 52)
        x, y, z, vx, vy, vz = orc.load_hump_maxwell(rng0, s << 40, 2052, npcl, n, 9.0, 4.0, vt)
        sp.append(dict(qm=qm, q=(1.0 if qm > 0 else -1.0) / ppc, x=x, y=y, z=z, vx=vx, vy=vy,
                       vz=vz))
    counts = orc.cell_counts(n, sp[0]["x"], sp[0]["y"], sp[0]["z"])
    assert counts.max() >= 5 * max(1, counts.min())
    _compare_step(ohp, orc, n, ppc, 2, TOL, TOL, sp=sp)


def test_errors_are_reported(ohp, orc):
    n = (8, 8, 8)
    cfg = ohp.SimConfig(n=n, qm=[-1.0], q_macro=[-0.1], capacity=[10])
    with ohp.Simulation(cfg) as sim:
        big = np.zeros(11)
        with pytest.raises(ohp.OhpicError) as e:
            sim.load_particles(0, big, big, big, big, big, big)
        assert e.value.code == -3
        bad = np.array([8.0])
        with pytest.raises(ohp.OhpicError) as e:
            sim.load_particles(0, bad, bad * 0, bad * 0, bad * 0, bad * 0, bad * 0)
        assert e.value.code == -1 and "outside" in str(e.value)
        ok = np.array([1.5])
        sim.load_particles(0, ok, ok, ok, ok * 0, ok * 0, ok * 0)
        sim.gather_push()
        with pytest.raises(ohp.OhpicError) as e:
            sim.step(1)
        assert e.value.code == -4
        sim.boundary()
        sim.step(1)


# ------------------------------------------------------------------ edge cases
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("n", [(20, 12, 10), (4, 3, 10), (2, 5, 9)])
def test_partial_tiles_match_oracle(ohp, orc, precision, n):
    # grid lengths that 8-cell tiles do not divide: the last tile of every
    # axis is partly outside the box; axes of 2-4 cells put a tile's E halo
    # and deposit block more than one period past the box (their periodic
    # indices take the division path)
    sp = two_species(orc, n, 4)
    osim = orc.OracleSim(n, sp, 0.1, (0, 0, 0.5), 0)
    prec = ohp.FP64 if precision == "fp64" else ohp.MIXED
    with sim_for(sp, n, tile=(8, 8, 8), precision=prec) as sim:
        sim.step(2)
        osim.step(2)
        rho = sim.grid(ohp.FIELD_RHO)
        counts = sim.cell_counts()
        g = [by_pid(sim.store_particles(s, True)) for s in range(2)]
    tol = TOL if precision == "fp64" else 1e-3
    assert rel_l2(rho, osim.rho) <= tol
    ref_counts = np.zeros(n[0] * n[1] * n[2], np.int32)
    for s in range(2):
        assert len(g[s]["x"]) == len(osim.sp[s]["x"])
        assert np.all((g[s]["x"] >= 0) & (g[s]["x"] < n[0]) & (g[s]["z"] < n[2]))
        ref_counts += orc.cell_counts(n, osim.sp[s]["x"], osim.sp[s]["y"], osim.sp[s]["z"])
    if precision == "fp64":
        assert np.array_equal(counts, ref_counts)
    else:
        assert counts.sum() == ref_counts.sum()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_results_do_not_depend_on_tile(ohp, orc, precision):
    # particles, cell counts and rho/phi/E are the same bit for bit whatever
    # the tile shape: the deposit is integer, and the mixed form keeps
    # positions at 2^-24 cells relative to the tile origin for every tile
    n = (16, 16, 16)
    sp = two_species(orc, n, 8)
    prec = ohp.FP64 if precision == "fp64" else ohp.MIXED
    out = []
    for tile in ((4, 4, 4), (8, 8, 8), (8, 8, 4)):
        with sim_for(sp, n, tile=tile, precision=prec) as sim:
            sim.step(3)
            grids = [sim.grid(w) for w in (ohp.FIELD_RHO, ohp.FIELD_PHI, ohp.FIELD_EX,
                                           ohp.FIELD_EY, ohp.FIELD_EZ)]
            parts = [by_pid(sim.store_particles(s, True)) for s in range(2)]
            out.append((grids, parts, sim.cell_counts()))
    g0, p0, c0 = out[0]
    for g, p, c in out[1:]:
        for a, b in zip(g, g0):
            assert np.array_equal(a, b)
        for s in range(2):
            for k in p0[s]:
                assert np.array_equal(p[s][k], p0[s][k]), k
        assert np.array_equal(c, c0)


def test_empty_species_and_empty_instance(ohp, orc):
    n = (16, 16, 16)
    sp = two_species(orc, n, 2)
    empty = [dict(sp[0]), dict(sp[1])]
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        empty[1][k] = empty[1][k][:0]
    osim = orc.OracleSim(n, empty, 0.1, (0, 0, 0.5), 0)
    with sim_for(empty, n) as sim:
        assert sim.particle_count(1) == 0
        sim.step(3)
        osim.step(3)
        assert rel_l2(sim.grid(ohp.FIELD_PHI), osim.phi) <= 1e-11
        assert len(sim.store_particles(1)["x"]) == 0
        d = sim.diagnostics()
        assert d["ke_new"][1] == 0.0 and d["ke_old"][1] == 0.0
    # an instance without particles inside a group steps as well
    cfg = ohp.SimConfig(n=(16, 16, 16), qm=(-1.0, 0.01), q_macro=(-0.5, 0.5),
                        capacity=(0, 0), track_ids=True)
    with ohp.Simulation(cfg) as nothing, sim_for(sp, n) as full, sim_for(sp, n) as ref:
        ref.step(2)
        with ohp.Group() as grp:
            grp.attach(nothing)
            grp.attach(full)
            grp.step(2)
        assert np.array_equal(full.grid(ohp.FIELD_PHI), ref.grid(ohp.FIELD_PHI))
        assert np.all(nothing.grid(ohp.FIELD_RHO) == 0.0)
        assert nothing.diagnostics()["step"] == 2


def test_run_from_reference_parameters_matches_oracle(ohp, orc):
    # the config-0 run described in the reference's parameters.toml layout:
    # the device-seeded state and one step equal the oracle's
    import os
    from ohhelppic3d_amd import params
    spec = params.load_parameters(os.path.join(os.path.dirname(__file__), "golden",
                                               "parameters_config0.toml"))
    sp = two_species(orc, spec.cfg.n, 16)
    osim = orc.OracleSim(spec.cfg.n, sp, spec.cfg.dt, spec.cfg.b0, 0)
    with spec.create() as sim:
        sim.step(1)
        osim.step(1)
        assert rel_l2(sim.grid(ohp.FIELD_PHI), osim.phi) <= TOL
        assert rel_l2(sim.grid(ohp.FIELD_RHO), osim.rho) <= TOL
