"""Pins the CPU oracle: reference golden vectors, the reference's own PCG build,
closed forms, invariants and an independent FFT library (numpy's pocketfft).

Every check is chosen so that a plausible slip (dropped term, wrong sign,
swapped weight, off-by-one node, wrong normalisation) fails it.
"""
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
RNG = np.random.default_rng(20240611)


# ----------------------------------------------------------------------------- PCG
def _kat_lines(orc, seed=(42, 54), rounds=5):
    """Re-create lib/pcg-c/test-high/check-base.c's output for pcg32."""
    r = orc.Pcg32(*seed)
    out = []
    number = "A23456789TJQK"
    suit = "hcds"
    for rnd in range(1, rounds + 1):
        out.append(f"Round {rnd}:")
        vals = [r.random() for _ in range(6)]
        out.append("  32bit:" + "".join(f" 0x{v:08x}" for v in vals))
        r.advance(-6)
        vals = [r.random() for _ in range(6)]
        out.append("  Again:" + "".join(f" 0x{v:08x}" for v in vals))
        out.append("  Coins: " + "".join("H" if r.boundedrand(2) else "T" for _ in range(65)))
        out.append("  Rolls:" + "".join(f" {r.boundedrand(6) + 1}" for _ in range(33)))
        cards = list(range(52))
        for i in range(52, 1, -1):
            ch = r.boundedrand(i)
            cards[ch], cards[i - 1] = cards[i - 1], cards[ch]
        s = "  Cards:"
        for i, c in enumerate(cards):
            s += f" {number[c // 4]}{suit[c % 4]}"
            if (i + 1) % 22 == 0:
                s += "\n\t"
        out.append(s)
        out.append("")
    return "\n".join(out)


def test_pcg32_kat_matches_reference_expected_output(orc):
    # lib/pcg-c/test-high/expected/check-pcg32.out (seeds 42, 54), committed as data
    with open(os.path.join(HERE, "golden", "pcg32_kat_seed42_54.txt")) as f:
        expected = f.read()
    body = expected[expected.index("Round 1:"):].rstrip("\n")
    got = _kat_lines(orc).rstrip("\n")
    assert got.replace("\t", "").split() == body.replace("\t", "").split()


def test_pcg32_matches_reference_build(orc):
    ref = orc.ref_pcg_lib()
    if ref is None:
        pytest.skip("oracle/_ref/libpcgref.so not built (reference not mounted)")
    import ctypes as C
    st = orc.Pcg32State()
    ref.pcg_setseq_64_srandom_r(C.byref(st), 42, 52)
    mine = orc.Pcg32(42, 52)
    for _ in range(1000):
        assert ref.pcg32_random_double_r(C.byref(st)) == mine.double()
    ref.pcg_setseq_64_advance_r(C.byref(st), 123456789)
    mine.advance(123456789)
    for _ in range(100):
        assert ref.pcg_setseq_64_xsh_rr_32_random_r(C.byref(st)) == mine.random()


def test_pcg32_advance_equals_stepping(orc):
    a = orc.Pcg32(42, 52)
    b = a.copy()
    for _ in range(777):
        a.random()
    b.advance(777)
    assert [a.random() for _ in range(10)] == [b.random() for _ in range(10)]


def test_pcg32_double_is_u32_times_2m32(orc):
    a = orc.Pcg32(42, 52)
    b = a.copy()
    for _ in range(200):
        assert a.double() == math.ldexp(b.random(), -32)


# -------------------------------------------------------------------------- loaders
def test_uniform_maxwell_loader(orc):
    n = (8, 4, 6)
    rng = orc.Pcg32(42, 52)
    check = rng.copy()
    x, y, z, vx, vy, vz = orc.load_uniform_maxwell(rng, 20000, n, 1.5)
    # draw order: random_fix draw, then x, y, z, then four velocity draws
    check.double()
    assert x[0] == check.double() * 8 and y[0] == check.double() * 4
    assert z[0] == check.double() * 6
    assert np.all((x >= 0) & (x < 8) & (y >= 0) & (y < 4) & (z >= 0) & (z < 6))
    v = np.concatenate([vx, vy, vz])
    assert abs(v.mean()) < 0.03 and abs(v.std() - 1.5) < 0.03
    # the stream continues where the loader stopped (1 + 7 draws per particle)
    c2 = orc.Pcg32(42, 52)
    c2.advance(1 + 7 * 20000)
    assert c2.random() == rng.random()


def test_maxwell_velocities_follow_reference_formula_exactly(orc):
    """velocity_maxwell_distribution3d.f90:40-54 draws rands(1..4) and sets
    v1 = 2 sigma sqrt(-2 log r1) cos(2 pi r2), v2 = ... sin(2 pi r2),
    v3 = 2 sigma sqrt(-2 log r3) sin(2 pi r4), clipped at 6 sigma.  The
    loader's resolution: sigma = v_th / 2 (so the std is v_th), log of 1 - r
    (a PCG double can be 0), no clip.  The first particles' velocities equal
    that formula on the same PCG draws bit for bit."""
    vth = 1.7
    rng = orc.Pcg32(42, 52)
    check = rng.copy()
    n = (8, 4, 6)
    x, y, z, vx, vy, vz = orc.load_uniform_maxwell(rng, 5000, n, vth)
    pi = 3.14159265358979  # science_constants.f90:4
    sigma = vth / 2.0
    check.double()  # random_fix
    for i in range(4):
        for _ in range(3):
            check.double()  # x, y, z
        r = [check.double() for _ in range(4)]
        v1 = 2.0 * sigma * math.sqrt(-2.0 * math.log(1.0 - r[0])) * math.cos(2.0 * pi * r[1])
        v2 = 2.0 * sigma * math.sqrt(-2.0 * math.log(1.0 - r[0])) * math.sin(2.0 * pi * r[1])
        v3 = 2.0 * sigma * math.sqrt(-2.0 * math.log(1.0 - r[2])) * math.sin(2.0 * pi * r[3])
        assert (vx[i], vy[i], vz[i]) == (v1, v2, v3), i
    # no clip: among 15000 components some exceed 3 standard deviations
    # (6 sigma of the reference's formula)
    v = np.concatenate([vx, vy, vz])
    assert np.max(np.abs(v)) > 3.0 * vth


def test_hump_loader_density_profile(orc):
    n = (32, 32, 32)
    rng0 = orc.Pcg32(42, 52)
    x, y, z, *_ = orc.load_hump_maxwell(rng0, 0, 2052, 200000, n, 9.0, 4.0, 1.0)
    r2 = (x - 16) ** 2 + (y - 16) ** 2 + (z - 16) ** 2
    inner = np.mean(r2 < 4.0 ** 2)
    # the hump raises the inner-sphere fraction far above its volume fraction
    vol = 4 / 3 * np.pi * 4.0 ** 3 / 32 ** 3
    assert inner > 4 * vol
    # determinism: particle i depends only on (seed, i)
    x2, *_ = orc.load_hump_maxwell(rng0, 2052 * 1000, 2052, 10, n, 9.0, 4.0, 1.0)
    assert np.array_equal(x2, x[1000:1010])


# -------------------------------------------------------------------------- wrap
def test_apply_periodic_special_values(orc):
    n = 16
    assert orc.apply_periodic(n, 3.25) == 3.25
    assert orc.apply_periodic(n, -0.5) == 15.5
    assert orc.apply_periodic(n, 16.0) == 0.0
    assert orc.apply_periodic(n, 16.75) == 0.75
    tiny = 5e-324
    assert orc.apply_periodic(n, -tiny) == 0.0  # n - tiny rounds to n, then n -> 0
    below = np.nextafter(16.0, 0.0)
    assert orc.apply_periodic(n, below) == below
    r = orc.apply_periodic(n, -0.0)
    assert r == 0.0 and math.copysign(1, r) == -1.0  # -0 < 0 is false
    assert orc.apply_periodic(n, -16.0 + 0.25) == 0.25


# -------------------------------------------------------------------------- Boris
def test_boris_electric_only(orc):
    E = np.array([0.3, -0.2, 0.7])
    pos, vel = orc.boris_move([1, 2, 3], [0.1, 0.2, 0.3], -1.0,
                              np.concatenate([E, [0, 0, 0]]), 0.1)
    assert np.allclose(vel, np.array([0.1, 0.2, 0.3]) + (-1.0) * E * 0.1, rtol=0, atol=1e-15)
    assert np.allclose(pos, np.array([1, 2, 3]) + vel * 0.1, rtol=0, atol=1e-15)


def test_boris_magnetic_rotation_angle(orc):
    qm, B, dt = -1.0, 0.5, 0.1
    _, vel = orc.boris_move([0, 0, 0], [1.0, 0, 0], qm, [0, 0, 0, 0, 0, B], dt)
    theta = 2 * math.atan(qm * B * dt / 2)
    # dv/dt = qm v x B rotates (1,0,0) towards -y for qm*B > 0
    assert np.allclose(vel, [math.cos(theta), -math.sin(theta), 0.0], atol=1e-15)
    assert abs(np.linalg.norm(vel) - 1.0) < 1e-15


def test_boris_exb_drift(orc):
    qm, dt = 1.0, 0.05
    E = np.array([0.0, 0.2, 0.0])
    B = np.array([0.0, 0.0, 1.0])
    vd = np.cross(E, B) / B.dot(B)
    pos, vel = np.zeros(3), vd + np.array([0.0, 0.3, 0.0])
    vs = []
    for _ in range(4000):
        pos, vel = orc.boris_move(pos, vel, qm, np.concatenate([E, B]), dt)
        vs.append(vel)
    vmean = np.mean(vs[-int(2 * np.pi / dt) * 10:], axis=0)
    assert np.allclose(vmean, vd, atol=2e-3)


# ------------------------------------------------------------------ gather/deposit
def _halo_field(nelem, n, vals):
    """t_OhField-shaped values(nelem, -1:n+1, ...) filled periodically."""
    lb = [-1, -1, -1]
    ub = [n[0] + 1, n[1] + 1, n[2] + 1]
    shape = (nelem, n[0] + 3, n[1] + 3, n[2] + 3)
    f = np.zeros(shape)
    for ix in range(-1, n[0] + 2):
        for iy in range(-1, n[1] + 2):
            for iz in range(-1, n[2] + 2):
                f[:, ix + 1, iy + 1, iz + 1] = vals(ix % n[0], iy % n[1], iz % n[2])
    return f, lb, ub


def test_gather_uniform_exact_and_weights_mirrored(orc):
    n = (4, 5, 6)
    f, lb, ub = _halo_field(6, n, lambda i, j, k: np.arange(6) + 1.0)
    for pos in RNG.uniform(0, 1, (20, 3)) * np.array(n):
        r = orc.interp(f, lb, ub, pos, 0)
        assert np.allclose(r, np.arange(6) + 1.0, rtol=2e-16 * 8, atol=0)
    # linear field in x: the mirrored weights give the value at the mirrored
    # position i0 + (1 - r); standard weights give it at i0 + r
    g, lb, ub = _halo_field(1, (8, 4, 4), lambda i, j, k: np.array([float(i)]))
    for xx in [1.25, 2.5, 3.9, 6.01]:
        i0 = int(xx)
        rr = xx - i0
        got_ref = orc.interp(g, lb, ub, [xx, 1.5, 2.5], 0)[0]
        got_std = orc.interp(g, lb, ub, [xx, 1.5, 2.5], 1)[0]
        assert abs(got_ref - (i0 + (1 - rr))) < 1e-14
        assert abs(got_std - xx) < 1e-14


def test_periodic_gather_matches_halo_interp(orc):
    n = (5, 4, 6)
    g = n[0] * n[1] * n[2]
    e = [RNG.standard_normal(g) for _ in range(3)]

    def vals(i, j, k):
        idx = i + n[0] * (j + n[1] * k)
        return np.array([e[0][idx], e[1][idx], e[2][idx]])

    f, lb, ub = _halo_field(3, n, vals)
    for wo in (0, 1):
        for pos in RNG.uniform(0, 1, (50, 3)) * np.array(n):
            a = orc.interp(f, lb, ub, pos, wo)
            b = orc.gather_periodic(n, e, pos, wo)
            assert np.allclose(a, b, rtol=0, atol=1e-14)


def test_deposit_single_particle_weights(orc):
    n = (4, 4, 4)
    rho = orc.deposit_periodic(n, np.array([1.25]), np.array([2.0]), np.array([3.5]), 2.0, 0)
    r = rho.reshape(4, 4, 4)  # [z][y][x]
    # reference order: lower x node (1) gets r = 0.25, upper (2) gets 0.75;
    # y exactly on node 2 -> lower gets r = 0, upper (3) gets 1;
    # z node 3 lower gets 0.5, upper wraps to node 0 and gets 0.5
    assert r[3, 3, 1] == pytest.approx(2.0 * 0.25 * 1.0 * 0.5)
    assert r[3, 3, 2] == pytest.approx(2.0 * 0.75 * 1.0 * 0.5)
    assert r[0, 3, 2] == pytest.approx(2.0 * 0.75 * 1.0 * 0.5)
    assert r[3, 2, 1] == 0.0
    assert rho.sum() == pytest.approx(2.0)


def test_deposit_conserves_charge_and_is_adjoint_of_gather(orc):
    n = (6, 5, 7)
    g = n[0] * n[1] * n[2]
    npcl = 500
    x, y, z = (RNG.uniform(0, m, npcl) for m in n)
    e = [RNG.standard_normal(g) for _ in range(3)]
    for wo in (0, 1):
        rho = orc.deposit_periodic(n, x, y, z, 0.37, wo)
        assert abs(rho.sum() - 0.37 * npcl) < 1e-12
        # sum_p q E(x_p) == sum_i rho_i E_i  (same weights in both directions)
        lhs = np.zeros(3)
        for p in range(npcl):
            lhs += 0.37 * orc.gather_periodic(n, e, (x[p], y[p], z[p]), wo)
        rhs = np.array([rho @ e[0], rho @ e[1], rho @ e[2]])
        assert np.allclose(lhs, rhs, rtol=1e-12, atol=1e-12)


# -------------------------------------------------------------------------- FFT
def test_r2hc_impulse_response_reference_kat(orc):
    # test/test_mpi_fftw3_solver.f90:53-97: n = 63, impulse at z = 1,
    # expected cos(2 pi k/63) for k <= 31 and sin(2 pi k/63) above
    n = 63
    a = np.zeros(n)
    a[1] = 1.0
    out = orc.r2hc_1d(a)
    k = np.arange(n)
    expected = np.where(k <= (n + 1) // 2 - 1, np.cos(k / n * 2 * np.pi),
                        np.sin(k / n * 2 * np.pi))
    assert np.sum(np.abs(out - expected)) <= 1e-8
    assert np.max(np.abs(out - expected)) <= 1e-14
    # the same through the 3-D axis transform on a 1 x 1 x 63 array
    out3 = orc.r2r_axis((1, 1, 63), a, 2, 0)
    assert np.max(np.abs(out3 - expected)) <= 1e-14


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 16, 24, 30, 48, 63, 64, 96])
def test_r2hc_matches_numpy_and_roundtrip(orc, n):
    a = RNG.standard_normal(n)
    X = np.fft.fft(a)
    hc = orc.r2hc_1d(a)
    exp = np.zeros(n)
    exp[: n // 2 + 1] = X[: n // 2 + 1].real
    for k in range(1, (n + 1) // 2):
        exp[n - k] = X[k].imag
    assert np.max(np.abs(hc - exp)) <= 1e-13 * max(1, n)
    back = orc.hc2r_1d(hc)
    assert np.max(np.abs(back - n * a)) <= 1e-13 * n


def test_wave_number_table(orc):
    # squared modified wave number, symmetric in k <-> n-k, zero at DC
    for n in (6, 7, 32):
        w = np.array([orc.wave_number_sq(k, n) for k in range(n)])
        assert w[0] == 0.0
        assert np.allclose(w[1:], w[1:][::-1], rtol=0, atol=1e-15)
        assert np.allclose(w, (2 * np.sin(np.pi * np.arange(n) / n)) ** 2, rtol=1e-13, atol=0)


def _laplacian(n, phi):
    p = phi.reshape(n[2], n[1], n[0])
    lap = -6 * p
    for ax in range(3):
        lap = lap + np.roll(p, 1, axis=ax) + np.roll(p, -1, axis=ax)
    return lap.reshape(-1)


@pytest.mark.parametrize("n", [(8, 8, 8), (6, 10, 12), (16, 4, 9), (1, 1, 7)])
def test_poisson_inverts_7_point_laplacian(orc, n):
    g = n[0] * n[1] * n[2]
    rho = RNG.standard_normal(g)
    phi = orc.poisson(n, rho)
    # -lap(phi) = rho - mean(rho) pins the Green's function, the sign,
    # the DC removal and the 1/N normalisation together
    res = -_laplacian(n, phi) - (rho - rho.mean())
    assert np.max(np.abs(res)) <= 1e-12 * np.max(np.abs(rho))
    assert abs(phi.mean()) <= 1e-13
    # independent library FFT (pocketfft) gives the same phi
    R = np.fft.fftn(rho.reshape(n[2], n[1], n[0]))
    kk = [np.arange(m) for m in (n[2], n[1], n[0])]
    lam = sum(np.meshgrid(*[(2 * np.sin(np.pi * k / m)) ** 2 for k, m in
                            zip(kk, (n[2], n[1], n[0]))], indexing="ij"))
    lam[0, 0, 0] = 1.0
    P = R / lam
    P[0, 0, 0] = 0
    phi_np = np.fft.ifftn(P).real.reshape(-1)
    assert np.max(np.abs(phi - phi_np)) <= 1e-12 * np.max(np.abs(phi_np))


def test_poisson_single_mode_eigenfunction(orc):
    n = (16, 12, 8)
    k = (3, 2, 1)
    iz, iy, ix = np.meshgrid(*[np.arange(m) for m in (n[2], n[1], n[0])], indexing="ij")
    th = 2 * np.pi * (k[0] * ix / n[0] + k[1] * iy / n[1] + k[2] * iz / n[2])
    rho = np.cos(th).reshape(-1)
    lam = sum((2 * np.sin(np.pi * k[a] / n[a])) ** 2 for a in range(3))
    phi = orc.poisson(n, rho)
    assert np.max(np.abs(phi - rho / lam)) <= 1e-13


def test_efield_single_mode_central_difference(orc):
    # phi = cos(theta i) -> E_x(i) = 0.5 (phi(i-1) - phi(i+1)) = sin(theta) sin(theta i)
    n = (16, 8, 4)
    th = 2 * np.pi * 3 / 16
    iz, iy, ix = np.meshgrid(*[np.arange(m) for m in (n[2], n[1], n[0])], indexing="ij")
    phi = np.cos(th * ix).reshape(-1)
    ex, ey, ez = orc.efield(n, phi)
    assert np.max(np.abs(ex - (np.sin(th) * np.sin(th * ix)).reshape(-1))) <= 1e-15
    assert np.max(np.abs(ey)) == 0 and np.max(np.abs(ez)) == 0
    thy = 2 * np.pi / 8
    phi = np.cos(thy * iy).reshape(-1)
    ex, ey, ez = orc.efield(n, phi)
    assert np.max(np.abs(ey - (np.sin(thy) * np.sin(thy * iy)).reshape(-1))) <= 1e-15
    assert np.max(np.abs(ex)) == 0


# -------------------------------------------------------------------------- step
def _two_species(orc, n, ppc, seed=(42, 52), b=0.5, vth=(1.0, 0.1)):
    rng = orc.Pcg32(*seed)
    g = n[0] * n[1] * n[2]
    sp = []
    for qm, q, vt in ((-1.0, -1.0 / ppc, vth[0]), (0.01, 1.0 / ppc, vth[1])):
        x, y, z, vx, vy, vz = orc.load_uniform_maxwell(rng, ppc * g, n, vt)
        sp.append(dict(qm=qm, q=q, x=x, y=y, z=z, vx=vx, vy=vy, vz=vz))
    return sp


def test_step_order_deposit_solve_push(orc):
    """mainstep's order (ohhelppic3d.f90:311-358): rho^n is the deposit of
    the positions before the step, phi and E solve that rho, and each
    particle is pushed (gather + Boris) in that E and wrapped twice.  One
    orc_step equals that composition of the oracle's own pinned operators
    bit for bit (a push before the deposit, or in the previous step's E,
    would not)."""
    n = (8, 8, 8)
    sp = _two_species(orc, n, 2)
    sim = orc.OracleSim(n, sp, 0.1, (0, 0, 0.5), 0)
    sim.step(1)
    rho = np.zeros(n[0] * n[1] * n[2])
    for d in sp:
        orc.deposit_periodic(n, d["x"], d["y"], d["z"], d["q"], 0, rho)
    assert np.array_equal(sim.rho, rho)
    phi = orc.poisson(n, rho)
    assert np.array_equal(sim.phi, phi)
    e = orc.efield(n, phi)
    for k, a in zip(("ex", "ey", "ez"), e):
        assert np.array_equal(getattr(sim, k), a)
    for s, d in enumerate(sp):
        for i in range(0, len(d["x"]), 37):
            pos = (d["x"][i], d["y"][i], d["z"][i])
            g = orc.gather_periodic(n, e, pos, 0)
            eb = np.array([g[0], g[1], g[2], 0.0, 0.0, 0.5])
            p, v = orc.boris_move(pos, (d["vx"][i], d["vy"][i], d["vz"][i]), d["qm"], eb, 0.1)
            for a, k in enumerate("xyz"):
                w = orc.apply_periodic(n[a], orc.apply_periodic(n[a], p[a]))
                assert sim.sp[s][k][i] == w, (s, i, k)
                assert sim.sp[s]["v" + k][i] == v[a], (s, i, k)


def test_step_conserves_momentum_without_b(orc):
    n = (8, 8, 8)
    sp = _two_species(orc, n, 4)
    sim = orc.OracleSim(n, sp, 0.1, (0, 0, 0), 0)
    p0 = sum((d["q"] / d["qm"]) * np.array([d["vx"].sum(), d["vy"].sum(), d["vz"].sum()])
             for d in sp)
    sim.step(3)
    p1 = sum((d["q"] / d["qm"]) * np.array([s["vx"].sum(), s["vy"].sum(), s["vz"].sum()])
             for d, s in zip(sp, sim.sp))
    assert np.max(np.abs(p1 - p0)) < 1e-11
    assert abs(sim.net_charge) < 1e-11
    for s in sim.sp:
        assert np.all((s["x"] >= 0) & (s["x"] < 8))


def test_step_energy_nearly_conserved(orc):
    n = (8, 8, 8)
    sp = _two_species(orc, n, 8)
    sim = orc.OracleSim(n, sp, 0.1, (0, 0, 0.5), 0)
    tot = []
    for _ in range(30):
        sim.step(1)
        tot.append(0.5 * (sim.ke_old.sum() + sim.ke_new.sum()) + sim.field_energy)
    tot = np.array(tot)
    assert np.max(np.abs(tot - tot[0])) / tot[0] < 0.02


def test_cell_counts(orc):
    n = (4, 3, 2)
    x = np.array([0.0, 3.999, 1.5, 1.5])
    y = np.array([0.0, 2.5, 0.2, 0.2])
    z = np.array([0.0, 1.0, 1.9, 1.9])
    c = orc.cell_counts(n, x, y, z)
    assert c.sum() == 4
    assert c[0] == 1 and c[3 + 4 * (2 + 3 * 1)] == 1 and c[1 + 4 * (0 + 3 * 1)] == 2


def _state(sim):
    return ([sim.rho, sim.phi, sim.ex, sim.ey, sim.ez] +
            [s[k] for s in sim.sp for k in ("x", "y", "z", "vx", "vy", "vz")])


def test_threaded_step_on_one_thread_is_the_serial_step(orc):
    # the threaded form runs the same per-particle and per-line functions; on
    # one thread even its sums come out in the serial order
    n = (12, 8, 10)
    a = orc.OracleSim(n, _two_species(orc, n, 5), 0.1, (0, 0, 0.5), 0)
    b = orc.OracleSim(n, _two_species(orc, n, 5), 0.1, (0, 0, 0.5), 0)
    a.step(3)
    assert b.step_threads(3, 1) == 1
    for u, v in zip(_state(a), _state(b)):
        assert np.array_equal(u, v)
    assert np.array_equal(a.ke_new, b.ke_new) and a.field_energy == b.field_energy


@pytest.mark.parametrize("threads", [2, 3, 8])
def test_threaded_step_matches_serial(orc, threads):
    # several threads: only the order of the charge sums changes
    n = (16, 12, 8)
    a = orc.OracleSim(n, _two_species(orc, n, 7), 0.1, (0, 0, 0.5), 0)
    b = orc.OracleSim(n, _two_species(orc, n, 7), 0.1, (0, 0, 0.5), 0)
    a.step(3)
    assert b.step_threads(3, threads) == threads
    for u, v in zip(_state(a), _state(b)):
        assert np.linalg.norm(u - v) <= 1e-12 * np.linalg.norm(u)
    assert np.allclose(a.ke_new, b.ke_new, rtol=1e-12, atol=0)
    assert abs(a.field_energy - b.field_energy) <= 1e-12 * abs(a.field_energy)
