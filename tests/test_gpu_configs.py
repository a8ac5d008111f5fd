"""The batched-ensemble and serving configurations at their own shapes,
checked against the FP64 oracle.

* The 128-instance ensemble (each axis from {16,24,32,48,64}, 8-64 particles
  per cell per species, seed 2024) stepped as one group -- one persistent
  particle launch per step, steps replayed from a captured hipGraph in
  sequences of 8 -- with size-independent checks on every member, and two of
  its members re-run alone, equal to the group's bit for bit.
* Members of that ensemble (a 64-cell axis, a radix-3 axis, the smallest;
  particles per cell scaled down so the oracle finishes in seconds) stepped
  as a group against OracleSim from the same seeded state: FP64 within
  1e-12, the mixed form within 1e-5, with and without graph replay and the
  persistent launch.
* A serving mix shaped like the serving configuration (32^3, 64^3 and 96^3
  requests, 16-32 particles per cell, overlapping arrivals, shortened step
  counts): every served run equals the same run alone bit for bit, and a
  64^3 request served in FP64 equals the oracle's run of it within 1e-11 (in
  the mixed form within 1e-5).
The reference runs one simulation per MPI job (ohhelppic3d.f90:286-381);
each member's step is that step."""
import numpy as np
import pytest

from helpers import by_pid, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-12
TOL_MIXED = 1e-5


@pytest.fixture(scope="module")
def ohp():
    import ohhelppic3d_amd as o
    if not o.available():
        pytest.fail("libohpic.so is not built: the GPU path has no fallback")
    return o


def oracle_for(orc, spec, weight_order):
    """OracleSim of an instance: the same seeded load as ohpic_seed_uniform_maxwell
    (one PCG32 stream (seed, 52), electrons then ions)."""
    rng = orc.Pcg32(spec.seed, 52)
    sp = []
    for s, (qm, vt) in enumerate(((-1.0, spec.vth_e), (0.01, spec.vth_e / 10.0))):
        x, y, z, vx, vy, vz = orc.load_uniform_maxwell(rng, spec.counts[s], spec.n, vt)
        q = (-1.0 / spec.ppc[0]) if s == 0 else (1.0 / spec.ppc[1])
        sp.append(dict(qm=qm, q=q, x=x, y=y, z=z, vx=vx, vy=vy, vz=vz))
    return orc.OracleSim(spec.n, sp, spec.dt, (0.0, 0.0, spec.omega_ce), weight_order)


def compare_with_oracle(ohp, orc, sim, osim, tol, exact_counts):
    err = {"rho": rel_l2(sim.grid(ohp.FIELD_RHO), osim.rho),
           "phi": rel_l2(sim.grid(ohp.FIELD_PHI), osim.phi),
           "Ex": rel_l2(sim.grid(ohp.FIELD_EX), osim.ex),
           "Ey": rel_l2(sim.grid(ohp.FIELD_EY), osim.ey),
           "Ez": rel_l2(sim.grid(ohp.FIELD_EZ), osim.ez)}
    n = osim.n
    ref_counts = np.zeros(n[0] * n[1] * n[2], np.int32)
    for s in range(2):
        g, r = by_pid(sim.store_particles(s, True)), osim.sp[s]
        assert np.array_equal(g["pid"], np.arange(len(r["x"])))
        for k in ("x", "y", "z", "vx", "vy", "vz"):
            err[f"{k}{s}"] = rel_l2(g[k], r[k])
        ref_counts += orc.cell_counts(n, r["x"], r["y"], r["z"])
    bad = int(np.sum(sim.cell_counts() != ref_counts))
    for k, v in err.items():
        assert v <= tol, (k, v)
    if exact_counts:
        assert bad == 0
    else:
        # the mixed form's cells differ only for particles within float
        # rounding of a face
        assert bad <= 1e-5 * ref_counts.sum() + 4
    return err


# ------------------------------------------------------------------ ensemble
def test_ensemble128_group_full_size(ohp):
    from ohhelppic3d_amd import ensemble
    specs = ensemble.sample_ensemble(128, seed=2024)
    assert len(specs) == 128
    axes = {a for s in specs for a in s.n}
    assert axes == set(ensemble.AXES)
    tile, chunks = (8, 8, 4), 1
    sims = [s.create(tile, chunks, ohp.MIXED) for s in specs]
    try:
        with ohp.Group(use_graph=True, persistent=True, batch=8) as grp:
            for sim in sims:
                grp.attach(sim)
            grp.step(1)
            d0 = [sim.diagnostics() for sim in sims]
            grp.step(8)
            st = grp.stats()
        assert st["steps"] == 9 and st["captures"] >= 1
        for i, (spec, sim) in enumerate(zip(specs, sims)):
            d = sim.diagnostics()
            assert d["step"] == 9, i
            for s in range(2):
                assert sim.particle_count(s) == spec.counts[s], (i, s)
            assert int(sim.cell_counts().sum()) == spec.particles, i
            # neutral box: sum rho at the charge grid's rounding
            assert abs(d["net_charge"]) <= 1e-9 * spec.cells, (i, d["net_charge"])
            # total energy kept within 1% over the 8 steps; the kinetic energy
            # within 5% of the loaded Maxwellians' (the field takes its
            # thermal-noise share, up to ~1/(3 ppc) of it, in the first
            # plasma period)
            e0 = d0[i]["total_energy"]
            assert np.isfinite(d["total_energy"])
            assert abs(d["total_energy"] - e0) <= 0.01 * e0, (i, d["total_energy"], e0)
            ke, ke0 = d["ke_new"].sum(), d0[i]["ke_old"].sum()
            assert abs(ke - ke0) <= 0.05 * ke0, (i, ke, ke0)
            assert d["repacks"] <= 1, i
        # two members again, alone: equal to the group's result bit for bit
        for i in (0, int(np.argmax([s.cells for s in specs]))):
            with specs[i].create(tile, chunks, ohp.MIXED) as alone:
                alone.step(9)
                for f in (ohp.FIELD_RHO, ohp.FIELD_PHI, ohp.FIELD_EX):
                    assert np.array_equal(alone.grid(f), sims[i].grid(f)), (i, f)
                assert np.array_equal(alone.cell_counts(), sims[i].cell_counts()), i
    finally:
        for sim in sims:
            sim.close()


def _ensemble_members(ppc_cap=12):
    """Three members of the 128-instance ensemble: one with a 64-cell axis,
    one with a radix-3 (24 or 48) axis, and the smallest; particles per cell
    capped so the oracle steps them in seconds."""
    from dataclasses import replace
    from ohhelppic3d_amd import ensemble
    specs = ensemble.sample_ensemble(128, seed=2024)
    a = next(s for s in specs if 64 in s.n and max(s.n) * min(s.n) <= 64 * 32)
    b = next(s for s in specs if (24 in s.n or 48 in s.n) and 64 not in s.n)
    c = min(specs, key=lambda s: s.cells)
    out = []
    for s in (a, b, c):
        out.append(replace(s, ppc=tuple(min(p, ppc_cap) for p in s.ppc)))
    return out


@pytest.mark.parametrize("graph,persistent", [(True, True), (False, False)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_ensemble_members_match_oracle(ohp, orc, precision, graph, persistent):
    prec = ohp.FP64 if precision == "fp64" else ohp.MIXED
    members = _ensemble_members()
    wo = ohp.WEIGHTS_STANDARD  # the ensemble's weight order
    steps = 4
    sims = [s.create((8, 8, 4), 1, prec, track_ids=True, weight_order=wo) for s in members]
    try:
        with ohp.Group(use_graph=graph, persistent=persistent, batch=steps) as grp:
            for sim in sims:
                grp.attach(sim)
            grp.step(steps)  # one launch sequence of 4 steps
        for spec, sim in zip(members, sims):
            osim = oracle_for(orc, spec, wo)
            osim.step(steps)
            err = compare_with_oracle(ohp, orc, sim, osim,
                                      TOL if prec == ohp.FP64 else TOL_MIXED,
                                      exact_counts=prec == ohp.FP64)
            print(precision, spec.n, spec.ppc, "worst rel L2",
                  f"{max(err.values()):.2e}")
    finally:
        for sim in sims:
            sim.close()


# ------------------------------------------------------------------- serving
def _serving_mix(serve):
    R = serve.Request
    # grids of the serving configuration (weights 4:3:2 of 32^3/64^3/96^3),
    # arrivals in group steps so the mix is reproducible
    return [R(0, 0, (32, 32, 32), (16, 24), 12, 1.0, 0.5, 101),
            R(1, 0, (64, 64, 64), (20, 16), 8, 1.4, 0.3, 102),
            R(2, 2, (32, 32, 32), (32, 20), 9, 0.6, 0.9, 103),
            R(3, 3, (96, 96, 96), (16, 16), 6, 1.9, 0.1, 104),
            R(4, 5, (64, 64, 64), (16, 28), 10, 0.8, 0.6, 105),
            R(5, 6, (32, 32, 32), (24, 16), 7, 1.2, 0.0, 106)]


def _snap(ohp, sim):
    return {"fields": [sim.grid(f) for f in (ohp.FIELD_RHO, ohp.FIELD_PHI, ohp.FIELD_EX,
                                            ohp.FIELD_EY, ohp.FIELD_EZ)],
            "parts": [by_pid(sim.store_particles(s, True)) for s in range(2)],
            "counts": sim.cell_counts(), "step": sim.diagnostics()["step"]}


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_serving_mix_matches_alone_and_oracle(ohp, orc, precision):
    from ohhelppic3d_amd import serve
    prec = ohp.FP64 if precision == "fp64" else ohp.MIXED
    reqs = _serving_mix(serve)
    kw = dict(tile=(8, 8, 8), chunks=1, track_ids=True, precision=prec)
    mix = {}
    loop = serve.ServingLoop(**kw)
    rep = loop.run(reqs, clock="steps",
                   on_complete=lambda r, sim: mix.__setitem__(r.rid, _snap(ohp, sim)))
    assert rep.summary()["done"] == len(reqs), [(o.rid, o.status, o.error)
                                               for o in rep.outcomes]
    assert rep.mean_active > 1.5  # runs overlapped on the device
    for r in reqs:
        with r.spec.create(loop.tile, loop.chunks, prec, track_ids=True,
                           weight_order=loop.weight_order) as sim:
            sim.step(r.steps)
            a = _snap(ohp, sim)
            if r.rid == 1:
                # the 64^3 request against the oracle's run of it
                osim = oracle_for(orc, r.spec, loop.weight_order)
                osim.step(r.steps)
                err = compare_with_oracle(ohp, orc, sim, osim,
                                          1e-11 if prec == ohp.FP64 else TOL_MIXED,
                                          exact_counts=prec == ohp.FP64)
                print(precision, "served 64^3 request vs oracle, worst rel L2",
                      f"{max(err.values()):.2e}")
        m = mix[r.rid]
        assert a["step"] == m["step"] == r.steps
        for k, (fa, fm) in enumerate(zip(a["fields"], m["fields"])):
            assert np.array_equal(fa, fm), (r.rid, "field", k)
        for s in range(2):
            for key in ("pid", "x", "y", "z", "vx", "vy", "vz"):
                assert np.array_equal(a["parts"][s][key], m["parts"][s][key]), (r.rid, s, key)
        assert np.array_equal(a["counts"], m["counts"]), r.rid
