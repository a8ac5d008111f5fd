"""Bins short of free slots move between steps instead of overflowing into
a full re-bin.

An expanding Gaussian hump of nearly free-streaming particles (q/m 1e-6)
fills the bins around it a little every step.  At slack 0.1 the bins there
get tight: the run rebalances (each tight bin moves, with fresh slack, into
the relocation pool at the end of the particle buffers) and never repacks,
and its fields, cell counts and particles equal, bit for bit, those of the
same run at slack 0.6, where no bin gets tight, and the FP64 oracle's.  Over
160 steps the moved bins fill the pool, and the run compacts every species
into the other buffer instead of re-binning it.  Bin layout is storage only: the reference has
no counterpart (its particles live in per-process arrays that
ohhelppic3d.f90:311-358 steps in place)."""
import numpy as np
import pytest

from helpers import by_pid, rel_l2, sim_for

pytestmark = pytest.mark.gpu

N = (32, 32, 32)
PPC = 8
STEPS = 48
SLACK = 0.1
TOL = 1e-12


@pytest.fixture(scope="module")
def ohp():
    import ohhelppic3d_amd as o
    if not o.available():
        pytest.fail("libohpic.so is not built: the GPU path has no fallback")
    return o


def expanding_hump(orc):
    """Two species, both nearly free-streaming humps (amplitude 9, sigma 3)
    with thermal speeds 0.3 and 0.15 cells per unit time."""
    npcl = PPC * N[0] * N[1] * N[2]
    sp = []
    for s, (qm, q, vt) in enumerate(((-1e-6, -1e-3, 0.3), (1e-6, 1e-3, 0.15))):
        rng0 = orc.Pcg32(42, 52)
        x, y, z, vx, vy, vz = orc.load_hump_maxwell(rng0, s << 40, 2052, npcl, N, 9.0, 3.0, vt)
        sp.append(dict(qm=qm, q=q, x=x, y=y, z=z, vx=vx, vy=vy, vz=vz))
    return sp


def snapshot(ohp, sim):
    return {"fields": [sim.grid(f) for f in (ohp.FIELD_RHO, ohp.FIELD_PHI, ohp.FIELD_EX,
                                            ohp.FIELD_EY, ohp.FIELD_EZ)],
            "parts": [by_pid(sim.store_particles(s, True)) for s in range(2)],
            "counts": sim.cell_counts(), "diag": sim.diagnostics()}


def assert_same(a, b):
    for k, (fa, fb) in enumerate(zip(a["fields"], b["fields"])):
        assert np.array_equal(fa, fb), ("field", k)
    for s in range(2):
        for key in ("pid", "x", "y", "z", "vx", "vy", "vz"):
            assert np.array_equal(a["parts"][s][key], b["parts"][s][key]), (s, key)
    assert np.array_equal(a["counts"], b["counts"])


def run(ohp, sp, precision, slack, steps=STEPS):
    with sim_for(sp, N, b0=(0.0, 0.0, 0.0), weight_order=1, precision=precision,
                 slack=slack) as sim:
        sim.step(steps)
        return snapshot(ohp, sim)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tight_bins_are_rebalanced_not_repacked(ohp, orc, precision):
    prec = ohp.FP64 if precision == "fp64" else ohp.MIXED
    sp = expanding_hump(orc)
    tight = run(ohp, sp, prec, SLACK)
    wide = run(ohp, sp, prec, 0.6)
    print(precision, "rebalances", tight["diag"]["rebalances"], "repacks",
          tight["diag"]["repacks"], "| wide:", wide["diag"]["rebalances"],
          wide["diag"]["repacks"])
    assert tight["diag"]["rebalances"] >= 1
    assert tight["diag"]["repacks"] == 0
    assert wide["diag"]["rebalances"] == 0 and wide["diag"]["repacks"] == 0
    assert_same(tight, wide)
    if prec == ohp.FP64:
        osim = orc.OracleSim(N, sp, 0.1, (0.0, 0.0, 0.0), 1)
        osim.step(STEPS)
        assert rel_l2(tight["fields"][0], osim.rho) <= TOL
        assert rel_l2(tight["fields"][2], osim.ex) <= TOL
        for s in range(2):
            for k in ("x", "y", "z", "vx", "vy", "vz"):
                assert rel_l2(tight["parts"][s][k], osim.sp[s][k]) <= TOL, (s, k)
        ref = sum(orc.cell_counts(N, osim.sp[s]["x"], osim.sp[s]["y"], osim.sp[s]["z"])
                  for s in range(2))
        assert np.array_equal(tight["counts"], ref)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_bins_beyond_the_device_list_are_found_from_the_counts(ohp, orc, monkeypatch,
                                                               precision):
    # OHPIC_TIGHT_LIST_LIMIT=0: every step's listed bins count as more than the
    # device list holds, so each relocation finds them from the downloaded bin
    # counts instead; the bins it moves, and so the run, are the same
    prec = ohp.FP64 if precision == "fp64" else ohp.MIXED
    sp = expanding_hump(orc)
    listed = run(ohp, sp, prec, SLACK)
    monkeypatch.setenv("OHPIC_TIGHT_LIST_LIMIT", "0")
    counted = run(ohp, sp, prec, SLACK)
    monkeypatch.delenv("OHPIC_TIGHT_LIST_LIMIT")
    for key in ("rebalances", "compactions", "repacks", "step"):
        assert counted["diag"][key] == listed["diag"][key], key
    assert counted["diag"]["rebalances"] >= 1 and counted["diag"]["repacks"] == 0
    assert_same(counted, listed)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pool_exhaustion_compacts_instead_of_repacking(ohp, orc, precision):
    # over 160 steps the moved bins fill the relocation pool: every species is
    # then compacted into the other buffer (fresh pools), never re-binned
    prec = ohp.FP64 if precision == "fp64" else ohp.MIXED
    steps = 160
    sp = expanding_hump(orc)
    tight = run(ohp, sp, prec, SLACK, steps)
    wide = run(ohp, sp, prec, 0.6, steps)
    d = tight["diag"]
    print(precision, "rebalances", d["rebalances"], "compactions", d["compactions"],
          "repacks", d["repacks"])
    assert d["compactions"] >= 1 and d["repacks"] == 0 and d["step"] == steps
    assert_same(tight, wide)
    if prec == ohp.FP64:
        osim = orc.OracleSim(N, sp, 0.1, (0.0, 0.0, 0.0), 1)
        osim.step(steps)
        assert rel_l2(tight["fields"][0], osim.rho) <= TOL
        for s in range(2):
            for k in ("x", "vx"):
                assert rel_l2(tight["parts"][s][k], osim.sp[s][k]) <= TOL, (s, k)


def test_group_member_rebalances_between_sequences(ohp, orc):
    # the tight bins are found inside graph-replayed sequences of 4 steps and
    # moved after each sequence; the member equals the run alone
    sp = expanding_hump(orc)
    alone = run(ohp, sp, ohp.FP64, SLACK)
    other = [dict(d) for d in sp]
    with sim_for(sp, N, b0=(0.0, 0.0, 0.0), weight_order=1, slack=SLACK) as a, \
            sim_for(other, N, b0=(0.0, 0.0, 0.0), weight_order=1, slack=0.6) as b:
        with ohp.Group(use_graph=True, persistent=True, batch=4) as grp:
            grp.attach(a)
            grp.attach(b)
            grp.step(STEPS)
        got = snapshot(ohp, a)
        wide = snapshot(ohp, b)
    print("group member rebalances", got["diag"]["rebalances"], "repacks",
          got["diag"]["repacks"])
    assert got["diag"]["step"] == STEPS
    assert got["diag"]["rebalances"] >= 1 and got["diag"]["repacks"] == 0
    assert_same(got, alone)
    assert_same(wide, alone)


def test_group_member_compacts_between_sequences(ohp, orc):
    # 160 steps in sequences of 4: the member's pool runs out and it is
    # compacted between sequences (its parity and charge grids flip outside
    # the group's graph); it equals the run alone, which compacts after single
    # steps
    steps = 160
    sp = expanding_hump(orc)
    alone = run(ohp, sp, ohp.MIXED, SLACK, steps)
    with sim_for(sp, N, b0=(0.0, 0.0, 0.0), weight_order=1, precision=ohp.MIXED,
                 slack=SLACK) as a, \
            sim_for([dict(d) for d in sp], N, b0=(0.0, 0.0, 0.0), weight_order=1,
                    precision=ohp.MIXED, slack=0.6) as b:
        with ohp.Group(use_graph=True, persistent=True, batch=4) as grp:
            grp.attach(a)
            grp.attach(b)
            grp.step(steps)
        got = snapshot(ohp, a)
        wide = snapshot(ohp, b)
    d = got["diag"]
    print("group member rebalances", d["rebalances"], "compactions", d["compactions"],
          "repacks", d["repacks"])
    assert d["step"] == steps and d["compactions"] >= 1 and d["repacks"] == 0
    assert_same(got, alone)
    assert_same(wide, alone)
