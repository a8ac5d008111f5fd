"""GPU checks of instance groups (ohpic_group_*): instances of different,
non-cubic grid sizes advanced together give, member by member, the same
fields, particles and cell counts as the same instances stepped alone with
ohpic_step -- bit for bit, since the charge deposit is integer and every
per-particle update depends only on that particle and the fields.  The
kinetic-energy sums are floating-point sums over particles whose order inside
a bin is not fixed, so they are compared to 1e-12 relative."""
import numpy as np
import pytest

from helpers import by_pid, sim_for

pytestmark = pytest.mark.gpu

# (grid, particles per cell per species, seed); axis lengths from {16,24,32,48}
CASES = [((16, 24, 32), 6, 1), ((32, 16, 16), 10, 2), ((24, 24, 48), 3, 3),
         ((16, 16, 16), 16, 4)]
KE_TOL = 1e-12


@pytest.fixture(scope="module")
def ohp():
    import ohhelppic3d_amd as o
    if not o.available():
        pytest.fail("libohpic.so is not built: the GPU path has no fallback")
    return o


def make(ohp, n, ppc, seed, precision, tile=(4, 4, 4), chunks=0, b0z=0.5, vth=(1.0, 0.1)):
    g = n[0] * n[1] * n[2]
    npart = ppc * g
    cfg = ohp.SimConfig(n=n, qm=(-1.0, 0.01), q_macro=(-1.0 / ppc, 1.0 / ppc),
                        capacity=(npart, npart), dt=0.1, b0=(0.0, 0.0, b0z),
                        precision=precision, tile=tile, chunks=chunks, track_ids=True)
    sim = ohp.Simulation(cfg)
    sim.seed_uniform_maxwell([npart, npart], vth, seed=(seed, 52))
    return sim


def snapshot(ohp, sim):
    out = {"fields": [sim.grid(f) for f in (ohp.FIELD_RHO, ohp.FIELD_PHI, ohp.FIELD_EX,
                                            ohp.FIELD_EY, ohp.FIELD_EZ)],
           "parts": [by_pid(sim.store_particles(s, True)) for s in range(sim.nspecies)],
           "counts": sim.cell_counts(-1),
           "diag": sim.diagnostics()}
    return out


def assert_same(a, b, what):
    for k, (fa, fb) in enumerate(zip(a["fields"], b["fields"])):
        assert np.array_equal(fa, fb), (what, "field", k)
    for s, (pa, pb) in enumerate(zip(a["parts"], b["parts"])):
        for key in ("pid", "x", "y", "z", "vx", "vy", "vz"):
            assert np.array_equal(pa[key], pb[key]), (what, s, key)
    assert np.array_equal(a["counts"], b["counts"]), what
    da, db = a["diag"], b["diag"]
    assert da["step"] == db["step"], what
    for key in ("field_energy", "field_energy_e2", "net_charge"):
        assert da[key] == db[key], (what, key)
    for s in range(len(a["parts"])):
        for key in ("ke_old", "ke_new"):
            x, y = da[key][s], db[key][s]
            assert abs(x - y) <= KE_TOL * max(abs(y), 1e-300), (what, key, s)


def run_alone(ohp, precision, steps, **kw):
    snaps = []
    for n, ppc, seed in CASES:
        with make(ohp, n, ppc, seed, precision, **kw) as sim:
            sim.step(steps)
            snaps.append(snapshot(ohp, sim))
    return snaps


def run_group(ohp, precision, steps, use_graph=True, persistent=True, batch=None, **kw):
    sims = [make(ohp, n, ppc, seed, precision, **kw) for n, ppc, seed in CASES]
    try:
        with ohp.Group(use_graph=use_graph, persistent=persistent, batch=batch) as grp:
            for s in sims:
                grp.attach(s)
            assert len(grp) == len(CASES)
            grp.step(steps)
            stats = grp.stats()
            snaps = [snapshot(ohp, s) for s in sims]
        return snaps, stats
    finally:
        for s in sims:
            s.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_group_equals_alone(ohp, precision):
    prec = ohp.FP64 if precision == "fp64" else ohp.MIXED
    alone = run_alone(ohp, prec, 3)
    grouped, stats = run_group(ohp, prec, 3)
    assert stats["steps"] == 3 and stats["captures"] >= 1
    for i, (a, g) in enumerate(zip(alone, grouped)):
        assert_same(g, a, f"{precision} member {i}")


def test_group_graph_equals_direct_launches(ohp):
    with_graph, st_g = run_group(ohp, ohp.MIXED, 4, use_graph=True, tile=(8, 8, 8), chunks=2)
    direct, st_d = run_group(ohp, ohp.MIXED, 4, use_graph=False, tile=(8, 8, 8), chunks=2)
    assert st_g["captures"] >= 1 and st_d["captures"] == 0
    # one thread: every sequence of the graph group is a replay, none of the
    # other's is
    assert st_g["replayed"] >= 1 and st_g["direct"] == 0
    assert st_d["replayed"] == 0 and st_d["direct"] >= 1
    for i, (a, b) in enumerate(zip(with_graph, direct)):
        assert_same(a, b, f"member {i}")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_group_persistent_launch_equals_per_item_launch(ohp, precision):
    # one resident grid claiming work items from a counter vs one workgroup
    # per work item: same particles, fields and counts
    prec = ohp.FP64 if precision == "fp64" else ohp.MIXED
    queued, _ = run_group(ohp, prec, 3, persistent=True, tile=(4, 4, 4), chunks=2)
    per_item, _ = run_group(ohp, prec, 3, persistent=False, tile=(4, 4, 4), chunks=2)
    for i, (a, b) in enumerate(zip(queued, per_item)):
        assert_same(a, b, f"{precision} member {i}")


@pytest.mark.parametrize("use_graph", [True, False])
def test_group_step_sequences_equal_alone(ohp, use_graph):
    # 10 steps in sequences of 4: two 4-step launches, then two single steps
    alone = run_alone(ohp, ohp.MIXED, 10)
    grouped, stats = run_group(ohp, ohp.MIXED, 10, use_graph=use_graph, batch=4)
    assert stats["steps"] == 10
    if use_graph:
        assert stats["captures"] >= 2  # the 4-step and the 1-step sequence
    for i, (a, g) in enumerate(zip(alone, grouped)):
        assert_same(g, a, f"member {i}")


def test_group_captures_while_other_threads_create_runs(ohp):
    # the group steps, capturing again before every sequence (a member leaves
    # and rejoins), while two threads create, seed, step and destroy contexts;
    # a capture is only taken while no other thread is inside a library call,
    # so no capture fails, no other thread's call fails, and the members equal
    # the same runs stepped alone
    import threading
    import time
    stop = threading.Event()
    errors, made = [], [0, 0]

    def churn(i):
        # a pause after each context leaves moments with no thread inside the
        # library, in which captures go through
        while not stop.is_set():
            try:
                with make(ohp, (16, 16, 24), 4, 100 + i, ohp.MIXED) as s:
                    s.step(1)
                made[i] += 1
            except Exception as e:  # reported below, not raised in the thread
                errors.append(repr(e))
                return
            time.sleep(0.01)

    sims = [make(ohp, n, ppc, seed, ohp.MIXED) for n, ppc, seed in CASES]
    threads = [threading.Thread(target=churn, args=(i,)) for i in range(2)]
    try:
        with ohp.Group(batch=2) as grp:
            for s in sims:
                grp.attach(s)
            for t in threads:
                t.start()
            seqs = 0
            # at least 10 sequences, and until each thread has made 3 contexts
            while seqs < 200 and (seqs < 10 or min(made) < 3) and not errors:
                grp.detach(sims[0])
                grp.attach(sims[0])
                grp.step(2)
                seqs += 1
            stop.set()
            for t in threads:
                t.join()
            stats = grp.stats()
            last = grp.last_error()
            grouped = [snapshot(ohp, s) for s in sims]
    finally:
        stop.set()
        for t in threads:
            if t.is_alive():
                t.join()
        for s in sims:
            s.close()
    print("contexts made by the other threads", made, "| sequences replayed",
          stats["replayed"], "direct", stats["direct"], "captures", stats["captures"])
    assert not errors, errors
    assert last == "", last
    assert min(made) >= 1
    assert stats["steps"] == 2 * seqs and stats["replayed"] + stats["direct"] == seqs
    assert stats["captures"] == stats["replayed"]  # every capture is replayed once
    alone = run_alone(ohp, ohp.MIXED, 2 * seqs)
    for i, (a, g) in enumerate(zip(alone, grouped)):
        assert_same(g, a, f"member {i}")


def test_two_groups_stepped_from_two_threads(ohp):
    # two groups, each needing a capture before every sequence, stepped at
    # once from two threads: a capture is only taken while the other thread
    # is outside the library (with both stepping back to back, that is
    # seldom, and those sequences are enqueued directly), so both groups'
    # members equal the runs alone
    import threading
    halves = [CASES[:2], CASES[2:]]
    seqs = 8
    errors = []
    out = [None, None]

    def drive(h):
        sims = [make(ohp, n, ppc, seed, ohp.MIXED) for n, ppc, seed in halves[h]]
        try:
            with ohp.Group(batch=2) as grp:
                for s in sims:
                    grp.attach(s)
                for _ in range(seqs):
                    grp.detach(sims[0])
                    grp.attach(sims[0])
                    grp.step(2)
                out[h] = ([snapshot(ohp, s) for s in sims], grp.stats(), grp.last_error())
        except Exception as e:  # reported below, not raised in the thread
            errors.append(repr(e))
        finally:
            for s in sims:
                s.close()

    threads = [threading.Thread(target=drive, args=(h,)) for h in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    alone = run_alone(ohp, ohp.MIXED, 2 * seqs)
    for h in range(2):
        snaps, stats, last = out[h]
        print("group", h, "sequences replayed", stats["replayed"], "direct", stats["direct"])
        assert last == "", last
        assert stats["steps"] == 2 * seqs and stats["replayed"] + stats["direct"] == seqs
        assert stats["captures"] == stats["replayed"]
        for i, g in enumerate(snaps):
            assert_same(g, alone[2 * h + i], f"group {h} member {i}")


def drifting_pair(ohp):
    # every particle starts in tile 0 and moves one cell into tile 1, whose
    # bin capacity is the minimum: the step overflows and is redone as a list
    n = (16, 8, 8)
    rng = np.random.default_rng(5)
    m = 3000
    x = rng.uniform(7.0, 8.0, m)
    y = rng.uniform(0, 8, m)
    z = rng.uniform(0, 8, m)
    zero = np.zeros(m)
    sp = [dict(qm=-1.0, q=-1e-3, x=x, y=y, z=z, vx=np.full(m, 10.0), vy=zero, vz=zero),
          dict(qm=0.01, q=1e-3, x=x.copy(), y=y.copy(), z=z.copy(), vx=zero, vy=zero, vz=zero)]
    return sim_for(sp, n, tile=(4, 4, 4), slack=0.01)


def test_group_member_overflow_is_redone(ohp):
    with drifting_pair(ohp) as a, make(ohp, (16, 16, 24), 4, 9, ohp.FP64) as b:
        a.step(3)
        b.step(3)
        ref = [snapshot(ohp, a), snapshot(ohp, b)]
    with drifting_pair(ohp) as a, make(ohp, (16, 16, 24), 4, 9, ohp.FP64) as b:
        with ohp.Group() as grp:
            grp.attach(a)
            grp.attach(b)
            grp.step(3)
            got = [snapshot(ohp, a), snapshot(ohp, b)]
            assert got[0]["diag"]["repacks"] >= 1
            assert got[0]["diag"]["repacks"] == ref[0]["diag"]["repacks"]
    for i in range(2):
        assert_same(got[i], ref[i], f"member {i}")


def test_overflow_inside_a_step_sequence_is_redone(ohp):
    # the overflowing member keeps its state for the rest of the sequence,
    # then the step is redone as a list and its remaining steps run alone
    with drifting_pair(ohp) as a, make(ohp, (16, 16, 24), 4, 9, ohp.FP64) as b:
        a.step(8)
        b.step(8)
        ref = [snapshot(ohp, a), snapshot(ohp, b)]
    with drifting_pair(ohp) as a, make(ohp, (16, 16, 24), 4, 9, ohp.FP64) as b:
        with ohp.Group(batch=4) as grp:
            grp.attach(a)
            grp.attach(b)
            grp.step(8)
            assert grp.stats()["steps"] == 8
            got = [snapshot(ohp, a), snapshot(ohp, b)]
            assert got[0]["diag"]["step"] == 8 and got[1]["diag"]["step"] == 8
            assert got[0]["diag"]["repacks"] >= 1
    for i in range(2):
        assert_same(got[i], ref[i], f"member {i}")


def test_group_detach_then_alone(ohp):
    n, ppc, seed = CASES[0]
    with make(ohp, n, ppc, seed, ohp.MIXED) as ref:
        ref.step(4)
        want = snapshot(ohp, ref)
    with make(ohp, n, ppc, seed, ohp.MIXED) as a, \
            make(ohp, CASES[1][0], CASES[1][1], CASES[1][2], ohp.MIXED) as b:
        grp = ohp.Group()
        grp.attach(a)
        grp.attach(b)
        grp.step(2)
        grp.detach(a)
        assert len(grp) == 1
        grp.step(1)       # b alone in the group
        a.step(1)         # a through the single-instance path
        grp.attach(a)
        grp.step(1)
        grp.close()
        got = snapshot(ohp, a)
    assert_same(got, want, "detached member")


def test_group_attach_errors(ohp):
    with make(ohp, (16, 16, 16), 2, 1, ohp.MIXED) as a, \
            make(ohp, (16, 16, 16), 2, 2, ohp.MIXED, tile=(8, 8, 8)) as b, \
            make(ohp, (16, 16, 16), 2, 3, ohp.FP64) as c:
        with ohp.Group() as grp:
            grp.attach(a)
            with pytest.raises(ohp.OhpicError) as e:
                grp.attach(b)
            assert e.value.code == -1
            with pytest.raises(ohp.OhpicError) as e:
                grp.attach(c)
            assert e.value.code == -1
            with pytest.raises(ohp.OhpicError) as e:
                grp.attach(a)
            assert e.value.code == -4
            with pytest.raises(ohp.OhpicError):
                grp.detach(b)
        # destroying the group detached a: it steps alone again
        a.step(1)


def test_failed_member_is_detached_and_others_continue(ohp):
    # FP64: particles that move more than two box lengths in one step are an
    # error (the reference's wrap is applied twice)
    with make(ohp, (16, 16, 16), 2, 1, ohp.FP64) as ok, \
            make(ohp, (16, 16, 16), 2, 2, ohp.FP64, vth=(5000.0, 0.1)) as hot, \
            make(ohp, (16, 16, 16), 2, 1, ohp.FP64) as ref:
        ref.step(2)
        want = snapshot(ohp, ref)
        with ohp.Group() as grp:
            grp.attach(ok)
            grp.attach(hot)
            with pytest.raises(ohp.OhpicError) as e:
                grp.step(1)
            assert "member 1" in str(e.value) and "box lengths" in str(e.value)
            assert grp.contains(ok) and not grp.contains(hot) and len(grp) == 1
            grp.step(1)
            with pytest.raises(ohp.OhpicError) as e:
                hot.step(1)
            assert e.value.code == -4
            with pytest.raises(ohp.OhpicError):
                grp.attach(hot)
        assert_same(snapshot(ohp, ok), want, "surviving member")


def test_destroying_a_member_detaches_it(ohp):
    a = make(ohp, (16, 16, 16), 2, 1, ohp.MIXED)
    b = make(ohp, (16, 24, 16), 2, 2, ohp.MIXED)
    with ohp.Group() as grp:
        grp.attach(a)
        grp.attach(b)
        grp.step(1)
        a.close()
        assert len(grp) == 1
        grp.step(1)
        assert b.diagnostics()["step"] == 2
    b.close()


def test_autotune_picks_a_timed_candidate(ohp):
    cfgs = []
    for n, ppc, _ in CASES[:3]:
        g = n[0] * n[1] * n[2]
        cfgs.append(ohp.SimConfig(n=n, qm=(-1.0, 0.01), q_macro=(-1.0 / ppc, 1.0 / ppc),
                                  capacity=(ppc * g, ppc * g), precision=ohp.MIXED))
    r = ohp.autotune(cfgs, sample=3, steps=1)
    assert r["candidates"] == 9
    assert r["tile"] in ((4, 4, 4), (8, 8, 8), (8, 8, 4))
    assert r["chunks"] in (1, 2, 4)
    assert r["particle_steps_per_s"] > 0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_group_with_256_point_z_lines_equals_alone(ohp, precision):
    # members whose Poisson z pass runs the 256-point register kernel beside
    # members on the radix-4 LDS stages, in one group: each equals its run
    # alone bit for bit
    prec = ohp.FP64 if precision == "fp64" else ohp.MIXED
    cases = [((8, 8, 256), 2, 7), ((16, 16, 16), 4, 8), ((4, 12, 256), 3, 9)]
    alone = []
    for n, ppc, seed in cases:
        with make(ohp, n, ppc, seed, prec) as sim:
            sim.step(3)
            alone.append(snapshot(ohp, sim))
    sims = [make(ohp, n, ppc, seed, prec) for n, ppc, seed in cases]
    try:
        with ohp.Group() as grp:
            for s in sims:
                grp.attach(s)
            grp.step(3)
        for i, (a, s) in enumerate(zip(alone, sims)):
            assert_same(snapshot(ohp, s), a, f"{precision} member {i}")
    finally:
        for s in sims:
            s.close()
