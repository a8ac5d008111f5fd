"""GPU checks of the serving loop: each run's result is the same alone and in
a mix of runs that arrive, wait for memory, overlap on the device and leave
at different steps (bit for bit: fields, particles by id, cell counts); a run
larger than the budget is rejected and a run that fails mid-step is dropped
while the others finish; a run holds no more device memory than the
admission estimate."""
import numpy as np
import pytest

from helpers import by_pid

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ohp():
    import ohhelppic3d_amd as o
    if not o.available():
        pytest.fail("libohpic.so is not built: the GPU path has no fallback")
    return o


def requests(serve):
    R = serve.Request
    return [R(0, 0, (16, 16, 16), (6, 4), 5, 1.0, 0.5, 11),
            R(1, 0, (24, 16, 32), (3, 5), 7, 1.5, 0.2, 12),
            R(2, 1, (16, 32, 16), (4, 4), 3, 0.7, 0.9, 13),
            R(3, 3, (32, 32, 32), (2, 3), 4, 2.0, 0.0, 14),
            R(4, 3, (16, 16, 24), (8, 2), 6, 0.5, 0.7, 15)]


def snapshot(ohp, sim):
    return {"fields": [sim.grid(f) for f in (ohp.FIELD_RHO, ohp.FIELD_PHI, ohp.FIELD_EX,
                                            ohp.FIELD_EY, ohp.FIELD_EZ)],
            "parts": [by_pid(sim.store_particles(s, True)) for s in range(2)],
            "counts": sim.cell_counts(-1), "step": sim.diagnostics()["step"]}


def test_serving_mix_equals_alone(ohp):
    from ohhelppic3d_amd import serve
    reqs = requests(serve)
    kw = dict(tile=(4, 4, 4), chunks=1, track_ids=True)
    probe = serve.ServingLoop(**kw)
    fps = [probe.footprint(r) for r in reqs]
    # room for about three runs at once: later arrivals wait for memory
    budget = sum(sorted(fps)[-3:])
    mix = {}
    loop = serve.ServingLoop(budget_bytes=budget, **kw)
    rep = loop.run(reqs, clock="steps",
                   on_complete=lambda r, sim: mix.__setitem__(r.rid, snapshot(ohp, sim)))
    summ = rep.summary()
    assert summ["done"] == len(reqs), [(o.rid, o.status, o.error) for o in rep.outcomes]
    assert rep.mean_active > 1.5           # runs overlapped on the device
    # no other thread makes HIP calls: every launch sequence is a graph replay
    assert summ["graph_sequences"] >= 1 and summ["direct_sequences"] == 0
    assert rep.peak_bytes <= budget
    waited = [o for o in rep.outcomes if o.admitted > o.arrival]
    assert waited                          # some run waited for memory
    for r in reqs:
        alone = {}
        one = serve.Request(**{**r.to_dict(), "arrival": 0.0, "n": tuple(r.n),
                               "ppc": tuple(r.ppc)})
        serve.ServingLoop(**kw).run(
            [one], clock="steps",
            on_complete=lambda q, sim: alone.__setitem__(q.rid, snapshot(ohp, sim)))
        a, m = alone[r.rid], mix[r.rid]
        assert a["step"] == m["step"] == r.steps
        for k, (fa, fm) in enumerate(zip(a["fields"], m["fields"])):
            assert np.array_equal(fa, fm), (r.rid, "field", k)
        for s in range(2):
            for key in ("pid", "x", "y", "z", "vx", "vy", "vz"):
                assert np.array_equal(a["parts"][s][key], m["parts"][s][key]), (r.rid, s, key)
        assert np.array_equal(a["counts"], m["counts"]), r.rid


def test_wall_clock_serving_creates_runs_while_stepping(ohp):
    # wall clock: runs arriving while others step are created by worker
    # threads and join later; each result still equals the run alone
    from ohhelppic3d_amd import serve
    R = serve.Request
    reqs = [R(0, 0.0, (32, 32, 32), (8, 8), 60, 1.0, 0.5, 21),
            R(1, 0.02, (16, 24, 16), (4, 6), 30, 1.5, 0.2, 22),
            R(2, 0.04, (24, 16, 32), (5, 3), 25, 0.7, 0.9, 23),
            R(3, 0.06, (16, 16, 16), (6, 4), 40, 2.0, 0.0, 24)]
    kw = dict(tile=(8, 8, 8), chunks=1, track_ids=True)
    mix = {}
    rep = serve.ServingLoop(creators=2, **kw).run(
        reqs, clock="wall", on_complete=lambda r, sim: mix.__setitem__(r.rid, snapshot(ohp, sim)))
    summ = rep.summary()
    assert summ["done"] == len(reqs), [(o.rid, o.status, o.error) for o in rep.outcomes]
    print("sequences replayed", summ["graph_sequences"], "direct", summ["direct_sequences"])
    for r in reqs:
        one = serve.Request(**{**r.to_dict(), "arrival": 0.0, "n": tuple(r.n),
                               "ppc": tuple(r.ppc)})
        alone = {}
        serve.ServingLoop(creators=0, **kw).run(
            [one], clock="steps",
            on_complete=lambda q, sim: alone.__setitem__(q.rid, snapshot(ohp, sim)))
        a, m = alone[r.rid], mix[r.rid]
        assert a["step"] == m["step"] == r.steps
        for k, (fa, fm) in enumerate(zip(a["fields"], m["fields"])):
            assert np.array_equal(fa, fm), (r.rid, "field", k)
        for sp in range(2):
            for key in ("pid", "x", "vx"):
                assert np.array_equal(a["parts"][sp][key], m["parts"][sp][key]), (r.rid, sp, key)


def test_serving_rejects_oversized_and_survives_failures(ohp):
    from ohhelppic3d_amd import serve
    R = serve.Request
    reqs = [R(0, 0, (16, 16, 16), (4, 4), 6, 1.0, 0.5, 1),
            # NaN thermal speed: the first push yields non-finite positions -> fails
            R(1, 0, (16, 16, 16), (2, 2), 6, float("nan"), 0.5, 2),
            R(2, 2, (24, 24, 24), (3, 3), 4, 1.0, 0.5, 3),
            R(3, 0, (512, 512, 512), (64, 64), 1, 1.0, 0.5, 4)]
    loop = serve.ServingLoop(budget_bytes=4 * 10 ** 9, tile=(8, 8, 8), chunks=1)
    rep = loop.run(reqs, clock="steps")
    st = {o.rid: o for o in rep.outcomes}
    assert st[3].status == "rejected" and "budget" in st[3].error
    assert st[1].status == "failed" and "box lengths" in st[1].error
    assert st[0].status == "done" and st[2].status == "done"
    assert st[0].time_to_completion == 6 and st[2].finished == 6


def test_footprint_estimate_bounds_allocation(ohp):
    from ohhelppic3d_amd import serve
    loop = serve.ServingLoop(tile=(8, 8, 8), chunks=1)
    for n, ppc in (((32, 32, 32), (16, 64)), ((64, 48, 16), (40, 8)), ((24, 24, 24), (8, 8))):
        r = serve.Request(0, 0, n, ppc, 1, 1.0, 0.5, 1)
        est = loop.footprint(r)
        sim = r.spec.create((8, 8, 8), 1)
        try:
            sim.step(1)
            got = sim.memory_footprint()
        finally:
            sim.close()
        assert got <= est, (n, ppc, got, est)
        assert got >= 0.7 * est, (n, ppc, got, est)
