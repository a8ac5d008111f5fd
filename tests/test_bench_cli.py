"""CPU checks of bench.py's contract: the refusal of more than one GPU, the
headline workload, and the CPU-baseline record (the oracle timed on the
host, a short sample here)."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_headline_workload_is_256_cubed_64_ppc_fp64():
    import bench
    wl = bench.WORKLOADS["uniform256_ppc64_fp64"]
    assert wl["n"] == (256, 256, 256) and wl["ppc"] == 64 and wl["precision"] == 0
    assert bench.WORKLOADS["uniform256_ppc64_mixed"]["precision"] == 1
    assert bench.WORKLOADS["uniform32_ppc16_fp64"]["n"] == (32, 32, 32)
    # the FP64 drop-in headline keeps the reference's mirrored CIC weights
    assert wl["wo"] == bench.W_REF


def test_mixed_runs_are_compared_with_fp64_of_the_same_load_and_weights():
    import bench
    seen = set()
    for m, f in bench.MIXED_PAIRS:
        wm, wf = bench.WORKLOADS[m], bench.WORKLOADS[f]
        assert wm["precision"] == 1 and wf["precision"] == 0
        assert wm["fp64"] == f
        for k in ("n", "ppc", "hump", "wo"):
            assert wm[k] == wf[k], (m, k)
        seen.add(wm["wo"])
    # both weight orders are timed, and the configuration-2 hump load
    assert seen == {bench.W_REF, bench.W_STD}
    assert any(bench.WORKLOADS[m]["hump"] == 9.0 for m, _ in bench.MIXED_PAIRS)


def test_cpu_baseline_record(orc):
    import bench
    rec = bench.run_cpu_baseline(0.2)
    json.dumps(rec)
    assert rec["kind"] == "port" and rec["unit"] == "particle-steps/s"
    assert rec["value"] > 0 and rec["single_thread_value"] > 0
    assert rec["cores"] >= 1 and rec["ranks"] == 1
    assert "32^3" in rec["sample"]
    assert rec["threads"] == rec["cores"]
    host = rec["host"]
    assert host["nproc"] >= 1 and 1 <= host["affinity"] <= host["nproc"]
    assert rec["cpu_model"] and rec["cpu_model"] != "unknown"
    # the headline's density on a scaled grid beside the 32^3 sample
    assert rec["scaled_headline_density"]["value"] > 0
    assert "64^3" in rec["scaled_headline_density"]["sample"]
