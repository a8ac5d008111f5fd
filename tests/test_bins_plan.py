"""Host logic of the bin re-layout (ohpic_bins.hpp: bin_capacity, bin_tight,
plan_rebalance), compiled alone with g++ and checked on 20000 random bin
layouts: windows disjoint and in order, each holding a tight bin and every
tight bin inside one, new capacities multiples of 64 that hold the used
counts with the slack and leave no bin tight, contiguous inside a window and
within its span, bins outside the windows untouched, and a whole-buffer
re-bin asked for only when the whole buffer cannot hold the counts."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "ohhelppic3d_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_plan_rebalance_properties(tmp_path):
    exe = str(tmp_path / "bins_plan_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I" + CSRC,
                    os.path.join(HERE, "bins_plan_check.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe, "20000"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    counts = dict(zip(r.stdout.split()[0:8:2], map(int, r.stdout.split()[1:8:2])))
    # every outcome the planner has is exercised
    assert counts["none"] > 0 and counts["windows"] > 1000 and counts["whole"] > 0
