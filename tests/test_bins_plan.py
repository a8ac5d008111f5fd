"""Host logic of the bin relocation (ohpic_bins.hpp: bin_capacity,
bin_tight, bin_listed, plan_relocation), compiled alone with g++ and checked
on 20000 random bin layouts: every tight bin is listed; the listed bins move,
in the given order, packed one after another from the pool's first free
element and within the buffer, with capacities that are multiples of 64,
hold the used counts with the slack and leave the bin unlisted; a full
re-bin is asked for only when the pool cannot hold them; and a bin whose
capacity was just set (with the slack or twice it) is never listed."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "ohhelppic3d_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_plan_relocation_properties(tmp_path):
    exe = str(tmp_path / "bins_plan_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I" + CSRC,
                    os.path.join(HERE, "bins_plan_check.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe, "20000"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    counts = dict(zip(r.stdout.split()[0:8:2], map(int, r.stdout.split()[1:8:2])))
    # every outcome the planner has is exercised
    assert counts["none"] > 0 and counts["moves"] > 1000 and counts["whole"] > 0
