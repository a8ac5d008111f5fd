"""Input/output files of ohhelppic3d_amd/fortran/ohpic_fortran_check."""
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "ohhelppic3d_amd", "fortran")
EXE = os.path.join(FDIR, "ohpic_fortran_check")

REC = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("vx", "<f8"), ("vy", "<f8"),
                ("vz", "<f8"), ("pid", "<i8"), ("nid", "<i4"), ("spec", "<i4")])


def build():
    subprocess.run(["make", "-s", "-C", FDIR], check=True)


def write_input(path, n, sp, dt, b0, nsteps, kcheck):
    recs = []
    pid = 0
    for s, d in enumerate(sp):
        r = np.zeros(len(d["x"]), REC)
        for k in ("x", "y", "z", "vx", "vy", "vz"):
            r[k] = d[k]
        r["pid"] = np.arange(pid, pid + len(r))
        pid += len(r)
        r["spec"] = s + 1
        recs.append(r)
    with open(path, "wb") as f:
        np.array(list(n) + [len(sp), nsteps, kcheck], "<i4").tofile(f)
        np.array([len(d["x"]) for d in sp], "<i8").tofile(f)
        np.array([d["qm"] for d in sp] + [d["q"] for d in sp] + [dt] + list(b0), "<f8").tofile(f)
        np.concatenate(recs).tofile(f)
    return np.concatenate(recs)


def read_output(path, n, kcheck):
    ext = lambda lo, hi: [m + lo + hi for m in n]  # noqa: E731
    with open(path, "rb") as f:
        out = {}
        nout = int(np.fromfile(f, "<i8", 1)[0])
        out["parts"] = np.fromfile(f, REC, nout)
        e = ext(0, 1)
        out["rho"] = np.fromfile(f, "<f8", e[0] * e[1] * e[2]).reshape(e[2], e[1], e[0])
        e = ext(1, 2)
        out["phi"] = np.fromfile(f, "<f8", e[0] * e[1] * e[2]).reshape(e[2], e[1], e[0])
        out["eb"] = np.fromfile(f, "<f8", 6 * e[0] * e[1] * e[2]).reshape(e[2], e[1], e[0], 6)
        out["probe"] = np.fromfile(f, REC, kcheck)
        out["ebi"] = np.fromfile(f, "<f8", 6 * kcheck).reshape(kcheck, 6)
        e = ext(0, 1)
        out["rho2"] = np.fromfile(f, "<f8", e[0] * e[1] * e[2]).reshape(e[2], e[1], e[0])
        out["spectrum"] = np.fromfile(f, "<f8", 63)
        nout2 = int(np.fromfile(f, "<i8", 1)[0])
        out["parts_fused"] = np.fromfile(f, REC, nout2)
    return out
