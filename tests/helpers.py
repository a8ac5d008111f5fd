"""Shared set-ups for the parity tests (seeded two-species plasmas)."""
import numpy as np


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


def two_species(orc, n, ppc, seed=(42, 52), vth=(1.0, 0.1), qm=(-1.0, 0.01)):
    """Config-0 physics: electrons q/m=-1, q=-1/ppc; ions q/m=0.01, q=+1/ppc;
    uniform positions and Maxwellian velocities from one PCG32 stream."""
    rng = orc.Pcg32(*seed)
    g = n[0] * n[1] * n[2]
    sp = []
    for s, (qm_s, vt) in enumerate(zip(qm, vth)):
        q = (1.0 if qm_s > 0 else -1.0) / ppc
        x, y, z, vx, vy, vz = orc.load_uniform_maxwell(rng, ppc * g, n, vt)
        sp.append(dict(qm=qm_s, q=q, x=x, y=y, z=z, vx=vx, vy=vy, vz=vz))
    return sp


def sim_for(sp, n, dt=0.1, b0=(0.0, 0.0, 0.5), weight_order=0, track_ids=True, **kw):
    from ohhelppic3d_amd import SimConfig, Simulation
    cfg = SimConfig(n=n, qm=[d["qm"] for d in sp], q_macro=[d["q"] for d in sp],
                    capacity=[len(d["x"]) for d in sp], dt=dt, b0=b0,
                    weight_order=weight_order, track_ids=track_ids, **kw)
    sim = Simulation(cfg)
    for s, d in enumerate(sp):
        sim.load_particles(s, d["x"], d["y"], d["z"], d["vx"], d["vy"], d["vz"],
                           pid=np.arange(len(d["x"]), dtype=np.int64))
    return sim


def by_pid(st):
    o = np.argsort(st["pid"], kind="stable")
    return {k: v[o] for k, v in st.items()}


def energy_histories(ohp, orc, precision, steps, n=(16, 16, 16), ppc=16):
    """Total and field energy per step of the GPU run and of the FP64 oracle
    from the same seeded state (config-0 physics)."""
    sp = two_species(orc, n, ppc)
    osim = orc.OracleSim(n, sp, 0.1, (0, 0, 0.5), 0)
    tot_g, tot_o, fe_g, fe_o = [], [], [], []
    with sim_for(sp, n, track_ids=False, precision=precision) as sim:
        for _ in range(steps):
            sim.step(1)
            osim.step(1)
            d = sim.diagnostics()
            tot_g.append(d["total_energy"])
            fe_g.append(d["field_energy"])
            tot_o.append(0.5 * (osim.ke_old.sum() + osim.ke_new.sum()) + osim.field_energy)
            fe_o.append(osim.field_energy)
    return (np.array(tot_g), np.array(tot_o), np.array(fe_g), np.array(fe_o))
