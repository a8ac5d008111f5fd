"""Run parameters in the reference's ``parameters.toml`` layout.

Mirrors ``new_Parameters`` (src/parameters/parameters.f90:76-150): the same
sections and keys, the same derived macro charge
``q_macro = wp^2 / (q/m) / nmacro_particles_per_grid`` (:140-149), and the
same fail-fast behaviour on a missing or invalid key (the reference aborts,
toml_wrapper.f90:101-124; here ``ParameterError`` names the key).

Only what the periodic per-timestep loop uses is accepted: periodic outer
boundaries on every axis, the 'boris' mover, isotropic thermal speeds, no
flow, and species initialised 'random-uniform' or 'none'.  Keys of
subsystems outside that loop (unit conversion, snapshots, MPI node counts,
injection, objects) are read past and listed in ``RunSpec.ignored``.

Initial state (the parity contract's loader): ``nmacro_particles_per_grid *
nx * ny * nz`` particles per 'random-uniform' species, uniform positions and
Maxwellian velocities with standard deviation ``thermal_velocity_para``, drawn
from the PCG32 stream seeded (42, 52) as at ohhelppic3d.f90:111.  The
reference parses ``cyclotron_frequency`` but never applies a magnetic field;
here it sets a uniform B0 along z with that cyclotron frequency for species 1.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

from . import _lib as L
from .pic import SimConfig, Simulation, charge_per_macro_particle

try:
    import tomllib as _toml  # Python >= 3.11
except ModuleNotFoundError:  # pragma: no cover - depends on the interpreter
    import tomli as _toml


class ParameterError(ValueError):
    """A missing, malformed or unsupported key in the run parameters."""


@dataclass
class RunSpec:
    cfg: SimConfig
    counts: List[int]            # particles per species at initialisation
    vth: List[float]             # Maxwellian standard deviation per species
    nstep: int
    seed: Tuple[int, int] = (42, 52)
    name: str = "main"
    ignored: List[str] = field(default_factory=list)

    def create(self) -> Simulation:
        """A device simulation seeded with the initial state."""
        sim = Simulation(self.cfg)
        try:
            sim.seed_uniform_maxwell(self.counts, self.vth, seed=self.seed)
        except Exception:
            sim.close()
            raise
        return sim


def _get(tbl: dict, *keys, default=None, required=True):
    cur = tbl
    path = ".".join(keys)
    for k in keys:
        if not isinstance(cur, dict) or k not in cur:
            if required and default is None:
                raise ParameterError(f"missing key '{path}'")
            return default
        cur = cur[k]
    return cur


def _num_list(tbl, *keys, n: Optional[int] = None, required=True) -> Optional[List[float]]:
    v = _get(tbl, *keys, required=required)
    if v is None:
        return None
    if not isinstance(v, list) or not all(isinstance(x, (int, float)) for x in v):
        raise ParameterError(f"'{'.'.join(keys)}' must be an array of numbers")
    if n is not None and len(v) < n:
        raise ParameterError(f"'{'.'.join(keys)}' has {len(v)} entries, nspecies is {n}")
    return [float(x) for x in v]


def parse_parameters(text: str, precision: int = L.FP64,
                     weight_order: int = L.WEIGHTS_REFERENCE,
                     tile: Sequence[int] = (0, 0, 0), device: int = 0) -> RunSpec:
    """RunSpec from the text of a parameters.toml."""
    try:
        t = _toml.loads(text)
    except Exception as e:  # malformed TOML
        raise ParameterError(f"not valid TOML: {e}") from None
    ignored = [k for k in ("meta", "continuaus", "output", "object") if k in t]

    n = []
    for k in ("nx", "ny", "nz"):
        v = _get(t, "system", k)
        if not isinstance(v, int) or v < 1:
            raise ParameterError(f"'system.{k}' must be a positive integer")
        n.append(v)
    dt = float(_get(t, "system", "dt"))
    nstep = int(_get(t, "system", "nstep"))
    if not dt > 0 or nstep < 0:
        raise ParameterError("'system.dt' must be positive and 'system.nstep' non-negative")

    mover = _get(t, "system", "method", "particle_mover", default="boris")
    if mover != "boris":  # particle_mover_factory.f90:20-31
        raise ParameterError(f"'system.method.particle_mover' = {mover!r}: only 'boris' exists")

    bc = _get(t, "system", "outer_boundary", "boundary_communication", required=False)
    bf = _get(t, "system", "outer_boundary", "boundary_type_for_electromagnetic_field",
              required=False)
    bp = _get(t, "system", "outer_boundary", "boundary_type_for_particle", required=False)
    for key, val in (("boundary_communication", bc),
                     ("boundary_type_for_electromagnetic_field", bf)):
        if val is not None and any(x != "periodic" for x in val):
            raise ParameterError(f"'system.outer_boundary.{key}' = {val}: only periodic "
                                 "boundaries are on the GPU path")
    if bp is not None and any(x != "periodic" for row in bp for x in row):
        raise ParameterError("'system.outer_boundary.boundary_type_for_particle': only "
                             "periodic particle boundaries are on the GPU path")
    if "mpi" in t.get("system", {}):
        ignored.append("system.mpi")

    ns = _get(t, "plasma", "nspecies")
    if not isinstance(ns, int) or not 1 <= ns <= L.MAX_SPECIES:
        raise ParameterError(f"'plasma.nspecies' must be in [1, {L.MAX_SPECIES}]")
    qm = _num_list(t, "plasma", "charge_to_mass_ratio", n=ns)[:ns]
    wp = _num_list(t, "plasma", "plasma_frequency", n=ns)[:ns]
    vpara = _num_list(t, "plasma", "thermal_velocity_para", n=ns)[:ns]
    vperp = _num_list(t, "plasma", "thermal_velocity_perp", n=ns, required=False)
    if vperp is not None and any(abs(a - b) > 1e-12 * max(abs(a), 1.0)
                                 for a, b in zip(vpara, vperp[:ns])):
        raise ParameterError("'plasma.thermal_velocity_perp' differs from '_para': only "
                             "isotropic Maxwellians are loaded")
    flow = _num_list(t, "plasma", "flow_velocity", required=False)
    if flow is not None and any(v != 0.0 for v in flow):
        raise ParameterError("'plasma.flow_velocity': drifting species are not loaded")
    if any(q == 0.0 for q in qm):
        raise ParameterError("'plasma.charge_to_mass_ratio' has a zero entry")
    wce = float(_get(t, "plasma", "cyclotron_frequency", default=0.0, required=False) or 0.0)

    nmacro = _get(t, "system", "nmacro_particles_per_grid")
    if not isinstance(nmacro, list) or len(nmacro) < ns or \
            not all(isinstance(v, int) and v >= 0 for v in nmacro[:ns]):
        raise ParameterError("'system.nmacro_particles_per_grid' needs a non-negative "
                             "integer per species")
    nmacro = nmacro[:ns]
    init = _get(t, "plasma", "initialization", "plasma_initialization")
    if not isinstance(init, list) or len(init) < ns:
        raise ParameterError("'plasma.initialization.plasma_initialization' needs an entry "
                             "per species")
    for s, kind in enumerate(init[:ns]):
        if kind not in ("random-uniform", "none"):
            raise ParameterError(f"species {s + 1}: initialisation {kind!r} is not loaded "
                                 "(random-uniform or none)")
    if "injection" in t.get("plasma", {}):
        ignored.append("plasma.injection")

    cells = n[0] * n[1] * n[2]
    counts = [nmacro[s] * cells if init[s] == "random-uniform" else 0 for s in range(ns)]
    q_macro = [float(v) if nmacro[s] > 0 else 0.0 for s, v in
               enumerate(charge_per_macro_particle(wp, qm, [max(m, 1) for m in nmacro]))]
    b0z = wce / abs(qm[0]) if wce else 0.0
    cfg = SimConfig(n=tuple(n), qm=tuple(qm), q_macro=tuple(q_macro), capacity=tuple(counts),
                    dt=dt, b0=(0.0, 0.0, b0z), precision=precision, weight_order=weight_order,
                    tile=tuple(tile), device=device)
    return RunSpec(cfg=cfg, counts=counts, vth=vpara, nstep=nstep,
                   name=str(t.get("simulation_name", "main")), ignored=ignored)


def load_parameters(path: str, **kw) -> RunSpec:
    """RunSpec from a parameters.toml file (see parse_parameters)."""
    with open(path, "r", encoding="utf-8") as f:
        return parse_parameters(f.read(), **kw)
