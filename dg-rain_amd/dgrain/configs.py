"""Synthetic workloads (seeded): meshes, initial states and parameters.

Conventions where the reference has no code (its 2D step has no tide, rain or
per-run CFL; boundary_conditions.c:10-83, time_evolution.c:399):
  * z is depth below the datum, positive down; bed elevation b maps to z = -b.
  * Manning n is per node (the mesh format's 5th column,
    extract_mesh_attributes.c:66).
  * Tide: edges on x = 0 carry BdryPrescribed code 2 and are driven by
    zeta_b(t) = datum + A sin(2 pi t / T) with the momentum sentinel
    (compute2DL.c:155-156).  The datum is lowered by D = 1.6 m
    (zeta' = zeta + D, z' = z - D; depth H is unchanged) so that zeta_b stays
    positive: the reference ignores a negative prescribed zeta
    (compute2DL.c:153-154).
  * Still water: zeta = max(0, H0 - z) at nodes (the pattern of
    initialize.c:54-62), midpoints by averaging (initialize.c:117-123).
"""
from dataclasses import dataclass

import numpy as np

from . import mesh as M
from . import native

H0 = 1e-3
TIDE_DATUM_SHIFT = 1.6


@dataclass
class Workload:
    name: str
    mesh: M.Mesh
    zeta: np.ndarray
    qx: np.ndarray
    qy: np.ndarray
    params: dict
    nsteps: int

    @property
    def num_el(self):
        return self.mesh.num_el

    def solver_params(self):
        return native.default_params(**self.params)


def dam_break(N=100, L=1000.0, nsteps=1000, wetdry=1):
    """Plain 2D dam break: flat bed z=0, h=10 m for x<L/2 else 5 m, walls,
    n=0.02, limiter on, SSP-RK2 at CFL 0.3."""
    mesh = M.grid_mesh(N, L, z=0.0, n=0.02)
    zeta_nodes = np.where(mesh.x < L / 2, 10.0, 5.0)
    zeta = M.nodal_to_midpoints(mesh, zeta_nodes)
    zero = np.zeros_like(zeta)
    params = dict(cfl=0.3, wetdry=wetdry)
    return Workload("dam_break_%dx%d" % (N, N), mesh, zeta, zero, zero.copy(), params, nsteps)


def coastal_bed(x, y, L, nbumps, rng, scale=1.0):
    """Linear slope from -10 m at x=0 to +2 m at x=L plus Gaussian bumps."""
    b = -10.0 + 12.0 * x / L
    cx = rng.uniform(0, L, nbumps)
    cy = rng.uniform(0, L, nbumps)
    amp = rng.uniform(-0.5, 0.5, nbumps)
    rad = rng.uniform(100.0, 500.0, nbumps) * scale
    for k in range(nbumps):
        b = b + amp[k] * np.exp(-((x - cx[k]) ** 2 + (y - cy[k]) ** 2) / rad[k] ** 2)
    return b


def coastal(N=512, L=1.0e4, nsteps=2000, seed=2, nbumps=32, tide_amp=1.5, rain_mm_h=0.0, bump_scale=1.0,
            start="flood", level=2.5, tide_phase=0.0):
    """Coastal inundation driven by a tide on x=0, wet/dry and limiter on.

    Bed: slope from -10 m (x=0) to +2 m (x=L) plus `nbumps` Gaussian bumps;
    Manning n lognormal (median 0.025, sigma 0.3) per node.

    start="still": still water at level 0 over that bed (the land above 0 m
        starts dry).  The reference's wet/dry scheme cannot hold this shoreline
        at rest: the compiled reference and the oracle both stop on a NaN flux
        within 3 steps (tests/test_oracle_vs_reference.py).
    start="flood" (default): still water at +`level` m, every element wet, the
        tide oscillating about that level.  Same mesh, bed, friction, forcing
        and code paths (wet/dry status, PDop, limiter run every stage); this is
        the variant the reference runs to 2000 steps.
    """
    rng = np.random.default_rng(seed)
    x, y = M.grid_nodes(N, L)
    bed = coastal_bed(x, y, L, nbumps, rng, bump_scale)
    depth = -bed
    n = 0.025 * np.exp(0.3 * rng.standard_normal(x.size))
    if start == "still":
        D, surface = TIDE_DATUM_SHIFT, 0.0
    elif start == "flood":
        D, surface = 0.0, float(level)
    else:
        raise ValueError(start)
    zeta_nodes = np.maximum(surface, H0 - depth) + D
    z = depth - D
    mesh = M.from_elements(x, y, z, n, M.grid_elements(N), out=M.side_pairs(N, "left"))
    zeta = M.nodal_to_midpoints(mesh, zeta_nodes)
    zero = np.zeros_like(zeta)
    params = dict(cfl=0.3, wetdry=1, tide_code=2, tide_datum=D + surface, tide_amp=tide_amp, tide_period=43200.0,
                  tide_phase=tide_phase, rain=rain_mm_h / 1000.0 / 3600.0)
    return Workload("coastal_%s_%dx%d" % (start, N, N), mesh, zeta, zero, zero.copy(), params, nsteps)


def ensemble(n=64, seed=3, nmin=64, nmax=512, nsteps=500, L_per_cell=1.0e4 / 512, level=3.5):
    """Batched ensemble of `n` coastal members of uneven sizes: N x N grids
    with N = floor(exp(U(ln nmin, ln nmax))) from `seed`; per member a tide
    amplitude U(0.5, 2.5) m, uniform rain U(0, 50) mm/h and its own seeded
    bumps and Manning n field.  The cell size is that of the 512 x 512 coastal
    mesh, so every member has the same minimum edge.  Members start fully wet
    at `level` (3.5 m: above every bump crest of the 64 seeded members, the
    highest 2.9 m; the reference's wet/dry scheme stops on a NaN flux where a
    crest starts at the waterline).  Returns the workloads."""
    rng = np.random.default_rng(seed)
    Ns = np.floor(np.exp(rng.uniform(np.log(nmin), np.log(nmax), n))).astype(int)
    amps = rng.uniform(0.5, 2.5, n)
    rains = rng.uniform(0.0, 50.0, n)
    seeds = rng.integers(0, 2**31 - 1, n)
    out = []
    for i in range(n):
        N = int(Ns[i])
        wl = coastal(N=N, L=N * L_per_cell, nsteps=nsteps, seed=int(seeds[i]), tide_amp=float(amps[i]),
                     rain_mm_h=float(rains[i]), level=level)
        wl.name = "ensemble_member_%d_%dx%d" % (i, N, N)
        out.append(wl)
    return out


def large(N=2048, L=4.0e4, nsteps=1000, seed=5):
    """The large single-GPU run: N x N over L x L (2048: 8,388,608 triangles),
    the coastal bathymetry statistics scaled to the domain (128 bumps, radii
    x4), tide 1.5 m plus rain 20 mm/h, flood start."""
    wl = coastal(N=N, L=L, nsteps=nsteps, seed=seed, nbumps=128, rain_mm_h=20.0, bump_scale=L / 1.0e4, level=3.5)
    wl.name = "large_%dx%d" % (N, N)
    return wl


def dry_bed_dam_break(N=16, L=10.0, nsteps=100):
    """Dam break onto a dry flat bed (the pattern of initialize.c:65-73):
    H = 0.005 m for x <= L/2, H0 elsewhere; wet/dry on, walls."""
    mesh = M.grid_mesh(N, L, z=0.0, n=0.02)
    zeta = M.nodal_to_midpoints(mesh, np.where(mesh.x <= L / 2, 0.005, H0))
    zero = np.zeros_like(zeta)
    return Workload("dry_bed_dam_break_%d" % N, mesh, zeta, zero, zero.copy(), dict(cfl=0.3, wetdry=1), nsteps)


def lake_at_rest(N=8, L=1000.0, depth=1.0, zeta0=0.1, bumps=True, seed=0):
    """Still water over a (bumpy) bed with walls: the state must not move."""
    rng = np.random.default_rng(seed)
    x, y = M.grid_nodes(N, L)
    z = np.full(x.size, depth)
    if bumps:
        z = z - 0.3 * np.exp(-((x - 0.4 * L) ** 2 + (y - 0.6 * L) ** 2) / (0.2 * L) ** 2)
        z = z + 0.05 * rng.uniform(size=x.size)
    mesh = M.from_elements(x, y, z, np.full(x.size, 0.03), M.grid_elements(N))
    zeta = np.full(3 * mesh.num_el, zeta0)
    zero = np.zeros_like(zeta)
    return Workload("lake_at_rest_%d" % N, mesh, zeta, zero, zero.copy(), dict(cfl=0.3, wetdry=0), 50)


WORKLOADS = {
    "dam_break": dam_break,
    "coastal": coastal,
    "lake_at_rest": lake_at_rest,
    "dry_bed_dam_break": dry_bed_dam_break,
    "ensemble": ensemble,
    "large": large,
}
