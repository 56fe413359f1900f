"""Synthetic scenes for the benchmark configurations and the parity tests.

Every generator is seeded through ``numpy.random.Generator(PCG64(seed))`` and
builds its scenes through the public API (Mesh -> FemDomain -> contact pair ->
SceneSpec), so the packed arrays are exactly what ``ExplicitLoop`` would
upload for a one-body loop with the same domains.

Configurations (sizes as in the benchmark definition):
  small     20x20 grid on [0.3,0.7]^2, Elastic(2e4, .2, 40), g=(0,-1), plane y=0, dt=1e-4
  large     1000x1000 jittered grid, Neo-Hookean(1e5, .3, 1000), g=(0,-9.8), plane y=0, CFL dt
  j2voxel   512x512 block, J2(1e5, .3, 1000, sy=500, H=1e3), v0=(0,-2), 256x256 value-noise voxels
  batched   4096 scenes, side log-uniform in [8,128], Elastic/NH/J2 mix, one height field each
  serving   Poisson stream, side in [16,256], plane / height field / voxel grounds, horizon 60..600
  springs   4096 spring lattices (structural, shear and bending springs), side log-uniform in
            [8,128], one height field each
"""

import math

import numpy as np

from .bcs import Gravity, InitialVel
from .definitions import MaterialType
from .domains import FemDomain, GroundDomain, HeightFieldDomain, VoxelGridDomain
from .femcontact import contactHelper
from .femesher import structured_grid
from .materials import Elastic, HyperElastic, MisesPlastic, SolidProp
from .mesh import Mesh
from .scene import MAT_SPRING, pack_domain, scene_stable_time
from .springmass import SpringMassDomain


def rng_for(seed):
    return np.random.Generator(np.random.PCG64(seed))


def one_body_spec(domain, ground=None, dt=None, damping=0.0, horizon=0):
    """SceneSpec of a one-body loop [ground, domain] with the loop's stable dt."""
    pairs = []
    if ground is not None:
        c = contactHelper(ground, domain, ground.type | domain.type)
        pairs.append((ground, c))
    if dt is None:
        dt = scene_stable_time(domain, [c for _, c in pairs], has_plane=isinstance(ground, GroundDomain))
    return pack_domain(domain, pairs, dt=dt, damping=damping, horizon=horizon)


def make_material(kind, E, nu, rho, sigma_y=None, H=None):
    if kind == MaterialType.ELASTIC:
        return Elastic(E, nu, rho)
    if kind == MaterialType.NEOHOOKEAN:
        return HyperElastic(E, nu, rho)
    return MisesPlastic(E, nu, rho, sigma_y, H)


def sinusoid_heightfield(rng, n=1024, lo=-0.5, hi=1.5, amplitude=0.05):
    """Sum of 3 sinusoids, total amplitude bound ``amplitude``, wavelengths U[0.1, 1]."""
    x = np.linspace(lo, hi, n)
    lam = rng.uniform(0.1, 1.0, size=3)
    phi = rng.uniform(0.0, 2.0 * math.pi, size=3)
    h = sum((amplitude / 3.0) * np.sin(2.0 * math.pi * x / lam[i] + phi[i]) for i in range(3))
    return HeightFieldDomain(2, h.astype(np.float32), lb=[lo, 0.0], ub=[hi, 1.0])


def value_noise_voxels(rng, nx, ny, lb, ub, spacing=8, threshold=0.5, floor_rows=0):
    """Random lattice every ``spacing`` cells, bilinear interpolation, thresholded."""
    gx = nx // spacing + 2
    gy = ny // spacing + 2
    lat = rng.uniform(0.0, 1.0, size=(gx, gy))
    ii = np.arange(nx) / spacing
    jj = np.arange(ny) / spacing
    i0 = np.floor(ii).astype(int)
    j0 = np.floor(jj).astype(int)
    ti = (ii - i0)[:, None]
    tj = (jj - j0)[None, :]
    v = (
        lat[i0][:, j0] * (1 - ti) * (1 - tj)
        + lat[i0 + 1][:, j0] * ti * (1 - tj)
        + lat[i0][:, j0 + 1] * (1 - ti) * tj
        + lat[i0 + 1][:, j0 + 1] * ti * tj
    )
    occ = (v > threshold).astype(np.int32)
    if floor_rows > 0:
        occ[:, :floor_rows] = 1
    return VoxelGridDomain(2, nx, ny, lb, ub, occupancy_np=occ)


def grid_domain(side, lower, extent, material, bcs, initials=(), jitter=0.0, seed=0):
    conn, coords = structured_grid(side, lower=lower, extent=extent, jitter=jitter, seed=seed)
    return FemDomain(Mesh(2, conn, coords), SolidProp(material), list(bcs), initials=list(initials))


# ------------------------------------------------------------------ configurations


def small_config():
    """20x20 Elastic block over a plane; fixed dt 1e-4, run with advance() 600 times."""
    dom = grid_domain(20, (0.3, 0.3), 0.4, Elastic(2e4, 0.2, 40.0), [Gravity([0.0, -1.0])])
    ground = GroundDomain(2, [0.0, 0.0], [0.0, 1.0])
    return one_body_spec(dom, ground, dt=1e-4)


def large_config(side=1000, seed=0):
    """Jittered 1000x1000 Neo-Hookean sheet resting on a plane at y=0 (bottom row penetrates)."""
    dom = grid_domain(side, (0.0, 0.0), 1.0, HyperElastic(1e5, 0.3, 1000.0), [Gravity([0.0, -9.8])],
                      jitter=0.1, seed=seed)
    ground = GroundDomain(2, [0.0, 0.0], [0.0, 1.0])
    return one_body_spec(dom, ground)


def j2_voxel_config(side=512, vox=256, seed=0):
    """J2 block falling at 2 m/s onto a value-noise voxel ground."""
    rng = rng_for(seed)
    mat = MisesPlastic(1e5, 0.3, 1000.0, 500.0, 1e3)
    dom = grid_domain(side, (0.25, 0.55), 0.5, mat, [Gravity([0.0, -9.8])], initials=[InitialVel("ALL", [0.0, -2.0])])
    ground = value_noise_voxels(rng, vox, vox, [0.0, 0.0], [1.0, 0.55])
    return one_body_spec(dom, ground)


# Density and placement of the mixed-scene workloads (both unspecified by the
# configurations, so chosen here).  Heavier blocks dropped from higher up invert
# elements on the rough height fields: with rho = 1000 and a 0.3 m drop 23 of 96
# sampled scenes went NaN (Neo-Hookean stress is log J, NaN once J < 0); with
# rho = 100 and a drop of 0.05 to 0.15 m, 119 of the 4096 scenes were NaN after
# 7200 substeps.  rho = 10 with each block's bottom 5 mm above the highest ground
# point under it kept all 4096 finite over 12000 substeps, with about 62% of the
# scenes in ground contact and half of the J2 scenes yielding once they landed.
MIX_RHO = 10.0
MIX_GAP = 0.005
# the spring lattices start with their lowest row at y = 0.1, above every height field
SPRING_Y0 = 0.1
# density of the two-block FEM-FEM contact loops
CONTACT_RHO = 100.0


def random_material(rng, rho=MIX_RHO, kinds=(MaterialType.ELASTIC, MaterialType.NEOHOOKEAN, MaterialType.MISES)):
    kind = kinds[int(rng.integers(0, len(kinds)))]
    E = float(rng.uniform(1e4, 1e5))
    nu = float(rng.uniform(0.2, 0.4))
    return make_material(kind, E, nu, rho, 5e-3 * E, 1e-2 * E)


def ground_top(ground, x_lo=0.3, x_hi=0.7):
    """Highest point of a ground under the x range [x_lo, x_hi]: the plane's height at
    the range (a horizontal plane's y), the highest height sample, or the top of the
    highest occupied voxel."""
    if isinstance(ground, GroundDomain):
        return float(ground.point[1])
    if isinstance(ground, HeightFieldDomain):
        xs = np.linspace(ground.lb[0], ground.ub[0], ground.height.shape[0])
        under = (xs >= x_lo) & (xs <= x_hi)
        return float(ground.height[under].max() if under.any() else ground.height.max())
    occ = np.asarray(ground.occ)
    dx = float(ground.dx)
    cx = float(ground.lb[0]) + (np.arange(occ.shape[0]) + 0.5) * dx  # cell centres
    ci = (cx >= x_lo - dx) & (cx <= x_hi + dx)
    rows = np.flatnonzero(occ[ci].any(axis=0))
    return float(ground.lb[1] + (rows.max() + 1) * float(ground.dz)) if rows.size else float(ground.lb[1])


def batched_config(n_scenes=4096, seed=0, side_lo=8, side_hi=128, rho=MIX_RHO, gap=MIX_GAP, damping=0.0):
    """Uneven scenes: side log-uniform in [side_lo, side_hi], one height field each; each
    block's bottom ``gap`` above the highest height sample under it."""
    rng = rng_for(seed)
    specs = []
    for _ in range(n_scenes):
        s = int(round(side_lo * (side_hi / side_lo) ** rng.uniform()))
        mat = random_material(rng, rho=rho)
        hf = sinusoid_heightfield(rng)
        dom = grid_domain(s, (0.3, ground_top(hf) + gap), 0.4, mat, [Gravity([0.0, -9.8])])
        specs.append(one_body_spec(dom, hf, damping=damping))
    return specs


def serving_job(rng, side_lo=16, side_hi=256, horizon_lo=60, horizon_hi=600, gap=0.0):
    """One request: random size, material, ground kind and horizon; the block's bottom
    ``gap`` above the highest point of its ground under it (0: touching it, so that a
    horizon of 60 to 600 substeps covers ground contact)."""
    s = int(rng.integers(side_lo, side_hi + 1))
    mat = random_material(rng)
    kind = int(rng.integers(0, 3))
    if kind == 0:
        ground = GroundDomain(2, [0.0, 0.0], [0.0, 1.0])
    elif kind == 1:
        ground = sinusoid_heightfield(rng)
    else:
        ground = value_noise_voxels(rng, 128, 8, [-0.5, 0.0], [1.5, 0.05], spacing=4, floor_rows=2)
    dom = grid_domain(s, (0.3, ground_top(ground) + gap), 0.4, mat, [Gravity([0.0, -9.8])])
    horizon = int(rng.integers(horizon_lo, horizon_hi + 1))
    return one_body_spec(dom, ground, horizon=horizon)


def spring_lattice(side, lower, spacing, prop, bcs, initials=(), jitter=0.0, seed=0, bending=True):
    """SpringMassDomain on a (side+1)^2 node lattice: structural springs along both axes,
    shear springs along both diagonals and (optionally) bending springs skipping one node.
    Node (i, j) has id i * (side + 1) + j, at lower + spacing * (i, j) (+ jitter)."""
    n = side + 1
    rng = rng_for(seed)
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    xy = np.stack([i.reshape(-1), j.reshape(-1)], 1).astype(np.float64) * spacing + np.asarray(lower, np.float64)
    if jitter > 0.0:
        xy += rng.uniform(-jitter * spacing, jitter * spacing, xy.shape)
    ids = np.arange(n * n).reshape(n, n)
    pairs = [np.stack([ids[:-1, :], ids[1:, :]], -1).reshape(-1, 2), np.stack([ids[:, :-1], ids[:, 1:]], -1).reshape(-1, 2),
             np.stack([ids[:-1, :-1], ids[1:, 1:]], -1).reshape(-1, 2), np.stack([ids[1:, :-1], ids[:-1, 1:]], -1).reshape(-1, 2)]
    if bending:
        pairs += [np.stack([ids[:-2, :], ids[2:, :]], -1).reshape(-1, 2), np.stack([ids[:, :-2], ids[:, 2:]], -1).reshape(-1, 2)]
    conn = np.concatenate(pairs).astype(np.int32)
    return SpringMassDomain(2, xy.astype(np.float32), conn, list(prop), list(bcs), initials=list(initials))


def spring_batched_config(n_scenes=4096, seed=0, side_lo=8, side_hi=128, spacing=0.01):
    """Uneven spring lattices: side log-uniform in [side_lo, side_hi] at a fixed spacing (a
    spring's force per unit length is k / l0, so the spacing, not the side, sets the stable
    step), random stiffness, light damping, one height field each."""
    rng = rng_for(seed)
    specs = []
    for _ in range(n_scenes):
        s = int(round(side_lo * (side_hi / side_lo) ** rng.uniform()))
        k = float(rng.uniform(50.0, 500.0))
        m = 0.01
        c = 0.02 * math.sqrt(k * m)
        width = s * spacing
        dom = spring_lattice(s, (0.5 - 0.5 * width, SPRING_Y0), spacing, [k, c, m], [Gravity([0.0, -9.8])])
        specs.append(one_body_spec(dom, sinusoid_heightfield(rng)))
    return specs


def contact_pair_group(rng, side_lo=8, side_hi=32, dt_fraction=0.35):
    """Two stacked blocks on a plane in one loop: the lower one resting on the plane,
    the upper one falling onto it with a random offset; both with counter-clockwise
    elements (outward boundary normals), a random material each and friction.
    Returns (packed scenes, FEM-FEM group) as ExplicitLoop would build them, with
    dt at ``dt_fraction`` of the loop's stable time (the reference's adaptive step
    lets two bodies pressing on each other through two work items diverge)."""
    from .definitions import ContactType
    from .femcontact import ContactFlexFlex
    from .flexcontact import build_flex_group
    from .scene import add_flex_record

    def block(side, lower, extent, mat, initials=(), friction=0.0):
        conn, coords = structured_grid(side, lower=lower, extent=extent)
        conn = np.ascontiguousarray(conn[:, [0, 2, 1]])
        return FemDomain(Mesh(2, conn, coords), SolidProp(mat), [Gravity([0.0, -9.8])], initials=list(initials),
                         friction=friction)

    s1 = int(round(side_lo * (side_hi / side_lo) ** rng.uniform()))
    s2 = int(round(side_lo * (side_hi / side_lo) ** rng.uniform()))
    w1, w2 = 0.3, float(rng.uniform(0.15, 0.3))
    # Elastic and J2: an impact can invert elements, where Neo-Hookean stress is NaN (log J < 0)
    kinds = (MaterialType.ELASTIC, MaterialType.MISES)
    lower = block(s1, (0.2, 0.0), w1, random_material(rng, rho=CONTACT_RHO, kinds=kinds),
                  friction=float(rng.uniform(0.0, 0.4)))
    upper = block(s2, (0.2 + float(rng.uniform(-0.1, 0.2)), w1 + 0.002), w2, random_material(rng, rho=CONTACT_RHO, kinds=kinds),
                  initials=[InitialVel("ALL", [0.0, -float(rng.uniform(0.1, 0.5))])])
    plane = GroundDomain(2, [0.0, 0.0], [0.0, 1.0])
    pair = ContactFlexFlex(upper, lower, ContactType.FLEXFLEX)
    grounds = {k: [(plane, contactHelper(plane, b, plane.type | b.type))] for k, b in enumerate((lower, upper))}
    group = build_flex_group([lower, upper], [pair], True, lower.nnodes + upper.nnodes)
    st = min([1.0, 1e-3, lower.stableTime, upper.stableTime, pair.stableTime] + [c.stableTime for (_, c), in grounds.values()])
    dt = dt_fraction * st
    specs = []
    for k, b in enumerate((lower, upper)):
        spec = add_flex_record(pack_domain(b, grounds[k], dt=dt), group.flex_cap(k))
        specs.append(spec)
    return specs, group


def contact_batched_config(n_groups=1024, seed=0, side_lo=8, side_hi=32):
    """Many independent two-block loops with FEM-FEM contact: a list of (scenes, group).
    Every loop takes the smallest loop dt, so that all share one hash-rebuild schedule."""
    rng = rng_for(seed)
    groups = [contact_pair_group(rng, side_lo, side_hi) for _ in range(n_groups)]
    dt = min(s.dt for specs, _ in groups for s in specs)
    for specs, _ in groups:
        for s in specs:
            s.dt = dt
    return groups


# Algorithmic bytes of the FEM-FEM contact pass per substep: per work item its record
# (body, node, slot, target, pair, penalty, friction: 28 B), its cache read and
# written (8 B) and the node's position and velocity (16 B); per slot the (first,
# count) range written (8 B).  Candidate segments and contributions depend on the
# contact state and are not counted.
CONTACT_ITEM_BYTES = 52
CONTACT_SLOT_BYTES = 8


def algorithmic_bytes(spec):
    """E*B_e + N*B_n with B_e = 32 (64 for J2, 12 for a spring: two node ids and the rest
    length) and B_n = 40 bytes."""
    be = 64 if spec.mat_type == MaterialType.MISES else (12 if spec.mat_type == MAT_SPRING else 32)
    return spec.n_elems * be + spec.n_nodes * 40
