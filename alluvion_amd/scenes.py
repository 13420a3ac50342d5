"""Deterministic synthetic scenes for the solver path.

Random numbers come from SplitMix64 (Steele, Lea, Flood 2014) mapped to
uniform doubles in [0, 1) by the top 53 bits, so every scene is reproducible
from its seed on any machine.  Lattice index order follows the reference's
create_fluid_block mode 0 (runner.hpp:970-995): z fastest, then x, then y.

Fluid blocks start one kernel radius h inside the container walls (first
lattice layer at distance h), the placement that the reference's scene
helpers produce by dropping every particle closer than h to a boundary
(compute_fluid_block_internal, runner.hpp:1040-1061).  Closer placement puts
wall particles far above rest density under the volume-map boundary term
(a corner particle at 2r starts at 2.6 rho0) and the first step explodes.

Scenes:
* ``dam_break_small``: 16,384 particles (32 x 32 x 16 lattice, r = 2.5 mm,
  spacing 2r, jitter +-0.01 r) in a 0.4 m cube container with a 64^3 volume
  map; DFSPH at dt = 2e-4 s.
* ``dam_break_rigid``: 1,048,576 particles (128 x 128 x 64) in a 1.0 m cube
  with four dynamic rigid bodies (two spheres r = 0.05 m, two cubes of edge
  0.08 m, 128^3 volume maps) seeded above the fluid.
* ``periodic_random``: N particles uniform in the periodic unit cube
  [-0.5, 0.5)^3 with uniform random velocities (+-0.5 m/s).
* ``dam_break_pellets``: a fluid block in a cube container whose walls are
  two layers of pellets, with a dynamic sphere sampled by pellets on its
  surface (Fibonacci lattice); ISPH with particle-sampled boundaries.
"""
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

_MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n successive SplitMix64 outputs (uint64) from the given seed."""
    with np.errstate(over="ignore"):
        state = np.uint64(seed) + np.arange(1, n + 1, dtype=np.uint64) * np.uint64(
            0x9E3779B97F4A7C15)
        z = state
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, n: int) -> np.ndarray:
    """n uniform doubles in [0, 1)."""
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


@dataclass
class BoundarySpec:
    kind: str  # "box", "box_shell", "sphere"
    params: Tuple[float, ...]
    sign: float
    resolution: Tuple[int, int, int]
    mass: float = 0.0
    inertia: Tuple[float, float, float] = (1.0, 1.0, 1.0)
    x: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    q: Tuple[float, float, float, float] = (0.0, 0.0, 0.0, 1.0)
    restitution: float = 1.0
    friction: float = 0.0
    # collision vertices in the body frame, (n, 3); None = no contacts
    collision_vertices: Optional[np.ndarray] = None
    # sample the shape's signed distance onto the boundary's map nodes and
    # take the reference's grid path (as for MeshDistance) instead of the
    # analytic one
    sampled: bool = False


@dataclass
class Scene:
    name: str
    x: np.ndarray  # (N, 3) float32
    v: np.ndarray  # (N, 3) float32
    particle_radius: float = 0.0025
    kernel_radius: float = 0.01
    density0: float = 1000.0
    particle_mass: float = 1.25e-4
    viscosity: float = 0.01
    boundary_viscosity: float = 0.01
    dfsph_factor_epsilon: float = 1e-6
    gravity: Tuple[float, float, float] = (0.0, -9.81, 0.0)
    wrap_length: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    grid_offset: Tuple[int, int, int] = (-21, -21, -21)
    grid_res: Tuple[int, int, int] = (42, 42, 42)
    max_num_neighbors: int = 64
    max_num_particles_per_cell: int = 64
    dt: float = 2e-4
    cfl: float = 0.4
    # time-step range; None = dt (fixed step)
    min_dt: Optional[float] = None
    max_dt: Optional[float] = None
    density_change_tolerance: float = 1e-3
    density_error_tolerance: float = 1e-3
    min_divergence_solve: int = 1
    max_divergence_solve: int = 100
    min_density_solve: int = 2
    max_density_solve: int = 100
    steps: int = 1000
    boundaries: List[BoundarySpec] = field(default_factory=list)
    pile_gravity: Tuple[float, float, float] = (0.0, -9.81, 0.0)
    # rigid contacts (Const::contact_tolerance, Consti::max_num_contacts)
    contact_tolerance: float = 0.0
    max_num_contacts: int = 10000
    # boundary representation (Pile's VolumeMethod): "volume_map", or
    # "pellets" (particle-sampled boundaries: each body's collision
    # vertices are its pellets; ISPH only)
    volume_method: str = "volume_map"
    max_num_pellets: int = 10000
    # hint_identical_sequence runs [begin, end) of identical bodies
    identical_sequences: List[Tuple[int, int]] = field(default_factory=list)

    @property
    def n(self) -> int:
        return int(self.x.shape[0])


def lattice_block(nx: int, ny: int, nz: int, radius: float, box_min: float,
                  seed: int, jitter: float) -> np.ndarray:
    """Jittered lattice in the reference's create_fluid_block mode-0 order."""
    n = nx * ny * nz
    p = np.arange(n, dtype=np.int64)
    j = (p % (nx * nz)) // nz  # x
    k = p // (nx * nz)  # y
    l = p % nz  # z
    d = 2.0 * radius
    start = box_min + d
    pos = np.stack([start + d * j, start + d * k, start + d * l], axis=1)
    u = uniform(seed, 3 * n).reshape(n, 3)
    pos = pos + (2.0 * u - 1.0) * jitter
    return pos.astype(np.float32)


def dam_break_small(seed: int = 1) -> Scene:
    r, h = 0.0025, 0.01
    x = lattice_block(32, 32, 16, r, -0.2 + h - 2 * r, seed, 0.01 * r)
    return Scene(
        name="dam_break_16k",
        x=x,
        v=np.zeros_like(x),
        grid_offset=(-21, -21, -21),
        grid_res=(42, 42, 42),
        steps=1000,
        boundaries=[BoundarySpec("box", (0.4, 0.4, 0.4, 0.0), -1.0, (64, 64, 64))],
    )


def adaptive_step(scene: Scene, cfl: float = 0.4, max_dt: float = 2e-3,
                  min_dt: float = 1e-6) -> Scene:
    """The same scene with the CFL-limited time step (Solver::update_dt,
    solver.hpp:153-169) instead of a fixed one."""
    scene.cfl = cfl
    scene.dt = max_dt
    scene.max_dt = max_dt
    scene.min_dt = min_dt
    return scene


SPHERE_RADIUS, BOX_EDGE = 0.05, 0.08
SPHERE_DENSITY, BOX_DENSITY = 500.0, 800.0


def shape_distance(kind: str, params, p: np.ndarray) -> np.ndarray:
    """Signed distance (negative inside, body frame) of points p [n, 3] to
    an analytic shape, in FP64: dg/box_distance.hpp:23-28,
    dg/box_shell_distance.hpp:36-59, dg/sphere_distance.hpp:18-21."""
    d = np.asarray(p, np.float64)

    def box(half, outset):
        q = np.abs(d) - np.asarray(half, np.float64)
        return (np.linalg.norm(np.maximum(q, 0.0), axis=1) + np.minimum(q.max(axis=1), 0.0)
                - outset)

    if kind == "sphere":
        return np.linalg.norm(d, axis=1) - params[0]
    half = 0.5 * np.asarray(params[:3], np.float64)
    if kind == "box":
        return box(half, params[3])
    inner = box(half, params[4])
    outer = box(half + params[3], params[4])
    return np.maximum(outer, -inner)


def shape_aabb(kind: str, params):
    """Bounding box (body frame) of an analytic shape, as its Distance
    computes it."""
    if kind == "sphere":
        r = params[0]
        return (-r, -r, -r), (r, r, r)
    half = 0.5 * np.asarray(params[:3], np.float64)
    grow = params[3] if kind == "box" else params[3] + params[4]
    lo = tuple(float(v) for v in -half - grow)
    hi = tuple(float(v) for v in half + grow)
    return lo, hi


def body_distance(p: np.ndarray, body: BoundarySpec) -> np.ndarray:
    """Signed distance (negative inside) of points p [n, 3] to an unrotated
    sphere or box body: the closed forms of dg/sphere_distance.hpp:18-21 and
    dg/box_distance.hpp:23-28."""
    d = p.astype(np.float64) - np.asarray(body.x, np.float64)
    if body.kind == "sphere":
        return np.linalg.norm(d, axis=1) - body.params[0]
    q = np.abs(d) - 0.5 * np.asarray(body.params[:3], np.float64)
    outside = np.linalg.norm(np.maximum(q, 0.0), axis=1)
    return outside + np.minimum(q.max(axis=1), 0.0)


def floating_bodies(x_range, z_range, y_top, body_seed, resolution) -> List[BoundarySpec]:
    """Two spheres (r = 0.05 m, 500 kg/m^3) and two cubes (edge 0.08 m,
    800 kg/m^3) floating at their equilibrium depth in a fluid whose top
    lattice layer is at y_top: a sphere's centre on the surface (half
    submerged), a cube's 0.3 edge below it (80 % submerged).  The fluid's
    x range is cut into four slots, sphere, cube, sphere, cube; each centre
    is seeded uniformly (SplitMix64) inside its slot with one kernel radius
    of clearance to the slot's and the fluid's sides."""
    h = 0.01
    u = uniform(body_seed, 8).reshape(4, 2)
    slot = (x_range[1] - x_range[0]) / 4.0
    bodies = []
    for b in range(4):
        sphere = b % 2 == 0
        half = SPHERE_RADIUS if sphere else 0.5 * BOX_EDGE
        room_x = slot - 2.0 * (half + h)
        room_z = (z_range[1] - z_range[0]) - 2.0 * (half + h)
        if room_x <= 0 or room_z <= 0:
            raise ValueError("the fluid block is too small to float the bodies")
        cx = x_range[0] + (b + 0.5) * slot + (u[b, 0] - 0.5) * room_x
        cz = 0.5 * (z_range[0] + z_range[1]) + (u[b, 1] - 0.5) * room_z
        cy = y_top if sphere else y_top - 0.3 * BOX_EDGE
        res = (resolution,) * 3
        if sphere:
            m = SPHERE_DENSITY * 4.0 / 3.0 * np.pi * SPHERE_RADIUS ** 3
            inertia = (0.4 * m * SPHERE_RADIUS ** 2,) * 3
            bodies.append(BoundarySpec("sphere", (SPHERE_RADIUS,), 1.0, res, m, inertia,
                                       (cx, cy, cz)))
        else:
            m = BOX_DENSITY * BOX_EDGE ** 3
            inertia = (m * BOX_EDGE ** 2 / 6.0,) * 3
            bodies.append(BoundarySpec("box", (BOX_EDGE, BOX_EDGE, BOX_EDGE, 0.0), 1.0, res, m,
                                       inertia, (cx, cy, cz)))
    return bodies


def dam_break_rigid(seed: int = 2, body_seed: int = 3, nx: int = 128, ny: int = 128,
                    nz: int = 64, body_resolution: int = 128,
                    floating: Optional[bool] = None) -> Scene:
    """The 1 M-particle dam break with four dynamic bodies.

    floating (the default whenever the fluid block is wide enough): the
    bodies float partly submerged in the fluid's top (``floating_bodies``);
    lattice particles closer than h to a body are dropped, as the reference's
    scene helpers drop particles closer than h to a boundary, and the lattice
    continues upward in its own index order until exactly nx * ny * nz
    particles remain (the first nx * ny * nz lattice sites, jitter included,
    are the plain block's).  Otherwise the bodies are seeded in the air above
    the fluid."""
    r, h = 0.0025, 0.01
    d = 2 * r
    box_min = -0.5 + h - 2 * r
    lo = box_min + d  # first lattice layer
    n = nx * ny * nz
    x_range, z_range = (lo, lo + (nx - 1) * d), (lo, lo + (nz - 1) * d)
    if floating is None:
        floating = (x_range[1] - x_range[0]) >= 4 * (2 * SPHERE_RADIUS + 2 * h) and \
            (z_range[1] - z_range[0]) >= 2 * SPHERE_RADIUS + 2 * h
    if floating:
        bodies = floating_bodies(x_range, z_range, lo + (ny - 1) * d, body_seed, body_resolution)
        extra = 1
        while True:
            x = lattice_block(nx, ny + extra, nz, r, box_min, seed, 0.01 * r)
            keep = np.ones(x.shape[0], bool)
            for b in bodies:
                keep &= body_distance(x, b) >= h
            if keep.sum() >= n:
                x = x[keep][:n]
                break
            extra += 1
    else:
        x = lattice_block(nx, ny, nz, r, box_min, seed, 0.01 * r)
        u = uniform(body_seed, 12).reshape(4, 3)
        bodies = []
        for b in range(4):
            pos = (-0.35 + 0.7 * u[b, 0], 0.2 + 0.2 * u[b, 1], -0.35 + 0.7 * u[b, 2])
            res = (body_resolution,) * 3
            if b < 2:
                m = SPHERE_DENSITY * 4.0 / 3.0 * np.pi * SPHERE_RADIUS ** 3
                inertia = (0.4 * m * SPHERE_RADIUS ** 2,) * 3
                bodies.append(BoundarySpec("sphere", (SPHERE_RADIUS,), 1.0, res, m, inertia, pos))
            else:
                m = BOX_DENSITY * BOX_EDGE ** 3
                inertia = (m * BOX_EDGE ** 2 / 6.0,) * 3
                bodies.append(BoundarySpec("box", (BOX_EDGE, BOX_EDGE, BOX_EDGE, 0.0), 1.0, res,
                                           m, inertia, pos))
    container = BoundarySpec("box", (1.0, 1.0, 1.0, 0.0), -1.0, (64, 64, 64))
    return Scene(
        name="dam_break_1m_rigid",
        x=x,
        v=np.zeros_like(x),
        grid_offset=(-51, -51, -51),
        grid_res=(102, 102, 102),
        steps=500,
        boundaries=[container] + bodies,
    )


def periodic_random(n: int, seed: int = 4, vseed: int = 5) -> Scene:
    x = (uniform(seed, 3 * n).reshape(n, 3) - 0.5).astype(np.float32)
    v = ((uniform(vseed, 3 * n).reshape(n, 3) - 0.5)).astype(np.float32)
    return Scene(
        name="periodic_random_%d" % n,
        x=x,
        v=v,
        gravity=(0.0, 0.0, 0.0),
        wrap_length=(1.0, 1.0, 1.0),
        grid_offset=(-50, -50, -50),
        grid_res=(100, 100, 100),
        steps=100,
        boundaries=[],
    )


def build_solver(scene: Scene, kind: str = "df", dp=None, max_particles: Optional[int] = None,
                 log=None):
    """Creates Depot, Pile and solver for a scene on the GPU path.

    Returns (dp, pile, solver, cn, cni).  The particle state is uploaded.
    """
    from . import Depot, _alluvion as _al

    if dp is None:
        dp = Depot(np.float32)
    cn, cni = dp.create_cn()
    cn.set_kernel_radius(scene.kernel_radius)
    cn.set_particle_attr(scene.particle_radius, scene.particle_mass, scene.density0)
    cn.gravity = dp.f3(*scene.gravity)
    cn.viscosity = scene.viscosity
    cn.boundary_viscosity = scene.boundary_viscosity
    cn.dfsph_factor_epsilon = scene.dfsph_factor_epsilon
    if scene.wrap_length[1] > 0:
        cn.set_wrap_length(scene.wrap_length[1])
    cn.wrap_length_x = scene.wrap_length[0]
    cn.wrap_length_z = scene.wrap_length[2]
    cni.grid_res = _al.uint3(*scene.grid_res)
    cni.grid_offset = _al.int3(*scene.grid_offset)
    cni.max_num_neighbors_per_particle = scene.max_num_neighbors
    cni.max_num_particles_per_cell = scene.max_num_particles_per_cell
    runner = dp.Runner()
    cn.contact_tolerance = scene.contact_tolerance
    pellets = scene.volume_method == "pellets"
    method = _al.VolumeMethod.pellets if pellets else _al.VolumeMethod.volume_map
    pile = dp.Pile(dp, runner, scene.max_num_contacts, volume_method=method,
                   max_num_pellets=scene.max_num_pellets, cn=cn, cni=cni)
    for b in scene.boundaries:
        if b.sampled:
            lo, hi = shape_aabb(b.kind, b.params)
            dist = dp.MapDistance.create(
                dp.f3(*lo), dp.f3(*hi),
                lambda pts, kind=b.kind, params=b.params:
                    shape_distance(kind, params, pts).astype(np.float32))
        elif b.kind == "box":
            dist = dp.BoxDistance.create(dp.f3(*b.params[:3]), b.params[3])
        elif b.kind == "box_shell":
            dist = dp.BoxShellDistance.create(dp.f3(*b.params[:3]), b.params[3], b.params[4])
        else:
            dist = dp.SphereDistance.create(b.params[0])
        if pellets:
            pts = np.zeros((0, 3), np.float32) if b.collision_vertices is None else \
                np.asarray(b.collision_vertices, np.float32).reshape(-1, 3)
            pile.add_pellets(dist, _al.uint3(*b.resolution), sign=b.sign, pellets=pts,
                             mass=b.mass, restitution=b.restitution, friction=b.friction,
                             inertia_tensor=dp.f3(*b.inertia), x=dp.f3(*b.x), q=dp.f4(*b.q))
            continue
        mesh = None
        if b.collision_vertices is not None:
            mesh = _al.Mesh()
            mesh.set_vertices(np.asarray(b.collision_vertices, np.float32).reshape(-1, 3))
        pile.add(dist, _al.uint3(*b.resolution), sign=b.sign, collision_mesh=mesh,
                 mass=b.mass, restitution=b.restitution, friction=b.friction,
                 inertia_tensor=dp.f3(*b.inertia), x=dp.f3(*b.x), q=dp.f4(*b.q))
        if log:
            log("boundary %s %s built" % (b.kind, b.resolution))
    pile.set_gravity(dp.f3(*scene.pile_gravity))
    for b0, b1 in scene.identical_sequences:
        pile.hint_identical_sequence(b0, b1)
    n_max = scene.n if max_particles is None else max_particles
    cls = {"df": dp.SolverDf, "ii": dp.SolverIi, "i": dp.SolverI}[kind]
    solver = cls(runner, pile, dp, n_max, cn=cn, cni=cni)
    solver.num_particles = scene.n
    solver.dt = scene.dt
    solver.max_dt = scene.dt if scene.max_dt is None else scene.max_dt
    solver.min_dt = scene.dt if scene.min_dt is None else scene.min_dt
    solver.cfl = scene.cfl
    solver.density_error_tolerance = scene.density_error_tolerance
    solver.min_density_solve = scene.min_density_solve
    solver.max_density_solve = scene.max_density_solve
    if kind == "df":
        solver.density_change_tolerance = scene.density_change_tolerance
        solver.min_divergence_solve = scene.min_divergence_solve
        solver.max_divergence_solve = scene.max_divergence_solve
    if pellets:  # pellet rows follow the fluid's; the step places them
        rows = n_max + scene.max_num_pellets
        for var, src in ((solver.particle_x, scene.x), (solver.particle_v, scene.v)):
            full = np.zeros((rows, 3), np.float32)
            full[:scene.n] = src
            dp.coat(var).set(full)
    else:
        dp.coat(solver.particle_x).set(scene.x)
        dp.coat(solver.particle_v).set(scene.v)
    return dp, pile, solver, cn, cni


def box_wall_pellets(edge: float, radius: float, layers: int = 2) -> np.ndarray:
    """Pellets of a cube container of inner edge `edge` centred at the origin:
    the lattice points (spacing 2r, offset r from the inner faces) of the
    `layers` outermost shells outside the inner cube, in x-major order."""
    d = 2.0 * radius
    m = int(round(edge / d))
    k = np.arange(-layers, m + layers)
    g = (k + 0.5) * d - 0.5 * edge
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pts = np.stack([X, Y, Z], -1).reshape(-1, 3)
    inside = np.all(np.abs(pts) < 0.5 * edge, axis=1)
    return pts[~inside].astype(np.float32)


def sphere_pellets(sphere_radius: float, spacing: float) -> np.ndarray:
    """Fibonacci-lattice pellets on a sphere (Gonzalez 2010, "Measurement of
    areas on a sphere using Fibonacci and latitude-longitude lattices"),
    about one per spacing^2 of surface."""
    n = max(8, int(round(4.0 * np.pi * sphere_radius ** 2 / spacing ** 2)))
    i = np.arange(n) + 0.5
    z = 1.0 - 2.0 * i / n
    phi = np.pi * (1.0 + 5.0 ** 0.5) * i
    rxy = np.sqrt(1.0 - z * z)
    return (sphere_radius * np.stack([rxy * np.cos(phi), z, rxy * np.sin(phi)], 1)).astype(
        np.float32)


def dam_break_pellets(nx: int = 16, ny: int = 16, nz: int = 8, seed: int = 7, edge: float = 0.14,
                      sphere: bool = True, sphere_x: Optional[Tuple[float, float, float]] = None
                      ) -> Scene:
    """A fluid block in the corner of a cube container of inner edge `edge`
    whose walls are two pellet layers (the first fluid layer r from the
    walls, so fluid and pellets pack at the lattice spacing 2r), and
    (optionally) a dynamic sphere of radius 0.02 m sampled by surface
    pellets, by default one spacing above the fluid; ISPH."""
    r, h = 0.0025, 0.01
    lo = -0.5 * edge - r  # first lattice layer at lo + 2r = r inside the wall
    x = lattice_block(nx, ny, nz, r, lo, seed, 0.01 * r)
    walls = box_wall_pellets(edge, r)
    bodies = [BoundarySpec("box", (edge, edge, edge, 0.0), -1.0, (4, 4, 4), restitution=0.8,
                           collision_vertices=walls)]
    if sphere:
        R = 0.02
        m = 0.6 * 1000.0 * 4.0 / 3.0 * np.pi * R ** 3
        I = 0.4 * m * R * R
        top = lo + 2 * r * ny
        sx = (0.01, top + 2 * r + R, 0.0) if sphere_x is None else sphere_x
        bodies.append(BoundarySpec("sphere", (R,), 1.0, (4, 4, 4), mass=m, inertia=(I, I, I),
                                   x=tuple(sx), restitution=0.8, friction=0.1,
                                   collision_vertices=sphere_pellets(R, 2 * r)))
    cells = int(np.ceil((0.5 * edge + 3 * r * 2) / h)) + 1
    return Scene(name="dam_break_pellets", x=x, v=np.zeros_like(x), grid_offset=(-cells,) * 3,
                 grid_res=(2 * cells,) * 3, boundaries=bodies, volume_method="pellets",
                 max_num_pellets=sum(b.collision_vertices.shape[0] for b in bodies) + 64,
                 dt=2e-4, min_density_solve=2, max_density_solve=20)


def dam_break_box(n_target: int, edge: float, seed: int, map_res: int = 64) -> Scene:
    """A dam break of about `n_target` particles in a cube container of edge
    `edge`: a square lattice column (x-z side up to the container's inner
    width) filled to the needed height, first layer at h from the walls."""
    r, h = 0.0025, 0.01
    fit = int(np.floor((edge - 2 * h) / (2 * r) + 1e-9)) + 1  # lattice points across
    side = max(2, min(fit, int(round(n_target ** (1.0 / 3.0)))))
    ny = max(2, min(fit, int(np.ceil(n_target / (side * side)))))
    x = lattice_block(side, ny, side, r, -0.5 * edge + h - 2 * r, seed, 0.01 * r)
    c = int(np.ceil(0.5 * edge / h - 1e-9)) + 1
    return Scene(
        name="dam_break_box_%d" % x.shape[0],
        x=x,
        v=np.zeros_like(x),
        grid_offset=(-c, -c, -c),
        grid_res=(2 * c, 2 * c, 2 * c),
        steps=500,
        boundaries=[BoundarySpec("box", (edge, edge, edge, 0.0), -1.0, (map_res,) * 3)],
    )


def batched_dam_breaks(num_scenes: int = 256, seed: int = 6, n_range=(4000, 120000),
                       edge_range=(0.2, 0.6), map_res: int = 64) -> List[Scene]:
    """Independent dam breaks: particle counts log-uniform in n_range, cube
    edges uniform in edge_range (SplitMix64 from `seed`)."""
    u = uniform(seed, 2 * num_scenes).reshape(num_scenes, 2)
    lo, hi = np.log(n_range[0]), np.log(n_range[1])
    out = []
    for k in range(num_scenes):
        n = int(round(np.exp(lo + (hi - lo) * u[k, 0])))
        edge = edge_range[0] + (edge_range[1] - edge_range[0]) * u[k, 1]
        out.append(dam_break_box(n, edge, seed * 1000 + k, map_res))
    return out


def build_batch(scenes: List[Scene], dp=None, log=None):
    """One batch solver for several scenes that share the constants of the
    first one (each keeps its own particles, grid and container).  Returns
    (dp, piles, solver, cn, cni); every scene's state is uploaded into its
    range of the particle arrays."""
    from . import Depot, _alluvion as _al

    if dp is None:
        dp = Depot(np.float32)
    s0 = scenes[0]
    cn, cni = dp.create_cn()
    cn.set_kernel_radius(s0.kernel_radius)
    cn.set_particle_attr(s0.particle_radius, s0.particle_mass, s0.density0)
    cn.gravity = dp.f3(*s0.gravity)
    cn.viscosity = s0.viscosity
    cn.boundary_viscosity = s0.boundary_viscosity
    cn.dfsph_factor_epsilon = s0.dfsph_factor_epsilon
    cni.max_num_neighbors_per_particle = s0.max_num_neighbors
    cni.max_num_particles_per_cell = s0.max_num_particles_per_cell
    runner = dp.Runner()
    piles = []
    for k, sc in enumerate(scenes):
        pile = dp.Pile(dp, runner, 0, cn=cn, cni=cni)
        for b in sc.boundaries:
            if b.kind == "box":
                dist = dp.BoxDistance.create(dp.f3(*b.params[:3]), b.params[3])
            elif b.kind == "box_shell":
                dist = dp.BoxShellDistance.create(dp.f3(*b.params[:3]), b.params[3], b.params[4])
            else:
                dist = dp.SphereDistance.create(b.params[0])
            pile.add(dist, _al.uint3(*b.resolution), sign=b.sign, mass=b.mass,
                     inertia_tensor=dp.f3(*b.inertia), x=dp.f3(*b.x), q=dp.f4(*b.q))
        piles.append(pile)
        if log and (k % 32 == 0 or k == len(scenes) - 1):
            log("batch: %d piles built" % (k + 1))
    solver = dp.SolverDf(dp, piles, [sc.n for sc in scenes],
                         [_al.uint3(*sc.grid_res) for sc in scenes],
                         [_al.int3(*sc.grid_offset) for sc in scenes], cn=cn, cni=cni)
    solver.dt = s0.dt
    solver.max_dt = s0.dt if s0.max_dt is None else s0.max_dt
    solver.min_dt = s0.dt if s0.min_dt is None else s0.min_dt
    solver.cfl = s0.cfl
    solver.density_change_tolerance = s0.density_change_tolerance
    solver.density_error_tolerance = s0.density_error_tolerance
    solver.min_density_solve = s0.min_density_solve
    solver.max_density_solve = s0.max_density_solve
    solver.min_divergence_solve = s0.min_divergence_solve
    solver.max_divergence_solve = s0.max_divergence_solve
    X = dp.coat(solver.particle_x)
    V = dp.coat(solver.particle_v)
    for k, sc in enumerate(scenes):
        first, count = solver.scene_range(k)
        assert count == sc.n
        X.set(sc.x, offset=first)
        V.set(sc.v, offset=first)
    return dp, piles, solver, cn, cni
