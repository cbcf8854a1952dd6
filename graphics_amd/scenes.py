"""Synthetic dam-break workloads (identical bytes for the GPU path and the oracle).

The reference builds its scenes in C++ (SphWithIO::addFluidCuboid,
SphWithIO.cpp:149-194, and SphWithIO::addRigidCuboid -> makeCubeSampleParts,
SphWithIO.cpp:489-522, 642-659).  The same samplers live in libsphgpu.so
(sphgpu_scene_lattice / sphgpu_scene_box); this module only lays scenes out.
Jitter is uniform in +-1 % of the spacing per axis from PCG32 seeded with
1000 + workload number.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import sphgpu as sg


@dataclass
class Scene:
    name: str
    params: dict
    phases: list = field(default_factory=list)         # (rest_density, visc_alpha)
    fluids: list = field(default_factory=list)         # (scene, phase, xyz[n,3])
    solids: list = field(default_factory=list)         # (scene, xyz[n,3], visc_alpha)
    domains: list = field(default_factory=list)        # (min[3], max[3]) per scene
    rigids: list = field(default_factory=list)         # (scene, body-frame xyz[n,3], visc_alpha, pose)
    rigid_mass: list = field(default_factory=list)     # (mass, inertia[3]) per rigid, for a host step
    beta: object = None                                # per-particle rest-density factor by add order (IISph DF_*)
    velocity: object = None                            # initial velocity (N, 3) by add order (default 0)

    @property
    def num_fluid(self) -> int:
        return sum(len(f[2]) for f in self.fluids)

    @property
    def num_boundary(self) -> int:
        return sum(len(s[1]) for s in self.solids) + sum(len(r[1]) for r in self.rigids)

    @property
    def num_scenes(self) -> int:
        return len(self.domains)


def make_simulation(sc: Scene, **overrides) -> sg.Simulation:
    p = dict(sc.params)
    p.update(overrides)
    p["num_scenes"] = sc.num_scenes
    p["space_min"], p["space_max"] = sc.domains[0]
    sim = sg.Simulation(sg.default_params(**p))
    for s, (mn, mx) in enumerate(sc.domains):
        sim.set_scene_domain(s, mn, mx)
    for rho0, visc in sc.phases:
        sim.add_phase(rho0, visc)
    first = 0
    for s, ph, xyz in sc.fluids:
        vel = None if sc.velocity is None else np.asarray(sc.velocity)[first:first + len(xyz)]
        sim.add_fluid(xyz, vel, phase=ph, scene=s)
        first += len(xyz)
    for s, xyz, visc in sc.solids:
        sim.add_solid(xyz, visc_alpha=visc, scene=s)
    for s, xyz, visc, pose in sc.rigids:
        k = sim.add_rigid(xyz, visc_alpha=visc, scene=s)
        sim.set_solid_pose(k, *pose)
    if sc.beta is not None:
        sim.set_fluid_beta(0, sc.beta)
    return sim


def pose(angle_y=0.0, translation=(0.0, 0.0, 0.0), com=None, v=(0.0, 0.0, 0.0), w=(0.0, 0.0, 0.0)):
    """(T[16] OpenGL column-major, com, v, w) for a rotation about +y by
    angle_y followed by a translation; com defaults to the translation."""
    c, s = np.cos(angle_y), np.sin(angle_y)
    T = np.array([c, 0.0, -s, 0.0,  0.0, 1.0, 0.0, 0.0,  s, 0.0, c, 0.0,
                  translation[0], translation[1], translation[2], 1.0])
    return (T, tuple(translation) if com is None else tuple(com), tuple(v), tuple(w))


def dam_with_box(solver=sg.IISPH, n=(12, 10, 12), box=(0.5, 0.4, 0.4), spacing=0.02, half=(0.05, 0.05, 0.05),
                 centre=(0.31, 0.1, 0.14), **kw) -> Scene:
    """Small dam break with a dynamic rigid box (its surface samples in body
    coordinates, posed at `centre`) standing in the fluid's path."""
    sc = small_dam(solver, n=n, box=box, spacing=spacing, **kw)
    sc.name = f"dam_with_box_{solver}"
    local = sg.box_walls((0.0, 0.0, 0.0), half, spacing, True)
    sc.rigids.append((0, local, 0.01, pose(0.0, centre)))
    return sc


def rigid_world(sc: Scene, gravity=None):
    """A host rigid-body world (graphics_amd.rigid_host) holding the scene's
    rigids at their initial poses; solid indices follow the static solids."""
    from .rigid_host import RigidWorld
    w = RigidWorld(sc.params.get("gravity", (0.0, -9.81, 0.0)) if gravity is None else gravity)
    for k, ((_, _, _, pz), (mass, inertia)) in enumerate(zip(sc.rigids, sc.rigid_mass)):
        T = np.asarray(pz[0], np.float64)
        R = np.stack([T[0:3], T[4:7], T[8:11]], axis=1)
        w.add(len(sc.solids) + k, mass, inertia, pz[1], R=R, v=pz[2], w=pz[3])
    return w


def _layers(xyz, spacing, level, amp, length, x0):
    """Heavy-layer mask: below level + amp * spacing * sin(2 pi (x - x0) / length)."""
    iface = level + amp * spacing * np.sin(2.0 * np.pi * (xyz[:, 0] - x0) / length)
    return xyz[:, 1] < iface


def _add_box_body(sc: Scene, s: int, half, centre, spacing, visc, density):
    from .rigid_host import box_inertia
    local = sg.box_walls((0.0, 0.0, 0.0), half, spacing, True)
    mass = density * 8.0 * half[0] * half[1] * half[2]
    sc.rigids.append((s, local, visc, pose(0.0, centre)))
    sc.rigid_mass.append((mass, box_inertia(mass, half)))
    return len(local)


def layers_with_boxes(solver=sg.IISPH, n=(20, 12, 10), box=(0.44, 0.5, 0.24), spacing=0.02, jitter=0.01,
                      seed=1006, rho=(1000.0, 3000.0), amp=0.5, halves=((0.05, 0.04, 0.05),),
                      body_density=2000.0, gap=2.0, **kw) -> Scene:
    """Two immiscible layers (rho[1] below, rho[0] on top, interface
    perturbed by amp * spacing) with box bodies of density body_density
    placed `gap` spacings above the top fluid layer, in a row along x."""
    sc = small_dam(solver, n=n, box=box, spacing=spacing, jitter=jitter, seed=seed, **kw)
    sc.name = f"layers_with_boxes_{solver}"
    sc.phases[:] = [(rho[0], 0.01), (rho[1], 0.01)]
    xyz = sc.fluids[0][2]
    heavy = _layers(xyz, spacing, spacing + 0.5 * n[1] * spacing, amp, n[0] * spacing, spacing)
    sc.fluids[:] = [(0, 1, xyz[heavy]), (0, 0, xyz[~heavy])]
    top = spacing * n[1] + gap * spacing
    slot = box[0] / len(halves)
    for k, h in enumerate(halves):
        centre = (slot * (k + 0.5), top + h[1], 0.5 * box[2])
        _add_box_body(sc, 0, h, centre, spacing, 0.01, body_density)
    return sc


def iisph_c5(seed=1005, scale=1.0) -> Scene:
    """Multiphase IISPH with rigid coupling: two immiscible layers at 1000 and
    3000 kg/m^3 (599 x 96 x 139 = 7,993,056 particles at 0.01 m, the heavy
    layer below, interface perturbed by 0.5 spacing), and 20 box bodies of
    2,000-20,000 surface samples each (log-uniform from the seed), density
    2000 kg/m^3, starting just above the free surface in a 10 x 2 grid, in a
    closed 6.0 x 2.0 x 1.4 m tank.  The pressure solve follows the
    number-density force between the phases (multiphase_consistent): with
    IISph_M's own-mass solve the interface does not hold.  `scale` < 1
    shrinks the fluid height and box sizes for tests."""
    s = 0.01
    rng = np.random.default_rng(seed)
    p = dict(solver=sg.IISPH, spacing_r=s, smooth_radius_h=2 * s, dt=1e-3, cfl=0.4, dt_min=1e-5,
             gravity=(0.0, -9.81, 0.0), sound_speed=20.0, nu=1e-3, omega=0.5,
             density_tol=1e-3, min_iter_density=2, max_iter_density=100, multiphase_consistent=1)
    size = (6.0, 2.0, 1.4)
    ny = max(4, int(round(96 * scale)))
    nx, nz = int(round(size[0] / s)) - 1, int(round(size[2] / s)) - 1
    sc = Scene(f"iisph_layers_boxes_{nx * ny * nz}", p)
    sc.phases[:] = [(1000.0, 0.01), (3000.0, 0.01)]
    _tank(sc, 0, size, s, True, 0.01)
    sc.domains.append(((-s, -s, -s), tuple(t + s for t in size)))
    xyz = sg.lattice((nx, ny, nz), (s, s, s), s, 0.01, seed, 0)
    heavy = _layers(xyz, s, s + 0.5 * ny * s, 0.5, nx * s, s)
    sc.fluids[:] = [(0, 1, xyz[heavy]), (0, 0, xyz[~heavy])]
    top = s * ny + 2.0 * s
    for k in range(20):
        target = math.exp(rng.uniform(math.log(2000.0), math.log(20000.0))) * scale ** 2
        edge = min(0.56, s * math.sqrt(target / 6.0))
        h = 0.5 * edge
        centre = (0.6 * (k % 10) + 0.3, top + h, 0.7 * (k // 10) + 0.35)
        _add_box_body(sc, 0, (h, h, h), centre, s, 0.01, 2000.0)
    return sc


def _tank(scene: Scene, s: int, size, spacing, have_top, visc, origin=(0.0, 0.0, 0.0)):
    half = [0.5 * float(t) for t in size]
    center = [origin[k] + half[k] for k in range(3)]
    walls = sg.box_walls(center, half, spacing, have_top)
    scene.solids.append((s, walls, visc))
    return walls


def small_dam(solver=sg.DFSPH, n=(10, 10, 10), box=(0.4, 0.3, 0.3), spacing=0.02, jitter=0.01,
              seed=1001, have_top=True, offset=1.0, **kw) -> Scene:
    """A small dam break for parity tests (seconds on the oracle); the fluid
    lattice starts `offset` spacings from the walls."""
    p = dict(solver=solver, spacing_r=spacing, smooth_radius_h=2 * spacing, dt=1e-3,
             gravity=(0.0, -9.81, 0.0), sound_speed=20.0, gamma=7.0, nu=1e-3, cfl=0.0,
             density_tol=1e-4, divergence_tol=1.0, min_iter_density=2, max_iter_density=100,
             min_iter_div=1, max_iter_div=30, omega=0.5, warm_start=0)
    if solver == sg.WCSPH:
        p["dt"] = 1e-4
    if solver == sg.IISPH:
        p["density_tol"] = 1e-3
    if solver == sg.PCISPH:      # PciSph.h:35 and m_TH.densityErro_eta (SphBase.cpp:221)
        p.update(density_tol=0.01, min_iter_density=3, max_iter_density=10)
    p.update(kw)
    sc = Scene(f"small_dam_{solver}", p)
    sc.phases.append((1000.0, 0.01))
    _tank(sc, 0, box, spacing, have_top, 0.01)
    top = box[1] if have_top else box[1] + 0.2
    sc.domains.append(((0.0, 0.0, 0.0), (box[0], top, box[2])))
    o = offset * spacing
    xyz = sg.lattice(n, (o, o, o), spacing, jitter, seed, 0)
    sc.fluids.append((0, 0, xyz))
    return sc


def two_layer(solver=sg.IISPH, n=(12, 16, 12), box=(0.4, 0.5, 0.4), spacing=0.02, jitter=0.01,
              seed=1005, rho=(1000.0, 3000.0), amp=0.5, **kw) -> Scene:
    """Two immiscible layers at rest densities rho[1] (below) and rho[0] (on
    top); the interface height is perturbed by amp * spacing * sin(2 pi x / L)."""
    sc = small_dam(solver, n=n, box=box, spacing=spacing, jitter=jitter, seed=seed, **kw)
    sc.name = f"two_layer_{solver}"
    sc.phases[:] = [(rho[0], 0.01), (rho[1], 0.01)]
    xyz = sc.fluids[0][2]
    length = n[0] * spacing
    level = spacing + 0.5 * n[1] * spacing
    iface = level + amp * spacing * np.sin(2.0 * np.pi * (xyz[:, 0] - spacing) / length)
    heavy = xyz[:, 1] < iface
    sc.fluids[:] = [(0, 1, xyz[heavy]), (0, 0, xyz[~heavy])]
    return sc


def wcsph_c1(seed=1001) -> Scene:
    """WCSPH 3-D dam break: 25^3 fluid at 0.02 m in a 1.0x0.6x0.6 m open-top box."""
    s = 0.02
    p = dict(solver=sg.WCSPH, spacing_r=s, smooth_radius_h=2 * s, dt=1e-4,
             gravity=(0.0, -9.81, 0.0), sound_speed=20.0, gamma=7.0, cfl=0.0)
    sc = Scene("wcsph_dam_15625", p)
    sc.phases.append((1000.0, 0.01))
    _tank(sc, 0, (1.0, 0.6, 0.6), s, False, 0.01)
    sc.domains.append(((0.0, 0.0, 0.0), (1.0, 1.0, 0.6)))
    sc.fluids.append((0, 0, sg.lattice((25, 25, 25), (s, s, s), s, 0.01, seed, 0)))
    return sc


def dfsph_c2(seed=1002, n=100) -> Scene:
    """DFSPH dam break: n^3 fluid (1,000,000 at n=100) at 0.01 m in a closed
    4x2x2 m box.  Density tolerance 0.01 % average, divergence tolerance 0.1 %
    of rho0 per second; both loops may run 100 iterations (the reference breaks
    the divergence loop after 30, DFSPH.cpp:772, where the plain solve has
    mostly not reached this tolerance yet)."""
    s = 0.01
    p = dict(solver=sg.DFSPH, spacing_r=s, smooth_radius_h=2 * s, dt=1e-3, cfl=0.4, dt_min=1e-5,
             gravity=(0.0, -9.81, 0.0), sound_speed=20.0, nu=1e-3,
             density_tol=1e-4, divergence_tol=1.0, min_iter_density=2, max_iter_density=100,
             min_iter_div=1, max_iter_div=100)
    sc = Scene(f"dfsph_dam_{n ** 3}", p)
    sc.phases.append((1000.0, 0.01))
    _tank(sc, 0, (4.0, 2.0, 2.0), s, True, 0.01)
    sc.domains.append(((0.0, 0.0, 0.0), (4.0, 2.0, 2.0)))
    sc.fluids.append((0, 0, sg.lattice((n, n, n), (s, s, s), s, 0.01, seed, 0)))
    return sc


def pcisph_dam(seed=1002, n=100) -> Scene:
    """PCISPH dam break on the DFSPH dam's geometry (n^3 = 1,000,000 fluid at
    0.01 m, closed 4x2x2 m box), dt 5e-4 with CFL 0.4, eta 1 %, 3-10
    iterations (PciSph.h:35, SphBase.cpp:221); the fluid starts 1.5 spacings
    from the walls (one spacing starts it 10 % compressed, which the
    single-particle stiffness overshoots)."""
    s = 0.01
    p = dict(solver=sg.PCISPH, spacing_r=s, smooth_radius_h=2 * s, dt=5e-4, cfl=0.4, dt_min=1e-5,
             gravity=(0.0, -9.81, 0.0), sound_speed=20.0, nu=1e-3,
             density_tol=0.01, min_iter_density=3, max_iter_density=10)
    sc = Scene(f"pcisph_dam_{n ** 3}", p)
    sc.phases.append((1000.0, 0.01))
    _tank(sc, 0, (4.0, 2.0, 2.0), s, True, 0.01)
    sc.domains.append(((0.0, 0.0, 0.0), (4.0, 2.0, 2.0)))
    sc.fluids.append((0, 0, sg.lattice((n, n, n), (1.5 * s, 1.5 * s, 1.5 * s), s, 0.01, seed, 0)))
    return sc


def dfsph_c3(seed=1003, n=126) -> Scene:
    """DFSPH, two facing n^3 blocks (4,000,752 at n=126) at 0.01 m in a closed
    6x3x3 m box, tolerances as dfsph_c2; the learned warm start trains on the
    first 200 steps."""
    s = 0.01
    sc = dfsph_c2(seed, 1)
    sc.name = f"dfsph_two_blocks_{2 * n ** 3}"
    sc.params["learn_steps"] = 200
    sc.solids.clear()
    sc.fluids.clear()
    sc.domains.clear()
    _tank(sc, 0, (6.0, 3.0, 3.0), s, True, 0.01)
    sc.domains.append(((0.0, 0.0, 0.0), (6.0, 3.0, 3.0)))
    ext = (n - 1) * s
    a = sg.lattice((n, n, n), (s, s, s), s, 0.01, seed, 0)
    b = sg.lattice((n, n, n), (6.0 - s - ext, s, 3.0 - s - ext), s, 0.01, seed, 1)
    sc.fluids.append((0, 0, a))
    sc.fluids.append((0, 0, b))
    return sc


def iisph_batched_c4(seed=1004, scenes=64, nmin=10_000, nmax=500_000, scale=1.0) -> Scene:
    """Batched IISPH: `scenes` dam breaks, fluid counts log-uniform in [nmin, nmax],
    box aspect ratio uniform in [1, 3] (`scale` shrinks counts for tests)."""
    s = 0.01
    rng = np.random.default_rng(seed)
    p = dict(solver=sg.IISPH, spacing_r=s, smooth_radius_h=2 * s, dt=1e-3, cfl=0.4, dt_min=1e-5,
             gravity=(0.0, -9.81, 0.0), sound_speed=20.0, nu=1e-3, omega=0.5,
             density_tol=1e-3, min_iter_density=2, max_iter_density=100)
    sc = Scene(f"iisph_batched_{scenes}", p)
    sc.phases.append((1000.0, 0.01))
    for k in range(scenes):
        nk = math.exp(rng.uniform(math.log(nmin), math.log(nmax))) * scale
        aspect = rng.uniform(1.0, 3.0)
        m = max(2, int(round(nk ** (1.0 / 3.0))))
        width = (m + 3) * s
        size = (aspect * width + 2 * s, 1.5 * width, width)
        _tank(sc, k, size, s, True, 0.01)
        # the walls lie on the tank faces: the domain extends one spacing beyond
        sc.domains.append(((-s, -s, -s), tuple(t + s for t in size)))
        sc.fluids.append((k, 0, sg.lattice((m, m, m), (s, s, s), s, 0.01, seed, k)))
    return sc


WORKLOADS = {
    "wcsph_c1": wcsph_c1,
    "dfsph_c2": dfsph_c2,
    "dfsph_c3": dfsph_c3,
    "iisph_c4": iisph_batched_c4,
    "iisph_c5": iisph_c5,
    "pcisph_dam": pcisph_dam,
}
