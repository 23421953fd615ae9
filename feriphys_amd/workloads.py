"""Synthetic workloads for the springy step (numbers and choices are the documented benchmark ones).

Each workload is described by plain arguments (constructor arguments of the reference API, a
Config, obstacle triangles and a seeded initial-state jitter), so the same description can be fed
to the GPU path (feriphys_amd.springy) and to the CPU oracle.

Choices (the reference's literal constants diverge at dt = 1/600, so the cloth uses these):
  * Spring constants are given per unit strain: the reference scales a nominal constant by
    1/rest_length (springy_mesh.rs:83-84), so nominal = k * L_class makes the effective stiffness
    k at the class's rest length (tensile L = s, shear L = s*sqrt(2), bend L = 2s).
  * The reference adds Config.gravity as a force (springy_mesh.rs:194-196), so an acceleration g
    is passed as the force m*g.
  * The cloth lies in the xz plane (rows along +z) and "top" is the last row.
  * Jitter: uniform +-1e-3 m on every coordinate of every non-pinned vertex, applied after
    construction (rest lengths stay those of the exact grid), from numpy's PCG64 with the seed.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

F32 = np.float32


@dataclass
class ClothSpec:
    rows: int
    cols: int
    spacing: float
    position: Tuple[float, float, float]
    point_mass: float
    tk: float
    tc: float
    sk: float
    sc: float
    bk: float
    bc: float
    pins: List[int]
    plane: int = 1  # xz

    def args(self):
        return (self.rows, self.cols, self.spacing, self.position, self.point_mass, self.tk, self.tc, self.sk,
                self.sc, self.bk, self.bc, self.pins)


@dataclass
class ConfigSpec:
    integration: int = 1
    dt: float = 0.001
    gravity: Tuple[float, float, float] = (0.0, -10.0, 0.0)
    wind: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    lift_coefficient: float = 1.0
    drag_coefficient: float = 1.0
    coefficient_of_restitution: float = 0.95
    coefficient_of_friction: float = 0.3

    def f32(self) -> "ConfigSpec":
        """All values rounded to f32 (the reference's storage)."""
        f = lambda v: float(F32(v))
        return ConfigSpec(self.integration, f(self.dt), tuple(f(x) for x in self.gravity),
                          tuple(f(x) for x in self.wind), f(self.lift_coefficient), f(self.drag_coefficient),
                          f(self.coefficient_of_restitution), f(self.coefficient_of_friction))


@dataclass
class Scene:
    cloths: List[ClothSpec]
    config: ConfigSpec
    obstacles: List[Tuple[np.ndarray, np.ndarray]] = field(default_factory=list)  # (positions (n,3) f32, indices)
    jitter: float = 1e-3
    seed: int = 0
    steps: int = 600
    name: str = ""


def per_strain(k: float, length: float) -> float:
    """Nominal constant giving effective stiffness/damping k at rest length `length` (f32)."""
    return float(F32(F32(k) * F32(length)))


def cloth_spec(rows, cols, spacing=0.05, mass=0.01, tensile=(500.0, 2.0), shear=(300.0, 1.0), bend=(50.0, 0.5),
               pins=(), position=(0.0, 0.0, 0.0), plane=1) -> ClothSpec:
    s = float(F32(spacing))
    ls, lsh, lb = s, float(F32(s * math.sqrt(2.0))), float(F32(2.0 * s))
    return ClothSpec(rows, cols, s, tuple(float(F32(p)) for p in position), float(F32(mass)),
                     per_strain(tensile[0], ls), per_strain(tensile[1], ls),
                     per_strain(shear[0], lsh), per_strain(shear[1], lsh),
                     per_strain(bend[0], lb), per_strain(bend[1], lb), [int(p) for p in pins], plane)


def gravity_force(mass: float, g: float = 9.81):
    return (0.0, float(F32(F32(-g) * F32(mass))), 0.0)


def jitter_records(records: np.ndarray, pins: Sequence[int], amount: float, seed: int) -> np.ndarray:
    """Adds seeded uniform +-amount to the positions of non-pinned vertices (columns 1..3)."""
    rec = np.array(records, dtype=F32, copy=True)
    rng = np.random.default_rng(seed)
    d = rng.uniform(-amount, amount, size=(rec.shape[0], 3)).astype(F32)
    if len(pins):
        d[np.asarray(list(pins), dtype=np.int64)] = 0.0
    rec[:, 1:4] = (rec[:, 1:4] + d).astype(F32)
    return rec


def floor_obstacle(y: float = -2.0, half: float = 1.0):
    """The cloth demo's floor (demos/cloth.rs:250-259): a square at height y, normal +y."""
    p = np.array([[-half, y, half], [half, y, half], [half, y, -half], [-half, y, -half]], F32)
    return p, np.array([0, 1, 2, 0, 2, 3], np.uint64)


def _rotation(ax, ay, az):
    cx, sx, cy, sy, cz, sz = math.cos(ax), math.sin(ax), math.cos(ay), math.sin(ay), math.cos(az), math.sin(az)
    rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return rz @ ry @ rx


def icosphere(subdivisions: int):
    """Unit icosphere: vertices (float64) and outward counter-clockwise triangles (int64)."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
         [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
         [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    verts = [np.array(p, float) / np.linalg.norm(p) for p in v]
    for _ in range(subdivisions):
        cache = {}
        nf = []

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[k] = len(verts) - 1
            return cache[k]

        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        f = nf
    return np.array(verts), np.array(f, np.int64)


# Rotation of the collider sphere chosen so that no face normal has a tiny component on the axis the
# reference's containment test drops (that test picks the axis by SIGNED maximum,
# collidable_mesh.rs:133-154); the smallest such component over the 320 faces is 0.0228.
SPHERE_ROTATION = (5.34088472, 1.83004511, 1.68085871)


def sphere_obstacle(radius: float, centre, subdivisions: int = 2, rotation=SPHERE_ROTATION):
    v, f = icosphere(subdivisions)
    v = v @ _rotation(*rotation).T
    pos = (v * radius + np.asarray(centre, float)).astype(F32)
    return pos, f.reshape(-1).astype(np.uint64)


def config0(steps: int = 600, seed: int = 0, rows: int = 32, cols: int = 32, pins: str = "corners") -> Scene:
    """One square cloth of 32x32 vertices, spacing 0.05 m, 0.01 kg per vertex; tensile (500, 2),
    shear (300, 1), bend (50, 0.5); the two top corners pinned; gravity 9.81 m/s^2, wind (1, 0, 0.5)
    m/s, drag 0.5, lift 0.2; RK4 at dt = 1/600 s; +-1e-3 m jitter.

    pins="corners" is the specified case.  The reference's integrators amplify every undamped mode
    (the frozen-force RK4 step has determinant 1 + h^2 K / 2), and the out-of-plane modes of a
    tensioned cloth have no damping (dampers act along struts), so a cloth hanging from two corners
    blows up after ~270 steps in the reference method itself.  pins="edge" pins the whole top row,
    which keeps the tension low enough to stay bounded for 600 steps."""
    pins = ([(rows - 1) * cols, rows * cols - 1] if pins == "corners"
            else [(rows - 1) * cols + c for c in range(cols)])
    c = cloth_spec(rows, cols, pins=pins)
    cfg = ConfigSpec(1, 1.0 / 600.0, gravity_force(0.01), (1.0, 0.0, 0.5), 0.2, 0.5, 0.95, 0.3).f32()
    return Scene([c], cfg, [], 1e-3, seed, steps, f"cloth{rows}x{cols}")


def config1(n: int = 1024, steps: int = 600, seed: int = 0, subdivisions: int = 2) -> Scene:
    """One n x n grid cloth (1024: 1,048,576 vertices), config-0 strut classes; top edge pinned; one
    icosphere collider of radius 0.3 x cloth width whose top is 0.5 m below the cloth (restitution
    0.3, friction 0.2); RK4 at dt = 1/600 s; +-1e-3 m jitter."""
    pins = [(n - 1) * n + c for c in range(n)]
    cl = cloth_spec(n, n, pins=pins)
    width = (n - 1) * cl.spacing
    r = 0.3 * width
    sphere = sphere_obstacle(r, (0.0, -(0.5 + r), 0.0), subdivisions)
    cfg = ConfigSpec(1, 1.0 / 600.0, gravity_force(0.01), (1.0, 0.0, 0.5), 0.2, 0.5, 0.3, 0.2).f32()
    return Scene([cl], cfg, [sphere], 1e-3, seed, steps, f"cloth{n}x{n}+sphere")


def config2(n_cloths: int = 4096, seed: int = 0, smin: int = 8, smax: int = 256, steps: int = 100) -> Scene:
    """n_cloths independent grid cloths: rows and cols each drawn log-uniform in [smin, smax]
    (about 21M vertices for 4096); per-cloth stiffness scale uniform in [0.5, 2] on the config-0
    tensile/shear/bend constants (damping unchanged); pins: 1-4 random corners or one full edge;
    per-cloth wind with direction uniform on the upper hemisphere and magnitude uniform in [0, 3]
    m/s; config-0 mass, spacing, gravity, drag and lift; RK4 at dt = 1/600 s.  Each cloth is a
    mesh with its own Config (one reference Simulation per cloth)."""
    rng = np.random.default_rng(seed)
    lo, hi = math.log(smin), math.log(smax + 1)
    cloths, winds = [], []
    for k in range(n_cloths):
        R = int(min(smax, max(smin, math.floor(math.exp(rng.uniform(lo, hi))))))
        C = int(min(smax, max(smin, math.floor(math.exp(rng.uniform(lo, hi))))))
        sk = float(rng.uniform(0.5, 2.0))
        corners = [0, C - 1, (R - 1) * C, R * C - 1]
        if rng.uniform() < 0.5:
            npin = int(rng.integers(1, 5))
            pins = sorted(int(x) for x in rng.choice(corners, npin, replace=False))
        else:
            edge = int(rng.integers(0, 4))
            pins = ([c for c in range(C)] if edge == 0 else [(R - 1) * C + c for c in range(C)] if edge == 1
                    else [r * C for r in range(R)] if edge == 2 else [r * C + C - 1 for r in range(R)])
        cloths.append(cloth_spec(R, C, tensile=(500.0 * sk, 2.0), shear=(300.0 * sk, 1.0), bend=(50.0 * sk, 0.5),
                                 pins=pins))
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        d[1] = abs(d[1])
        winds.append(tuple(float(F32(x)) for x in d * rng.uniform(0.0, 3.0)))
    base = ConfigSpec(1, 1.0 / 600.0, gravity_force(0.01), (0.0, 0.0, 0.0), 0.2, 0.5, 0.95, 0.3).f32()
    sc = Scene(cloths, base, [], 1e-3, seed, steps, f"{n_cloths} cloths")
    sc.mesh_winds = winds
    return sc


def springy_cube():
    """The reference's springy cube demo (demos/spring_mass_damper.rs:230-272, graphics/forms.rs:91-146,
    :329-342): a unit cube of 8 vertices and 12 triangles, total mass 80, struts at the stiff
    defaults (999999, 700, springy_mesh.rs:12-13) plus six diagonals by add_strut, torsional springs
    configured (7000, 7000; :16-17) but not applied in the step; a 2 x 2 floor square at y = -2.
    Returns (positions, indices, mass, k, c, torsion, extra struts, floor positions, floor indices)."""
    pos = np.array([[-0.5, -0.5, 0.5], [0.5, -0.5, 0.5], [0.5, 0.5, 0.5], [-0.5, 0.5, 0.5],
                    [-0.5, -0.5, -0.5], [0.5, -0.5, -0.5], [0.5, 0.5, -0.5], [-0.5, 0.5, -0.5]], np.float32)
    idx = [0, 1, 2, 2, 3, 0, 1, 5, 6, 6, 2, 1, 7, 6, 5, 5, 4, 7, 4, 0, 3, 3, 7, 4, 4, 5, 1, 1, 0, 4, 3, 2, 6, 6, 7, 3]
    extra = [(1, 3), (2, 5), (4, 6), (0, 7), (0, 5), (2, 7)]
    floor = np.array([[-1.0, -2.0, 1.0], [1.0, -2.0, 1.0], [1.0, -2.0, -1.0], [-1.0, -2.0, -1.0]], np.float32)
    return pos, idx, 80.0, 999999.0, 700.0, (7000.0, 7000.0), extra, floor, [0, 1, 2, 0, 2, 3]


def surrogate_training_scene(n_cloths: int = 128, seed: int = 0, smin: int = 12, smax: int = 48,
                             sphere: bool = False) -> Scene:
    """Cloths whose plain trajectories train the surrogate: rows and cols each log-uniform in
    [smin, smax], one full edge pinned (chosen at random), config-0 stiffness, mass, spacing,
    gravity, drag and lift; per-cloth wind with direction uniform on the upper hemisphere and
    magnitude uniform in [0, 3] m/s; RK4 at dt = 1/600 s; +-1e-3 m jitter.  Corner pins are left
    out: a cloth hanging from corners diverges in the reference method itself (see config0).
    sphere: one icosphere collider (radius 0.3-2 m, apex 0.1-0.6 m below the cloths, which all
    start centred on the origin) that every cloth falls onto."""
    rng = np.random.default_rng(seed)
    lo, hi = math.log(smin), math.log(smax + 1)
    cloths, winds = [], []
    for k in range(n_cloths):
        R = int(min(smax, max(smin, math.floor(math.exp(rng.uniform(lo, hi))))))
        C = int(min(smax, max(smin, math.floor(math.exp(rng.uniform(lo, hi))))))
        edge = int(rng.integers(0, 4))
        pins = ([c for c in range(C)] if edge == 0 else [(R - 1) * C + c for c in range(C)] if edge == 1
                else [r * C for r in range(R)] if edge == 2 else [r * C + C - 1 for r in range(R)])
        cloths.append(cloth_spec(R, C, pins=pins))
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        d[1] = abs(d[1])
        winds.append(tuple(float(F32(x)) for x in d * rng.uniform(0.0, 3.0)))
    base = ConfigSpec(1, 1.0 / 600.0, gravity_force(0.01), (0.0, 0.0, 0.0), 0.2, 0.5, 0.95, 0.3).f32()
    obstacles = []
    if sphere:
        r = float(rng.uniform(0.3, 2.0))
        obstacles = [sphere_obstacle(r, (0.0, -(float(rng.uniform(0.1, 0.6)) + r), 0.0), 2)]
        base = ConfigSpec(1, 1.0 / 600.0, gravity_force(0.01), (0.0, 0.0, 0.0), 0.2, 0.5, 0.3, 0.2).f32()
    sc = Scene(cloths, base, obstacles, 1e-3, seed, 600, f"{n_cloths} training cloths")
    sc.mesh_winds = winds
    return sc


def mesh_configs(scene: Scene):
    """The Config of every mesh: the scene's, with the per-cloth wind where the scene has one."""
    winds = getattr(scene, "mesh_winds", None)
    out = []
    for k in range(len(scene.cloths)):
        c = scene.config
        w = winds[k] if winds is not None else c.wind
        out.append(ConfigSpec(c.integration, c.dt, c.gravity, w, c.lift_coefficient, c.drag_coefficient,
                              c.coefficient_of_restitution, c.coefficient_of_friction))
    return out


def icosphere_fast(subdivisions: int):
    """icosphere() with the subdivision vectorised (edge midpoints by numpy unique): vertices
    (float64, unit radius) and outward counter-clockwise triangles (int64)."""
    v, f = icosphere(0)
    for _ in range(subdivisions):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        key = np.sort(e, axis=1)
        uniq, inv = np.unique(key, axis=0, return_inverse=True)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = v.shape[0]
        v = np.concatenate([v, mid])
        m = inv.reshape(3, -1).T + base  # midpoints of (v0v1, v1v2, v2v0) per face
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        ab, bc, ca = m[:, 0], m[:, 1], m[:, 2]
        f = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1), np.stack([c, ca, bc], 1),
                            np.stack([ab, bc, ca], 1)])
    return v, f


@dataclass
class SurfaceSpec:
    """A SpringyMesh from a triangle list (SpringyMesh::new) plus bending struts across every
    shared edge (add_strut between the two opposite vertices, in strut order)."""
    positions: np.ndarray  # (n, 3) f32
    indices: np.ndarray    # (3F,) uint64
    mass: float            # total
    k: float
    c: float
    bk: float
    bc: float


def bend_pairs(v01: np.ndarray, faces: np.ndarray, face_verts: np.ndarray) -> np.ndarray:
    """For every strut with two adjacent faces (strut order), the two vertices opposite it."""
    both = (faces[:, 0] >= 0) & (faces[:, 1] >= 0)
    s = v01[both].astype(np.int64)
    out = []
    for side in (0, 1):
        fv = face_verts[faces[both, side]].astype(np.int64)
        is_end = (fv == s[:, :1]) | (fv == s[:, 1:2])
        out.append(fv[~is_end].reshape(-1))
    return np.stack(out, 1)


def config4(n_meshes: int = 64, seed: int = 0, levels=None) -> "SurfaceScene":
    """Seeded icospheres, subdivided to ~2M vertices over 64 meshes of uneven size (levels 4-7),
    vertices perturbed radially by +-2%; tensile struts on triangle edges and bending struts across
    shared edges; 0.005 kg per vertex; per-triangle wind/drag/lift; one ground-plane collider at
    y = 0.  Spring constants per unit strain at the mesh's mean edge length (tensile 100 N/m,
    0.3 N s/m; bend 10 N/m, 0.05 N s/m), chosen so that the explicit step is stable at dt = 1/600 s
    for this mass and a vertex degree of 5-7."""
    rng = np.random.default_rng(seed)
    if levels is None:
        # 8 x 163,842 + 14 x 40,962 + 20 x 10,242 + 22 x 2,562 = 2,145,408 vertices (levels 7, 6, 5, 4)
        levels = [7] * 8 + [6] * 14 + [5] * 20 + [4] * 22
        levels = levels[:n_meshes]
    levels = list(levels)
    rng.shuffle(levels)
    meshes = []
    cache = {}
    for k, lv in enumerate(levels):
        if lv not in cache:
            cache[lv] = icosphere_fast(lv)
        v, f = cache[lv]
        r = 0.004 * 2 ** lv * rng.uniform(0.8, 1.25)  # edge length ~5 mm x 0.8..1.25
        noise = rng.uniform(-0.02, 0.02, size=(v.shape[0], 1))
        centre = np.array([(k % 8) * 3.0, r + rng.uniform(0.05, 0.2), (k // 8) * 3.0])
        pos = (v * r * (1.0 + noise) + centre).astype(F32)
        e = np.linalg.norm(pos[f[:, 0]].astype(np.float64) - pos[f[:, 1]], axis=1).mean()
        meshes.append(SurfaceSpec(pos, f.reshape(-1).astype(np.uint64), float(F32(0.005 * v.shape[0])),
                                  per_strain(100.0, e), per_strain(0.3, e), per_strain(10.0, 2 * e),
                                  per_strain(0.05, 2 * e)))
    ground = (np.array([[-1e3, 0.0, 1e3], [1e3, 0.0, 1e3], [1e3, 0.0, -1e3], [-1e3, 0.0, -1e3]], F32),
              np.array([0, 1, 2, 0, 2, 3], np.uint64))
    cfg = ConfigSpec(1, 1.0 / 600.0, gravity_force(0.005), (1.0, 0.0, 0.5), 0.2, 0.5, 0.3, 0.2).f32()
    return SurfaceScene(meshes, cfg, [ground], f"{len(meshes)} icospheres")


def surface_training_scene(n_meshes: int = 400, seed: int = 0, levels=(2, 3, 4, 5)) -> "SurfaceScene":
    """Icospheres whose plain trajectories train the irregular-mesh surrogate: the config4 recipe
    (radial +-2 % noise, edge and bend struts per unit strain at the mean edge length, 0.005 kg per
    vertex, ground plane at y = 0) with levels drawn from `levels`, drop heights 0.05-0.5 m and a
    per-scene wind of the config4 direction scaled by a factor in [0, 3] (the same wind for all
    meshes: one Config)."""
    rng = np.random.default_rng(seed)
    meshes, cache = [], {}
    for k in range(n_meshes):
        lv = int(rng.choice(levels))
        if lv not in cache:
            cache[lv] = icosphere_fast(lv)
        v, f = cache[lv]
        r = 0.004 * 2 ** lv * rng.uniform(0.8, 1.25)
        noise = rng.uniform(-0.02, 0.02, size=(v.shape[0], 1))
        centre = np.array([(k % 20) * 3.0, r + rng.uniform(0.05, 0.5), (k // 20) * 3.0])
        pos = (v * r * (1.0 + noise) + centre).astype(F32)
        e = np.linalg.norm(pos[f[:, 0]].astype(np.float64) - pos[f[:, 1]], axis=1).mean()
        meshes.append(SurfaceSpec(pos, f.reshape(-1).astype(np.uint64), float(F32(0.005 * v.shape[0])),
                                  per_strain(100.0, e), per_strain(0.3, e), per_strain(10.0, 2 * e),
                                  per_strain(0.05, 2 * e)))
    ground = (np.array([[-1e3, 0.0, 1e3], [1e3, 0.0, 1e3], [1e3, 0.0, -1e3], [-1e3, 0.0, -1e3]], F32),
              np.array([0, 1, 2, 0, 2, 3], np.uint64))
    w = float(rng.uniform(0.0, 3.0))
    cfg = ConfigSpec(1, 1.0 / 600.0, gravity_force(0.005), (w, 0.0, 0.5 * w), 0.2, 0.5, 0.3, 0.2).f32()
    return SurfaceScene(meshes, cfg, [ground], f"{n_meshes} training icospheres")


@dataclass
class SurfaceScene:
    meshes: List[SurfaceSpec]
    config: ConfigSpec
    obstacles: list
    name: str = ""


def build_surface_gpu(scene: "SurfaceScene", threads: int = 8):
    """The SurfaceScene as a feriphys_amd Simulation."""
    from concurrent.futures import ThreadPoolExecutor

    from . import springy as sp

    def make(spec):
        m = sp.SpringyMesh(spec.positions, spec.indices, spec.mass, spec.k, spec.c)
        v01, _, faces = m.struts()
        _, fv = m.faces()
        m.add_struts(bend_pairs(v01, faces, fv).astype(np.uint64), spec.bk, spec.bc)
        return m

    with ThreadPoolExecutor(max_workers=threads) as ex:
        meshes = list(ex.map(make, scene.meshes))
    sim = sp.Simulation(meshes, [sp.CollidableMesh(p, i) for p, i in scene.obstacles])
    c = scene.config
    sim.config = sp.Config(sp.Integration(c.integration), c.dt, c.gravity, c.wind, c.lift_coefficient,
                           c.drag_coefficient, c.coefficient_of_restitution, c.coefficient_of_friction)
    return sim


# ---- building a scene on either side ----------------------------------------------------------

def build_gpu(scene: Scene, threads: int = 8):
    """The scene as a feriphys_amd Simulation (records jittered, configs set).  Cloths are built on
    a thread pool (the host builder runs outside the GIL)."""
    from concurrent.futures import ThreadPoolExecutor

    from . import springy as sp

    def make(k):
        c = scene.cloths[k]
        cl = sp.Cloth(*c.args(), plane=c.plane)
        if scene.jitter:
            cl.mesh.set_points(jitter_records(cl.mesh.get_points(), c.pins, scene.jitter, scene.seed + k))
        return cl.mesh

    if len(scene.cloths) > 1 and threads > 1:
        with ThreadPoolExecutor(max_workers=threads) as ex:
            meshes = list(ex.map(make, range(len(scene.cloths))))
    else:
        meshes = [make(k) for k in range(len(scene.cloths))]
    obs = [sp.CollidableMesh(p, i) for p, i in scene.obstacles]
    sim = sp.Simulation(meshes, obs)
    to_cfg = lambda c: sp.Config(sp.Integration(c.integration), c.dt, c.gravity, c.wind, c.lift_coefficient,
                                 c.drag_coefficient, c.coefficient_of_restitution, c.coefficient_of_friction)
    if getattr(scene, "mesh_winds", None) is not None:
        sim.set_mesh_configs([to_cfg(c) for c in mesh_configs(scene)])
    else:
        sim.config = to_cfg(scene.config)
    return sim
