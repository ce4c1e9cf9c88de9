"""Seeded synthetic scenes for the raw-physics step, as flat world descriptors.

A world is a batch of independent scenes. Each body references a *shape*: a
triangle corner stream of float32 xyz (three corners per triangle), exactly
what the reference's obj_parse returns (src/render/obj.cpp:45-107). The hull
builder widens it to double and multiplies by the body's scale, as
examples_util_create_single_convex_hull_collider_array does
(src/examples/examples_util.cpp:5-24).

Workloads (sizes as named in the benchmark configuration):
  * ``stack_scene``      — 10 unit cubes stacked on a 50x1x50 static floor.
  * ``batched_uniform``  — N scenes of a static floor + 8 cubes + 8 42-vertex
                           icospheres dropped from a 4x4 grid.

Random numbers: SplitMix64 seeds a xoshiro256** stream per scene with seed
0x5EED0000 + scene index; u = (x >> 11) * 2^-53.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

# --------------------------------------------------------------------------
# Shapes
# --------------------------------------------------------------------------
# res/cube.obj: 8 vertices and 12 triangles (face lines, 1-based). The corner
# order fixes the hull's vertex and face order, which the step's results
# depend on (first-max tie-breaks), so the asset is reproduced as data.
_CUBE_V = [(1.0, 1.0, -1.0), (1.0, -1.0, -1.0), (1.0, 1.0, 1.0), (1.0, -1.0, 1.0),
           (-1.0, 1.0, -1.0), (-1.0, -1.0, -1.0), (-1.0, 1.0, 1.0), (-1.0, -1.0, 1.0)]
_CUBE_F = [(5, 3, 1), (3, 8, 4), (7, 6, 8), (2, 8, 6), (1, 4, 2), (5, 2, 6),
           (5, 7, 3), (3, 7, 8), (7, 5, 6), (2, 4, 8), (1, 3, 4), (5, 1, 2)]

# res/ico.obj: the 12-vertex icosahedron (vertex values as written in the file).
_ICO_V = [(0.000000, -1.000000, 0.000000), (0.723600, -0.447215, 0.525720),
          (-0.276385, -0.447215, 0.850640), (-0.894425, -0.447215, 0.000000),
          (-0.276385, -0.447215, -0.850640), (0.723600, -0.447215, -0.525720),
          (0.276385, 0.447215, 0.850640), (-0.723600, 0.447215, 0.525720),
          (-0.723600, 0.447215, -0.525720), (0.276385, 0.447215, -0.850640),
          (0.894425, 0.447215, 0.000000), (0.000000, 1.000000, 0.000000)]
_ICO_F = [(1, 2, 3), (2, 1, 6), (1, 3, 4), (1, 4, 5), (1, 5, 6), (2, 6, 11), (3, 2, 7), (4, 3, 8),
          (5, 4, 9), (6, 5, 10), (2, 11, 7), (3, 7, 8), (4, 8, 9), (5, 9, 10), (6, 10, 11),
          (7, 11, 12), (8, 7, 12), (9, 8, 12), (10, 9, 12), (11, 10, 12)]


def _corners(verts, faces) -> np.ndarray:
    out = np.empty((len(faces) * 3, 3), dtype=np.float32)
    for t, f in enumerate(faces):
        for k in range(3):
            out[3 * t + k] = verts[f[k] - 1]
    # no negative zeros: the reference's vertex hash keys on the bit pattern
    out += np.float32(0.0)
    return out


_CUBE = None


def cube_corners() -> np.ndarray:
    """Corner stream of res/cube.obj (36 corners, 8 unique vertices)."""
    global _CUBE
    if _CUBE is None:
        _CUBE = _corners(_CUBE_V, _CUBE_F)
    return _CUBE


def prism_corners(n: int, radius: float = 1.0, half_height: float = 0.5) -> np.ndarray:
    """Corner stream of a right prism over a regular n-gon (2n vertices): the
    caps as triangle fans, which the hull builder merges into two n-vertex
    faces, and each side as two triangles."""
    a = [2.0 * math.pi * k / n for k in range(n)]
    verts = [(radius * math.cos(t), half_height, radius * math.sin(t)) for t in a]
    verts += [(radius * math.cos(t), -half_height, radius * math.sin(t)) for t in a]
    faces = []
    for k in range(1, n - 1):
        faces.append((1, k + 2, k + 1))              # top cap, normal +y
        faces.append((n + 1, n + k + 1, n + k + 2))  # bottom cap, normal -y
    for k in range(n):
        k1 = (k + 1) % n
        faces.append((k + 1, k1 + 1, n + k1 + 1))
        faces.append((k + 1, n + k1 + 1, n + k + 1))
    return _corners(verts, faces)


_ICO_CACHE = {}


def icosphere_corners(subdivisions: int) -> np.ndarray:
    if subdivisions not in _ICO_CACHE:
        _ICO_CACHE[subdivisions] = _icosphere_corners(subdivisions)
    return _ICO_CACHE[subdivisions]


def _icosphere_corners(subdivisions: int) -> np.ndarray:
    """Unit-radius icosphere: res/ico.obj's icosahedron with every triangle
    split at its edge midpoints `subdivisions` times, vertices projected to
    the unit sphere (12 / 42 / 162 vertices for 0 / 1 / 2 subdivisions)."""
    verts = [tuple(float(np.float32(c)) for c in v) for v in _ICO_V]
    if subdivisions > 0:
        verts = [tuple(c / math.sqrt(sum(x * x for x in v)) for c in v) for v in verts]
    faces = [tuple(f) for f in _ICO_F]
    for _ in range(subdivisions):
        mid = {}

        def midpoint(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                va, vb = verts[a - 1], verts[b - 1]
                m = [(va[i] + vb[i]) * 0.5 for i in range(3)]
                n = math.sqrt(sum(x * x for x in m))
                verts.append(tuple(x / n for x in m))
                mid[key] = len(verts)
            return mid[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nf += [(a, ab, ca), (ab, b, bc), (ca, bc, c), (ab, bc, ca)]
        faces = nf
    return _corners(verts, faces)


def hull_volume(corners: np.ndarray, scale) -> float:
    """Volume of the closed triangle mesh (divergence theorem), scaled."""
    p = corners.astype(np.float64) * np.asarray(scale, dtype=np.float64)
    a, b, c = p[0::3], p[1::3], p[2::3]
    return float(abs(np.einsum("ij,ij->i", a, np.cross(b, c)).sum()) / 6.0)


# --------------------------------------------------------------------------
# Random numbers
# --------------------------------------------------------------------------
_M64 = (1 << 64) - 1


class Rng:
    """xoshiro256** seeded by SplitMix64."""

    def __init__(self, seed: int):
        s = seed & _M64
        self.s = []
        for _ in range(4):
            s = (s + 0x9E3779B97F4A7C15) & _M64
            z = s
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
            self.s.append(z ^ (z >> 31))

    def next(self) -> int:
        s = self.s
        r = (((s[1] * 5) & _M64) << 7 | ((s[1] * 5) & _M64) >> 57) & _M64
        r = (r * 9) & _M64
        t = (s[1] << 17) & _M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = ((s[3] << 45) | (s[3] >> 19)) & _M64
        return r

    def uniform(self, a: float = 0.0, b: float = 1.0) -> float:
        return a + (b - a) * ((self.next() >> 11) * (1.0 / 9007199254740992.0))

    def normal(self, sigma: float = 1.0) -> float:
        u1 = max(self.uniform(), 1e-300)
        u2 = self.uniform()
        return sigma * math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)

    def quaternion(self):
        """Uniform random rotation (Shoemake), as x, y, z, w."""
        u1, u2, u3 = self.uniform(), self.uniform(), self.uniform()
        a, b = math.sqrt(1.0 - u1), math.sqrt(u1)
        return (a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2),
                b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3))


PI_F = 3.14159265358979  # include/gm.h:8


def quaternion_new(axis, degrees):
    """quaternion_new (src/quaternion.cpp:18-31): axis-angle in degrees."""
    ln = math.sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2])
    if ln != 0.0:
        axis = (axis[0] / ln, axis[1] / ln, axis[2] / ln)
    rad = PI_F * degrees / 180.0
    s = math.sin(rad / 2.0)
    return (axis[0] * s, axis[1] * s, axis[2] * s, math.cos(rad / 2.0))


# --------------------------------------------------------------------------
# World descriptor
# --------------------------------------------------------------------------
# Constraint_Type / PBD_Axis_Type numeric values (src/physics/pbd.h:5-20)
POSITIONAL, COLLISION, MUTUAL_ORIENTATION, HINGE, SPHERICAL = range(5)
POS_X, NEG_X, POS_Y, NEG_Y, POS_Z, NEG_Z = range(6)


class ConstraintDesc(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("body1", ctypes.c_uint32), ("body2", ctypes.c_uint32),
                ("axis", ctypes.c_int32 * 4), ("limited", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("r1_lc", ctypes.c_double * 3), ("r2_lc", ctypes.c_double * 3),
                ("compliance", ctypes.c_double), ("distance", ctypes.c_double * 3),
                ("limits", ctypes.c_double * 4)]


_P = ctypes.c_void_p


class WorldDescC(ctypes.Structure):
    """Layout of orc_world_desc / rp_world_desc."""
    _fields_ = [("num_shapes", ctypes.c_uint32), ("shape_corner_offset", _P), ("shape_corners", _P),
                ("num_scenes", ctypes.c_uint32), ("scene_body_offset", _P), ("body_shape", _P),
                ("body_scale", _P), ("body_position", _P), ("body_rotation", _P),
                ("body_linear_velocity", _P), ("body_angular_velocity", _P), ("body_mass", _P),
                ("body_fixed", _P), ("body_material", _P), ("scene_constraint_offset", _P),
                ("constraints", _P), ("large_scene_bodies", ctypes.c_uint32), ("shape_sphere_radius", _P),
                ("body_collider_offset", _P), ("collider_shape", _P)]


class StateC(ctypes.Structure):
    """Layout of orc_state / rp_state: ten optional output arrays."""
    _fields_ = [(n, _P) for n in ("position", "rotation", "linear_velocity", "angular_velocity",
                                  "previous_position", "previous_rotation", "previous_linear_velocity",
                                  "previous_angular_velocity", "active", "deactivation_time")]


STATE_FIELDS = (("position", 3), ("rotation", 4), ("linear_velocity", 3), ("angular_velocity", 3),
                ("previous_position", 3), ("previous_rotation", 4), ("previous_linear_velocity", 3),
                ("previous_angular_velocity", 3), ("active", 1), ("deactivation_time", 1))


def alloc_state(nb: int) -> dict:
    st = {n: np.zeros((nb, k), dtype=np.float64) for n, k in STATE_FIELDS if n != "active"}
    st["active"] = np.zeros((nb, 1), dtype=np.int32)
    return st


def state_struct(st: dict) -> StateC:
    return StateC(*[st[n].ctypes.data if st.get(n) is not None else None for n, _ in STATE_FIELDS])


@dataclass
class World:
    """A batch of scenes as flat arrays (the world descriptor's contents)."""
    shapes: list = field(default_factory=list)           # list of (n,3) float32 corner streams
    scene_body_offset: list = field(default_factory=lambda: [0])
    body_shape: list = field(default_factory=list)
    body_scale: list = field(default_factory=list)
    body_position: list = field(default_factory=list)
    body_rotation: list = field(default_factory=list)
    body_linear_velocity: list = field(default_factory=list)
    body_angular_velocity: list = field(default_factory=list)
    body_mass: list = field(default_factory=list)
    body_fixed: list = field(default_factory=list)
    body_material: list = field(default_factory=list)
    scene_constraint_offset: list = field(default_factory=lambda: [0])
    constraints: list = field(default_factory=list)      # ConstraintDesc
    name: str = ""
    large_scene_bodies: int = 0                          # device large-scene threshold (0: library default)
    sphere_radius: list = field(default_factory=list)    # per shape: > 0 for a sphere collider
    body_colliders: list = field(default_factory=list)   # per body: tuple of shapes (compound) or None

    # -- building -----------------------------------------------------------
    def sphere(self, radius: float) -> int:
        """A sphere collider shape (collider_sphere_create); its radius is a
        float as in the reference (collider.h:29)."""
        r = float(np.float32(radius))
        for i, (s, sr) in enumerate(zip(self.shapes, self.sphere_radius)):
            if sr == r and len(s) == 0:
                return i
        self.shapes.append(np.zeros((0, 3), dtype=np.float32))
        self.sphere_radius.append(r)
        return len(self.shapes) - 1

    def shape(self, corners: np.ndarray) -> int:
        while len(self.sphere_radius) < len(self.shapes):
            self.sphere_radius.append(0.0)
        for i, s in enumerate(self.shapes):
            if s is corners or (s.shape == corners.shape and np.array_equal(s, corners)):
                return i
        self.shapes.append(np.ascontiguousarray(corners, dtype=np.float32))
        self.sphere_radius.append(0.0)
        return len(self.shapes) - 1

    def add_body(self, shape: int, scale, position, rotation, mass=1.0, fixed=False,
                 material=(0.5, 0.3, 0.0), linear_velocity=(0.0, 0.0, 0.0),
                 angular_velocity=(0.0, 0.0, 0.0)) -> int:
        """`shape` is one shape index, or a sequence of them for a compound
        body (one collider per entry, in order)."""
        if isinstance(shape, (list, tuple)):
            self.body_colliders.append(tuple(int(k) for k in shape))
            shape = shape[0]
        else:
            self.body_colliders.append(None)
        self.body_shape.append(shape)
        self.body_scale.append(tuple(scale))
        self.body_position.append(tuple(position))
        self.body_rotation.append(tuple(rotation))
        self.body_linear_velocity.append(tuple(linear_velocity))
        self.body_angular_velocity.append(tuple(angular_velocity))
        self.body_mass.append(float(mass))
        self.body_fixed.append(1 if fixed else 0)
        self.body_material.append(tuple(material))
        return len(self.body_shape) - 1 - self.scene_body_offset[-1]

    def add_constraint(self, c: ConstraintDesc):
        self.constraints.append(c)

    def end_scene(self):
        self.scene_body_offset.append(len(self.body_shape))
        self.scene_constraint_offset.append(len(self.constraints))

    @property
    def num_scenes(self) -> int:
        return len(self.scene_body_offset) - 1

    @property
    def num_bodies(self) -> int:
        return len(self.body_shape)

    def dynamic_bodies(self) -> int:
        return int(self.num_bodies - sum(self.body_fixed))

    # -- flattening ---------------------------------------------------------
    def arrays(self) -> dict:
        offs = [0]
        for s in self.shapes:
            offs.append(offs[-1] + s.shape[0])
        a = {
            "shape_corner_offset": np.asarray(offs, dtype=np.uint32),
            "shape_corners": np.ascontiguousarray(np.concatenate(self.shapes, axis=0).reshape(-1), dtype=np.float32),
            "scene_body_offset": np.asarray(self.scene_body_offset, dtype=np.uint32),
            "body_shape": np.asarray(self.body_shape, dtype=np.uint32),
            "body_scale": np.asarray(self.body_scale, dtype=np.float64).reshape(-1),
            "body_position": np.asarray(self.body_position, dtype=np.float64).reshape(-1),
            "body_rotation": np.asarray(self.body_rotation, dtype=np.float64).reshape(-1),
            "body_linear_velocity": np.asarray(self.body_linear_velocity, dtype=np.float64).reshape(-1),
            "body_angular_velocity": np.asarray(self.body_angular_velocity, dtype=np.float64).reshape(-1),
            "body_mass": np.asarray(self.body_mass, dtype=np.float64),
            "body_fixed": np.asarray(self.body_fixed, dtype=np.int32),
            "body_material": np.asarray(self.body_material, dtype=np.float64).reshape(-1),
            "scene_constraint_offset": np.asarray(self.scene_constraint_offset, dtype=np.uint32),
        }
        if self.constraints:
            a["constraints"] = (ConstraintDesc * len(self.constraints))(*self.constraints)
        if any(r > 0.0 for r in self.sphere_radius):
            a["shape_sphere_radius"] = np.asarray(self.sphere_radius, dtype=np.float32)
        if any(c is not None for c in self.body_colliders):
            offs, shapes = [0], []
            for b, c in enumerate(self.body_colliders):
                shapes.extend(c if c is not None else (self.body_shape[b],))
                offs.append(len(shapes))
            a["body_collider_offset"] = np.asarray(offs, dtype=np.uint32)
            a["collider_shape"] = np.asarray(shapes, dtype=np.uint32)
        return a

    def desc(self):
        """Returns (WorldDescC, keepalive) for the C ABI."""
        a = self.arrays()
        cons = a.get("constraints")
        d = WorldDescC(
            len(self.shapes), a["shape_corner_offset"].ctypes.data, a["shape_corners"].ctypes.data,
            self.num_scenes, a["scene_body_offset"].ctypes.data, a["body_shape"].ctypes.data,
            a["body_scale"].ctypes.data, a["body_position"].ctypes.data, a["body_rotation"].ctypes.data,
            a["body_linear_velocity"].ctypes.data, a["body_angular_velocity"].ctypes.data,
            a["body_mass"].ctypes.data, a["body_fixed"].ctypes.data, a["body_material"].ctypes.data,
            a["scene_constraint_offset"].ctypes.data if cons is not None else None,
            ctypes.addressof(cons) if cons is not None else None, self.large_scene_bodies,
            a["shape_sphere_radius"].ctypes.data if "shape_sphere_radius" in a else None,
            a["body_collider_offset"].ctypes.data if "body_collider_offset" in a else None,
            a["collider_shape"].ctypes.data if "collider_shape" in a else None)
        return d, a


# --------------------------------------------------------------------------
# Workloads
# --------------------------------------------------------------------------
def _floor(w: World, material=(0.5, 0.3, 0.0)):
    """50x1x50 static box at (0,-2,0), top face at y=-1 (src/examples/stack.cpp:48-51)."""
    cube = w.shape(cube_corners())
    w.add_body(cube, (50.0, 1.0, 50.0), (0.0, -2.0, 0.0), quaternion_new((0.0, 1.0, 0.0), 0.0), fixed=True,
               material=material)


def stack_scene(w: World | None = None, seed: int = 0, num_cubes: int = 10) -> World:
    """Plain-method workload: `num_cubes` unit cubes (8-vertex hull, half-extent
    1, mass 1) stacked on the static floor; jitter +-0.01 on x/z and +-2 deg
    yaw; friction 0.5/0.3, restitution 0."""
    w = w or World(name="stack")
    rng = Rng(0x5EED0000 + seed)
    _floor(w)
    cube = w.shape(cube_corners())
    for k in range(num_cubes):
        x = rng.uniform(-0.01, 0.01)
        z = rng.uniform(-0.01, 0.01)
        yaw = rng.uniform(-2.0, 2.0)
        w.add_body(cube, (1.0, 1.0, 1.0), (x, 2.01 * k, z), quaternion_new((0.0, 1.0, 0.0), yaw), mass=1.0,
                   material=(0.5, 0.3, 0.0))
    w.end_scene()
    return w


GRID_SPACING = 3.6  # > 2 * sqrt(3) * 1.0 + 0.1: no two bounding spheres overlap at t=0


def drop_scene(w: World, seed: int, n_cubes: int = 8, n_icos: int = 8, ico_subdiv: int = 1,
               grid: int = 4) -> World:
    """One batched-uniform scene: static floor + `n_cubes` cubes (half-extent r)
    + `n_icos` icospheres (radius r), r ~ U(0.5, 1.0), mass = hull volume
    (density 1), dropped from a grid x grid lattice with their lowest point
    U(1, 6) above the floor top, uniformly random orientation,
    v ~ N(0, 0.5)^3, w = 0."""
    rng = Rng(0x5EED0000 + seed)
    _floor(w)
    cube = w.shape(cube_corners())
    ico = w.shape(icosphere_corners(ico_subdiv))
    unit_volume = {cube: hull_volume(w.shapes[cube], (1.0, 1.0, 1.0)),
                   ico: hull_volume(w.shapes[ico], (1.0, 1.0, 1.0))}
    n = n_cubes + n_icos
    for k in range(n):
        is_cube = k < n_cubes
        r = rng.uniform(0.5, 1.0)
        gx, gz = k % grid, (k // grid) % grid
        layer = k // (grid * grid)
        x = (gx - (grid - 1) / 2.0) * GRID_SPACING
        z = (gz - (grid - 1) / 2.0) * GRID_SPACING
        # lowest point of the body 1..6 above the floor top (y = -1)
        bound = math.sqrt(3.0) * r if is_cube else r
        y = -1.0 + bound + rng.uniform(1.0, 6.0) + 8.0 * layer
        q = rng.quaternion()
        v = (rng.normal(0.5), rng.normal(0.5), rng.normal(0.5))
        shape = cube if is_cube else ico
        scale = (r, r, r)
        mass = unit_volume[shape] * r * r * r
        w.add_body(shape, scale, (x, y, z), q, mass=mass, material=(0.5, 0.3, 0.0), linear_velocity=v)
    w.end_scene()
    return w


def batched_uniform(num_scenes: int, first_seed: int = 0, **kw) -> World:
    w = World(name=f"batched_uniform_{num_scenes}")
    for s in range(num_scenes):
        drop_scene(w, first_seed + s, **kw)
    return w


def box_corners(half, offset=(0.0, 0.0, 0.0)) -> np.ndarray:
    """cube.obj's corner stream scaled by `half` and moved by `offset`: a box
    collider placed off the body origin (a part of a compound body)."""
    c = cube_corners().astype(np.float64) * np.asarray(half, dtype=np.float64) + np.asarray(offset)
    return c.astype(np.float32)


def mixed_scene(w: World, seed: int, n_spheres: int = 5, n_cubes: int = 3, n_dumbbells: int = 2) -> World:
    """Sphere colliders and compound bodies: a static floor, spheres of radius
    U(0.5, 1.0), unit cubes, and dumbbells (two 0.5-half-extent boxes at
    x = +-0.75, one body with two hull colliders), dropped from a 4 x 4 lattice
    with lowest points U(1, 5) above the floor and random orientation and
    velocity N(0, 0.5). Masses are density-1 volumes (4/3 pi r^3 for spheres).
    Friction 0.5/0.3, restitution U(0, 0.3)."""
    rng = Rng(0x5EED0000 + seed)
    _floor(w)
    cube = w.shape(cube_corners())
    left = w.shape(box_corners((0.5, 0.5, 0.5), (-0.75, 0.0, 0.0)))
    right = w.shape(box_corners((0.5, 0.5, 0.5), (0.75, 0.0, 0.0)))
    n = n_spheres + n_cubes + n_dumbbells
    for k in range(n):
        gx, gz = k % 4, (k // 4) % 4
        layer = k // 16
        x = (gx - 1.5) * GRID_SPACING
        z = (gz - 1.5) * GRID_SPACING
        q = rng.quaternion()
        v = (rng.normal(0.5), rng.normal(0.5), rng.normal(0.5))
        e = rng.uniform(0.0, 0.3)
        if k < n_spheres:
            r = rng.uniform(0.5, 1.0)
            shape, bound, mass = w.sphere(r), r, 4.0 / 3.0 * math.pi * r ** 3
        elif k < n_spheres + n_cubes:
            shape, bound, mass = cube, math.sqrt(3.0), 8.0
        else:
            shape, bound, mass = (left, right), math.sqrt(1.25 ** 2 + 0.5), 2.0
        y = -1.0 + bound + rng.uniform(1.0, 5.0) + 8.0 * layer
        w.add_body(shape, (1.0, 1.0, 1.0), (x, y, z), q, mass=mass, material=(0.5, 0.3, e), linear_velocity=v)
    w.end_scene()
    return w


def spot_storm_scene(w: World, hulls, seed: int = 0, n: int = 2) -> World:
    """The reference's spot_storm example (src/examples/spot_storm.cpp:56-137):
    a static cube.obj x (50, 1, 50) floor at (0, -2, 0) (friction 0.5/0.5),
    and n^3 bodies of the 11 convex pieces `hulls` (corner streams of
    res/spot/spot-hull-1..11.obj) scaled by 2, mass 1, friction 0.8/0.8,
    restitution 0, on its lattice (y from 2, x and z from -n, gap 3.5) with
    random orientations (its generate_random_quaternion: axis U(0,1)^3,
    angle U(-180, 180) degrees, from this seed instead of rand())."""
    rng = Rng(0x5EED0000 + seed)
    floor = w.shape(cube_corners())
    w.add_body(floor, (50.0, 1.0, 50.0), (0.0, -2.0, 0.0), quaternion_new((0.0, 1.0, 0.0), 0.0), fixed=True,
               material=(0.5, 0.5, 0.0))
    pieces = tuple(w.shape(np.ascontiguousarray(h, dtype=np.float32)) for h in hulls)
    gap, y = 3.5, 2.0
    for i in range(n):
        y += gap
        x = -2.0 * (n / 2.0)
        for j in range(n):
            x += gap
            z = -2.0 * (n / 2.0)
            for k in range(n):
                z += gap
                axis = (rng.uniform(), rng.uniform(), rng.uniform())
                angle = -180.0 + rng.uniform() * 360.0
                w.add_body(pieces, (2.0, 2.0, 2.0), (x, y, z), quaternion_new(axis, angle), mass=1.0,
                           material=(0.8, 0.8, 0.0))
    w.end_scene()
    return w


def mirror_cube_scene(w: World, pieces) -> World:
    """The reference's mirror_cube example (src/examples/mirror_cube.cpp:50-68):
    a static cube.obj x (50, 1, 50) floor at (0, -2, 0) (friction 0.5/0.5)
    and one body of the two convex pieces `pieces` (corner streams of
    res/mirror_cube_collider1.obj and _collider2.obj), scale 1, at (0, 2, 0)
    rotated 33 degrees about (1, 1, 1), mass 1, friction 0.8/0.8."""
    floor = w.shape(cube_corners())
    w.add_body(floor, (50.0, 1.0, 50.0), (0.0, -2.0, 0.0), quaternion_new((0.0, 1.0, 0.0), 0.0), fixed=True,
               material=(0.5, 0.5, 0.0))
    shapes = tuple(w.shape(np.ascontiguousarray(p, dtype=np.float32)) for p in pieces)
    w.add_body(shapes, (1.0, 1.0, 1.0), (0.0, 2.0, 0.0), quaternion_new((1.0, 1.0, 1.0), 33.0), mass=1.0,
               material=(0.8, 0.8, 0.0))
    w.end_scene()
    return w


def batched_mixed(num_scenes: int, first_seed: int = 0, **kw) -> World:
    w = World(name=f"batched_mixed_{num_scenes}")
    for s in range(num_scenes):
        mixed_scene(w, first_seed + s, **kw)
    return w


def ramp_scene(w: World, seed: int) -> np.ndarray:
    """One differentiable-workload scene: a static box cube.obj x (10, 0.5, 4)
    tilted about z by U(10, 35) degrees at the origin, and a cube of
    half-extent 1 with mass U(0.5, 2) whose centre starts 1.5 + sqrt(3) above
    the centre of the ramp's top face (clear of it in any orientation), with
    v ~ N(0, 2)^3, w ~ N(0, 3)^3 and a uniformly random orientation; friction
    0.6/0.5 for both bodies, restitution 0. Returns the loss target: the foot
    of the ramp plus a U([-2,2] x [0,1] x [-2,2]) offset."""
    rng = Rng(0x5EED0000 + seed)
    cube = w.shape(cube_corners())
    angle = rng.uniform(10.0, 35.0)
    qr = quaternion_new((0.0, 0.0, 1.0), angle)
    w.add_body(cube, (10.0, 0.5, 4.0), (0.0, 0.0, 0.0), qr, fixed=True, material=(0.6, 0.5, 0.0))
    top = _qrot(qr, (0.0, 0.5, 0.0))
    normal = _qrot(qr, (0.0, 1.0, 0.0))
    lift = 1.5 + math.sqrt(3.0)
    x = tuple(top[i] + lift * normal[i] for i in range(3))
    mass = rng.uniform(0.5, 2.0)
    v = (rng.normal(2.0), rng.normal(2.0), rng.normal(2.0))
    om = (rng.normal(3.0), rng.normal(3.0), rng.normal(3.0))
    w.add_body(cube, (1.0, 1.0, 1.0), x, rng.quaternion(), mass=mass, material=(0.6, 0.5, 0.0), linear_velocity=v,
               angular_velocity=om)
    w.end_scene()
    foot = _qrot(qr, (-10.0, 0.5, 0.0))
    return np.array([foot[0] + rng.uniform(-2.0, 2.0), foot[1] + rng.uniform(0.0, 1.0),
                     foot[2] + rng.uniform(-2.0, 2.0)])


def ramps(num_scenes: int, first_seed: int = 0):
    """The differentiable workload: `num_scenes` ramp scenes; returns the
    world and the (num_scenes, 3) loss targets. Body 2s+1 is scene s's cube."""
    w = World(name=f"ramps_{num_scenes}")
    targets = np.stack([ramp_scene(w, first_seed + s) for s in range(num_scenes)])
    return w, targets


# Differentiable-workload parameters of a ramp scene's cube, with the
# adjoint's output for each: ("state", column of (x, q, v, w)) or
# ("param", column of (mass, static friction, dynamic friction)).
RAMP_PARAMS = {"vx": ("state", 7), "vy": ("state", 8), "vz": ("state", 9),
               "wx": ("state", 10), "wy": ("state", 11), "wz": ("state", 12),
               "mass": ("param", 0), "mu_s": ("param", 1), "mu_d": ("param", 2)}


def _perturb(w: World, b: int, name: str, delta: float) -> None:
    """Moves parameter `name` of body b (global index) by delta."""
    if name[0] in "vw" and len(name) == 2:
        arr = w.body_linear_velocity if name[0] == "v" else w.body_angular_velocity
        t = list(arr[b])
        t["xyz".index(name[1])] += delta
        arr[b] = tuple(t)
    elif name == "mass":
        w.body_mass[b] += delta
    else:
        m = list(w.body_material[b])
        m[0 if name == "mu_s" else 1] += delta
        w.body_material[b] = tuple(m)


def ramps_perturbed(num_scenes: int, first_seed: int, params, h: float):
    """Central-difference copies of ramps(num_scenes, first_seed): copy
    (s * len(params) + k) * 2 + j is scene s with parameter params[k] of its
    cube moved by +h (j = 0) or -h (j = 1). Returns the world and the loss
    target of every copy; the cube of copy c is body 2c + 1."""
    w = World(name=f"ramps_fd_{num_scenes}x{len(params)}")
    targets = []
    for s in range(num_scenes):
        for name in params:
            for sign in (1.0, -1.0):
                targets.append(ramp_scene(w, first_seed + s))
                _perturb(w, len(w.body_shape) - 1, name, sign * h)
    return w, np.stack(targets)


LARGE_SCALES = 64  # scale levels of the large scene (bodies of one level share a hull)


def large_scene(n: int, seed: int = 0, large_scene_bodies: int = 0, height: float = 40.0,
                ico_subdiv: int = 1) -> World:
    """The large single-scene workload: `n` dynamic bodies, 60% cubes and 40%
    42-vertex icospheres, scale U(0.5, 1.5) (quantised to LARGE_SCALES
    levels so that bodies share hulls), uniformly random orientation, at rest,
    seeded uniformly in a (2L) x height x (2L) volume above a static floor
    (cube.obj x (L + 10, 1, L + 10) at y = -2), L = 100 sqrt(n / 10^5) so the
    density matches 10^5 bodies in 200 x 40 x 200. Positions are rejection-
    sampled so that no two bounding spheres overlap: candidates are drawn in
    order and a candidate is kept if its bounding sphere overlaps no kept
    one (greedy in draw order, checked in batches against a k-d tree of the
    kept bodies and pairwise within the batch). Mass = hull volume
    (density 1); friction 0.5/0.3, restitution 0. `ico_subdiv` = 2 gives
    162-vertex icospheres instead. Draws come from numpy's PCG64 seeded with
    0x5EED0000 + seed (candidates) and + 1 (orientations)."""
    from scipy.spatial import cKDTree

    L = 100.0 * math.sqrt(n / 1.0e5)
    w = World(name=f"large_{n}", large_scene_bodies=large_scene_bodies)
    cube = w.shape(cube_corners())
    ico = w.shape(icosphere_corners(ico_subdiv))
    w.add_body(cube, (L + 10.0, 1.0, L + 10.0), (0.0, -2.0, 0.0), quaternion_new((0.0, 1.0, 0.0), 0.0),
               fixed=True, material=(0.5, 0.3, 0.0))
    unit_volume = {cube: hull_volume(w.shapes[cube], (1.0, 1.0, 1.0)),
                   ico: hull_volume(w.shapes[ico], (1.0, 1.0, 1.0))}
    cand = np.random.Generator(np.random.PCG64(0x5EED0000 + seed))
    orient = np.random.Generator(np.random.PCG64(0x5EED0000 + seed + 1))
    max_bound = math.sqrt(3.0) * 1.5
    # kept bodies: a k-d tree over the older ones (rebuilt every `rebuild`
    # bodies) and a small one over the newer ones
    kept_pos, kept_bound, kept_cube, kept_level = [], [], [], []
    old_n, old_tree, old_pos, old_bound = 0, None, None, None
    total, batch, rebuild = 0, 4096, 16384

    def reject(pos, bound, ok, tree, tpos, tbound):
        # a nearest neighbour closer than bound + 0.5 (the smallest bound)
        # overlaps; the rest are checked against every body within reach
        d, _ = tree.query(pos, k=1)
        ok &= ~(d < bound + 0.5)
        live = np.nonzero(ok)[0]
        if not len(live):
            return
        m = cKDTree(pos[live]).sparse_distance_matrix(tree, 2.0 * max_bound, output_type="ndarray")
        i, j = live[m["i"]], m["j"]
        hit = ((tpos[j] - pos[i]) ** 2).sum(axis=1) < (tbound[j] + bound[i]) ** 2
        ok[i[hit]] = False

    while total < n:
        u = cand.random((batch, 5))
        is_cube = u[:, 0] < 0.6
        level = np.minimum(LARGE_SCALES - 1, (u[:, 1] * LARGE_SCALES).astype(np.int64))
        r = 0.5 + level / (LARGE_SCALES - 1)
        bound = np.where(is_cube, math.sqrt(3.0) * r, r)
        pos = np.stack([-L + 2.0 * L * u[:, 2], -1.0 + bound + 0.5 + height * u[:, 3], -L + 2.0 * L * u[:, 4]],
                       axis=1)
        ok = np.ones(batch, dtype=bool)
        if old_tree is not None:
            reject(pos, bound, ok, old_tree, old_pos, old_bound)
        if total > old_n:
            new_pos = np.concatenate(kept_pos)[old_n:]
            new_bound = np.concatenate(kept_bound)[old_n:]
            reject(pos, bound, ok, cKDTree(new_pos), new_pos, new_bound)
        idx = np.nonzero(ok)[0]
        if len(idx) > 1:  # greedy in draw order among the batch's survivors
            pairs = cKDTree(pos[idx]).query_pairs(2.0 * max_bound, output_type="ndarray")
            if len(pairs):
                a, b = idx[pairs[:, 0]], idx[pairs[:, 1]]
                lo, hi = np.minimum(a, b), np.maximum(a, b)
                hit = ((pos[lo] - pos[hi]) ** 2).sum(axis=1) < (bound[lo] + bound[hi]) ** 2
                lo, hi = lo[hit], hi[hit]
                order = np.argsort(lo, kind="stable")
                for p, q in zip(lo[order].tolist(), hi[order].tolist()):
                    if ok[p]:
                        ok[q] = False
        idx = np.nonzero(ok)[0][: n - total]
        kept_pos.append(pos[idx])
        kept_bound.append(bound[idx])
        kept_cube.append(is_cube[idx])
        kept_level.append(level[idx])
        total += len(idx)
        if total - old_n >= rebuild:
            old_pos, old_bound = np.concatenate(kept_pos), np.concatenate(kept_bound)
            old_n, old_tree = total, cKDTree(old_pos)
    kept_pos, kept_cube, kept_level = (np.concatenate(kept_pos), np.concatenate(kept_cube),
                                       np.concatenate(kept_level))
    q = orient.random((n, 3))
    a, b = np.sqrt(1.0 - q[:, 0]), np.sqrt(q[:, 0])
    quat = np.stack([a * np.sin(2 * math.pi * q[:, 1]), a * np.cos(2 * math.pi * q[:, 1]),
                     b * np.sin(2 * math.pi * q[:, 2]), b * np.cos(2 * math.pi * q[:, 2])], axis=1)
    for i in range(n):
        shape = cube if kept_cube[i] else ico
        r = 0.5 + int(kept_level[i]) / (LARGE_SCALES - 1)
        w.add_body(shape, (r, r, r), tuple(kept_pos[i]), tuple(quat[i]), mass=unit_volume[shape] * r * r * r,
                   material=(0.5, 0.3, 0.0))
    w.end_scene()
    return w


def _qrot(q, v):
    """quaternion_apply_to_vec3 (src/quaternion.cpp:257-268), for placing links."""
    x, y, z, w = q
    ix = w * v[0] + y * v[2] - z * v[1]
    iy = w * v[1] + z * v[0] - x * v[2]
    iz = w * v[2] + x * v[1] - y * v[0]
    iw = -x * v[0] - y * v[1] - z * v[2]
    return ((ix * w) + (iw * -x) + (iy * -z) - (iz * -y),
            (iy * w) + (iw * -y) + (iz * -x) - (ix * -z),
            (iz * w) + (iw * -z) + (ix * -y) - (iy * -x))


def chain_scene(w: World, seed: int, length: int = 6, with_floor: bool = True, link_scale=(0.2, 0.5, 0.1),
                velocity_sigma: float = 1.0) -> World:
    """A static support with a chain of `length` box links hanging from it,
    joined alternately by limited hinge joints (aligned axis +z, limit axis
    +y, limits +-pi/2) and spherical joints (swing axis +x, twist axis +y,
    swing +-pi/4, twist +-pi/8), as in the reference's arm scene
    (src/examples/arm.cpp:44-110). Pivots are the link ends with seeded
    jitter; each link is moved so that its pivot meets the previous link's
    (reset_joint_distance, arm.cpp:32-42). Links start with v ~ N(0, sigma)."""
    rng = Rng(0x5EED0000 + seed)
    cube = w.shape(cube_corners())
    if with_floor:
        _floor(w)
    ident = quaternion_new((1.0, 0.0, 0.0), 0.0)
    sup_pos = (0.0, 1.0 + 2.0 * link_scale[1] * length + 0.5, 0.0)
    prev = w.add_body(cube, (0.2, 0.1, 0.1), sup_pos, ident, fixed=True, material=(0.5, 0.5, 0.0))
    prev_pos, prev_q, prev_half = sup_pos, ident, 0.1
    for k in range(length):
        jit = lambda: rng.uniform(-0.05, 0.05)  # noqa: E731
        r1 = (jit(), -prev_half + jit(), jit())
        r2 = (jit(), link_scale[1] + jit(), jit())
        yaw = rng.uniform(-10.0, 10.0)
        q = quaternion_new((0.0, 1.0, 0.0), yaw)
        # place at the pivot: x2 += (x1 + R1 r1) - (x2 + R2 r2) with x2 = x1 initially
        p1 = tuple(a + b for a, b in zip(prev_pos, _qrot(prev_q, r1)))
        p2 = tuple(a + b for a, b in zip(prev_pos, _qrot(q, r2)))
        pos = tuple(x + (a - b) for x, a, b in zip(prev_pos, p1, p2))
        v = (rng.normal(velocity_sigma), rng.normal(velocity_sigma), rng.normal(velocity_sigma))
        body = w.add_body(cube, link_scale, pos, q, mass=1.0, material=(0.6, 0.6, 0.0), linear_velocity=v)
        c = ConstraintDesc()
        c.body1, c.body2 = prev, body
        c.r1_lc[:] = r1
        c.r2_lc[:] = r2
        if k % 2 == 0:
            c.type = HINGE
            c.axis[:] = (POS_Z, POS_Z, POS_Y, POS_Y)
            c.limited = 1
            c.compliance = 0.0
            c.limits[:] = (-PI_F / 2.0, PI_F / 2.0, 0.0, 0.0)
        else:
            c.type = SPHERICAL
            c.axis[:] = (POS_X, POS_X, POS_Y, POS_Y)
            c.limits[:] = (-PI_F / 4.0, PI_F / 4.0, -PI_F / 8.0, PI_F / 8.0)
        w.add_constraint(c)
        prev, prev_pos, prev_q, prev_half = body, pos, q, link_scale[1]
    w.end_scene()
    return w


def chains(num_scenes: int, length: int = 6, first_seed: int = 0, **kw) -> World:
    w = World(name=f"chains_{num_scenes}x{length}")
    for s in range(num_scenes):
        chain_scene(w, first_seed + s, length=length, **kw)
    return w


def uneven_scene(w: World, seed: int, n_min: int = 4, n_max: int = 256, chain_fraction: float = 0.3,
                 max_bodies: int | None = None) -> World:
    """One batched-uneven scene: N = round(exp(U(ln n_min, ln n_max))) dynamic
    bodies (capped by `max_bodies`), each a cube (8 vertices), a 42-vertex or
    a 162-vertex icosphere with size r ~ U(0.5, 1.0) and mass = hull volume,
    on a jittered lattice above the floor (no initial overlap); per body
    mu_s = mu_d ~ U(0.2, 0.8) (one draw, so dynamic never exceeds static,
    entity.cpp:58-60) and restitution ~ U(0, 0.5); v ~ N(0, 0.5)^3. With
    probability `chain_fraction` the scene also carries a chain of U{2..12}
    box links hanging from a static support beside the pile, joined
    alternately by limited hinge and spherical joints (chain_scene)."""
    rng = Rng(0x5EED0000 + seed)
    n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
    if max_bodies is not None:
        n = min(n, max_bodies)
    has_chain = rng.uniform() < chain_fraction
    chain_len = 2 + int(rng.uniform() * 11) if has_chain else 0
    _floor(w)
    shapes = [w.shape(cube_corners()), w.shape(icosphere_corners(1)), w.shape(icosphere_corners(2))]
    unit_volume = {sh: hull_volume(w.shapes[sh], (1.0, 1.0, 1.0)) for sh in shapes}
    side = max(1, int(math.ceil(math.sqrt(min(n, 49)))))
    for k in range(n):
        kind = min(2, int(rng.uniform() * 3))
        shape = shapes[kind]
        r = rng.uniform(0.5, 1.0)
        gx, gz = k % side, (k // side) % side
        layer = k // (side * side)
        x = (gx - (side - 1) / 2.0) * GRID_SPACING + rng.uniform(-0.2, 0.2)
        z = (gz - (side - 1) / 2.0) * GRID_SPACING + rng.uniform(-0.2, 0.2)
        bound = math.sqrt(3.0) * r if kind == 0 else r
        y = -1.0 + bound + 1.0 + GRID_SPACING * layer + rng.uniform(0.0, 0.3)
        mu = rng.uniform(0.2, 0.8)
        e = rng.uniform(0.0, 0.5)
        q = rng.quaternion()
        v = (rng.normal(0.5), rng.normal(0.5), rng.normal(0.5))
        w.add_body(shape, (r, r, r), (x, y, z), q, mass=unit_volume[shape] * r * r * r, material=(mu, mu, e),
                   linear_velocity=v)
    if has_chain:
        # the chain hangs beside the pile, clear of the lattice
        base = len(w.body_shape) - w.scene_body_offset[-1]
        link = (0.2, 0.5, 0.1)
        ox = (side / 2.0 + 1.5) * GRID_SPACING
        sup_pos = (ox, 1.0 + 2.0 * link[1] * chain_len + 0.5, 0.0)
        ident = quaternion_new((1.0, 0.0, 0.0), 0.0)
        cube = shapes[0]
        prev = w.add_body(cube, (0.2, 0.1, 0.1), sup_pos, ident, fixed=True, material=(0.5, 0.5, 0.0))
        prev_pos, prev_q, prev_half = sup_pos, ident, 0.1
        for k in range(chain_len):
            jit = lambda: rng.uniform(-0.05, 0.05)  # noqa: E731
            r1 = (jit(), -prev_half + jit(), jit())
            r2 = (jit(), link[1] + jit(), jit())
            q = quaternion_new((0.0, 1.0, 0.0), rng.uniform(-10.0, 10.0))
            p1 = tuple(a + b for a, b in zip(prev_pos, _qrot(prev_q, r1)))
            p2 = tuple(a + b for a, b in zip(prev_pos, _qrot(q, r2)))
            pos = tuple(x + (a - b) for x, a, b in zip(prev_pos, p1, p2))
            v = (rng.normal(1.0), rng.normal(1.0), rng.normal(1.0))
            body = w.add_body(cube, link, pos, q, mass=1.0, material=(0.6, 0.6, 0.0), linear_velocity=v)
            c = ConstraintDesc()
            c.body1, c.body2 = prev, body
            c.r1_lc[:] = r1
            c.r2_lc[:] = r2
            if k % 2 == 0:
                c.type = HINGE
                c.axis[:] = (POS_Z, POS_Z, POS_Y, POS_Y)
                c.limited = 1
                c.limits[:] = (-PI_F / 2.0, PI_F / 2.0, 0.0, 0.0)
            else:
                c.type = SPHERICAL
                c.axis[:] = (POS_X, POS_X, POS_Y, POS_Y)
                c.limits[:] = (-PI_F / 4.0, PI_F / 4.0, -PI_F / 8.0, PI_F / 8.0)
            w.add_constraint(c)
            prev, prev_pos, prev_q, prev_half = body, pos, q, link[1]
        _ = base
    w.end_scene()
    return w


def batched_uneven(num_scenes: int, first_seed: int = 0, **kw) -> World:
    w = World(name=f"batched_uneven_{num_scenes}")
    for s in range(num_scenes):
        uneven_scene(w, first_seed + s, **kw)
    return w


def prisms_on_pillars(num_scenes: int, first_seed: int = 0) -> World:
    """Octagonal prisms (16 vertices, circumradius 0.47) dropped onto fixed
    cube pillars (half-size 0.4), half of them level and turned 0-8 degrees
    about y, half tilted 2-8 degrees, plus a decagonal prism (20 vertices) on
    the floor. The prisms come first in body order, so a level prism's
    octagon is the incident face clipped by the pillar's square, whose edges
    cut all eight corners: the polygon grows to 10-12 points in the clipping
    passes."""
    w = World(name=f"prisms_{num_scenes}")
    for sc in range(num_scenes):
        rng = Rng(0x5EED0000 + first_seed + sc)
        _floor(w)
        oct_ = w.shape(prism_corners(8, 0.47, 0.2))
        dec = w.shape(prism_corners(10, 0.6, 0.3))
        cube = w.shape(cube_corners())
        for k in range(4):
            x, z = 3.0 * (k % 2) - 1.5, 3.0 * (k // 2) - 1.5
            if k % 2 == 0:
                q = quaternion_new((0.0, 1.0, 0.0), rng.uniform(0.0, 8.0))
            else:
                q = quaternion_new((rng.uniform(-1.0, 1.0), 0.0, rng.uniform(-1.0, 1.0)), rng.uniform(2.0, 8.0))
            w.add_body(oct_, (1.0, 1.0, 1.0), (x + rng.uniform(-0.03, 0.03), 0.9 + rng.uniform(0.0, 0.3),
                                                z + rng.uniform(-0.03, 0.03)), q,
                       mass=hull_volume(w.shapes[oct_], (1.0, 1.0, 1.0)))
        w.add_body(dec, (1.0, 1.0, 1.0), (0.0, 0.0, 4.0), quaternion_new((1.0, 0.0, 0.0), 10.0),
                   mass=hull_volume(w.shapes[dec], (1.0, 1.0, 1.0)))
        for k in range(4):
            x, z = 3.0 * (k % 2) - 1.5, 3.0 * (k // 2) - 1.5
            w.add_body(cube, (0.4, 0.4, 0.4), (x, 0.0, z), quaternion_new((0.0, 1.0, 0.0), 0.0), fixed=True)
        w.end_scene()
    return w
