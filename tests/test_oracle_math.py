"""Pins of the oracle that come from mathematics rather than from the reference.

Each check is chosen so that a plausible restatement mistake (a dropped term,
a wrong sign, a transposed operand, a swapped index) fails it:
  * free fall follows the semi-implicit Euler closed form;
  * a cube's point-mass inertia is 2m*I and its spin stays about a fixed axis;
  * a box resting on the floor yields its four bottom corners as contacts;
  * the EPA depth of two overlapping boxes equals the SAT overlap;
  * equal-and-opposite contact impulses conserve linear momentum;
  * the broad phase equals an independent brute-force test;
  * hull topology obeys Euler's formula and closed face loops;
  * a resting body falls asleep after one second of quiet frames.
"""
import numpy as np

from _libs import SimWorld, scenes

DT = 1.0 / 60.0


def _single(shape_corners, scale=(1.0, 1.0, 1.0), pos=(0.0, 0.0, 0.0), q=(0.0, 0.0, 0.0, 1.0), v=(0, 0, 0),
            w=(0, 0, 0), mass=1.0, floor=False):
    world = scenes.World(name="single")
    if floor:
        scenes._floor(world)
    shp = world.shape(shape_corners)
    world.add_body(shp, scale, pos, q, mass=mass, linear_velocity=v, angular_velocity=w)
    world.end_scene()
    return world


def test_free_fall_closed_form():
    sim = SimWorld("orc", _single(scenes.cube_corners(), pos=(0.0, 100.0, 0.0)))
    frames, sub = 30, 20
    sim.step(frames)
    st = sim.state()
    n = frames * sub
    h = DT / sub
    y_exact = 100.0 - 10.0 * h * h * n * (n + 1) / 2.0
    v_exact = -10.0 * h * n
    # the velocity is rebuilt from position differences each substep
    # (pbd.cpp:633), which costs ~1e-9 relative; a dropped or mis-signed term
    # is off by O(1)
    assert abs(st["position"][0, 1] - y_exact) < 1e-8
    assert abs(st["linear_velocity"][0, 1] / v_exact - 1.0) < 1e-8
    assert np.all(st["position"][0, [0, 2]] == 0.0)


def test_cube_inertia_and_isotropic_spin():
    w0 = np.array([1.0, -2.0, 3.0])
    sim = SimWorld("orc", _single(scenes.cube_corners(), mass=1.5, w=tuple(w0)))
    _, inertia, inv_inertia, radius = sim.body_info()
    assert np.allclose(inertia[0].reshape(3, 3), 2.0 * 1.5 * np.eye(3), atol=1e-15)
    assert np.allclose(inv_inertia[0].reshape(3, 3), np.eye(3) / 3.0, atol=1e-15)
    assert abs(radius[0] - np.sqrt(3.0)) < 1e-15
    sim.step(60, gravity=0.0)
    st = sim.state()
    w1 = st["angular_velocity"][0]
    # the axis is preserved (no gyroscopic torque for an isotropic body) ...
    assert np.linalg.norm(np.cross(w1, w0)) / np.linalg.norm(w0) ** 2 < 1e-12
    # ... and the magnitude follows the linearised quaternion update exactly:
    # q' = normalize(q + h/2 w q) and w' = 2/h (q' q^-1).xyz give
    # |w'| = |w| / sqrt(1 + (h|w|/2)^2) per substep (pbd.cpp:561-568, 636-642)
    h = DT / 20
    mag = np.linalg.norm(w0)
    for _ in range(60 * 20):
        mag = mag / np.sqrt(1.0 + (h * mag / 2.0) ** 2)
    assert abs(np.linalg.norm(w1) / mag - 1.0) < 1e-10
    assert abs(np.linalg.norm(w1) / np.linalg.norm(w0) - 1.0) > 1e-4  # the decay is really there
    assert abs(np.linalg.norm(st["rotation"][0]) - 1.0) < 1e-14


def test_resting_box_four_corner_contacts():
    world = _single(scenes.cube_corners(), pos=(0.3, -0.01, -0.2), floor=True)
    sim = SimWorld("orc", world)
    c = sim.contacts(0, 0, 1)  # floor (fixed) vs cube
    assert len(c) == 4
    p_cube = c[:, 3:6]  # points on the cube (second body)
    # clipping truncates plane offsets and edge factors to float
    # (clipping.cpp:13,27-39); the incident face is the 100-wide floor top, so
    # a float edge factor places corners to ~6e-8 * 100
    xz = p_cube[:, [0, 2]] - [0.3, -0.2]
    xz = xz[np.lexsort((np.round(xz[:, 1], 3), np.round(xz[:, 0], 3)))]
    assert np.allclose(xz, [(-1.0, -1.0), (-1.0, 1.0), (1.0, -1.0), (1.0, 1.0)], atol=2e-5)
    assert np.allclose(p_cube[:, 1], -1.01, atol=1e-6)
    n = c[0, 6:9]
    assert np.allclose(np.abs(n), [0.0, 1.0, 0.0], atol=1e-6)
    depth = np.abs(np.einsum("ij,j->i", c[:, 0:3] - c[:, 3:6], n))
    assert np.allclose(depth, 0.01, atol=1e-6)


def test_epa_depth_equals_box_sat_overlap():
    world = scenes.World(name="boxes")
    cube = world.shape(scenes.cube_corners())
    world.add_body(cube, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0, 1.0))
    world.add_body(cube, (1.0, 1.0, 1.0), (1.93, 0.31, -0.22), (0.0, 0.0, 0.0, 1.0))
    world.end_scene()
    c = SimWorld("orc", world).contacts(0, 0, 1)
    assert len(c) >= 1
    n = c[0, 6:9]
    assert np.allclose(np.abs(n), [1.0, 0.0, 0.0], atol=1e-9)  # SAT: least overlap along x
    depth = np.abs(np.einsum("ij,j->i", c[:, 0:3] - c[:, 3:6], n))
    assert np.allclose(depth, 2.0 - 1.93, atol=1e-4)  # EPA stops within 1e-4 (epa.cpp:6,154)


def test_collision_conserves_momentum():
    world = scenes.World(name="headon")
    cube = world.shape(scenes.cube_corners())
    q = scenes.quaternion_new((0.3, 1.0, 0.2), 17.0)
    world.add_body(cube, (0.5, 0.5, 0.5), (-0.6, 0.0, 0.0), q, mass=2.0, linear_velocity=(3.0, 0.2, 0.0),
                   material=(0.4, 0.3, 0.2))
    world.add_body(cube, (0.7, 0.7, 0.7), (0.9, 0.05, 0.1), (0.0, 0.0, 0.0, 1.0), mass=3.0,
                   linear_velocity=(-2.0, 0.0, 0.1), material=(0.4, 0.3, 0.2))
    world.end_scene()
    sim = SimWorld("orc", world)
    m = np.array([2.0, 3.0])
    p0 = (m[:, None] * sim.state()["linear_velocity"]).sum(0)
    sim.step(12, gravity=0.0)
    st = sim.state()
    p1 = (m[:, None] * st["linear_velocity"]).sum(0)
    assert np.linalg.norm(st["linear_velocity"][0] - [3.0, 0.2, 0.0]) > 0.1  # they did collide
    assert np.allclose(p1, p0, atol=1e-11 * np.linalg.norm(p0))


def test_broad_phase_equals_brute_force():
    rng = np.random.default_rng(3)
    world = scenes.World(name="cloud")
    cube = world.shape(scenes.cube_corners())
    ico = world.shape(scenes.icosphere_corners(1))
    n = 60
    pos = rng.uniform(-6.0, 6.0, (n, 3))
    r = rng.uniform(0.3, 1.2, n)
    for k in range(n):
        world.add_body(cube if k % 2 else ico, (r[k],) * 3, tuple(pos[k]), (0.0, 0.0, 0.0, 1.0))
    world.end_scene()
    sim = SimWorld("orc", world)
    radius = sim.body_info()[3]
    expect = []
    for i in range(n):
        for j in range(i + 1, n):
            d = pos[i] - pos[j]
            dist = np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
            if dist <= radius[i] + radius[j] + 0.1:
                expect.append((i, j))
    assert np.array_equal(sim.broad_pairs(0), np.asarray(expect, dtype=np.uint32).reshape(-1, 2))
    assert np.allclose(radius[1::2], r[1::2] * np.sqrt(3.0), rtol=1e-15)  # cube radius = half-diagonal


def test_hull_topology_invariants():
    for corners, nv_expect, nf_expect in ((scenes.cube_corners(), 8, 6), (scenes.icosphere_corners(1), 42, 80),
                                          (scenes.icosphere_corners(2), 162, 320)):
        sim = SimWorld("orc", _single(corners))
        h = sim.hull(0)
        nv, nf = len(h["vertices"]), len(h["face_normals"])
        assert (nv, nf) == (nv_expect, nf_expect)
        edges = set()
        for f in range(nf):
            el = h["face_el"][h["face_off"][f]:h["face_off"][f + 1]]
            for a, b in zip(el, np.roll(el, -1)):
                edges.add((min(a, b), max(a, b)))
            # each face normal is a unit vector pointing away from the centre
            nrm = h["face_normals"][f]
            assert abs(np.linalg.norm(nrm) - 1.0) < 1e-14
            assert np.dot(nrm, h["vertices"][el].mean(0)) > 0.0
        assert nv - len(edges) + nf == 2  # Euler's formula on the merged faces


def test_resting_body_sleeps_after_one_second():
    world = _single(scenes.cube_corners(), pos=(0.0, -0.001, 0.0), floor=True)
    sim = SimWorld("orc", world)
    for f in range(200):
        sim.step(1)
        st = sim.state()
        if st["active"][1, 0] == 0:
            break
    else:
        raise AssertionError("resting cube never slept")
    # quiet frames accumulate dt per frame; sleep needs >= 1.0 s of them
    assert 60 <= f + 1 <= 90
    assert st["deactivation_time"][1, 0] >= 1.0


def test_sphere_and_compound_inertia_closed_forms():
    """A lone sphere is a solid ball, I = (2/5) m r^2 with r the float radius
    (collider.cpp:452-462); a compound spreads its mass over the vertices of
    all its hulls (collider.cpp:465-491): two half-extent-0.5 boxes at
    x = +-0.75 give 16 corner masses m/16 at (+-0.75 +- 0.5, +-0.5, +-0.5)."""
    w = scenes.World(name="inertia")
    r = 0.7
    sph = w.sphere(r)
    left = w.shape(scenes.box_corners((0.5, 0.5, 0.5), (-0.75, 0.0, 0.0)))
    right = w.shape(scenes.box_corners((0.5, 0.5, 0.5), (0.75, 0.0, 0.0)))
    w.add_body(sph, (1, 1, 1), (0, 0, 0), (0, 0, 0, 1), mass=3.0)
    w.add_body((left, right), (1, 1, 1), (0, 5, 0), (0, 0, 0, 1), mass=2.0)
    w.end_scene()
    _, inertia, _, radius = SimWorld("orc", w).body_info()
    rf = float(np.float32(r))
    assert np.allclose(inertia[0].reshape(3, 3), np.eye(3) * 0.4 * 3.0 * rf * rf, rtol=1e-15, atol=0)
    assert radius[0] == rf
    xs = np.array([x + d for x in (-0.75, 0.75) for d in (-0.5, 0.5)])
    pts = np.array([(x, y, z) for x in xs for y in (-0.5, 0.5) for z in (-0.5, 0.5)])
    m = 2.0 / len(pts)
    I = np.zeros((3, 3))
    for x, y, z in pts:
        I += m * np.array([[y * y + z * z, x * y, x * z], [x * y, x * x + z * z, y * z], [x * z, y * z, x * x + y * y]])
    assert np.allclose(inertia[1].reshape(3, 3), I, rtol=1e-12, atol=1e-15)
    assert np.isclose(radius[1], np.sqrt(1.25 ** 2 + 0.5), rtol=1e-15)


def test_sphere_rests_at_its_radius():
    """A ball dropped on the floor (top face y = -1) comes to rest with its
    centre at y = -1 + r, and its single contact sits at the ball's bottom."""
    w = scenes.World(name="ball")
    scenes._floor(w)
    r = 0.8
    w.add_body(w.sphere(r), (1, 1, 1), (0.2, 1.0, -0.1), (0, 0, 0, 1), mass=1.0, material=(0.5, 0.3, 0.0))
    w.end_scene()
    sim = SimWorld("orc", w)
    sim.step(180)
    y = sim.state()["position"][1, 1]
    assert abs(y - (-1.0 + float(np.float32(r)))) < 2e-3
    st = sim.state()
    st["position"][1, 1] -= 0.005
    sim.set_state(st)
    c = sim.contacts(0, 0, 1)
    assert len(c) == 1
    p_ball = c[0, 3:6]
    assert np.allclose(p_ball, st["position"][1] - np.array([0.0, float(np.float32(r)), 0.0]), atol=1e-9)
