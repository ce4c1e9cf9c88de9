"""The drop-in: rp_pbd_simulate (= the exported pbd_simulate_with_constraints)
stepping Entity structs that the reference's own code built.

oracle/_ref/libref.so holds the reference compiled from its sources; its
harness creates entities with entity_create / collider_convex_hull_create and
runs the reference's frame driver (stack.cpp:86-104), calling either the
reference's step or, here, the library's drop-in through a function pointer.
The two runs must agree bit for bit on contact-only scenes; scenes with joint
limits use the device asin/sin/cos and are held to a tolerance instead.
"""
import ctypes

import numpy as np
import pytest

from _libs import SimWorld, load, scenes, state_bits_equal

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(load("ref") is None, reason="reference build (oracle/_ref) not present")]


def _dropin_fn():
    lib = load("rp")
    return ctypes.cast(lib.rp_pbd_simulate, ctypes.c_void_p).value


def _run_pair(world, frames, worlds=False, **kw):
    ref = load("ref")
    ref.ref_world_step_with.restype = ctypes.c_uint32
    ref.ref_world_step_with.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_uint32,
                                        ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32, ctypes.c_double]
    a = SimWorld("ref", world)  # reference step
    b = SimWorld("ref", world)  # same entities, library step
    dt, sub, it = kw.get("dt", 1.0 / 60.0), kw.get("substeps", 20), kw.get("iters", 1)
    coll = 1 if kw.get("collisions", True) else 0
    fn = _dropin_fn()
    for f in range(frames):
        a.step(1, dt=dt, substeps=sub, iters=it, collisions=bool(coll))
        err = ref.ref_world_step_with(b.h, fn, dt, 1, sub, it, coll, 10.0)
        assert err == 0, f"frame {f}: error word {err:#x}"
    if worlds:
        return a, b
    return a.state(), b.state()


def _geometry(sim):
    """Every body's collider geometry as the entities hold it."""
    ref = load("ref")
    ref.ref_world_collider_geometry.restype = ctypes.c_uint32
    ref.ref_world_collider_geometry.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32]
    out = []
    for g in range(sim.world.num_bodies):
        n = ref.ref_world_collider_geometry(sim.h, g, None, 0)
        buf = np.zeros(n)
        ref.ref_world_collider_geometry(sim.h, g, buf.ctypes.data, n)
        out.append(buf)
    return out


def test_dropin_stack_bit_exact():
    sa, sb = _run_pair(scenes.stack_scene(seed=5), 120)
    assert not state_bits_equal(sa, sb)


def test_dropin_batched_scene_bit_exact():
    sa, sb = _run_pair(scenes.batched_uniform(1, first_seed=42), 150)
    assert not state_bits_equal(sa, sb)


def test_dropin_joint_chain_within_tolerance():
    sa, sb = _run_pair(scenes.chains(1, length=6, first_seed=3), 120, collisions=False)
    # joint limits call asin / sin / cos (pbd.cpp:177, quaternion.cpp:7-10);
    # the device libm may round differently from glibc in the last bit
    assert np.allclose(sa["position"], sb["position"], rtol=0, atol=1e-9)
    assert np.allclose(sa["rotation"], sb["rotation"], rtol=0, atol=1e-9)


def test_dropin_dt_nonpositive_is_noop():
    lib = load("rp")
    assert lib.rp_pbd_simulate(0.0, None, None, 20, 1, 1) == 0
    assert lib.rp_pbd_simulate(-1.0, None, None, 20, 1, 1) == 0


def test_dropin_spheres_and_compounds_bit_exact():
    """Entities whose colliders the reference built (collider_sphere_create,
    several convex hulls per entity) through the drop-in."""
    sa, sb = _run_pair(scenes.batched_mixed(1, first_seed=7), 120)
    assert not state_bits_equal(sa, sb)


def _writeback_scene():
    """Floor; a cube resting on it (asleep after about 75 frames); a sphere and
    a two-box compound falling onto it; a cube far above everything (in no
    broad pair, so never visited by the collision pass, yet moving)."""
    w = scenes.World(name="writeback")
    scenes._floor(w)
    cube = w.shape(scenes.cube_corners())
    w.add_body(cube, (1, 1, 1), (0.0, -0.001, 0.0), (0, 0, 0, 1))
    w.add_body(w.sphere(0.6), (1, 1, 1), (3.0, 2.0, 0.5), (0, 0, 0, 1), angular_velocity=(0.0, 0.0, -2.0))
    left = w.shape(scenes.box_corners((0.4, 0.4, 0.4), (-0.6, 0.0, 0.0)))
    right = w.shape(scenes.box_corners((0.4, 0.4, 0.4), (0.6, 0.0, 0.0)))
    w.add_body((left, right), (1, 1, 1), (-3.0, 1.5, 0.0), scenes.quaternion_new((0.3, 1.0, 0.2), 25.0))
    w.add_body(cube, (1, 1, 1), (0.0, 500.0, 0.0), (0, 0, 0, 1), angular_velocity=(0.4, 0.0, 0.3))
    w.end_scene()
    return w


@pytest.mark.parametrize("make,frames", [(_writeback_scene, 100),
                                         (lambda: scenes.batched_mixed(1, first_seed=11), 90)])
def test_dropin_writes_back_collider_geometry(make, frames):
    """After every call the entities' colliders hold what the reference's
    collision pass left in them, bit for bit: transformed hull vertices,
    transformed face normals and sphere centers at the pose after the last
    substep's integration for the entities the pass visited, and the frame
    driver's own update (the frame-start pose) for the others."""
    a, b = _run_pair(make(), frames, worlds=True)
    assert not state_bits_equal(a.state(), b.state())
    for g, (x, y) in enumerate(zip(_geometry(a), _geometry(b))):
        assert x.shape == y.shape
        bad = np.flatnonzero(x.view(np.uint64) != y.view(np.uint64))
        assert bad.size == 0, f"body {g}: {bad.size} values differ, first at {bad[0]}"
    if make is _writeback_scene:
        st = a.state()
        assert st["active"][1, 0] == 0, "the resting cube should be asleep"
        assert st["position"][4, 1] < 490.0, "the far cube should have fallen"


def test_dropin_collider_geometry_untouched_without_collisions():
    """With collisions off the reference never calls collider_update."""
    a, b = _run_pair(scenes.chains(1, length=4, first_seed=5), 5, worlds=True, collisions=False)
    for x, y in zip(_geometry(a), _geometry(b)):
        assert np.array_equal(x.view(np.uint64), y.view(np.uint64))


def test_model_matrices_match_entity_get_model_matrix():
    """rp_world_model_matrices against the reference's entity_get_model_matrix
    on the same states (scaled boxes, rotated bodies, spheres)."""
    ref = load("ref")
    ref.ref_world_model_matrix.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
    world = scenes.batched_uniform(2, first_seed=3)
    sim = SimWorld("ref", world)
    sim.step(30)
    from _libs import pkg
    gw = pkg.World(world)
    gw.set_state(sim.state())
    m = gw.model_matrices()
    mf = gw.model_matrices(as_float=True)
    for g in range(world.num_bodies):
        r = np.zeros(16)
        ref.ref_world_model_matrix(sim.h, g, r.ctypes.data)
        assert np.array_equal(m[g].reshape(-1).view(np.uint64), r.view(np.uint64)), g
        assert np.array_equal(mf[g].reshape(-1), r.astype(np.float32)), g


def test_dropin_large_scene_equals_batched_world():
    """More than 4096 entities in one call: the drop-in runs the large-scene
    path (grid broad phase, coloured solve), built from the entities' own
    bounding radii. Its frames equal the batched world's on the same scene
    bit for bit (that path's coloured solve is pinned against the oracle in
    test_gpu_large.py), and stay within the large-scene tolerance of the
    reference's sequential order."""
    w = scenes.large_scene(5000, seed=4)
    a, b = _run_pair(w, 2, worlds=True)
    g = SimWorld("rp", w)
    g.step(2)
    sb, sg, sa = b.state(), g.state(), a.state()
    bad = state_bits_equal(sb, sg)
    assert not bad, bad
    dx = np.abs(sa["position"] - sb["position"]).max()
    print(f"drop-in (coloured) vs reference (sequential) after 2 frames: max |dx| = {dx:.3e}")
    assert dx <= 1e-2
    assert np.all(sb["position"][1:, 1] > -1.0)


def _step_mode(sim, mode, frames, **kw):
    lib = sim.lib
    lib.ref_world_step_mode.restype = ctypes.c_uint32
    lib.ref_world_step_mode.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double, ctypes.c_uint32,
                                        ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32, ctypes.c_double]
    coll = 1 if kw.get("collisions", True) else 0
    assert lib.ref_world_step_mode(sim.h, mode, 1.0 / 60.0, frames, kw.get("substeps", 20), kw.get("iters", 1),
                                   coll, 10.0) == 0


@pytest.mark.skipif(load("ref_dropin") is None, reason="link-level drop-in harness (oracle/_ref) not built")
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_link_level_dropin_bit_exact(mode):
    """The reference's own objects without pbd.o, linked against
    librawphys.so (oracle/_ref/libref_dropin.so): the harness's calls to the
    mangled pbd_simulate_with_constraints (mode 0) and pbd_simulate (1), and
    to physics_simulate (2), run the library's step on entities the
    reference's code built, through the reference's frame driver
    (stack.cpp:86-104). Bit for bit against the reference's own step."""
    for world, frames in ((scenes.stack_scene(seed=6), 90), (scenes.batched_uniform(2, first_seed=77), 90)):
        a = SimWorld("ref", world)
        b = SimWorld("ref_dropin", world)
        _step_mode(a, 0 if mode == 0 else 1, frames)
        _step_mode(b, mode, frames)
        bad = state_bits_equal(a.state(), b.state())
        assert not bad, f"{world.name}, mode {mode}: {bad}"
    assert load("rp").rp_last_error() == 0


@pytest.mark.skipif(load("ref_dropin") is None, reason="link-level drop-in harness (oracle/_ref) not built")
def test_link_level_dropin_joint_chain():
    """External constraints through the linked pbd_simulate_with_constraints,
    built with the library's exported constraint constructors
    (pbd.cpp:18-79); joint limits use the device libm (tolerance)."""
    world = scenes.chains(1, length=6, first_seed=4)
    a = SimWorld("ref", world)
    b = SimWorld("ref_dropin", world)
    _step_mode(a, 0, 90, collisions=False)
    _step_mode(b, 0, 90, collisions=False)
    sa, sb = a.state(), b.state()
    assert np.allclose(sa["position"], sb["position"], rtol=0, atol=1e-9)
    assert np.allclose(sa["rotation"], sb["rotation"], rtol=0, atol=1e-9)


def _asset_scenes():
    from pathlib import Path

    z = np.load(Path(__file__).resolve().parent / "golden" / "assets.npz")
    assert np.array_equal(z["cube"], scenes.cube_corners())
    spot = scenes.spot_storm_scene(scenes.World(name="spot_storm"), [z[f"spot_hull_{k}"] for k in range(1, 12)])
    mirror = scenes.mirror_cube_scene(scenes.World(name="mirror_cube"),
                                      [z["mirror_cube_collider1"], z["mirror_cube_collider2"]])
    return spot, mirror


@pytest.mark.parametrize("which", ["spot_storm", "mirror_cube"])
def test_dropin_reference_asset_scenes_bit_exact(which):
    """The reference's spot_storm (8 bodies of 11 convex pieces, up to 529
    vertices each, from res/spot/*.obj) and mirror_cube (one body of two
    pieces) examples, with entities and colliders built by the reference's
    own code from the corner streams its obj_parse returns
    (tests/golden/assets.npz): the drop-in against the reference's step, bit
    for bit, through the landing."""
    spot, mirror = _asset_scenes()
    world, frames = (spot, 100) if which == "spot_storm" else (mirror, 120)
    sa, sb = _run_pair(world, frames)
    assert not state_bits_equal(sa, sb)
    assert sa["position"][1:, 1].min() < 1.0  # the bodies reached the floor
