"""GPU parity: the HIP step (librawphys.so) against the oracle, bit for bit.

The oracle (oracle/liboracle.so) is the C restatement of the reference's step,
itself pinned bit-exactly to the compiled reference and to committed golden
vectors (tests/test_oracle_*.py). Contact-only scenes must match to the last
bit: pair lists, contact sets and every body-state field after every frame.
"""
import ctypes

import numpy as np
import pytest

from _libs import SimWorld, load, scenes, state_bits_equal

pytestmark = pytest.mark.gpu


def _pair(world, frames, check_every=1, **kw):
    o = SimWorld("orc", world)
    g = SimWorld("rp", world)
    done = 0
    while done < frames:
        n = min(check_every, frames - done)
        o.step(n, **kw)
        g.step(n, **kw)
        done += n
        so, sg = o.state(), g.state()
        bad = state_bits_equal(so, sg)
        if bad:
            diff = np.abs(so["position"] - sg["position"]).max()
            pytest.fail(f"{world.name}: state fields {bad} differ after frame {done} (max |dx| = {diff:.3e})")
    for s in range(world.num_scenes):
        assert g.error(s) == 0, f"scene {s} error word {g.error(s):#x}"
    return o, g


def test_fp64_sqrt_div_correctly_rounded():
    lib = load("rp")
    rng = np.random.default_rng(1)
    n = 1 << 22
    a = np.exp(rng.uniform(-40, 40, n)) * rng.choice([1.0], n)
    b = rng.standard_normal(n) * np.exp(rng.uniform(-20, 20, n))
    out = np.zeros(5 * n)
    assert lib.rp_fp64_selftest(a.ctypes.data, b.ctypes.data, out.ctypes.data, n) == 0
    out = out.reshape(n, 5)
    assert np.array_equal(out[:, 0].view(np.uint64), np.sqrt(a).view(np.uint64)), "device sqrt not correctly rounded"
    assert np.array_equal(out[:, 1].view(np.uint64), (a / b).view(np.uint64)), "device division not correctly rounded"
    # asin/sin/cos (joint limits only): within 1 ulp of the host libm
    x = b / (np.abs(b) + 1.0)
    for k, ref in ((2, np.arcsin(x)), (3, np.sin(b)), (4, np.cos(b))):
        ulp = np.spacing(np.abs(ref)) + 1e-300
        assert np.all(np.abs(out[:, k] - ref) <= 2 * ulp)


def test_hull_topology_matches_oracle():
    w = scenes.batched_uniform(2)
    w2 = scenes.batched_uniform(1, ico_subdiv=2)
    for world in (w, w2):
        o, g = SimWorld("orc", world), SimWorld("rp", world)
        for b in range(world.num_bodies):
            ho, hg = o.hull(b), g.hull(b)
            for k in ho:
                assert np.array_equal(ho[k], hg[k]), (b, k)
        for a, b in zip(o.body_info(), g.body_info()):
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_broad_pairs_and_contacts_match_oracle():
    world = scenes.batched_uniform(4)
    o, g = SimWorld("orc", world), SimWorld("rp", world)
    o.step(40)
    g.set_state(o.state())
    for s in range(world.num_scenes):
        po, pg = o.broad_pairs(s), g.broad_pairs(s)
        assert np.array_equal(po, pg), s
        for i, j in po:
            co, cg = o.contacts(s, int(i), int(j)), g.contacts(s, int(i), int(j))
            assert co.shape == cg.shape, (s, i, j)
            assert np.array_equal(co.view(np.uint64), cg.view(np.uint64)), (s, i, j)


def test_stack_one_frame_bit_exact():
    _pair(scenes.stack_scene(), frames=3)


def test_stack_300_frames_bit_exact():
    _pair(scenes.stack_scene(seed=3), frames=300, check_every=30)


def test_batched_uniform_bit_exact():
    o, g = _pair(scenes.batched_uniform(24), frames=120, check_every=20)
    # this workload has GJK runs that cycle without separating or enclosing
    # the origin (the oracle runs them to the 100-iteration limit); the device
    # ends them early on a repeated state, with identical results
    assert g.lib.rp_world_gjk_cycle_exits(g.h) > 0


def test_batched_ico162_bit_exact():
    _pair(scenes.batched_uniform(3, ico_subdiv=2), frames=60, check_every=20)


def test_642_vertex_hulls_on_lane_groups_bit_exact():
    """Icospheres of 642 vertices beside cubes: a 16-lane group stages two
    such hulls in 30.8 KiB, so four groups take 123 KiB of a block's LDS,
    which gfx950 allows (160 KiB). The world runs its pairs on the fused
    16-lane schedule (large-hull pairs and cube pairs in their own lists)
    and matches the oracle bit for bit; split schedules are refused."""
    world = scenes.batched_uniform(2, first_seed=40, n_cubes=6, n_icos=3, ico_subdiv=3)
    g = SimWorld("rp", world)
    lim = g.lib.rp_world_lds_limit(g.h)
    print(f"LDS limit per block: {lim} bytes")
    if lim < 4 * 2 * 642 * 24:
        pytest.skip(f"the device gives a block {lim} bytes of LDS")
    assert g.lib.rp_world_narrow_group(g.h) > 1, "the large hulls should fit a lane-group schedule"
    # slices above 64 KiB run in the fused kernel only
    assert g.lib.rp_world_set_manifold_group(g.h, 1) != 0
    assert g.lib.rp_world_set_manifold_group(g.h, 16) != 0
    assert g.lib.rp_world_manifold_group(g.h) == 0
    g.close()
    _pair(world, frames=45, check_every=15)


@pytest.mark.parametrize("kind", ["narrow", "manifold", "fused"])
@pytest.mark.parametrize("lanes", [1, 4, 8, 16])
def test_every_narrow_schedule_bit_exact(kind, lanes):
    """The GJK + EPA kernel's and the manifold kernel's schedules (one lane per
    pair, or 4/8/16 lanes per pair with split support searches), and the
    GJK + EPA kernel building the manifolds itself (manifold schedule 0),
    give identical results."""
    setter = "rp_world_set_manifold_group" if kind == "manifold" else "rp_world_set_narrow_group"
    for world, frames in ((scenes.batched_uniform(6, first_seed=500), 90),
                          (scenes.batched_uniform(2, first_seed=600, ico_subdiv=2), 40)):
        o = SimWorld("orc", world)
        g = SimWorld("rp", world)
        if kind == "fused":
            assert g.lib.rp_world_set_manifold_group(g.h, 0) == 0
        # a group stages both hulls of its pair in LDS; schedules whose
        # (64 / lanes) groups would need more than the device's LDS per block
        # are refused
        max_nv = max(len(np.unique(c, axis=0)) for c in world.shapes)
        lds = (64 // lanes) * 2 * max_nv * 24 if lanes > 1 else 0
        if lds > g.lib.rp_world_lds_limit(g.h):
            assert getattr(g.lib, setter)(g.h, lanes) != 0
            continue
        assert getattr(g.lib, setter)(g.h, lanes) == 0
        for done in range(0, frames, 15):
            o.step(15)
            g.step(15)
            bad = state_bits_equal(o.state(), g.state())
            assert not bad, f"{kind} {lanes} lanes/pair, {world.name}, frame {done + 15}: {bad}"
        for s in range(world.num_scenes):
            assert g.error(s) == 0


def test_uneven_scenes_match_oracle():
    """Mixed 8/42/162-vertex hulls, per-body friction and restitution, uneven
    scene sizes; scenes carrying joint chains use the device libm in the joint
    limits (held to 1e-9), all others must match bit for bit."""
    world = scenes.batched_uneven(10, first_seed=21, max_bodies=48)
    o = SimWorld("orc", world)
    g = SimWorld("rp", world)
    o.step(45)
    g.step(45)
    so, sg = o.state(), g.state()
    off = world.scene_body_offset
    con = world.scene_constraint_offset
    exact = 0
    for s in range(world.num_scenes):
        sl = slice(off[s], off[s + 1])
        part_o = {k: v[sl] for k, v in so.items()}
        part_g = {k: v[sl] for k, v in sg.items()}
        if con[s + 1] == con[s]:
            assert not state_bits_equal(part_o, part_g), f"scene {s}"
            exact += 1
        else:
            assert np.allclose(part_o["position"], part_g["position"], rtol=0, atol=1e-9), f"scene {s}"
        assert g.error(s) == 0
    assert exact >= 5


# Scenes of the batched uneven workload itself (batched_uneven(1024) uses
# seeds 0..1023): six with more than 128 bodies (up to 252), seven with joint
# chains, every one with 8-, 42- and 162-vertex hulls.
UNEVEN_SEEDS = [6, 16, 27, 36, 70, 107, 0, 3, 26, 33, 41, 59]
# Joint limits call asin / sin / cos (pbd.cpp:177, quaternion.cpp:7-10); the
# device libm may differ from glibc in the last bit, and those differences
# propagate through the chain and whatever it touches. Positions and
# rotations within JOINT_TOL after 120 frames; velocities are position
# differences over h = 1/1200 (pbd.cpp:629-642), so within JOINT_TOL / h.
JOINT_TOL = 1e-9


def test_uneven_full_size_scenes_match_oracle():
    w = scenes.World(name="uneven_full")
    for sd in UNEVEN_SEEDS:
        scenes.uneven_scene(w, sd)
    assert max(np.diff(w.scene_body_offset)) > 250
    o = SimWorld("orc", w)
    g = SimWorld("rp", w)
    off, con = w.scene_body_offset, w.scene_constraint_offset
    for frame in (30, 60, 90, 120):
        o.step(30)
        g.step(30)
        so, sg = o.state(), g.state()
        worst = 0.0
        for s in range(w.num_scenes):
            sl = slice(off[s], off[s + 1])
            part_o = {k: v[sl] for k, v in so.items()}
            part_g = {k: v[sl] for k, v in sg.items()}
            if con[s + 1] == con[s]:
                bad = state_bits_equal(part_o, part_g)
                assert not bad, f"scene seed {UNEVEN_SEEDS[s]}, frame {frame}: {bad}"
                continue
            for k, tol in (("position", JOINT_TOL), ("rotation", JOINT_TOL),
                           ("linear_velocity", JOINT_TOL * 1200.0), ("angular_velocity", JOINT_TOL * 1200.0)):
                d = np.abs(part_o[k] - part_g[k]).max()
                worst = max(worst, d / tol)
                assert d <= tol, f"scene seed {UNEVEN_SEEDS[s]}, frame {frame}, {k}: {d:.3e}"
            assert np.array_equal(part_o["active"], part_g["active"])
        print(f"frame {frame}: chain scenes within {worst:.3f} of their tolerance")
    for s in range(w.num_scenes):
        assert g.error(s) == 0 and o.error(s) == 0


@pytest.mark.parametrize("lanes", [64, 32, 16])
def test_every_solve_schedule_bit_exact(lanes):
    """The solve's schedules (1, 2 or 4 scenes per wave) give identical results,
    with and without external constraints."""
    for world, frames, kw in ((scenes.batched_uniform(7, first_seed=700), 60, {}),
                              (scenes.chains(5, length=4, first_seed=800), 30, {"collisions": False})):
        o = SimWorld("orc", world)
        g = SimWorld("rp", world)
        assert g.lib.rp_world_set_solve_lanes(g.h, lanes) == 0
        o.step(frames, **kw)
        g.step(frames, **kw)
        so, sg = o.state(), g.state()
        if world.constraints:
            assert np.allclose(so["position"], sg["position"], rtol=0, atol=1e-9)
        else:
            assert not state_bits_equal(so, sg), f"{lanes} lanes/scene"
        for s in range(world.num_scenes):
            assert g.error(s) == 0


def test_graph_replays_match_eager_launches():
    """Frames replayed from a captured HIP graph equal eager launches and the
    oracle, over many replays with host reads and writes of the state between
    them (graph nodes must not depend on anything the first replay changed)."""
    world = scenes.batched_uniform(5, first_seed=900)
    o = SimWorld("orc", world)
    ge = SimWorld("rp", world)
    gg = SimWorld("rp", world)
    ge.lib.rp_world_set_graphs(ge.h, 0)
    gg.lib.rp_world_set_graphs(gg.h, 1)
    for k in range(8):
        o.step(7)
        ge.step(7)
        gg.step(7)
        so, se, sg = o.state(), ge.state(), gg.state()
        assert not state_bits_equal(so, se), f"eager, round {k}"
        assert not state_bits_equal(so, sg), f"graph, round {k}"
        gg.set_state(sg)  # host round trip between replays
    # a parameter change re-captures
    o.step(3, substeps=10)
    gg.step(3, substeps=10)
    assert not state_bits_equal(o.state(), gg.state())


def test_spheres_and_compounds_bit_exact():
    """Sphere colliders and two-hull compound bodies (general narrow path:
    one lane per pair over every collider pair) match the oracle bit for bit;
    the oracle matches the compiled reference on these scenes
    (tests/golden/traj_mixed.npz)."""
    world = scenes.batched_mixed(8, first_seed=40)
    o, g = _pair(world, frames=150, check_every=30)
    assert g.lib.rp_world_error(g.h, 0) == 0


@pytest.mark.parametrize("views", [2, 3, 4])
def test_views_bit_exact(views):
    """Graph replay in scene-range views on concurrent streams matches the
    oracle bit for bit; 9 scenes cut into unequal ranges."""
    world = scenes.batched_uniform(9, first_seed=900)
    o = SimWorld("orc", world)
    g = SimWorld("rp", world)
    assert g.lib.rp_world_set_views(g.h, views) == 0
    for done in range(0, 60, 20):
        o.step(20)
        g.step(20)
        bad = state_bits_equal(o.state(), g.state())
        assert not bad, f"{views} views, frame {done + 20}: {bad}"
    assert g.lib.rp_world_views(g.h) == views
    for s in range(world.num_scenes):
        assert g.error(s) == 0


@pytest.mark.parametrize("blocks", [1, 3])
def test_few_narrow_blocks_bit_exact(blocks):
    """With the narrow-phase grid capped at a few one-wave blocks, each lane
    group takes many live pairs in turn (grid-stride), reusing its LDS slice
    and working-memory slot: fused, split (GJK + EPA, then manifold) and
    size-class launches against the oracle."""
    uneven = scenes.batched_uneven(6, first_seed=41, max_bodies=48)
    for world, frames, narrow, manifold in ((scenes.batched_uniform(6, first_seed=500), 60, 8, 0),
                                            (scenes.batched_uniform(6, first_seed=500), 60, 8, 8),
                                            (scenes.batched_uniform(6, first_seed=500), 60, 1, 1),
                                            (uneven, 40, 8, 0)):
        o = SimWorld("orc", world)
        g = SimWorld("rp", world)
        assert g.lib.rp_world_set_narrow_group(g.h, narrow) == 0
        assert g.lib.rp_world_set_manifold_group(g.h, manifold) == 0
        assert g.lib.rp_world_set_narrow_block_cap(g.h, blocks) == 0
        for done in range(0, frames, 20):
            o.step(20)
            g.step(20)
            bad = state_bits_equal(o.state(), g.state())
            assert not bad, f"{world.name} narrow {narrow} manifold {manifold} blocks {blocks}, frame {done + 20}: {bad}"
        for s in range(world.num_scenes):
            assert g.error(s) == 0
    assert g.lib.rp_world_set_narrow_block_cap(g.h, 8193) != 0


def test_size_classes_change_nothing():
    """With hulls above 64 vertices, the fused schedule visits small-hull and
    large-hull pairs in two lists (LDS slices sized per list): the same bits
    as one list (manifold schedule 1), on uneven scenes, in views, and on a
    large scene of cubes and 162-vertex icospheres beside small scenes,
    which also matches the oracle."""
    big = scenes.large_scene(240, seed=3, large_scene_bodies=64, ico_subdiv=2)
    for seed in range(3):
        scenes.drop_scene(big, 1100 + seed)
    for world, frames, views in ((scenes.batched_uneven(8, first_seed=41, max_bodies=48), 40, 1),
                                 (scenes.batched_uneven(8, first_seed=41, max_bodies=48), 40, 3),
                                 (big, 30, 1)):
        for lanes in (8, 16):
            a, b = SimWorld("rp", world), SimWorld("rp", world)
            for w_, m in ((a, 0), (b, 1)):
                assert w_.lib.rp_world_set_narrow_group(w_.h, lanes) == 0
                assert w_.lib.rp_world_set_manifold_group(w_.h, m) == 0
                assert w_.lib.rp_world_set_views(w_.h, views) == 0
            a.step(frames)
            b.step(frames)
            bad = state_bits_equal(a.state(), b.state())
            assert not bad, f"{world.name}, {lanes} lanes, {views} views: {bad}"
            for s in range(world.num_scenes):
                assert a.error(s) == b.error(s)
    # the large scene is solved class by class (coloured): the oracle solves
    # it in the device's class order, the small scenes in list order
    o, g = SimWorld("orc", big), SimWorld("rp", big)
    o.lib.orc_world_set_solve_order.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32]
    o.lib.orc_world_solve_order_status.restype = ctypes.c_uint32
    for f in range(30):
        g.step(1)
        n = g.lib.rp_world_item_classes(g.h, 0, None, 0)
        cls = np.zeros(max(n, 1), dtype=np.int32)
        g.lib.rp_world_item_classes(g.h, 0, cls.ctypes.data, n)
        o.lib.orc_world_set_solve_order(o.h, 0, cls.ctypes.data, n)
        o.step(1)
    assert o.lib.orc_world_solve_order_status(ctypes.c_void_p(o.h)) == 0
    bad = state_bits_equal(o.state(), g.state())
    assert not bad, f"{big.name} against the oracle: {bad}"


@pytest.mark.parametrize("lanes,manifold", [(8, 0), (16, 0), (8, 1)])
def test_prisms_on_pillars_bit_exact(lanes, manifold):
    """Octagonal prisms resting on cube pillars: with 16 + 8-vertex pairs in
    slices sized for two 20-vertex hulls, the group clips in 8-point LDS
    buffers; a level prism's octagon clipped by the pillar's square grows to
    10-12 points, so the clipping is redone by lane 0 in working memory. The
    contacts stay bit-exact on every schedule."""
    world = scenes.prisms_on_pillars(6, first_seed=60)
    o, g = SimWorld("orc", world), SimWorld("rp", world)
    assert g.lib.rp_world_set_narrow_group(g.h, lanes) == 0
    assert g.lib.rp_world_set_manifold_group(g.h, manifold) == 0
    for done in range(0, 90, 30):
        o.step(30)
        g.step(30)
        bad = state_bits_equal(o.state(), g.state())
        assert not bad, f"{lanes}/{manifold}, frame {done + 30}: {bad}"
    for s in range(world.num_scenes):
        assert g.error(s) == 0


def test_long_run_default_schedule_bit_exact():
    """600 frames of 64 scenes on the default path (three concurrent views,
    graph replay, fused lane-group narrow phase, solve with integration):
    falling, resting and sleeping bodies, bit for bit at frames 300 and 600."""
    world = scenes.batched_uniform(64, first_seed=700)
    o, g = SimWorld("orc", world), SimWorld("rp", world)
    for done in (300, 600):
        o.step(300)
        g.step(300)
        bad = state_bits_equal(o.state(), g.state())
        assert not bad, f"frame {done}: {bad}"
    assert g.lib.rp_world_views(g.h) == 3
    for s in range(world.num_scenes):
        assert g.error(s) == 0


def _large_beside_small():
    w = scenes.large_scene(300, seed=1, large_scene_bodies=64)
    for seed in range(3):
        scenes.drop_scene(w, 1000 + seed)
    return w


def test_views_change_nothing():
    """Views only reorder independent scenes' work: joint chains (external
    constraints in every range) and a large scene (grid broad phase,
    coloured solve) beside small ones give the same bits with 4 views as
    with one."""
    for world, frames in ((scenes.batched_uneven(8, first_seed=31, max_bodies=48), 40),
                          (_large_beside_small(), 30)):
        a, b = SimWorld("rp", world), SimWorld("rp", world)
        assert a.lib.rp_world_set_views(a.h, 4) == 0
        assert b.lib.rp_world_set_views(b.h, 1) == 0
        a.step(frames)
        b.step(frames)
        assert a.lib.rp_world_views(a.h) == 4 and b.lib.rp_world_views(b.h) == 1
        bad = state_bits_equal(a.state(), b.state())
        assert not bad, f"{world.name}: {bad}"
        for s in range(world.num_scenes):
            assert a.error(s) == b.error(s)


def test_reference_asset_scenes_bit_exact():
    """spot_storm and mirror_cube (tests/golden/assets.npz: the reference's
    collider assets as its obj_parse returns them) on the batched device
    path against the oracle, bit for bit, with the error words clear."""
    from test_gpu_dropin import _asset_scenes

    for world, frames in zip(_asset_scenes(), (90, 120)):
        _pair(world, frames=frames, check_every=30)
