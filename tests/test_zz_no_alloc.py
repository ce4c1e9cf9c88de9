"""rp_world_step_noalloc: the step without allocation (lane-group working
memory in a static device array, narrow grids capped to its slots, eager
launches) gives rp_world_step's bits, on two calls in a row, the second on a
smaller world, and with the array's contents left by the first call."""
import numpy as np
import pytest

from _libs import SimWorld, scenes, state_bits_equal

pytestmark = pytest.mark.gpu


def _noalloc(sim: SimWorld, frames: int) -> None:
    rc = sim.lib.rp_world_step_noalloc(sim.h, 1.0 / 60.0, frames, 20, 1, 1, 10.0)
    assert rc == 0, sim.lib.rp_last_error_string()


@pytest.mark.parametrize("narrow,manifold", [(8, 0), (4, 4), (16, 0)])
def test_noalloc_step_equals_step(narrow, manifold):
    # first call: 48 scenes (more live pairs than the array's lane groups
    # early on, so groups take several pairs in turn); second call: a
    # smaller world of 5 scenes after the first left its contents
    for world, frames in ((scenes.batched_uniform(48, first_seed=900), 90),
                          (scenes.batched_uniform(5, first_seed=950), 60)):
        a, b = SimWorld("rp", world), SimWorld("rp", world)
        for w_ in (a, b):
            assert w_.lib.rp_world_set_narrow_group(w_.h, narrow) == 0
            assert w_.lib.rp_world_set_manifold_group(w_.h, manifold) == 0
        for done in range(0, frames, 30):
            a.step(30)
            _noalloc(b, 30)
            bad = state_bits_equal(a.state(), b.state())
            assert not bad, f"{world.name}, narrow {narrow} manifold {manifold}, frame {done + 30}: {bad}"
        for s in range(world.num_scenes):
            assert a.error(s) == 0 and b.error(s) == 0


def test_noalloc_step_against_oracle_with_a_large_scene():
    """A large scene (grid broad phase, coloured solve) beside small ones;
    the small scenes against the oracle, the whole world against the
    allocating step."""
    w = scenes.large_scene(300, seed=1, large_scene_bodies=64)
    for seed in range(3):
        scenes.drop_scene(w, 1200 + seed)
    a, b, o = SimWorld("rp", w), SimWorld("rp", w), SimWorld("orc", w)
    a.step(20)
    _noalloc(b, 20)
    o.step(20)
    sa, sb, so = a.state(), b.state(), o.state()
    assert not state_bits_equal(sa, sb)
    small = slice(w.scene_body_offset[1], w.scene_body_offset[-1])
    assert not state_bits_equal({k: v[small] for k, v in sb.items()}, {k: v[small] for k, v in so.items()})


def test_noalloc_dt_nonpositive_is_noop():
    world = scenes.batched_uniform(1)
    b = SimWorld("rp", world)
    st = b.state()
    assert b.lib.rp_world_step_noalloc(b.h, 0.0, 5, 20, 1, 1, 10.0) == 0
    assert not state_bits_equal(st, b.state())
