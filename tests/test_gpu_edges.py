"""Edge cases of the batched step against the oracle, bit for bit: ragged
scene sizes in one world (a floor alone, one body, a few, a full scene),
scenes of fixed bodies only, several position iterations, collisions off,
one substep and a longer time step, and a world of one body."""
from __future__ import annotations

import numpy as np
import pytest

from _libs import SimWorld, scenes, state_bits_equal

pytestmark = pytest.mark.gpu


def _compare(world, frames, check_every=10, **kw):
    o = SimWorld("orc", world)
    g = SimWorld("rp", world)
    done = 0
    while done < frames:
        n = min(check_every, frames - done)
        o.step(n, **kw)
        g.step(n, **kw)
        done += n
        bad = state_bits_equal(o.state(), g.state())
        assert not bad, f"{world.name}: {bad} differ after frame {done} ({kw})"
    for s in range(world.num_scenes):
        assert g.error(s) == 0, f"scene {s} error word {g.error(s):#x}"
    o.close()
    g.close()


def _ragged_world():
    """Scenes of 1, 2, 3, 6 and 17 bodies, and one of fixed bodies only."""
    w = scenes.World(name="ragged")
    for seed, (nc, ni) in enumerate(((0, 0), (1, 0), (1, 1), (3, 2), (8, 8))):
        scenes.drop_scene(w, 700 + seed, n_cubes=nc, n_icos=ni)
    cube = w.shape(scenes.cube_corners())
    scenes._floor(w)
    for k in range(4):  # fixed cubes resting on and above the floor
        w.add_body(cube, (1.0, 1.0, 1.0), (3.0 * k, 0.5 * k, 0.0), scenes.quaternion_new((0.0, 1.0, 0.0), 10.0 * k),
                   fixed=True)
    w.end_scene()
    return w


def test_ragged_scene_sizes_bit_exact():
    _compare(_ragged_world(), frames=120, check_every=30)


@pytest.mark.parametrize("kw", [{"iters": 4}, {"collisions": False}, {"substeps": 1, "dt": 1.0 / 30.0},
                                {"substeps": 7, "dt": 1.0 / 45.0}], ids=["iters4", "no_collisions", "one_substep",
                                                                        "seven_substeps"])
def test_step_parameters_bit_exact(kw):
    _compare(scenes.batched_uniform(4, first_seed=800), frames=60, check_every=20, **kw)


def test_single_body_world_bit_exact():
    w = scenes.World(name="one_body")
    cube = w.shape(scenes.cube_corners())
    w.add_body(cube, (0.7, 0.7, 0.7), (0.0, 3.0, 0.0), scenes.quaternion_new((1.0, 0.0, 0.0), 30.0),
               linear_velocity=(0.5, 0.0, -0.25), angular_velocity=(0.0, 2.0, 1.0))
    w.end_scene()
    _compare(w, frames=60, check_every=20)


def test_scene_of_fixed_bodies_stays_put():
    """A scene whose bodies are all fixed matches the oracle and never moves
    (fixed bodies are never written, pbd.cpp:702-711)."""
    w = scenes.World(name="all_fixed")
    cube = w.shape(scenes.cube_corners())
    scenes._floor(w)
    w.add_body(cube, (1.0, 1.0, 1.0), (0.0, -0.5, 0.0), scenes.quaternion_new((0.0, 1.0, 0.0), 0.0), fixed=True)
    w.end_scene()
    g = SimWorld("rp", w)
    before = g.state()
    g.step(30)
    moved = state_bits_equal(before, g.state(), fields=("position", "rotation"))
    assert not moved, f"fixed bodies changed {moved}"
    assert g.error(0) == 0
    g.close()
    _compare(w, frames=30, check_every=10)
