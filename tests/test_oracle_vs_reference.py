"""The oracle against the reference compiled from its own sources, live.

Runs only where oracle/_ref/libref.so was built (it needs the reference
sources at build time); elsewhere the committed golden vectors pin the oracle.
"""
import numpy as np
import pytest

from _libs import SimWorld, load, scenes, state_bits_equal

pytestmark = pytest.mark.skipif(load("ref") is None, reason="reference build (oracle/_ref) not present")


def _compare(world, frames, every=1, **kw):
    o, r = SimWorld("orc", world), SimWorld("ref", world)
    done = 0
    while done < frames:
        n = min(every, frames - done)
        o.step(n, **kw)
        r.step(n, **kw)
        done += n
        bad = state_bits_equal(o.state(), r.state())
        assert not bad, f"{world.name}: frame {done}: {bad}"


@pytest.mark.parametrize("seed", [11, 12])
def test_batched_scenes_bit_exact(seed):
    _compare(scenes.batched_uniform(3, first_seed=1000 * seed), 240, every=10)


def test_joint_chains_bit_exact():
    _compare(scenes.chains(2, length=8, first_seed=77), 200, every=20)
    _compare(scenes.chains(2, length=5, first_seed=78), 200, every=50, collisions=False, substeps=50, iters=50)


def test_prisms_on_pillars_bit_exact():
    """Octagonal and decagonal prisms (8- and 10-vertex cap faces) resting on
    cube pillars: clipping polygons of up to a dozen points."""
    _compare(scenes.prisms_on_pillars(3, first_seed=60), 90, every=30)


def test_substep_and_iteration_counts():
    w = scenes.batched_uniform(2, first_seed=5)
    _compare(w, 30, every=10, substeps=1, iters=1)
    _compare(w, 30, every=10, substeps=7, iters=3)


def test_random_pair_contacts_bit_exact():
    """Contacts of random overlapping cube / icosphere poses."""
    rng = np.random.default_rng(7)
    w = scenes.World(name="pairs")
    cube = w.shape(scenes.cube_corners())
    ico = w.shape(scenes.icosphere_corners(1))
    n = 24
    for k in range(n):
        shp = cube if k % 3 else ico
        r = 0.5 + 0.5 * rng.random()
        pos = rng.uniform(-1.2, 1.2, 3)
        q = rng.standard_normal(4)
        q /= np.linalg.norm(q)
        w.add_body(shp, (r, r, r), tuple(pos), tuple(q), mass=1.0)
    w.end_scene()
    o, r = SimWorld("orc", w), SimWorld("ref", w)
    hits = 0
    for i in range(n):
        for j in range(i + 1, n):
            a, b = o.contacts(0, i, j), r.contacts(0, i, j)
            assert a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64)), (i, j)
            hits += len(a) > 0
    assert hits > 30


def test_edge_worlds_bit_exact():
    """The edge-case worlds of the device tests (ragged scene sizes, fixed
    bodies only, one body, several iterations, collisions off, other substep
    counts), so that the oracle those tests check against is pinned on them."""
    from test_gpu_edges import _ragged_world

    _compare(_ragged_world(), 120, every=30)
    w = scenes.batched_uniform(4, first_seed=800)
    for kw in ({"iters": 4}, {"collisions": False}, {"substeps": 1, "dt": 1.0 / 30.0},
               {"substeps": 7, "dt": 1.0 / 45.0}):
        _compare(w, 60, every=20, **kw)
