"""The oracle against the golden vectors recorded from the compiled reference.

tests/golden/*.npz hold the reference's own outputs (oracle/_ref, built from
the reference sources; see make_golden.py) on seeded inputs. The oracle must
reproduce every recorded state, hull and contact bit for bit.
"""
import numpy as np
import pytest

from _libs import ROOT, SimWorld, scenes, state_bits_equal
from golden.make_golden import CASES, CONTACT_DROP, STATE_KEYS, input_digest

GOLDEN = ROOT / "tests" / "golden"


def _load(name):
    return np.load(GOLDEN / name, allow_pickle=False)


@pytest.mark.parametrize("case", sorted(CASES))
def test_trajectory_matches_reference(case):
    build, frames, kw = CASES[case]
    world = build()
    fx = _load(f"traj_{case}.npz")
    assert bytes(fx["digest"]).hex() == input_digest(world), "scene builder changed: regenerate fixtures"
    sim = SimWorld("orc", world)
    done = 0
    for f in fx["frames"]:
        sim.step(int(f) - done, **kw)
        done = int(f)
        st = sim.state()
        expected = {k: fx[f"f{f}_{k}"] for k in STATE_KEYS}
        bad = state_bits_equal(st, expected)
        assert not bad, f"{case}: frame {f}: {bad}"
    for s in range(world.num_scenes):
        assert sim.error(s) == 0


def test_hulls_match_reference():
    fx = _load("hulls.npz")
    w = scenes.World(name="hulls")
    cube = w.shape(scenes.cube_corners())
    shapes = {"cube": (cube, (1.0, 1.0, 1.0)), "floor": (cube, (50.0, 1.0, 50.0)), "box": (cube, (0.2, 0.5, 0.1)),
              "ico12": (w.shape(scenes.icosphere_corners(0)), (1.0, 1.0, 1.0)),
              "ico42": (w.shape(scenes.icosphere_corners(1)), (0.75, 0.75, 0.75)),
              "ico162": (w.shape(scenes.icosphere_corners(2)), (0.6, 0.6, 0.6))}
    for shp, scale in shapes.values():
        w.add_body(shp, scale, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0, 1.0), mass=2.0)
    w.end_scene()
    assert bytes(fx["digest"]).hex() == input_digest(w)
    sim = SimWorld("orc", w)
    info = sim.body_info()
    for b, name in enumerate(shapes):
        h = sim.hull(b)
        for k, v in h.items():
            assert np.array_equal(v, fx[f"{name}_{k}"]), (name, k)
        for i, key in enumerate(("inverse_mass", "inertia", "inverse_inertia", "radius")):
            got = np.atleast_1d(info[i][b])
            assert np.array_equal(got.view(np.uint64), fx[f"{name}_{key}"].view(np.uint64)), (name, key)


def test_pairs_and_contacts_match_reference():
    fx = _load("contacts.npz")
    world = scenes.batched_uniform(6, first_seed=300)
    assert bytes(fx["digest"]).hex() == input_digest(world)
    sim = SimWorld("orc", world)
    st = sim.state()
    st["position"][:] = fx["position"]
    st["rotation"][:] = fx["rotation"]
    sim.set_state(st)
    total = 0
    for s in range(world.num_scenes):
        pairs = sim.broad_pairs(s)
        assert np.array_equal(pairs, fx[f"s{s}_pairs"]), s
        got = [sim.contacts(s, int(i), int(j)) for i, j in pairs]
        counts = np.asarray([len(c) for c in got], dtype=np.int64)
        assert np.array_equal(counts, fx[f"s{s}_counts"]), s
        if got:
            allc = np.concatenate(got, axis=0)
            assert np.array_equal(allc.view(np.uint64), fx[f"s{s}_contacts"].view(np.uint64)), s
        total += int(counts.sum())
    assert total > 20  # the poses include resting and stacked contacts
