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
    _pair(scenes.batched_uniform(24), frames=120, check_every=20)


def test_batched_ico162_bit_exact():
    _pair(scenes.batched_uniform(3, ico_subdiv=2), frames=60, check_every=20)
