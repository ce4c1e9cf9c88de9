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


def _run_pair(world, frames, **kw):
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
    return a.state(), b.state()


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
