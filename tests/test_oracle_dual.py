"""The oracle's dual-number build (oracle/libdual.so): forward-mode
derivatives of the step, the checker of the device adjoint.

* Its value parts equal the double oracle bit for bit.
* With the contact set frozen, its derivatives match FP64 central finite
  differences (h = 1e-6) of the same frozen map: the perturbed runs replay
  the base run's logged contacts (orc_world_contact_log). Horizons are kept
  short enough that no branch of the piecewise-smooth step (penetration test,
  static friction cone, dynamic friction min) flips within +-h.
* A body that never touches anything has a closed-form derivative:
  d x_final / d v0 = (frames x dt) I.
"""
import copy
import ctypes

import numpy as np
import pytest

from _libs import ROOT, SimWorld, _bind, load, scenes

P = ctypes.c_void_p
DUAL_SO = ROOT / "oracle" / "libdual.so"
PARAMS = [("vx", 0, 7), ("vy", 0, 8), ("vz", 0, 9), ("wx", 0, 10), ("wz", 0, 12),
          ("mass", 1, 0), ("mu_s", 2, 0), ("mu_d", 3, 0)]


def _libs():
    if not DUAL_SO.exists():
        pytest.skip("oracle/libdual.so not built")
    orc = load("orc")
    orc.orc_world_contact_log.argtypes = [P, ctypes.c_int]
    orc.orc_world_contact_log_copy.argtypes = [P, P]
    orc.orc_world_contact_log_error.argtypes = [P]
    dual = _bind(ctypes.CDLL(str(DUAL_SO)), "orc_")
    dual.orc_dual_seed.argtypes = [P, ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    dual.orc_dual_state.argtypes = [P, ctypes.c_uint32, ctypes.c_int, P]
    dual.orc_dual_set_freeze.argtypes = [ctypes.c_int]
    return orc, dual


def _perturbed(world, body, name, delta):
    w2 = copy.deepcopy(world)
    if name[0] in "vw":
        arr = w2.body_linear_velocity if name[0] == "v" else w2.body_angular_velocity
        t = list(arr[body])
        t["xyz".index(name[1])] += delta
        arr[body] = tuple(t)
    elif name == "mass":
        w2.body_mass[body] += delta
    else:
        m = list(w2.body_material[body])
        m[0 if name == "mu_s" else 1] += delta
        w2.body_material[body] = tuple(m)
    return w2


def dual_gradients(dual, world, targets, frames, freeze=True):
    """dL/dparam for L_s = |x_cube(s) - target_s|^2, per scene and parameter."""
    dual.orc_dual_set_freeze(1 if freeze else 0)
    d = SimWorld("orc", world, lib=dual)
    for slot, (_, kind, comp) in enumerate(PARAMS):
        for s in range(len(targets)):
            assert dual.orc_dual_seed(d.h, 2 * s + 1, slot, kind, comp) == 0
    d.step(frames)
    x = d.state()["position"]
    g = np.zeros((len(targets), len(PARAMS)))
    for slot in range(len(PARAMS)):
        for s in range(len(targets)):
            out = np.zeros(13)
            dual.orc_dual_state(d.h, 2 * s + 1, slot, out.ctypes.data)
            g[s, slot] = 2.0 * ((x[2 * s + 1] - targets[s]) * out[:3]).sum()
    return g, d.state()


def test_dual_values_equal_double_oracle():
    orc, dual = _libs()
    world, _ = scenes.ramps(4, first_seed=10)
    a = SimWorld("orc", world)
    b = SimWorld("orc", world, lib=dual)
    a.step(40)
    b.step(40)
    sa, sb = a.state(), b.state()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k


def test_frozen_dual_gradients_match_central_differences():
    orc, dual = _libs()
    frames, h = 30, 1e-6
    world, targets = scenes.ramps(3, first_seed=0)
    g, _ = dual_gradients(dual, world, targets, frames, freeze=True)
    base = SimWorld("orc", world)
    orc.orc_world_contact_log(base.h, 1)
    base.step(frames)
    checked = 0
    for s in range(len(targets)):
        for slot, (name, _, _) in enumerate(PARAMS):
            L = []
            for sign in (1.0, -1.0):
                sim = SimWorld("orc", _perturbed(world, 2 * s + 1, name, sign * h))
                orc.orc_world_contact_log_copy(sim.h, base.h)
                orc.orc_world_contact_log(sim.h, 2)
                sim.step(frames)
                assert orc.orc_world_contact_log_error(sim.h) == 0
                x = sim.state()["position"][2 * s + 1]
                L.append(float(((x - targets[s]) ** 2).sum()))
            fd = (L[0] - L[1]) / (2 * h)
            scale = max(abs(fd), abs(g[s, slot]))
            if scale < 1e-9:
                continue
            assert abs(g[s, slot] - fd) <= 2e-6 * scale, (s, name, g[s, slot], fd)
            checked += 1
    assert checked >= 12


def test_free_flight_derivative_closed_form():
    """A cube far from the ramp (its velocity pointing away) flies freely:
    x_final = x0 + sum of substep displacements, so dx/dv0 = T I exactly
    up to rounding of the sum of h's."""
    orc, dual = _libs()
    world, _ = scenes.ramps(1, first_seed=3)
    world.body_position[1] = (0.0, 40.0, 0.0)
    world.body_linear_velocity[1] = (0.5, 3.0, -0.5)
    d = SimWorld("orc", world, lib=dual)
    dual.orc_dual_set_freeze(1)
    for comp, slot in ((7, 0), (8, 1), (9, 2)):
        dual.orc_dual_seed(d.h, 1, slot, 0, comp)
    frames = 30
    d.step(frames)
    T = frames / 60.0
    for slot in range(3):
        out = np.zeros(13)
        dual.orc_dual_state(d.h, 1, slot, out.ctypes.data)
        expect = np.zeros(3)
        expect[slot] = T
        assert np.allclose(out[:3], expect, rtol=0, atol=1e-12), (slot, out[:3])
        assert np.allclose(out[7:10], np.eye(3)[slot], rtol=0, atol=1e-12)
