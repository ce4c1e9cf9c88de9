#!/usr/bin/env python3
"""Writes the golden vectors in this directory from the reference's own step.

The outputs come from oracle/_ref/libref.so: the reference sources compiled
in place (oracle/Makefile) and driven through its caller pattern by
oracle/ref_harness.cpp. Inputs are the seeded scenes of
raw-physics_amd/scenes.py; each fixture stores a digest of its inputs so that
a change to the scene builder is caught instead of silently compared.

Run from the repository root where the reference build exists:
    make -C oracle && python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
from _libs import SimWorld, load, scenes  # noqa: E402

STATE_KEYS = ("position", "rotation", "linear_velocity", "angular_velocity", "previous_position",
              "previous_rotation", "previous_linear_velocity", "previous_angular_velocity", "active",
              "deactivation_time")


def input_digest(world: scenes.World) -> str:
    a = world.arrays()
    h = hashlib.sha256()
    for k in sorted(a):
        if k == "constraints":
            h.update(bytes(a[k]))
        else:
            h.update(np.ascontiguousarray(a[k]).tobytes())
    return h.hexdigest()


# name -> (world builder, frames at which the state is recorded, step kwargs)
CASES = {
    "stack": (lambda: scenes.stack_scene(seed=0), [1, 2, 10, 60, 300], {}),
    "batched": (lambda: scenes.batched_uniform(4), [1, 30, 120, 600], {}),
    "ico162": (lambda: scenes.batched_uniform(2, first_seed=100, ico_subdiv=2), [1, 60, 200], {}),
    "chains_nocoll": (lambda: scenes.chains(2, length=6), [1, 60, 300], {"collisions": False}),
    "chains": (lambda: scenes.chains(2, length=6, first_seed=10), [1, 60, 300], {}),
    "iters4": (lambda: scenes.batched_uniform(2, first_seed=200), [1, 60], {"iters": 4, "substeps": 10}),
    "mixed": (lambda: scenes.batched_mixed(4, first_seed=1), [1, 30, 120, 300], {}),
}


def trajectory(name):
    build, frames, kw = CASES[name]
    world = build()
    sim = SimWorld("ref", world)
    out = {"digest": np.frombuffer(bytes.fromhex(input_digest(world)), dtype=np.uint8),
           "frames": np.asarray(frames, dtype=np.int64)}
    done = 0
    for f in frames:
        sim.step(f - done, **kw)
        done = f
        st = sim.state()
        for k in STATE_KEYS:
            out[f"f{f}_{k}"] = st[k]
    np.savez_compressed(HERE / f"traj_{name}.npz", **out)
    print(f"traj_{name}.npz: {world.num_bodies} bodies, frames {frames}")


def hulls():
    w = scenes.World(name="hulls")
    cube = w.shape(scenes.cube_corners())
    shapes = {"cube": (cube, (1.0, 1.0, 1.0)), "floor": (cube, (50.0, 1.0, 50.0)),
              "box": (cube, (0.2, 0.5, 0.1)),
              "ico12": (w.shape(scenes.icosphere_corners(0)), (1.0, 1.0, 1.0)),
              "ico42": (w.shape(scenes.icosphere_corners(1)), (0.75, 0.75, 0.75)),
              "ico162": (w.shape(scenes.icosphere_corners(2)), (0.6, 0.6, 0.6))}
    for name, (shp, scale) in shapes.items():
        w.add_body(shp, scale, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0, 1.0), mass=2.0)
    w.end_scene()
    sim = SimWorld("ref", w)
    out = {"digest": np.frombuffer(bytes.fromhex(input_digest(w)), dtype=np.uint8)}
    inv_mass, inertia, inv_inertia, radius = sim.body_info()
    for b, name in enumerate(shapes):
        for k, v in sim.hull(b).items():
            out[f"{name}_{k}"] = v
        out[f"{name}_inverse_mass"] = inv_mass[b:b + 1]
        out[f"{name}_inertia"] = inertia[b]
        out[f"{name}_inverse_inertia"] = inv_inertia[b]
        out[f"{name}_radius"] = radius[b:b + 1]
    np.savez_compressed(HERE / "hulls.npz", **out)
    print("hulls.npz:", list(shapes))


CONTACT_FRAMES = 90
CONTACT_DROP = 0.01


def contacts():
    """Broad-phase pairs and the contacts of every pair of 6 scenes, at the
    poses reached after CONTACT_FRAMES frames (landed and stacked bodies)
    lowered by CONTACT_DROP."""
    world = scenes.batched_uniform(6, first_seed=300)
    sim = SimWorld("ref", world)
    sim.step(CONTACT_FRAMES)
    st = sim.state()
    # end-of-frame poses have been projected out of contact; lower the dynamic
    # bodies by CONTACT_DROP so that resting and stacked pairs overlap
    dyn = np.asarray(world.body_fixed) == 0
    st["position"][dyn, 1] -= CONTACT_DROP
    sim.set_state(st)
    out = {"digest": np.frombuffer(bytes.fromhex(input_digest(world)), dtype=np.uint8),
           "position": st["position"], "rotation": st["rotation"]}
    for s in range(world.num_scenes):
        pairs = sim.broad_pairs(s)
        out[f"s{s}_pairs"] = pairs
        counts, pts = [], []
        for i, j in pairs:
            c = sim.contacts(s, int(i), int(j))
            counts.append(len(c))
            pts.append(c)
        out[f"s{s}_counts"] = np.asarray(counts, dtype=np.int64)
        out[f"s{s}_contacts"] = np.concatenate(pts, axis=0) if pts else np.zeros((0, 9))
    np.savez_compressed(HERE / "contacts.npz", **out)
    print("contacts.npz: 6 scenes")


if __name__ == "__main__":
    if load("ref") is None:
        sys.exit("oracle/_ref/libref.so is not built (needs the reference sources): make -C oracle")
    hulls()
    contacts()
    for name in CASES:
        trajectory(name)
