"""Gradient checks of the differentiable workload (ramp scenes: a cube thrown
onto a tilted static box) against FP64 central finite differences, all on
the device: the adjoint of a recorded run (rp_world_adjoint_*) beside
perturbed forward runs of the same step, batched as extra scenes.

The step is piecewise smooth: contacts appear and vanish, static friction
holds or slips (pbd.cpp:107-154), dynamic friction saturates (pbd.cpp:
646-711). The adjoint differentiates the run's own pieces with the contact
set frozen. Central differences measure the derivative only where no branch
flips within +-h; they are called reproducible where h and h/10 agree.
Where the differences are reproducible, the adjoint still differs from them
by the frozen-contact approximation: the contact points and normals of a
perturbed run move with the state, the frozen ones do not (about 1e-6
relative at 30 frames, up to a few 1e-6 on some entries).
Each scene also carries the adjoint's gradient amplification
(rp_world_adjoint_amplification): impacts of a tumbling cube multiply the
state adjoint by 1e2-1e4, after which no perturbation finite differences can
resolve stays on one piece.
"""
from __future__ import annotations

import numpy as np

from . import scenes

PARAMS = ("vx", "vy", "vz", "wx", "wz", "mass", "mu_s", "mu_d")
AMPLIFICATION_VALID = 10.0


def adjoint_ramps(pkg, world, targets, frames: int, params=PARAMS):
    """Adjoint gradients of L_s = |x_cube(s) - target_s|^2 after `frames`
    frames: (grads [scenes, params], amplification [scenes], recompute
    mismatches, scenes with error words)."""
    gw = pkg.World(world)
    try:
        if gw.adjoint_forward(frames) != 0:
            raise RuntimeError("adjoint_forward failed")
        st = gw.state()
        n = len(targets)
        dyn = 2 * np.arange(n) + 1
        dl = np.zeros((world.num_bodies, 13))
        dl[dyn, 0:3] = 2.0 * (st["position"][dyn] - targets)
        gs, gp = gw.adjoint_backward(dl)
        g = np.zeros((n, len(params)))
        for k, name in enumerate(params):
            kind, col = scenes.RAMP_PARAMS[name]
            g[:, k] = gs[dyn, col] if kind == "state" else gp[dyn, col]
        return g, gw.adjoint_amplification()[dyn], gw.adjoint_mismatches(), int((gw.errors() != 0).sum())
    finally:
        gw.close()


def fd_ramps(pkg, num_scenes: int, first_seed: int, frames: int, h: float, params=PARAMS) -> np.ndarray:
    """Central differences (L(p + h) - L(p - h)) / 2h of every scene and
    parameter, from perturbed copies stepped on the device."""
    w, t = scenes.ramps_perturbed(num_scenes, first_seed, params, h)
    gw = pkg.World(w)
    try:
        gw.step(frames)
        x = gw.state()["position"][1::2]
    finally:
        gw.close()
    L = ((x - t) ** 2).sum(axis=1).reshape(num_scenes, len(params), 2)
    return (L[..., 0] - L[..., 1]) / (2.0 * h)


def gradient_check(pkg, num_scenes: int, first_seed: int, frames: int, h: float = 1e-6, rtol: float = 1e-6,
                   params=PARAMS) -> dict:
    """Adjoint against central differences at h (and h/10 for
    reproducibility) on num_scenes ramp scenes. An entry is compared where
    the differences are reproducible (|fd(h) - fd(h/10)| <= rtol * scale,
    scale = max(|adjoint|, |fd(h)|, 1e-6 x the scene's largest adjoint
    entry)); it agrees where |adjoint - fd(h)| <= rtol * scale. A scene is
    valid where its amplification is at most AMPLIFICATION_VALID."""
    world, targets = scenes.ramps(num_scenes, first_seed)
    g, amp, mism, errs = adjoint_ramps(pkg, world, targets, frames, params)
    f1 = fd_ramps(pkg, num_scenes, first_seed, frames, h, params)
    f2 = fd_ramps(pkg, num_scenes, first_seed, frames, h / 10.0, params)
    scale = np.maximum(np.maximum(np.abs(g), np.abs(f1)), 1e-6 * np.abs(g).max(axis=1, keepdims=True))
    scale = np.maximum(scale, 1e-300)
    nonzero = scale > 1e-12
    stable = (np.abs(f1 - f2) <= rtol * scale) & nonzero
    agree = np.abs(g - f1) <= rtol * scale
    valid = amp <= AMPLIFICATION_VALID
    norms = np.abs(g).max(axis=1)
    vs = stable & valid[:, None]
    rel = np.abs(g - f1) / scale
    return {
        "scenes": num_scenes, "frames": frames, "h": h, "rtol": rtol,
        "grads": g, "fd": f1, "fd_h10": f2, "amplification": amp, "stable": stable, "agree": agree,
        "valid": valid, "rel": rel, "recompute_mismatches": mism, "error_scenes": errs,
        "summary": {
            "valid_scenes": int(valid.sum()),
            "fd_reproducible_entries": int(stable.sum()), "entries": int(stable.size),
            "agree_on_reproducible": int((agree & stable).sum()),
            "valid_reproducible_entries": int(vs.sum()),
            "agree_on_valid_reproducible": int((agree & vs).sum()),
            "agree_1e-5_on_valid_reproducible": int(((rel <= 1e-5) & vs).sum()),
            "max_rel_on_valid_reproducible": float(rel[vs].max()) if vs.any() else None,
            "grad_max_per_scene_percentiles": {p: float(np.percentile(norms, p)) for p in (50, 90, 99, 100)},
            "amplification_percentiles": {p: float(np.percentile(amp, p)) for p in (50, 90, 99, 100)},
        },
    }
