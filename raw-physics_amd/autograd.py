"""torch.autograd binding of the device step's adjoint (over the C ABI).

`ThrowRun.apply(v0, w0, mass, mu_s, mu_d, run)` steps a world whose scenes
each hold one dynamic body and returns that body's final position per scene;
its backward calls rp_world_adjoint_backward. Inputs are float64 tensors of
shape (S, 3) or (S,) on any device; gradients come back on the same device.
The gradient is that of the step with the forward pass's contacts frozen.
"""
from __future__ import annotations

import numpy as np
import torch

from . import World


class ThrowSetup:
    """A world of S scenes with one dynamic body each, its initial state and
    the run length. `dynamic` holds the dynamic body of every scene."""

    def __init__(self, world_desc, frames: int, dt=1.0 / 60.0, substeps=20, iters=1, gravity=10.0):
        self.gw = World(world_desc)
        self.frames, self.dt, self.substeps, self.iters, self.gravity = frames, dt, substeps, iters, gravity
        self.init = self.gw.state()
        fixed = np.asarray(world_desc.body_fixed)
        off = np.asarray(world_desc.scene_body_offset)
        self.dynamic = np.array([next(b for b in range(off[s], off[s + 1]) if not fixed[b])
                                 for s in range(len(off) - 1)])
        self.mass0 = np.asarray(world_desc.body_mass, dtype=np.float64).copy()
        self.material0 = np.asarray(world_desc.body_material, dtype=np.float64).reshape(-1, 3).copy()


class ThrowRun(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v0, w0, mass, mu_s, mu_d, run: ThrowSetup):
        dyn = run.dynamic
        st = {k: v.copy() for k, v in run.init.items()}
        st["linear_velocity"][dyn] = v0.detach().cpu().double().numpy()
        st["angular_velocity"][dyn] = w0.detach().cpu().double().numpy()
        m = run.mass0.copy()
        ms = run.material0[:, 0].copy()
        md = run.material0[:, 1].copy()
        m[dyn] = mass.detach().cpu().double().numpy()
        ms[dyn] = mu_s.detach().cpu().double().numpy()
        md[dyn] = mu_d.detach().cpu().double().numpy()
        gw = run.gw
        gw.set_state(st)
        if gw.set_body_params(m, ms, md):
            raise RuntimeError("set_body_params failed")
        err = gw.adjoint_forward(run.frames, run.dt, run.substeps, run.iters, run.gravity)
        if err:
            raise RuntimeError(f"adjoint_forward failed ({err:#x})")
        x = gw.state()["position"][dyn]
        ctx.run = run
        ctx.device = v0.device
        return torch.from_numpy(x).to(v0.device)

    @staticmethod
    def backward(ctx, gx):
        run = ctx.run
        dyn = run.dynamic
        dl = np.zeros((run.gw.nb, 13))
        dl[dyn, 0:3] = gx.detach().cpu().double().numpy()
        gs, gp = run.gw.adjoint_backward(dl)
        dev = ctx.device
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        return t(gs[dyn, 7:10]), t(gs[dyn, 10:13]), t(gp[dyn, 0]), t(gp[dyn, 1]), t(gp[dyn, 2]), None
