"""raw-physics on MI355X: the rigid-body XPBD step of felipeek/raw-physics as
hand-written HIP kernels for gfx950, behind the C ABI in include/rawphys.h.

Python is a thin ctypes layer over librawphys.so (built in-tree by build.py).
There is no CPU fallback: creating a World without the built library, or
without a HIP device, raises.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

from . import scenes

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "librawphys.so"
_P = ctypes.c_void_p
_lib = None


def library():
    """Loads librawphys.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python raw-physics_amd/build.py`")
    lib = ctypes.CDLL(str(LIB_PATH))
    u32, u64, f64, i32 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double, ctypes.c_int32
    sig = {
        "rp_world_create": (_P, [ctypes.POINTER(scenes.WorldDescC)]),
        "rp_world_destroy": (None, [_P]),
        "rp_world_num_bodies": (u32, [_P]),
        "rp_world_num_scenes": (u32, [_P]),
        "rp_world_step": (u32, [_P, f64, u32, u32, u32, i32, f64]),
        "rp_world_step_async": (u32, [_P, f64, u32, u32, u32, i32, f64]),
        "rp_world_synchronize": (u32, [_P]),
        "rp_world_step_noalloc": (u32, [_P, f64, u32, u32, u32, i32, f64]),
        "rp_world_get_state": (None, [_P, ctypes.POINTER(scenes.StateC)]),
        "rp_world_set_state": (None, [_P, ctypes.POINTER(scenes.StateC)]),
        "rp_world_error": (u32, [_P, u32]),
        "rp_world_errors": (u32, [_P, _P]),
        "rp_world_epa_failures": (u64, [_P]),
        "rp_world_gjk_cycle_exits": (u64, [_P]),
        "rp_world_stream": (_P, [_P]),
        "rp_world_model_matrices": (u32, [_P, _P, i32, i32]),
        "rp_obj_parse": (u32, [ctypes.c_char_p, _P, u32, ctypes.POINTER(u32)]),
        "rp_obj_parse_memory": (u32, [ctypes.c_char_p, ctypes.c_size_t, _P, u32, ctypes.POINTER(u32)]),
        "rp_world_body_info": (None, [_P, _P, _P, _P, _P]),
        "rp_world_broad_pairs": (u32, [_P, u32, _P, u32]),
        "rp_world_contacts": (u32, [_P, u32, u32, u32, _P, u32]),
        "rp_world_item_classes": (u32, [_P, u32, _P, u32]),
        "rp_world_hull": (None, [_P, u32, _P] + [_P] * 10),
        "rp_last_error": (u32, []),
        "rp_last_error_string": (ctypes.c_char_p, []),
        "rp_pbd_simulate": (u32, [f64, _P, _P, u32, u32, i32]),
        "physics_simulate": (None, [_P, f64]),
        "rp_fp64_selftest": (u32, [_P, _P, _P, u64]),
        "rp_world_set_profiling": (None, [_P, i32]),
        "rp_world_set_body_params": (u32, [_P, _P, _P, _P]),
        "rp_world_adjoint_forward": (u32, [_P, f64, u32, u32, u32, f64]),
        "rp_world_adjoint_backward": (u32, [_P, _P, _P, _P]),
        "rp_world_adjoint_mismatches": (u64, [_P]),
        "rp_world_adjoint_amplification": (u32, [_P, _P]),
        "rp_world_set_narrow_group": (i32, [_P, i32]),
        "rp_world_narrow_group": (i32, [_P]),
        "rp_world_set_manifold_group": (i32, [_P, i32]),
        "rp_world_manifold_group": (i32, [_P]),
        "rp_world_set_narrow_block_cap": (i32, [_P, i32]),
        "rp_world_lds_limit": (i32, [_P]),
        "rp_world_set_solve_lanes": (i32, [_P, i32]),
        "rp_world_set_colour_launches": (i32, [_P, i32]),
        "rp_world_solve_lanes": (i32, [_P]),
        "rp_world_set_graphs": (None, [_P, i32]),
        "rp_world_set_views": (i32, [_P, i32]),
        "rp_world_views": (i32, [_P]),
        "rp_world_kernel_times": (None, [_P, _P, _P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def last_error_string() -> str:
    return (library().rp_last_error_string() or b"").decode()


def load_obj(path=None, text=None) -> np.ndarray:
    """obj_parse (src/render/obj.cpp:7-151) of an OBJ file or of OBJ text: one
    record per triangle corner with fields position (3 floats), normal (3)
    and texture_coordinates (2). `rec["position"]` is the corner stream a
    World shape takes."""
    lib = library()
    dt = np.dtype([("position", np.float32, 3), ("normal", np.float32, 3), ("texture_coordinates", np.float32, 2)])
    n = ctypes.c_uint32()

    def call(buf, cap):
        if text is not None:
            raw = text.encode() if isinstance(text, str) else bytes(text)
            return lib.rp_obj_parse_memory(raw, len(raw), buf, cap, ctypes.byref(n))
        return lib.rp_obj_parse(str(path).encode(), buf, cap, ctypes.byref(n))

    if (path is None) == (text is None):
        raise ValueError("give exactly one of path and text")
    err = call(None, 0)
    if err:
        raise RuntimeError(f"obj_parse failed ({err:#x}): {last_error_string()}")
    out = np.zeros(n.value, dtype=dt)
    if n.value:
        err = call(out.ctypes.data, n.value)
        if err:
            raise RuntimeError(f"obj_parse failed ({err:#x}): {last_error_string()}")
    return out


class World:
    """A batch of scenes resident on the GPU (rp_world_*)."""

    def __init__(self, world: "scenes.World"):
        lib = library()
        self.lib = lib
        self.world = world
        self.desc, self._keep = world.desc()
        self.h = lib.rp_world_create(ctypes.byref(self.desc))
        if not self.h:
            raise RuntimeError("rp_world_create failed: " + last_error_string())
        self.nb = world.num_bodies
        self.ns = world.num_scenes

    def close(self):
        if getattr(self, "h", None):
            self.lib.rp_world_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, frames=1, dt=1.0 / 60.0, substeps=20, iters=1, collisions=True, gravity=10.0, sync=True):
        f = self.lib.rp_world_step if sync else self.lib.rp_world_step_async
        rc = f(self.h, dt, frames, substeps, iters, 1 if collisions else 0, gravity)
        if rc:
            raise RuntimeError("rp_world_step failed: " + last_error_string())

    def synchronize(self):
        if self.lib.rp_world_synchronize(self.h):
            raise RuntimeError("rp_world_synchronize failed: " + last_error_string())

    def state(self) -> dict:
        st = scenes.alloc_state(self.nb)
        s = scenes.state_struct(st)
        self.lib.rp_world_get_state(self.h, ctypes.byref(s))
        return st

    def set_state(self, st: dict):
        s = scenes.state_struct(st)
        self.lib.rp_world_set_state(self.h, ctypes.byref(s))

    def errors(self) -> np.ndarray:
        """Every scene's error word."""
        out = np.zeros(self.ns, dtype=np.uint32)
        if self.ns and self.lib.rp_world_errors(self.h, out.ctypes.data) != 0:
            raise RuntimeError("rp_world_errors failed: " + last_error_string())
        return out

    def epa_failures(self) -> int:
        return int(self.lib.rp_world_epa_failures(self.h))

    # -- gradients ------------------------------------------------------------
    def set_body_params(self, mass=None, mu_s=None, mu_d=None) -> int:
        """Per-body mass and friction (arrays of nb; None keeps the current)."""
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (mass, mu_s, mu_d)]
        ptrs = [None if a is None else a.ctypes.data for a in arrs]
        return int(self.lib.rp_world_set_body_params(self.h, *ptrs))

    def adjoint_forward(self, frames: int, dt=1.0 / 60.0, substeps=20, iters=1, gravity=10.0) -> int:
        """Steps `frames` frames keeping per-frame checkpoints for adjoint_backward."""
        return int(self.lib.rp_world_adjoint_forward(self.h, dt, frames, substeps, iters, gravity))

    def adjoint_backward(self, dl_dstate: np.ndarray):
        """dl_dstate (nb, 13): dL/d(x, q, v, w) at the end of the recorded run.
        Returns (grad_state (nb, 13) at its start, grad_param (nb, 3) for mass,
        static and dynamic friction)."""
        dl = np.ascontiguousarray(dl_dstate, dtype=np.float64).reshape(self.nb, 13)
        gs = np.zeros((self.nb, 13))
        gp = np.zeros((self.nb, 3))
        err = self.lib.rp_world_adjoint_backward(self.h, dl.ctypes.data, gs.ctypes.data, gp.ctypes.data)
        if err:
            raise RuntimeError(f"adjoint_backward failed ({err:#x}): {last_error_string()}")
        return gs, gp

    def adjoint_amplification(self) -> np.ndarray:
        """Per body: the backward pass's largest state adjoint at a frame
        start over the largest component of dl_dstate (gradient
        amplification; see rawphys.h)."""
        out = np.zeros(self.nb)
        err = self.lib.rp_world_adjoint_amplification(self.h, out.ctypes.data)
        if err:
            raise RuntimeError(f"adjoint_amplification failed ({err:#x}): {last_error_string()}")
        return out

    def adjoint_mismatches(self) -> int:
        return int(self.lib.rp_world_adjoint_mismatches(self.h))

    def item_classes(self, scene: int) -> np.ndarray:
        """Solve classes of a large scene's items in the last frame (colours
        0..62, 63 = the serial class; empty for a scene that is not large)."""
        n = int(self.lib.rp_world_item_classes(self.h, scene, None, 0))
        out = np.zeros(max(n, 1), dtype=np.int32)
        self.lib.rp_world_item_classes(self.h, scene, out.ctypes.data, n)
        return out[:n]

    def gjk_cycle_exits(self) -> int:
        return int(self.lib.rp_world_gjk_cycle_exits(self.h))

    def model_matrices(self, as_float: bool = False) -> np.ndarray:
        """entity_get_model_matrix of every body: (nb, 4, 4), float64 or float32."""
        out = np.zeros((self.nb, 4, 4), dtype=np.float32 if as_float else np.float64)
        err = self.lib.rp_world_model_matrices(self.h, out.ctypes.data, 1 if as_float else 0, 0)
        if err:
            raise RuntimeError(f"model_matrices failed ({err:#x}): {last_error_string()}")
        return out

    KERNEL_CLASSES = ("frame", "integrate", "transform", "narrow", "solve", "manifold")

    def set_narrow_group(self, lanes_per_pair: int):
        """GJK + EPA schedule: 1, 4, 8 or 16 lanes per body pair (same results)."""
        if self.lib.rp_world_set_narrow_group(self.h, lanes_per_pair) != 0:
            raise ValueError(f"narrow-phase schedule {lanes_per_pair} does not fit this world")

    def set_manifold_group(self, lanes_per_pair: int):
        """Contact-manifold schedule: 1, 4, 8 or 16 lanes per body pair in a
        kernel of its own, or 0 = built by the GJK + EPA kernel on the hulls it
        staged (same results)."""
        if self.lib.rp_world_set_manifold_group(self.h, lanes_per_pair) != 0:
            raise ValueError(f"manifold schedule {lanes_per_pair} does not fit this world")

    # (the one-lane GJK + EPA kernel, schedule (1, 1), reserves 1.8 GiB of
    # private memory for its EPA arrays and was never the fastest: it stays
    # available through set_narrow_group but is not tried)
    NARROW_SCHEDULES = ((8, 0), (16, 0), (4, 0), (8, 1), (16, 1), (4, 1), (8, 8))

    def autotune(self, frames: int = 6, schedules=NARROW_SCHEDULES, solve=(64, 32, 16), **step_kw) -> dict:
        """Picks the narrow-phase schedule (GJK + EPA lanes per pair, manifold
        lanes per pair or 0 for the fused kernel) and the solve schedule from
        per-kernel times: each trial runs `frames` frames from the current
        state in profiling mode, and the state is restored after each trial.
        A schedule is charged its GJK + EPA, manifold and vertex-transform
        kernels. All schedules give identical results.
        Returns {"narrow": {"g/m": ms}, "solve": {lanes: ms}}."""
        start = self.state()
        ok = []
        for g, m in schedules:
            if self.lib.rp_world_set_narrow_group(self.h, g) == 0 and self.lib.rp_world_set_manifold_group(self.h, m) == 0:
                ok.append((g, m))
        solve = [n for n in solve if self.lib.rp_world_set_solve_lanes(self.h, n) == 0]
        out = {"narrow": {}, "solve": {}}
        for t in range(max(len(ok), len(solve))):
            (g, m), n = ok[t % len(ok)], solve[t % len(solve)]
            self.set_narrow_group(g)
            self.set_manifold_group(m)
            self.set_solve_lanes(n)
            self.step(1, **step_kw)  # first launches of these variants
            self.set_state(start)
            self.set_profiling(True)
            self.step(frames, **step_kw)
            kt = self.kernel_times()
            self.set_profiling(False)
            self.set_state(start)
            out["narrow"].setdefault(f"{g}/{m}", kt["narrow"][0] + kt["manifold"][0] + kt["transform"][0])
            out["solve"].setdefault(n, kt["solve"][0])
        g, m = map(int, min(out["narrow"], key=out["narrow"].get).split("/"))
        self.set_narrow_group(g)
        self.set_manifold_group(m)
        self.set_solve_lanes(min(out["solve"], key=out["solve"].get))
        return out

    def set_views(self, views: int):
        """Concurrent scene-range views for graph replay: 0 = automatic, 1..4."""
        if self.lib.rp_world_set_views(self.h, views) != 0:
            raise ValueError(f"views must be 0..4, not {views}")

    @property
    def views(self) -> int:
        return int(self.lib.rp_world_views(self.h))

    def set_graphs(self, on: bool):
        """Replay each frame as a captured HIP graph (default) or launch eagerly."""
        self.lib.rp_world_set_graphs(self.h, 1 if on else 0)

    def set_colour_launches(self, classes: int):
        """Large scenes: colour classes solved one launch each before the
        one-block tail (0..25, default 6; identical results)."""
        if self.lib.rp_world_set_colour_launches(self.h, classes) != 0:
            raise ValueError(f"colour launches must be 0..25, not {classes}")

    def set_solve_lanes(self, lanes_per_scene: int):
        if self.lib.rp_world_set_solve_lanes(self.h, lanes_per_scene) != 0:
            raise ValueError(f"solve schedule {lanes_per_scene} lanes per scene is not available")

    @property
    def solve_lanes(self) -> int:
        return int(self.lib.rp_world_solve_lanes(self.h))

    @property
    def lds_limit(self) -> int:
        """Dynamic LDS bytes one narrow-phase block may use on the device."""
        return int(self.lib.rp_world_lds_limit(self.h))

    @property
    def narrow_group(self) -> int:
        return int(self.lib.rp_world_narrow_group(self.h))

    @property
    def manifold_group(self) -> int:
        return int(self.lib.rp_world_manifold_group(self.h))

    def set_profiling(self, on: bool):
        self.lib.rp_world_set_profiling(self.h, 1 if on else 0)

    def kernel_times(self) -> dict:
        """{class: (total ms, launches)} recorded in profiling mode."""
        ms = np.zeros(8)
        n = np.zeros(8, dtype=np.uint64)
        self.lib.rp_world_kernel_times(self.h, ms.ctypes.data, n.ctypes.data)
        return {c: (float(ms[i]), int(n[i])) for i, c in enumerate(self.KERNEL_CLASSES)}

    @property
    def stream(self) -> int:
        return int(self.lib.rp_world_stream(self.h) or 0)
