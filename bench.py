#!/usr/bin/env python3
"""Throughput of the raw-physics step on one MI355X.

Workload (batched uniform): 4096 independent scenes, each a static 50x1x50
floor plus 16 dynamic bodies (8 cubes with 8-vertex hulls, 8 icospheres with
42-vertex hulls, size U(0.5, 1.0), mass = hull volume), dropped from a 4x4
grid, dt = 1/60, 20 substeps, 1 position iteration, FP64, seeded synthetic
inputs, 600 frames. One "step" is one frame of all 4096 scenes (20 substeps
each). The batch is first stepped untimed to --start-frame (300), then
--steps frames are timed from there: a window deep in the run, with falling,
resting and sleeping scenes, rather than the first frames' free fall. The
whole 600-frame run from frame 0 is timed as well (`whole_run`).

Metric: body-substeps/s = dynamic bodies x substeps x frames / second, for
  * the batched method (all scenes resident in HBM, one launch sequence per
    frame for the whole batch)            -> `value`
  * the plain method (one scene per step call, state uploaded and written
    back every frame, as the drop-in pbd_simulate does)  -> `plain`
  * the reference CPU step on the host    -> `cpu_baseline`

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes
import importlib.util
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
METRIC = "body-substeps/s, 4096 scenes×16 bodies, batched vs plain; adjoint steps/s"


def _package():
    if "raw_physics_amd" in sys.modules:
        return sys.modules["raw_physics_amd"]
    pkg_dir = ROOT / "raw-physics_amd"
    spec = importlib.util.spec_from_file_location("raw_physics_amd", pkg_dir / "__init__.py",
                                                  submodule_search_locations=[str(pkg_dir)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["raw_physics_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _scene_slice(world, st: dict, first: int, count: int) -> dict:
    """The state rows of scenes [first, first + count) of a world."""
    off = world.scene_body_offset
    a, b = off[first], off[first + count]
    return {k: np.ascontiguousarray(v[a:b]) for k, v in st.items()}


def cpu_worker(args):
    """One process of the all-cores CPU leg: builds its scenes in the given
    reference build, loads their window-start state, reports ready, waits for
    the start line on stdin, steps its frames and prints the elapsed time."""
    sys.path.insert(0, str(ROOT / "tests"))
    from _libs import SimWorld  # test infrastructure: the CPU reference only

    sc = _package().scenes
    with np.load(args.cpu_worker, allow_pickle=False) as z:
        st = {k: z[k] for k in z.files}
    world = sc.batched_uniform(args.worker_scenes, first_seed=args.worker_first)
    sim = SimWorld(args.worker_kind, world)
    sim.set_state(st)
    print("ready", flush=True)
    sys.stdin.readline()
    t0 = time.perf_counter()
    sim.step(args.steps)
    print(json.dumps({"elapsed": time.perf_counter() - t0}), flush=True)


def cpu_baseline(scenes_mod, world, start_state: dict, start_frame: int, frames: int, budget_s: float):
    """Times the reference step on the host over a bounded sample of the
    timed window: the first N scenes of the batch, from the batch's
    window-start state (the GPU state is bit-identical to the reference's),
    for `frames` frames. Three legs: the reference compiled from its sources
    at the parity flags on one core, at its release flags (-O2 -ffast-math)
    on one core, and at its release flags on every host core available (up
    to 16), one process per core on its own scenes (the reference step is
    single-threaded and not reentrant). The fastest is the baseline.
    Falls back to the oracle's C restatement where the reference build is
    absent."""
    sys.path.insert(0, str(ROOT / "tests"))
    from _libs import SimWorld, load  # test infrastructure: the CPU reference only

    kinds = [k for k in ("ref", "ref_fast") if load(k) is not None] or ["orc"]
    legs = {}
    per_scene_frame = None
    for kind in kinds:
        probe_world = scenes_mod.batched_uniform(4, first_seed=0)
        probe = SimWorld(kind, probe_world)
        probe.set_state(_scene_slice(world, start_state, 0, 4))
        t0 = time.perf_counter()
        probe.step(min(frames, 5))
        per_scene_frame = (time.perf_counter() - t0) / min(frames, 5) / 4
        probe.close()
        n_scenes = max(1, min(world.num_scenes, int(budget_s / (per_scene_frame * frames))))
        w = scenes_mod.batched_uniform(n_scenes, first_seed=0)
        sim = SimWorld(kind, w)
        sim.set_state(_scene_slice(world, start_state, 0, n_scenes))
        t0 = time.perf_counter()
        sim.step(frames)
        dt = time.perf_counter() - t0
        sim.close()
        legs[kind] = {"value": w.dynamic_bodies() * 20 * frames / dt, "cores": 1, "scenes": n_scenes,
                      "seconds": dt}
    fast = "ref_fast" if "ref_fast" in kinds else kinds[-1]
    cores = min(16, len(os.sched_getaffinity(0)))
    if cores > 1:
        # per process the same sample size as the one-core leg of this build
        n_each = min(legs[fast]["scenes"], world.num_scenes // cores)
        with tempfile.TemporaryDirectory() as tmp:
            procs = []
            for c in range(cores):
                path = Path(tmp) / f"state{c}.npz"
                np.savez(path, **_scene_slice(world, start_state, c * n_each, n_each))
                procs.append(subprocess.Popen(
                    [sys.executable, str(ROOT / "bench.py"), "--cpu-worker", str(path), "--worker-kind", fast,
                     "--worker-first", str(c * n_each), "--worker-scenes", str(n_each), "--steps", str(frames)],
                    stdin=subprocess.PIPE, stdout=subprocess.PIPE, text=True))
            for pr in procs:
                assert pr.stdout.readline().strip() == "ready"
            t0 = time.perf_counter()
            for pr in procs:
                pr.stdin.write("go\n")
                pr.stdin.flush()
            el = [json.loads(pr.stdout.readline())["elapsed"] for pr in procs]
            wall = time.perf_counter() - t0
            for pr in procs:
                pr.wait()
        body_substeps = cores * n_each * 16 * 20 * frames
        legs[fast + "_all_cores"] = {"value": body_substeps / wall, "cores": cores, "scenes": cores * n_each,
                                     "seconds": wall, "slowest_process_s": max(el)}
    best = max(legs, key=lambda k: legs[k]["value"])
    b = legs[best]
    names = {"ref": "the reference at its parity flags (-O2 -fno-fast-math -ffp-contract=off)",
             "ref_fast": "the reference at its release flags (-O2 -ffast-math)",
             "orc": "the oracle's C restatement (parity flags)"}
    base = best.replace("_all_cores", "")
    return {"value": b["value"], "unit": "body-substeps/s", "cores": b["cores"],
            "kind": "port" if base == "orc" else "reference",
            "sample": f"{names[base]}, {b['scenes']} of the {world.num_scenes} scenes (seeds 0..{b['scenes'] - 1}) "
                      f"x frames {start_frame}-{start_frame + frames - 1} from the batch's frame-{start_frame} "
                      f"state, {b['cores']} process(es), {b['seconds']:.1f} s",
            "legs": legs}


FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector (and matrix) peak


def op_counts(scenes_mod, world, start_state: dict, n_scenes: int, frames: int):
    """Algorithmic FP64 work per phase on a sample of the timed window (the
    first n_scenes scenes from the window-start state, `frames` frames),
    counted by the oracle's op-counting build (oracle/libopcount.so: every
    add, subtract, multiply, divide and square root the restatement
    executes)."""
    sys.path.insert(0, str(ROOT / "tests"))
    from _libs import SimWorld, _bind  # measurement infrastructure only

    path = ROOT / "oracle" / "libopcount.so"
    if not path.exists():
        return None
    lib = _bind(ctypes.CDLL(str(path)), "orc_")
    lib.orc_opcount_get.argtypes = [ctypes.c_void_p]
    sample = scenes_mod.batched_uniform(n_scenes, first_seed=0)
    sim = SimWorld("orc", sample, lib=lib)
    sim.set_state(_scene_slice(world, start_state, 0, n_scenes))
    lib.orc_opcount_reset()
    pairs = contacts = 0
    for _ in range(frames):
        sim.step(1)
        pairs += lib.orc_world_last_narrow_pairs(sim.h)
        contacts += lib.orc_world_last_contacts(sim.h)
    ops = np.zeros(16, dtype=np.uint64)
    lib.orc_opcount_get(ops.ctypes.data)
    sim.close()
    names = ("frame", "integrate", "transform", "narrow", "position", "velocity_update", "velocity_solve", "setup",
             "manifold")
    d = {n: int(ops[i]) for i, n in enumerate(names)}
    d["solve"] = d["position"] + d["velocity_update"] + d["velocity_solve"]
    return {"ops": d, "scenes": n_scenes, "frames": frames, "narrow_pairs": int(pairs), "contacts": int(contacts),
            "body_substeps": sample.dynamic_bodies() * 20 * frames}


TRAFFIC_FILE = ROOT / "profiles" / "traffic_window_f300_batched_uniform_4096x16.json"


def committed_traffic(kernel_class: str, schedule: dict):
    """HBM bytes per launch of a kernel from the committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE passes, scaled by the copy-kernel
    calibration), when it covers this kernel under the run's schedules
    ({"narrow": lanes, "manifold": lanes or 0, "solve": lanes}) that shape
    it (the narrow and manifold schedules for the narrow-phase kernels, the
    solve schedule for the solve); else None."""
    if not TRAFFIC_FILE.exists():
        return None
    d = json.loads(TRAFFIC_FILE.read_text())
    k = d.get("kernels", {}).get(kernel_class)
    shaping = {"narrow": ("narrow", "manifold"), "manifold": ("narrow", "manifold"), "solve": ("solve",)}
    keys = shaping.get(kernel_class, tuple(schedule))
    if not k or any(d.get("schedule", {}).get(key) != schedule.get(key) for key in keys):
        return None
    return {"bytes_per_launch": k["bytes_per_launch"], "source": f"{TRAFFIC_FILE.relative_to(ROOT)}: {d['window']}"}


def adjoint_run(pkg, sc, n_scenes: int, frames: int, check_scenes: int = 256):
    """Differentiable workload: n_scenes cubes thrown onto tilted static boxes,
    `frames` frames; the checkpointed forward pass, then the backward pass of
    L = sum |x_final - target|^2. Adjoint steps/s = scenes x frames x
    substeps reversed per second of the backward pass (recomputation included)."""
    world, targets = sc.ramps(n_scenes)
    gw = pkg.World(world)
    init = gw.state()
    gw.adjoint_forward(2)  # first launches
    gw.set_state(init)
    gw.synchronize()
    t0 = time.perf_counter()
    err = gw.adjoint_forward(frames)
    gw.synchronize()
    t_fwd = time.perf_counter() - t0
    st = gw.state()
    dl = np.zeros((world.num_bodies, 13))
    dyn = np.arange(1, world.num_bodies, 2)
    dl[dyn, 0:3] = 2.0 * (st["position"][dyn] - targets)
    t0 = time.perf_counter()
    gs, gp = gw.adjoint_backward(dl)
    t_bwd = time.perf_counter() - t0
    out = {"metric": "adjoint steps/s", "unit": "scene-substeps reversed/s",
           "value": n_scenes * frames * 20 / t_bwd,
           "config": {"workload": f"differentiable: {n_scenes} scenes x (cube thrown onto a tilted static box), "
                                  f"{frames} frames, dt=1/60, 20 substeps, contact set frozen",
                      "scenes": n_scenes, "frames": frames},
           "forward_checkpointed_s": t_fwd, "backward_s": t_bwd, "forward_error": int(err),
           "recompute_mismatches": gw.adjoint_mismatches()}
    amp = gw.adjoint_amplification()[dyn]
    gw.close()
    from raw_physics_amd import gradcheck

    # gradient norms over the scenes whose adjoint stays within the validity
    # bound, and over all scenes (those include impacts that amplify the
    # adjoint beyond what a finite perturbation resolves)
    ok = amp <= gradcheck.AMPLIFICATION_VALID
    out["grad_norm_v0_valid_scenes"] = float(np.linalg.norm(gs[dyn[ok], 7:10])) if ok.any() else None
    out["grad_norm_mass_valid_scenes"] = float(np.linalg.norm(gp[dyn[ok], 0])) if ok.any() else None
    out["grad_norm_v0_all_scenes"] = float(np.linalg.norm(gs[dyn, 7:10]))
    out["grad_norm_mass_all_scenes"] = float(np.linalg.norm(gp[dyn, 0]))
    out["grads_finite"] = bool(np.isfinite(gs[dyn]).all() and np.isfinite(gp[dyn]).all())
    # Gradient check against FP64 central differences (h = 1e-6, and 1e-7 for
    # reproducibility) of perturbed runs on the device, on the first scenes
    # of the workload at several horizons (raw_physics_amd.gradcheck).
    out["amplification_valid"] = gradcheck.AMPLIFICATION_VALID
    out["valid_scenes"] = int((amp <= gradcheck.AMPLIFICATION_VALID).sum())
    out["gradient_check"] = {}
    for horizon in sorted({30, 60, frames}):
        r = gradcheck.gradient_check(pkg, min(check_scenes, n_scenes), 0, horizon)
        out["gradient_check"][f"{horizon}_frames"] = r["summary"]
    log("adjoint: " + json.dumps(out))
    return out


def _leg_state(gw, load):
    """The leg's start state: a saved one (profiling windows) or the world's own."""
    if load:
        with np.load(load, allow_pickle=False) as z:
            gw.set_state({k: z[k] for k in z.files})
        log(f"start state loaded from {load}")
    return gw.state()


def _leg_end(gw, save, profile):
    """Saves the end state and returns the per-kernel HIP-event times (or None)."""
    if save:
        np.savez(save, **gw.state())
        log(f"end state saved to {save}")
    if not profile:
        return None
    kt = gw.kernel_times()
    gw.set_profiling(False)
    return kt


def plain_run(pkg, sc, world, start_state: dict, n_scenes: int, frames: int, substeps: int, dt: float,
              eager=False, profile=False) -> dict:
    """The plain method: one scene per step call, each frame's state written
    back to the host (the drop-in pbd_simulate's cost model), every scene
    from its state in `start_state` (the batched window's start).
    eager: kernels launched one by one; profile: HIP events around every
    launch of the first scene's frames (a second pass, not timed)."""
    worlds = []
    for s in range(n_scenes):
        w1 = sc.World(name="plain")
        sc.drop_scene(w1, s)
        worlds.append(pkg.World(w1))
    for w1 in worlds:  # warm
        if eager:
            w1.set_graphs(False)
        w1.step(1)
    for s, w1 in enumerate(worlds):
        w1.set_state(_scene_slice(world, start_state, s, 1))
    t0 = time.perf_counter()
    for w1 in worlds:
        for _ in range(frames):
            w1.step(1, dt=dt, substeps=substeps)
            w1.state()  # write-back to the host, as the drop-in does every call
    pt = time.perf_counter() - t0
    out = {"value": 16 * n_scenes * substeps * frames / pt, "unit": "body-substeps/s",
           "ms_per_call": 1e3 * pt / (n_scenes * frames), "scenes": n_scenes, "frames": frames,
           "schedule": {"narrow": worlds[0].narrow_group, "manifold": worlds[0].manifold_group,
                        "solve": worlds[0].solve_lanes}}
    log(f"plain: {n_scenes} scenes x {frames} frames one scene per call -> {out['value']:.4g}")
    if profile:
        w1 = worlds[0]
        w1.set_state(_scene_slice(world, start_state, 0, 1))
        w1.set_profiling(True)
        w1.step(frames, dt=dt, substeps=substeps)
        out["kernel_times_ms"] = w1.kernel_times()
        w1.set_profiling(False)
    for w1 in worlds:
        w1.close()
    return out


def _plain_only(pkg, sc, args, substeps: int, dt: float) -> dict:
    """The plain leg alone, from a batched window-start state saved with
    --save-state (--load-state), or from frame 0."""
    world = sc.batched_uniform(args.plain_scenes)
    if args.load_state:
        with np.load(args.load_state, allow_pickle=False) as z:
            st = {k: z[k] for k in z.files}
        log(f"window start state loaded from {args.load_state}")
    else:
        gw = pkg.World(world)
        st = gw.state()
        gw.close()
    return plain_run(pkg, sc, world, st, args.plain_scenes, args.plain_frames, substeps, dt,
                     eager=args.eager, profile=args.leg_profile)


def large_run(pkg, sc, n: int, frames: int, eager=False, load=None, save=None, profile=False):
    """One large scene (grid broad phase, coloured solve): body-substeps/s.
    load / save: start from / keep a saved state (profiling windows deep in
    the run); eager: kernels launched one by one; profile: HIP events around
    every launch (the frame loop then waits for each frame)."""
    t0 = time.perf_counter()
    world = sc.large_scene(n)
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    gw = pkg.World(world)
    t_upload = time.perf_counter() - t0
    if eager:
        gw.set_graphs(False)
    init = _leg_state(gw, load)
    gw.step(2)  # first launches and graph capture
    gw.set_state(init)
    gw.synchronize()
    if profile:
        gw.set_profiling(True)
    el = 0.0
    for done in range(0, frames, 30):  # chunks with a progress line each
        k = min(30, frames - done)
        t0 = time.perf_counter()
        gw.step(k)
        el += time.perf_counter() - t0
        log(f"large scene: frames {done}-{done + k - 1} at {1e3 * el / (done + k):.1f} ms/frame so far")
    kt = _leg_end(gw, save, profile)
    st = gw.state()
    out = {"workload": f"large single scene: {n} dynamic bodies (60% cubes, 40% ico42, scale U(0.5,1.5)) above a "
                       f"static floor, frames 0-{frames - 1}, dt=1/60, 20 substeps; uniform-grid broad phase, "
                       "graph-coloured solve",
           "bodies": n, "frames": frames,
           "value": n * 20 * frames / el, "unit": "body-substeps/s", "ms_per_frame": 1e3 * el / frames,
           "scene_build_s": t_build, "upload_s": t_upload, "error_word": int(gw.errors()[0]),
           "lowest_body_y": float(st["position"][1:, 1].min()), "active_bodies": int(st["active"][1:].sum()),
           "start": load or "frame 0", "kernel_times_ms": kt}
    gw.close()
    log("large scene: " + json.dumps(out))
    return out


def uneven_run(pkg, sc, n_scenes: int, frames: int, tune_from: int, autotune=True, eager=False, load=None,
               save=None, profile=False):
    """Batched uneven workload: n_scenes scenes of 4-256 bodies with 8/42/162-
    vertex hulls, 30% carrying joint chains; the autotuner picks the
    schedules on frames from a mid-run state; body-substeps/s over `frames`
    frames from the start (graph replay). load / save / eager / profile as
    for large_run (autotune off keeps the world's default schedules)."""
    t0 = time.perf_counter()
    world = sc.batched_uneven(n_scenes)
    t_build = time.perf_counter() - t0
    gw = pkg.World(world)
    if eager:
        gw.set_graphs(False)
    init = _leg_state(gw, load)
    tune = None
    if autotune:
        gw.step(tune_from)
        tune = gw.autotune(frames=4)
        gw.set_state(init)
    gw.step(2)
    gw.set_state(init)
    gw.synchronize()
    if profile:
        gw.set_profiling(True)
    t0 = time.perf_counter()
    gw.step(frames, sync=False)
    gw.synchronize()
    el = time.perf_counter() - t0
    kt = _leg_end(gw, save, profile)
    dyn = world.dynamic_bodies()
    errs = gw.errors()
    out = {"workload": f"batched uneven: {n_scenes} scenes, 4-256 bodies (log-uniform), 8/42/162-vertex hulls, "
                       f"30% with joint chains, {frames} frames",
           "scenes": n_scenes, "dynamic_bodies": dyn, "constraints": len(world.constraints), "frames": frames,
           "value": dyn * 20 * frames / el, "unit": "body-substeps/s", "ms_per_frame": 1e3 * el / frames,
           "schedule": {"narrow": gw.narrow_group, "manifold": gw.manifold_group, "solve": gw.solve_lanes,
                        "views": gw.views},
           "autotune_ms": tune, "scene_build_s": t_build, "scenes_with_error_words": int((errs != 0).sum()),
           "start": load or "frame 0", "kernel_times_ms": kt}
    gw.close()
    log("uneven: " + json.dumps(out))
    return out


def copy_calibration(nbytes: int):
    """One hipMemcpy device to device of `nbytes` (the runtime's copy kernel
    reads and writes nbytes each)."""
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: {nbytes} bytes device to device")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60, help="timed frames of the whole batch")
    ap.add_argument("--warmup", type=int, default=5, help="untimed frames before the timed region")
    ap.add_argument("--start-frame", type=int, default=300,
                    help="frame the timed window starts at: the batch is stepped there untimed first "
                         "(frame 300 of the 600-frame run: a mix of falling, resting and sleeping scenes whose "
                         "per-frame cost matches the run's mean)")
    ap.add_argument("--run-frames", type=int, default=600, help="frames of the whole run timed from frame 0")
    ap.add_argument("--scenes", type=int, default=4096)
    ap.add_argument("--plain-scenes", type=int, default=16)
    ap.add_argument("--plain-frames", type=int, default=60)
    ap.add_argument("--cpu-budget", type=float, default=8.0, help="seconds of CPU baseline work per leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-profile", action="store_true", help="skip the per-kernel timing run")
    ap.add_argument("--op-scenes", type=int, default=8, help="scenes in the op-count sample")
    ap.add_argument("--ops-out", metavar="JSON", help="write the window's op counts (per phase and per launch)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tune-from", type=int, default=90, help="frame the autotuner's trials start from")
    ap.add_argument("--tune-frames", type=int, default=6)
    ap.add_argument("--narrow-lanes", type=int, default=None,
                    help="fixed narrow schedule (lanes per pair); skips the autotuner")
    ap.add_argument("--manifold-lanes", type=int, default=None,
                    help="fixed manifold schedule (lanes per pair; 0 = built in the narrow kernel); skips the autotuner")
    ap.add_argument("--solve-lanes", type=int, default=None,
                    help="fixed solve schedule (lanes per scene); skips the autotuner")
    ap.add_argument("--eager", action="store_true", help="launch each kernel eagerly instead of replaying frame graphs")
    ap.add_argument("--views", type=int, default=0,
                    help="concurrent scene-range views for graph replay (0 = the library's automatic choice)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint measurement")
    ap.add_argument("--adjoint-scenes", type=int, default=2048)
    ap.add_argument("--adjoint-frames", type=int, default=120)
    ap.add_argument("--no-uneven", action="store_true", help="skip the batched uneven measurement")
    ap.add_argument("--uneven-scenes", type=int, default=1024)
    ap.add_argument("--uneven-frames", type=int, default=600)
    ap.add_argument("--large-bodies", type=int, default=100000,
                    help="bodies of the large single scene (0: skip it)")
    ap.add_argument("--large-frames", type=int, default=300)
    ap.add_argument("--window-only", action="store_true",
                    help="only the batched window (profiling runs): no whole run, no other workloads")
    ap.add_argument("--only", choices=("uneven", "large", "adjoint", "plain"),
                    help="run only this workload's leg (profiling runs) and print its result")
    ap.add_argument("--save-state", metavar="NPZ",
                    help="save the body state: after the window's start frame (batched), or at the end of an --only leg")
    ap.add_argument("--load-state", metavar="NPZ",
                    help="start from a saved body state: the batched window, or an --only leg (profiling windows)")
    ap.add_argument("--leg-profile", action="store_true",
                    help="with --only: HIP-event times of every kernel launch of the leg")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run one device-to-device copy of MIB MiB (a kernel of known traffic that scales "
                         "the FETCH_SIZE / WRITE_SIZE counters of a profiled run)")
    # one process of the all-cores CPU leg (started by cpu_baseline)
    ap.add_argument("--cpu-worker", metavar="NPZ", help=argparse.SUPPRESS)
    ap.add_argument("--worker-kind", help=argparse.SUPPRESS)
    ap.add_argument("--worker-first", type=int, help=argparse.SUPPRESS)
    ap.add_argument("--worker-scenes", type=int, help=argparse.SUPPRESS)
    args = ap.parse_args()

    if args.cpu_worker:
        cpu_worker(args)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: scenes are independent, so a multi-GPU run is the same "
              "kernel on each GPU's share of the scenes; no figure at several GPUs is produced.",
              file=sys.stderr)
        sys.exit(2)
    if args.window_only:
        args.no_adjoint = args.no_uneven = args.no_plain = args.no_cpu = True
        args.large_bodies = 0

    pkg = _package()
    sc = pkg.scenes
    substeps, dt = 20, 1.0 / 60.0

    if args.copy_calibration:
        copy_calibration(args.copy_calibration << 20)
    if args.only:
        kw = {"eager": args.eager, "load": args.load_state, "save": args.save_state, "profile": args.leg_profile}
        leg = {"uneven": lambda: uneven_run(pkg, sc, args.uneven_scenes, args.uneven_frames, args.tune_from,
                                            autotune=not args.no_autotune, **kw),
               "large": lambda: large_run(pkg, sc, args.large_bodies, args.large_frames, **kw),
               "adjoint": lambda: adjoint_run(pkg, sc, args.adjoint_scenes, args.adjoint_frames),
               "plain": lambda: _plain_only(pkg, sc, args, substeps, dt)}[args.only]
        print(json.dumps({"only": args.only, "result": leg()}), flush=True)
        return
    t0 = time.perf_counter()
    world = sc.batched_uniform(args.scenes)
    log(f"built {args.scenes} scenes ({world.num_bodies} bodies) in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    gw = pkg.World(world)
    log(f"uploaded world in {time.perf_counter() - t0:.1f} s")
    init = gw.state()

    # schedule: the package's autotuner times each narrow-phase variant on a
    # few frames from a mid-run state (bodies landed, contacts present)
    tune = None
    if args.eager:
        gw.set_graphs(False)
    if args.views:
        gw.set_views(args.views)
    fixed = [a is not None for a in (args.narrow_lanes, args.manifold_lanes, args.solve_lanes)]
    if args.narrow_lanes is not None:
        gw.set_narrow_group(args.narrow_lanes)
    if args.manifold_lanes is not None:
        gw.set_manifold_group(args.manifold_lanes)
    if args.solve_lanes is not None:
        gw.set_solve_lanes(args.solve_lanes)
    if not (args.no_autotune or any(fixed)):
        gw.step(args.tune_from)
        tune = gw.autotune(frames=args.tune_frames)
        gw.set_state(init)
        log(f"autotune (kernel ms over {args.tune_frames} frames): {tune} -> narrow {gw.narrow_group} lanes/pair, "
            f"manifold {gw.manifold_group} lanes/pair, "
            f"solve {gw.solve_lanes} lanes/scene")

    # the whole run from frame 0 (graph replay), which also reaches the
    # window's start state
    dyn = world.dynamic_bodies()
    run = None
    gw.step(2)
    gw.set_state(init)
    gw.synchronize()
    if args.load_state:
        with np.load(args.load_state, allow_pickle=False) as z:
            gw.set_state({k: z[k] for k in z.files})
        start_state = gw.state()
        log(f"window start state loaded from {args.load_state}")
    elif not args.window_only and args.run_frames >= args.start_frame:
        t0 = time.perf_counter()
        gw.step(args.start_frame, dt=dt, substeps=substeps, sync=False)
        gw.synchronize()
        t_head = time.perf_counter() - t0
        start_state = gw.state()
        t0 = time.perf_counter()
        gw.step(args.run_frames - args.start_frame, dt=dt, substeps=substeps, sync=False)
        gw.synchronize()
        el = t_head + time.perf_counter() - t0
        run = {"frames": args.run_frames, "value": dyn * substeps * args.run_frames / el,
               "ms_per_frame": 1e3 * el / args.run_frames, "error_scenes": int((gw.errors() != 0).sum())}
        log(f"whole run: {args.run_frames} frames in {el:.3f} s -> {run['value']:.4g} body-substeps/s")
    else:
        gw.step(args.start_frame, dt=dt, substeps=substeps)
        start_state = gw.state()
    if args.save_state:
        np.savez(args.save_state, **start_state)
        log(f"window start state saved to {args.save_state}")

    # the timed window: frames start_frame .. start_frame + steps - 1
    if args.warmup > 0:
        gw.set_state(start_state)
        gw.step(args.warmup)
    gw.set_state(start_state)
    gw.synchronize()
    t0 = time.perf_counter()
    gw.step(args.steps, dt=dt, substeps=substeps, sync=False)
    gw.synchronize()
    elapsed = time.perf_counter() - t0
    value = dyn * substeps * args.steps / elapsed
    errs = gw.errors()
    narrow_group, manifold_group, solve_lanes, views = gw.narrow_group, gw.manifold_group, gw.solve_lanes, gw.views
    log(f"batched window: frames {args.start_frame}-{args.start_frame + args.steps - 1} in {elapsed:.3f} s -> "
        f"{value:.4g} body-substeps/s; error words nonzero in {int((errs != 0).sum())} scenes; "
        f"EPA non-convergence {gw.epa_failures()}")

    # Per-kernel durations: the same window again, with HIP events recorded
    # around every launch on the world's stream (eager launches, one view).
    ktimes = None
    if not args.no_profile:
        gw.set_state(start_state)
        gw.set_profiling(True)
        gw.step(args.steps, dt=dt, substeps=substeps)
        ktimes = gw.kernel_times()
        gw.set_profiling(False)
        log("kernel times (ms, launches): " + json.dumps(ktimes))
    gw.close()

    adjoint = None if args.no_adjoint else adjoint_run(pkg, sc, args.adjoint_scenes, args.adjoint_frames)
    large = None if not args.large_bodies else large_run(pkg, sc, args.large_bodies, args.large_frames)
    uneven = None if args.no_uneven else uneven_run(pkg, sc, args.uneven_scenes, args.uneven_frames, args.tune_from)

    plain = None
    if not args.no_plain:
        plain = plain_run(pkg, sc, world, start_state, args.plain_scenes, args.plain_frames, substeps, dt)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(sc, world, start_state, args.start_frame, args.steps, args.cpu_budget)
        log(f"cpu: {cpu}")

    roofline = None
    kernels = None
    if ktimes is not None:
        oc = op_counts(sc, world, start_state, min(args.op_scenes, args.scenes), args.steps)
        if oc is not None:
            scale = args.scenes / oc["scenes"]
            kernels = {}
            fused = manifold_group == 0  # the narrow kernel builds the manifolds too
            ops = dict(oc["ops"])
            if fused:
                ops["narrow"] = ops["narrow"] + ops["manifold"]
                ops["manifold"] = 0.0
            n_int, n_solve = ktimes["integrate"][1], ktimes["solve"][1]
            if 0 < n_int < n_solve:
                # the solve kernel integrates the next substep's bodies; the
                # integrate kernel runs once per frame
                moved = ops["integrate"] * (1.0 - n_int / n_solve)
                ops["integrate"] -= moved
                ops["solve"] += moved
            for cls in ("integrate", "transform", "narrow", "manifold", "solve"):
                ms, n = ktimes[cls]
                flops = ops[cls] * scale
                kernels[cls] = {"ms_total": ms, "launches": n, "avg_us": 1e3 * ms / max(n, 1),
                                "algorithmic_flops": flops, "flops_per_launch": flops / max(n, 1),
                                "tflops": flops / (ms * 1e-3) / 1e12 if ms > 0 else None}
            # the dominant kernel: largest share of the frame
            names = {"narrow": ("k_gjk_epa (GJK + EPA + contact manifold per live pair)" if fused
                                else "k_gjk_epa (GJK + EPA per live pair)"),
                     "manifold": "k_manifold (contact manifold per colliding pair)",
                     "solve": "k_solve (levelled position + velocity solve, next substep's integration)"}
            dom = max(names, key=lambda c: kernels[c]["ms_total"])
            k = kernels[dom]
            unit_ops = {"narrow": (ops["narrow"], oc["narrow_pairs"], "live pair"),
                        "manifold": (ops["manifold"], oc["narrow_pairs"], "live pair"),
                        "solve": (oc["ops"]["solve"], oc["body_substeps"], "body-substep")}[dom]
            traffic = committed_traffic(dom, {"narrow": narrow_group, "manifold": manifold_group,
                                              "solve": solve_lanes})
            # the compute roof ("mfma" in the record's vocabulary): MI355X's FP64
            # peak, the same for its vector and matrix units; the kernel's FP64
            # work is VALU (no matrix-shaped work in the step)
            roofline = {"bound": "mfma", "bound_detail": "FP64 compute roof (78.6 TFLOP/s, vector = matrix "
                        "peak); the kernel issues FP64 VALU, no MFMA; HBM under 8 % of 8 TB/s",
                        "kernel": names[dom],
                        "achieved": k["tflops"], "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": k["tflops"] / FP64_PEAK_TFLOPS if k["tflops"] else None,
                        "traffic": traffic["bytes_per_launch"] if traffic else None,
                        "traffic_source": traffic["source"] if traffic else None,
                        "flops_per_launch": k["flops_per_launch"],
                        "avg_launch_us": k["avg_us"],
                        "op_sample": f"{oc['scenes']} scenes x frames {args.start_frame}-"
                                     f"{args.start_frame + args.steps - 1} counted by the oracle's op-counting "
                                     f"build, scaled x{scale:g} to {args.scenes} scenes; "
                                     f"{unit_ops[0] / max(unit_ops[1], 1):.0f} FP64 ops per {unit_ops[2]}"}
            log("roofline: " + json.dumps(roofline))
            if args.ops_out:
                Path(args.ops_out).parent.mkdir(parents=True, exist_ok=True)
                Path(args.ops_out).write_text(json.dumps(
                    {"window": f"batched uniform {args.scenes} scenes, frames {args.start_frame}-"
                               f"{args.start_frame + args.steps - 1} from the frame-{args.start_frame} state",
                     "sample": {k2: oc[k2] for k2 in ("scenes", "frames", "narrow_pairs", "contacts",
                                                      "body_substeps")},
                     "scale": scale, "ops_sample": oc["ops"], "kernels": kernels,
                     "schedule": {"narrow": narrow_group, "manifold": manifold_group, "solve": solve_lanes}},
                    indent=1))

    window = f"frames {args.start_frame}-{args.start_frame + args.steps - 1}"
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "body-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded scenes)",
        "config": {"workload": f"batched uniform: {args.scenes} scenes x (floor + 8 cubes + 8 ico42), {window} of "
                               f"the {max(args.run_frames, args.start_frame + args.steps)}-frame run (stepped "
                               f"untimed to frame {args.start_frame} first), dt=1/60, 20 substeps, "
                               "1 position iteration",
                   "scenes": args.scenes, "dynamic_bodies": dyn, "start_frame": args.start_frame,
                   "frames": args.steps, "substeps": substeps,
                   "narrow_lanes_per_pair": narrow_group, "manifold_lanes_per_pair": manifold_group,
                   "solve_lanes_per_scene": solve_lanes, "views": views, "autotune_ms": tune},
        "whole_run": run,
        "plain": None if plain is None else {**plain, "window": window,
                                             "ratio_batched_over_plain": value / plain["value"]},
        "adjoint": adjoint,
        "large_scene": large,
        "uneven": uneven,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
