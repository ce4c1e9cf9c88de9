"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

oracle/Makefile `sanitize` builds the C restatement (oracle/_san/
liboracle_san.so) and the library's OBJ parser (rp_obj.cpp, with a stub for
the library's error slot: oracle/_san/librpobj_san.so) with
-fsanitize=address,undefined. A child process with libasan preloaded steps
scenes of every kind through the restatement (hull building, GJK, EPA,
clipping, compounds, spheres, joints, the solve-order mode) and parses the
300 synthetic OBJ texts of tests/golden/obj_cases.npz, checking the parser's
output against the reference's. Any sanitizer report fails the child.
"""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SAN = ROOT / "oracle" / "_san"

CHILD = r'''
import ctypes, sys
import numpy as np
sys.path.insert(0, sys.argv[1] + "/tests")
from _libs import SimWorld, _bind, scenes
orc = _bind(ctypes.CDLL(sys.argv[1] + "/oracle/_san/liboracle_san.so"), "orc_")
orc.orc_world_set_solve_order.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32]
worlds = [(scenes.stack_scene(seed=2), 60, {}), (scenes.batched_uniform(2, first_seed=9), 60, {}),
          (scenes.batched_uniform(1, first_seed=3, ico_subdiv=2), 30, {}), (scenes.batched_mixed(2), 40, {}),
          (scenes.chains(2, length=5), 40, {"collisions": False}), (scenes.prisms_on_pillars(1), 40, {}),
          (scenes.batched_uneven(2, first_seed=21, max_bodies=40), 40, {})]
for w, frames, kw in worlds:
    sim = SimWorld("orc", w, lib=orc)
    sim.step(frames, **kw)
    st = sim.state()
    assert np.all(np.isfinite(st["position"])), w.name
    for s in range(w.num_scenes):
        assert sim.error(s) == 0, (w.name, s)
    sim.close()
# the solve-order mode: every item in class 0 of a 2-body scene is a valid order
w = scenes.batched_uniform(1, first_seed=4)
sim = SimWorld("orc", w, lib=orc)
sim.step(30)
cls = np.zeros(1024, dtype=np.int32)
orc.orc_world_set_solve_order(sim.h, 0, cls.ctypes.data, 1024)
sim.step(1)
obj = ctypes.CDLL(sys.argv[1] + "/oracle/_san/librpobj_san.so")
obj.rp_obj_parse_memory.restype = ctypes.c_uint32
obj.rp_obj_parse_memory.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_uint32,
                                    ctypes.POINTER(ctypes.c_uint32)]
g = np.load(sys.argv[1] + "/tests/golden/obj_cases.npz")
text, toff, verts, voff = g["text"], g["text_offset"], g["vertices"], g["vertex_offset"]
for k in range(len(toff) - 1):
    src = bytes(text[toff[k]:toff[k + 1]])
    n = ctypes.c_uint32(0)
    assert obj.rp_obj_parse_memory(src, len(src), None, 0, ctypes.byref(n)) == 0, k
    out = np.zeros((max(n.value, 1), 8), dtype=np.float32)
    assert obj.rp_obj_parse_memory(src, len(src), out.ctypes.data, n.value, ctypes.byref(n)) == 0, k
    want = verts[voff[k]:voff[k + 1]]
    cols = 8 if g["has_texcoords"][k] else 6
    assert np.array_equal(out[:n.value, :cols].view(np.uint32), np.ascontiguousarray(want[:, :cols]).view(np.uint32)), k
# malformed inputs take the error path, not a crash
for bad in (b"f 1 2 3\n", b"v 0 0 0\nf 0 1 1\n", b"v 1 2\nf 1 1 1 1\n"):
    n = ctypes.c_uint32(0)
    obj.rp_obj_parse_memory(bad, len(bad), None, 0, ctypes.byref(n))
print("sanitized run ok")
'''


@pytest.mark.skipif(not (SAN / "liboracle_san.so").exists() or not (SAN / "librpobj_san.so").exists(),
                    reason="sanitizer builds (make -C oracle sanitize) not present")
def test_host_code_clean_under_asan_and_ubsan():
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    ubsan = subprocess.run(["gcc", "-print-file-name=libubsan.so"], capture_output=True, text=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=f"{asan}:{ubsan}", ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([sys.executable, "-c", CHILD, str(ROOT)], env=env, capture_output=True, text=True,
                       timeout=600)
    report = r.stdout[-4000:] + r.stderr[-4000:]
    assert r.returncode == 0, report
    assert "sanitized run ok" in r.stdout, report
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, report
