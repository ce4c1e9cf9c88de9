"""ctypes access to the three implementations of the world API used by tests.

* ``orc_``  oracle/liboracle.so      — the C restatement (the checker)
* ``ref_``  oracle/_ref/libref.so    — the reference compiled from its sources
                                       (only where /root/reference was present
                                       at build time)
* ``rp_``   raw-physics_amd/librawphys.so — the HIP product (GPU)
"""
from __future__ import annotations

import ctypes
import importlib.util
import os
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
PKG_DIR = ROOT / "raw-physics_amd"
ORACLE_SO = ROOT / "oracle" / "liboracle.so"
REF_SO = ROOT / "oracle" / "_ref" / "libref.so"
# the reference at its release flags (-O2 -ffast-math): CPU-baseline timing only
REF_FAST_SO = ROOT / "oracle" / "_ref" / "libref_fast.so"
# the reference's objects without pbd.o, linked against librawphys.so (its
# calls to pbd_simulate* bind to the library: the link-level drop-in)
REF_DROPIN_SO = ROOT / "oracle" / "_ref" / "libref_dropin.so"


def load_package():
    """Imports the product package directory (its name is not an identifier)."""
    if "raw_physics_amd" in sys.modules:
        return sys.modules["raw_physics_amd"]
    spec = importlib.util.spec_from_file_location("raw_physics_amd", PKG_DIR / "__init__.py",
                                                  submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["raw_physics_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


pkg = load_package()
from raw_physics_amd import scenes  # noqa: E402

_P = ctypes.c_void_p


def _bind(lib, prefix):
    f = lambda n: getattr(lib, prefix + n)  # noqa: E731
    f("world_create").restype = _P
    f("world_create").argtypes = [ctypes.POINTER(scenes.WorldDescC)]
    f("world_destroy").argtypes = [_P]
    f("world_step").argtypes = [_P, ctypes.c_double, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                ctypes.c_int32, ctypes.c_double]
    f("world_get_state").argtypes = [_P, ctypes.POINTER(scenes.StateC)]
    f("world_set_state").argtypes = [_P, ctypes.POINTER(scenes.StateC)]
    f("world_num_bodies").restype = ctypes.c_uint32
    f("world_num_bodies").argtypes = [_P]
    f("world_body_info").argtypes = [_P, _P, _P, _P, _P]
    f("world_broad_pairs").restype = ctypes.c_uint32
    f("world_broad_pairs").argtypes = [_P, ctypes.c_uint32, _P, ctypes.c_uint32]
    f("world_contacts").restype = ctypes.c_uint32
    f("world_contacts").argtypes = [_P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _P, ctypes.c_uint32]
    f("world_hull").argtypes = [_P, ctypes.c_uint32, _P] + [_P] * 10
    if hasattr(lib, prefix + "world_error"):
        f("world_error").restype = ctypes.c_uint32
        f("world_error").argtypes = [_P, ctypes.c_uint32]
    for n in ("world_epa_failures", "world_last_contacts", "world_last_narrow_pairs"):
        if hasattr(lib, prefix + n):
            f(n).restype = ctypes.c_uint64
            f(n).argtypes = [_P]
    return lib


_cache = {}


def load(kind: str):
    """Library for `kind`; None if that library is not built here."""
    if kind in _cache:
        return _cache[kind]
    if kind == "rp":
        lib = pkg.library()  # raises if librawphys.so is missing: no fallback
        _cache[kind] = lib
        return lib
    path = {"orc": ORACLE_SO, "ref": REF_SO, "ref_fast": REF_FAST_SO, "ref_dropin": REF_DROPIN_SO}[kind]
    if not path.exists():
        return None
    lib = _bind(ctypes.CDLL(str(path)), kind.split("_")[0] + "_")
    _cache[kind] = lib
    return lib


class SimWorld:
    """One world in one implementation (prefix `kind`)."""

    def __init__(self, kind: str, world: scenes.World, lib=None):
        self.kind = kind
        self.lib = lib or load(kind)
        self.p = kind.split("_")[0] + "_"
        self.world = world
        self.desc, self.keep = world.desc()
        self.h = self._f("world_create")(ctypes.byref(self.desc))
        if not self.h:
            raise RuntimeError(f"{kind}_world_create failed")
        self.nb = world.num_bodies

    def _f(self, n):
        return getattr(self.lib, self.p + n)

    def close(self):
        if self.h:
            self._f("world_destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, frames=1, dt=1.0 / 60.0, substeps=20, iters=1, collisions=True, gravity=10.0):
        self._f("world_step")(self.h, dt, frames, substeps, iters, 1 if collisions else 0, gravity)

    def state(self) -> dict:
        st = scenes.alloc_state(self.nb)
        s = scenes.state_struct(st)
        self._f("world_get_state")(self.h, ctypes.byref(s))
        return st

    def set_state(self, st: dict):
        s = scenes.state_struct(st)
        self._f("world_set_state")(self.h, ctypes.byref(s))

    def body_info(self):
        im = np.zeros(self.nb)
        it = np.zeros((self.nb, 9))
        iit = np.zeros((self.nb, 9))
        rad = np.zeros(self.nb)
        self._f("world_body_info")(self.h, im.ctypes.data, it.ctypes.data, iit.ctypes.data, rad.ctypes.data)
        return im, it, iit, rad

    def broad_pairs(self, scene: int):
        cap = 1 << 16
        out = np.zeros((cap, 2), dtype=np.uint32)
        n = self._f("world_broad_pairs")(self.h, scene, out.ctypes.data, cap)
        return out[:n].copy()

    def contacts(self, scene, i, j):
        cap = 256
        out = np.zeros((cap, 9))
        n = self._f("world_contacts")(self.h, scene, i, j, out.ctypes.data, cap)
        return out[:n].copy()

    def hull(self, body: int) -> dict:
        counts = np.zeros(6, dtype=np.uint32)
        self._f("world_hull")(self.h, body, counts.ctypes.data, *([None] * 10))
        nv, nf, nfe, nvf, nvn, nff = (int(c) for c in counts)
        h = {"vertices": np.zeros((nv, 3)), "face_normals": np.zeros((nf, 3)),
             "face_off": np.zeros(nf + 1, np.uint32), "face_el": np.zeros(max(nfe, 1), np.uint32),
             "vf_off": np.zeros(nv + 1, np.uint32), "vf": np.zeros(max(nvf, 1), np.uint32),
             "vn_off": np.zeros(nv + 1, np.uint32), "vn": np.zeros(max(nvn, 1), np.uint32),
             "ff_off": np.zeros(nf + 1, np.uint32), "ff": np.zeros(max(nff, 1), np.uint32)}
        order = ["vertices", "face_normals", "face_off", "face_el", "vf_off", "vf", "vn_off", "vn", "ff_off", "ff"]
        self._f("world_hull")(self.h, body, counts.ctypes.data, *[h[k].ctypes.data for k in order])
        h["face_el"] = h["face_el"][:nfe]
        h["vf"] = h["vf"][:nvf]
        h["vn"] = h["vn"][:nvn]
        h["ff"] = h["ff"][:nff]
        return h

    def error(self, scene: int) -> int:
        f = getattr(self.lib, self.p + "world_error", None)
        return int(f(self.h, scene)) if f else 0


def state_bits_equal(a: dict, b: dict, fields=None) -> list:
    """Names of state fields whose bit patterns differ."""
    bad = []
    for n, _ in scenes.STATE_FIELDS:
        if fields and n not in fields:
            continue
        if not np.array_equal(a[n].view(np.uint64) if a[n].dtype == np.float64 else a[n],
                              b[n].view(np.uint64) if b[n].dtype == np.float64 else b[n]):
            bad.append(n)
    return bad
