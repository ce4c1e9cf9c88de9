#!/usr/bin/env python3
"""Writes tests/golden/assets.npz: the triangle corner positions that the
reference's obj_parse (src/render/obj.cpp:7-151) returns for the collider
assets of two of its example scenes, as the input geometry of device tests
of those scenes (the reference directory is not present where the GPU tests
run):

* spot_storm (src/examples/spot_storm.cpp:35-45, 76-115): res/spot/
  spot-hull-1.obj ... spot-hull-11.obj, the 11 convex pieces of one body;
* mirror_cube (src/examples/mirror_cube.cpp:50-68): res/mirror_cube_collider1.obj
  and res/mirror_cube_collider2.obj, the 2 pieces of one body;
* res/cube.obj, the floor of both.

Each array is float32 (corners, 3): obj_parse's Vertex.position per corner,
in its corner order (one corner per index; its index array is the identity).
Outputs come from oracle/_ref/libref.so (oracle/Makefile). Run from the
repository root where the reference build exists:
    make -C oracle && python tests/golden/make_asset_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
from make_obj_golden import RES, ref_lib, ref_parse  # noqa: E402

FILES = {f"spot_hull_{k}": RES / "spot" / f"spot-hull-{k}.obj" for k in range(1, 12)}
FILES.update({"mirror_cube_collider1": RES / "mirror_cube_collider1.obj",
              "mirror_cube_collider2": RES / "mirror_cube_collider2.obj",
              "cube": RES / "cube.obj"})


def main():
    lib = ref_lib()
    out = {}
    for name, path in FILES.items():
        v = ref_parse(lib, path)
        out[name] = np.ascontiguousarray(v[:, :3])
        print(f"{name}: {len(v)} corners, {len(np.unique(v[:, :3], axis=0))} distinct positions")
    np.savez_compressed(HERE / "assets.npz", **out)


if __name__ == "__main__":
    main()
