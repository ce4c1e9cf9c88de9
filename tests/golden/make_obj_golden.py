#!/usr/bin/env python3
"""Writes tests/golden/obj_cases.npz: OBJ texts and what the reference's
obj_parse (src/render/obj.cpp:7-151, over its bundled tinyobjloader) returns
for them, plus digests of its output for the reference's own assets.

* Synthetic cases (seeded, written here): numbers in every form the loader's
  decimal parser accepts or half-accepts (signs, leading dots, exponents,
  more than seven fraction digits, trailing garbage, overflow to inf,
  underflow), relative face indices, "v//n" and "v/t" corners, convex and
  concave polygons (ear clipping), degenerate polygons, groups and objects
  between faces, line and point statements, comments, CRLF / CR line ends.
  Every case is one obj_parse can load without undefined behaviour: no index
  0, every corner has a texture coordinate when the file has any and a
  normal when the file has any.
* Reference assets (res/*.obj, res/spot/*.obj): the corner count and a
  SHA-256 of the returned positions and normals (and texture coordinates for
  files that have them), so the assets themselves need not be copied.

The outputs come from oracle/_ref/libref.so (oracle/Makefile compiles
src/render/obj.cpp with the reference's sources). Run from the repository
root where the reference build exists:
    make -C oracle && python tests/golden/make_obj_golden.py
"""
from __future__ import annotations

import ctypes
import hashlib
import random
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
REF_SO = ROOT / "oracle" / "_ref" / "libref.so"
RES = Path("/root/reference/res")


def ref_lib():
    lib = ctypes.CDLL(str(REF_SO))
    lib.ref_obj_parse.restype = ctypes.c_uint32
    lib.ref_obj_parse.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_uint32]
    return lib


def ref_parse(lib, path: Path) -> np.ndarray:
    n = lib.ref_obj_parse(str(path).encode(), None, 0)
    assert n != 0xFFFFFFFF, "obj_parse index array is not the identity"
    out = np.zeros((n, 8), dtype=np.float32)
    if n:
        lib.ref_obj_parse(str(path).encode(), out.ctypes.data, n)
    return out


def output_digest(verts: np.ndarray, with_texcoords: bool) -> bytes:
    cols = verts if with_texcoords else verts[:, :6]
    return hashlib.sha256(np.ascontiguousarray(cols).tobytes()).digest()


# ---------------------------------------------------------------------------
NUMBER_FORMS = [
    lambda r, x: f"{x:.6f}",
    lambda r, x: f"{x:.3f}",
    lambda r, x: f"{x:.12f}",            # more than seven fraction digits
    lambda r, x: f"{x:.17g}",
    lambda r, x: f"{x:e}",
    lambda r, x: f"{x:.9E}",
    lambda r, x: f"{x * 1e-5:.4e}",
    lambda r, x: f"{x:+.5f}",
    lambda r, x: (f"{x:.4f}".replace("0.", ".", 1) if abs(x) < 1 else f"{x:.4f}"),  # leading dot
    lambda r, x: f"{int(x)}",
    lambda r, x: f"{x:.2f}.7",             # trailing garbage after a number
    lambda r, x: f"{x:.3f}x",
]
SPECIAL_NUMBERS = ["-0", "+0", "0.", "-.", ".", "1e", "1e+", "e5", "abc", "1e40", "-1e39", "1e-50",
                   "3.4028235e38", "1.17549435e-38", "1.4e-45", "0.1", "0.30000000000000004",
                   "123456789012345678901234567890", "0.000000000000000000000000123", "+.5e1", "-.25E-1",
                   "7E0", "2.5e+003", "1.0000000596046448", "16777217", "-16777217.5"]


def number(r: random.Random, scale=3.0) -> str:
    if r.random() < 0.12:
        return r.choice(SPECIAL_NUMBERS)
    x = r.uniform(-scale, scale)
    return r.choice(NUMBER_FORMS)(r, x)


def vertex_line(r: random.Random, vals=None) -> str:
    sep = r.choice([" ", "  ", "\t", " \t "])
    nums = vals if vals is not None else [number(r) for _ in range(3)]
    extra = ""
    if r.random() < 0.1:
        extra = sep + number(r)  # a w coordinate (read as a colour, ignored)
    if r.random() < 0.05:
        nums = nums[:2]          # missing z: defaults to 0
    return "v" + sep + sep.join(nums) + extra


def polygon_2d(r: random.Random, n: int, concave: bool):
    """A simple polygon in a plane (star-shaped around the origin)."""
    angles = sorted(r.uniform(0, 2 * np.pi) for _ in range(n))
    pts = []
    for k, a in enumerate(angles):
        rad = r.uniform(0.5, 1.5)
        if concave and k % 2 == 1:
            rad *= 0.35
        pts.append((rad * np.cos(a), rad * np.sin(a)))
    if r.random() < 0.5:
        pts.reverse()
    return pts


def embed(r: random.Random, pts):
    """Places a 2D polygon on a plane; sometimes axis-aligned (exercises the
    projection-axis choice), sometimes collinear (degenerate)."""
    kind = r.choice(["xy", "yz", "xz", "tilted", "tilted", "line"])
    out = []
    for (a, b) in pts:
        if kind == "xy":
            out.append((a, b, 0.25))
        elif kind == "yz":
            out.append((-1.0, a, b))
        elif kind == "xz":
            out.append((a, 2.0, b))
        elif kind == "line":
            out.append((a, 2 * a, -a))
        else:
            out.append((a + 0.3 * b, 0.5 * a - b, 0.2 * a + 0.7 * b))
    return out


def synthetic_case(seed: int) -> str:
    r = random.Random(seed)
    has_vt = r.random() < 0.4
    has_vn = r.random() < 0.4
    lines = [f"# synthetic case {seed}", "mtllib none.mtl"]
    nv = nvt = nvn = 0
    for _ in range(r.randint(2, 6)):
        stmt = r.random()
        if stmt < 0.15:
            lines.append(r.choice(["g", "o"]) + " part" + str(r.randint(0, 9)))
        elif stmt < 0.2:
            lines.append(r.choice(["s off", "s 1", "usemtl mat", "# comment", "", "   ", "vp 0.5 0.5"]))
        # a batch of vertices then faces using them
        k = r.randint(3, 9)
        polys = []
        if r.random() < 0.6:
            n = r.randint(3, 8)
            pts = embed(r, polygon_2d(r, n, concave=r.random() < 0.5))
            for p in pts:
                lines.append(vertex_line(r, [repr(float(np.float32(c))) if r.random() < 0.5 else f"{c:.7f}"
                                             for c in p]))
            polys.append(list(range(nv + 1, nv + n + 1)))
            nv += n
        for _ in range(k):
            lines.append(vertex_line(r))
        nv += k
        for _ in range(k if has_vt else 0):
            lines.append("vt " + " ".join(number(r, 1.0) for _ in range(2)))
        nvt += k if has_vt else 0
        for _ in range(k if has_vn else 0):
            lines.append("vn " + " ".join(number(r, 1.0) for _ in range(3)))
        nvn += k if has_vn else 0
        for _ in range(r.randint(1, 5)):
            size = r.choice([3, 3, 3, 4, 4, 5, 6])
            polys.append([r.randint(1, nv) for _ in range(size)])
        for poly in polys:
            corners = []
            for vi in poly:
                vs = str(vi) if r.random() < 0.7 else str(vi - nv - 1)  # relative index
                ti = r.randint(1, nvt) if has_vt else 0
                ni = r.randint(1, nvn) if has_vn else 0
                if has_vt and has_vn:
                    corners.append(f"{vs}/{ti}/{ni}")
                elif has_vt:
                    corners.append(f"{vs}/{ti}")
                elif has_vn:
                    corners.append(f"{vs}//{ni}")
                else:
                    corners.append(vs)
            sep = r.choice([" ", "  ", "\t"])
            lines.append("f" + sep + sep.join(corners) + r.choice(["", " ", "\t"]))
        if r.random() < 0.1:
            lines.append(f"l {r.randint(1, nv)} {r.randint(1, nv)}")
        if r.random() < 0.1:
            lines.append(f"p {r.randint(1, nv)}")
        if r.random() < 0.1:
            lines.append("f 1 2")  # fewer than three corners: no triangle
    end = r.choice(["\n", "\r\n", "\r", "\n"])
    text = end.join(lines)
    if r.random() < 0.5:
        text += end
    return text


def main() -> None:
    lib = ref_lib()
    texts, outs, has_vt = [], [], []
    with tempfile.TemporaryDirectory() as d:
        for seed in range(300):
            text = synthetic_case(seed)
            path = Path(d) / f"case{seed}.obj"
            path.write_bytes(text.encode())
            verts = ref_parse(lib, path)
            texts.append(text.encode())
            outs.append(verts)
            has_vt.append("\nvt " in "\n" + text.replace("\r", "\n"))
    t_off = np.cumsum([0] + [len(t) for t in texts]).astype(np.int64)
    o_off = np.cumsum([0] + [len(o) for o in outs]).astype(np.int64)
    res_names, res_counts, res_digests, res_vt = [], [], [], []
    if RES.exists():
        for path in sorted(RES.glob("**/*.obj")):
            verts = ref_parse(lib, path)
            with_vt = "\nvt " in "\n" + path.read_text()
            res_names.append(str(path.relative_to(RES)))
            res_counts.append(len(verts))
            res_digests.append(np.frombuffer(output_digest(verts, with_vt), dtype=np.uint8))
            res_vt.append(with_vt)
    np.savez_compressed(
        HERE / "obj_cases.npz",
        text=np.frombuffer(b"".join(texts), dtype=np.uint8), text_offset=t_off,
        vertices=np.concatenate(outs) if outs else np.zeros((0, 8), np.float32), vertex_offset=o_off,
        has_texcoords=np.asarray(has_vt),
        res_name=np.asarray(res_names, dtype="U64"), res_count=np.asarray(res_counts, dtype=np.int64),
        res_digest=np.asarray(res_digests, dtype=np.uint8).reshape(-1, 32), res_texcoords=np.asarray(res_vt))
    print(f"{len(texts)} synthetic cases ({o_off[-1]} corners), {len(res_names)} reference assets")


if __name__ == "__main__":
    main()
