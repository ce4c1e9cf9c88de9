"""The C ABI of librawphys.so: it loads, exports every entry point that
include/rawphys.h declares (plus the reference's own C++ symbols for the
drop-in), and its struct mirrors have the reference's layout. No GPU needed:
nothing here calls into the device."""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

from _libs import ROOT, pkg

HEADER = ROOT / "include" / "rawphys.h"
LIB = ROOT / "raw-physics_amd" / "librawphys.so"
REF = Path("/root/reference")


def _declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    names = set(re.findall(r"\b(rp_\w+|physics_simulate)\s*\(", text))
    return sorted(n for n in names if not n.startswith("rp_world_desc"))


def _exported():
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB)], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_library_loads_without_a_device_call():
    assert LIB.exists(), "build librawphys.so first (python raw-physics_amd/build.py)"
    lib = pkg.library()
    assert lib.rp_last_error() == 0


def test_every_declared_entry_point_is_exported():
    names = _declared_functions()
    assert len(names) >= 20
    exported = _exported()
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_reference_cpp_symbols_are_exported():
    exported = _exported()
    # mangled pbd_simulate / pbd_simulate_with_constraints (src/physics/pbd.h:93-94)
    assert "_Z12pbd_simulatedPP6Entityjji" in exported
    assert "_Z29pbd_simulate_with_constraintsdPP6EntityP10Constraintjji" in exported


@pytest.mark.skipif(not (REF / "src" / "entity.h").exists(), reason="reference headers not present")
def test_struct_mirrors_match_reference_layout(tmp_path):
    fields = {
        "Entity": ("rp_entity", ["id", "mesh", "world_position", "world_rotation", "world_scale", "color",
                                 "colliders", "bounding_sphere_radius", "forces", "inverse_mass",
                                 "inertia_tensor", "inverse_inertia_tensor", "angular_velocity",
                                 "linear_velocity", "fixed", "active", "deactivation_time",
                                 "static_friction_coefficient", "dynamic_friction_coefficient",
                                 "restitution_coefficient", "previous_world_position", "previous_world_rotation",
                                 "previous_linear_velocity", "previous_angular_velocity"]),
        "Constraint": ("rp_constraint", ["type", "e1_id", "e2_id"]),
        "Collider": ("rp_collider", ["type"]),
        "Physics_Force": ("rp_physics_force", ["position", "force", "local_coords"]),
        "Collider_Convex_Hull": ("rp_convex_hull", ["vertices", "transformed_vertices", "faces",
                                                    "transformed_faces", "vertex_to_faces", "vertex_to_neighbors",
                                                    "face_to_neighbors"]),
        "Hinge_Joint_Constraint": ("rp_hinge_joint_constraint", ["r1_lc", "r2_lc", "compliance", "lambda_pos",
                                                                 "e1_aligned_axis", "e2_aligned_axis",
                                                                 "lambda_aligned_axes", "limited", "upper_limit",
                                                                 "lower_limit", "e1_limit_axis", "e2_limit_axis",
                                                                 "lambda_limit_axes"]),
        "Spherical_Joint_Constraint": ("rp_spherical_joint_constraint",
                                       ["r1_lc", "r2_lc", "lambda_pos", "lambda_swing", "swing_upper_limit",
                                        "swing_lower_limit", "e1_swing_axis", "e2_swing_axis", "lambda_twist",
                                        "twist_upper_limit", "twist_lower_limit", "e1_twist_axis",
                                        "e2_twist_axis"]),
    }
    lines = ['#include "entity.h"', '#include "physics/pbd.h"', f'#include "{HEADER}"', "#include <cstddef>"]
    for ref, (mine, fs) in fields.items():
        lines.append(f'static_assert(sizeof({ref}) == sizeof({mine}), "{ref} size");')
        for f in fs:
            lines.append(f'static_assert(offsetof({ref}, {f}) == offsetof({mine}, {f}), "{ref}.{f}");')
    lines.append('static_assert(offsetof(Constraint, collision_constraint) == offsetof(rp_constraint, u), "union");')
    lines.append('static_assert(offsetof(Collider, convex_hull) == offsetof(rp_collider, u), "union");')
    lines.append("int main() { return 0; }")
    src = tmp_path / "layout.cpp"
    src.write_text("\n".join(lines) + "\n")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-w", f"-I{REF}/include", f"-I{REF}/src",
                        "-DGLEW_NO_GLU", str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


def test_oracle_is_not_linked_into_the_product():
    out = subprocess.run(["nm", "-D", str(LIB)], capture_output=True, text=True, check=True).stdout
    assert "orc_" not in out
    deps = subprocess.run(["ldd", str(LIB)], capture_output=True, text=True).stdout
    assert "oracle" not in deps and "libref" not in deps
    assert ctypes.CDLL(str(LIB))  # loads on its own
