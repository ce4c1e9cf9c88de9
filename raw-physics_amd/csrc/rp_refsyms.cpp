// rp_refsyms.cpp — the constraint constructors of src/physics/pbd.cpp with
// the reference's C++ signatures (src/physics/pbd.h:96-102), so that
// librawphys.so replaces that whole translation unit: a program built from
// the reference drops pbd.cpp and links this library instead.
//
// The types are declared here under the reference's own names so that the
// exported symbols carry the reference's mangling; their layouts are the
// ones mirrored in include/rawphys.h (checked by static_assert below).
#include <stdint.h>
#include <string.h>

#include "../../include/rawphys.h"

#pragma pack(push, 1)
typedef union {
    struct { double x, y, z; };
    struct { double r, g, b; };
} vec3;  // include/gm.h:45-57
#pragma pack(pop)
typedef enum {
    PBD_POSITIVE_X_AXIS,
    PBD_NEGATIVE_X_AXIS,
    PBD_POSITIVE_Y_AXIS,
    PBD_NEGATIVE_Y_AXIS,
    PBD_POSITIVE_Z_AXIS,
    PBD_NEGATIVE_Z_AXIS
} PBD_Axis_Type;  // src/physics/pbd.h:5-12
struct Constraint;

static_assert(sizeof(vec3) == sizeof(rp_vec3), "vec3 layout");
static_assert(sizeof(PBD_Axis_Type) == sizeof(int32_t), "PBD_Axis_Type layout");

#define REFSYM __attribute__((visibility("default")))

static rp_vec3 V(vec3 v) { rp_vec3 r = {v.x, v.y, v.z}; return r; }
static rp_constraint* C(Constraint* c) { return reinterpret_cast<rp_constraint*>(c); }

// pbd.cpp:18-26
REFSYM void pbd_positional_constraint_init(Constraint* constraint, uint64_t e1_id, uint64_t e2_id, vec3 r1_lc,
                                           vec3 r2_lc, double compliance, vec3 distance) {
    rp_constraint* c = C(constraint);
    c->type = RP_POSITIONAL_CONSTRAINT;
    c->e1_id = e1_id;
    c->e2_id = e2_id;
    c->u.positional_constraint.r1_lc = V(r1_lc);
    c->u.positional_constraint.r2_lc = V(r2_lc);
    c->u.positional_constraint.compliance = compliance;
    c->u.positional_constraint.distance = V(distance);
}

// pbd.cpp:28-33
REFSYM void pbd_mutual_orientation_constraint_init(Constraint* constraint, uint64_t e1_id, uint64_t e2_id,
                                                   double compliance) {
    rp_constraint* c = C(constraint);
    c->type = RP_MUTUAL_ORIENTATION_CONSTRAINT;
    c->e1_id = e1_id;
    c->e2_id = e2_id;
    c->u.mutual_orientation_constraint.compliance = compliance;
}

// pbd.cpp:35-45
REFSYM void pbd_hinge_joint_constraint_unlimited_init(Constraint* constraint, uint64_t e1_id, uint64_t e2_id,
                                                      vec3 r1_lc, vec3 r2_lc, double compliance,
                                                      PBD_Axis_Type e1_aligned_axis, PBD_Axis_Type e2_aligned_axis) {
    rp_constraint* c = C(constraint);
    c->type = RP_HINGE_JOINT_CONSTRAINT;
    c->e1_id = e1_id;
    c->e2_id = e2_id;
    rp_hinge_joint_constraint& h = c->u.hinge_joint_constraint;
    h.r1_lc = V(r1_lc);
    h.r2_lc = V(r2_lc);
    h.compliance = compliance;
    h.e1_aligned_axis = e1_aligned_axis;
    h.e2_aligned_axis = e2_aligned_axis;
    h.limited = 0;
}

// pbd.cpp:47-62
REFSYM void pbd_hinge_joint_constraint_limited_init(Constraint* constraint, uint64_t e1_id, uint64_t e2_id, vec3 r1_lc,
                                                    vec3 r2_lc, double compliance, PBD_Axis_Type e1_aligned_axis,
                                                    PBD_Axis_Type e2_aligned_axis, PBD_Axis_Type e1_limit_axis,
                                                    PBD_Axis_Type e2_limit_axis, double lower_limit,
                                                    double upper_limit) {
    rp_constraint* c = C(constraint);
    c->type = RP_HINGE_JOINT_CONSTRAINT;
    c->e1_id = e1_id;
    c->e2_id = e2_id;
    rp_hinge_joint_constraint& h = c->u.hinge_joint_constraint;
    h.r1_lc = V(r1_lc);
    h.r2_lc = V(r2_lc);
    h.compliance = compliance;
    h.e1_aligned_axis = e1_aligned_axis;
    h.e2_aligned_axis = e2_aligned_axis;
    h.limited = 1;
    h.e1_limit_axis = e1_limit_axis;
    h.e2_limit_axis = e2_limit_axis;
    h.lower_limit = lower_limit;
    h.upper_limit = upper_limit;
}

// pbd.cpp:64-79
REFSYM void pbd_spherical_joint_constraint_init(Constraint* constraint, uint64_t e1_id, uint64_t e2_id, vec3 r1_lc,
                                                vec3 r2_lc, PBD_Axis_Type e1_swing_axis, PBD_Axis_Type e2_swing_axis,
                                                PBD_Axis_Type e1_twist_axis, PBD_Axis_Type e2_twist_axis,
                                                double swing_lower_limit, double swing_upper_limit,
                                                double twist_lower_limit, double twist_upper_limit) {
    rp_constraint* c = C(constraint);
    c->type = RP_SPHERICAL_JOINT_CONSTRAINT;
    c->e1_id = e1_id;
    c->e2_id = e2_id;
    rp_spherical_joint_constraint& s = c->u.spherical_joint_constraint;
    s.r1_lc = V(r1_lc);
    s.r2_lc = V(r2_lc);
    s.e1_swing_axis = e1_swing_axis;
    s.e2_swing_axis = e2_swing_axis;
    s.e1_twist_axis = e1_twist_axis;
    s.e2_twist_axis = e2_twist_axis;
    s.swing_lower_limit = swing_lower_limit;
    s.swing_upper_limit = swing_upper_limit;
    s.twist_lower_limit = twist_lower_limit;
    s.twist_upper_limit = twist_upper_limit;
}
