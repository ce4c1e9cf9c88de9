/*
 * rp_oracle.h — CPU restatement of raw-physics' rigid-body step (pbd_simulate).
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP product
 * path is compared against; it is never linked into, loaded by, or used as a
 * fallback of the product library. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.
 *
 * The restatement follows the reference's evaluation order operation by
 * operation (citations are to files of felipeek/raw-physics). It is pinned
 * by golden vectors produced by the compiled reference (oracle/_ref) and by
 * closed-form / invariant tests; see tests/.
 *
 * The world descriptor below is shared, field for field, by the product's
 * C ABI (include/rawphys.h: rp_world_desc) so that tests can build one
 * description and hand it to both.
 */
#ifndef RP_ORACLE_H
#define RP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Constraint types: numeric values of Constraint_Type (src/physics/pbd.h:14-20). */
enum {
    ORC_POSITIONAL_CONSTRAINT = 0,
    ORC_COLLISION_CONSTRAINT = 1,
    ORC_MUTUAL_ORIENTATION_CONSTRAINT = 2,
    ORC_HINGE_JOINT_CONSTRAINT = 3,
    ORC_SPHERICAL_JOINT_CONSTRAINT = 4
};

/* External constraint description (payloads of Constraint, src/physics/pbd.h:22-91).
 * body1/body2 are scene-local body indices. */
typedef struct {
    int32_t type;
    uint32_t body1, body2;
    int32_t axis[4];   /* hinge: e1_aligned, e2_aligned, e1_limit, e2_limit;
                          spherical: e1_swing, e2_swing, e1_twist, e2_twist (PBD_Axis_Type values) */
    int32_t limited;   /* hinge only */
    int32_t pad;
    double r1_lc[3], r2_lc[3];
    double compliance; /* positional, mutual orientation, hinge */
    double distance[3];/* positional */
    double limits[4];  /* hinge: lower, upper; spherical: swing lower, swing upper, twist lower, twist upper */
} orc_constraint_desc;

/* A batch of independent scenes. Shapes are triangle corner streams exactly as
 * obj_parse (src/render/obj.cpp:53-115) produces them: float xyz per corner,
 * three corners per triangle. A body's hull is built from its shape's corner
 * stream widened to double and multiplied by the body's scale, as
 * examples_util_create_single_convex_hull_collider_array does
 * (src/examples/examples_util.cpp:5-24). */
typedef struct {
    uint32_t num_shapes;
    const uint32_t* shape_corner_offset;  /* [num_shapes+1], in corners */
    const float* shape_corners;           /* [3 * total corners] */
    uint32_t num_scenes;
    const uint32_t* scene_body_offset;    /* [num_scenes+1] */
    const uint32_t* body_shape;           /* [nb] */
    const double* body_scale;             /* [3nb] */
    const double* body_position;          /* [3nb] */
    const double* body_rotation;          /* [4nb] x,y,z,w */
    const double* body_linear_velocity;   /* [3nb] */
    const double* body_angular_velocity;  /* [3nb] */
    const double* body_mass;              /* [nb] (ignored for fixed bodies) */
    const int32_t* body_fixed;            /* [nb] */
    const double* body_material;          /* [3nb] static friction, dynamic friction, restitution */
    const uint32_t* scene_constraint_offset; /* [num_scenes+1] or NULL */
    const orc_constraint_desc* constraints;  /* external constraints, scene-grouped */
    uint32_t large_scene_bodies;             /* the device path's large-scene threshold (unused here) */
    /* Spheres and compound bodies (NULL: every shape is a hull, every body
     * has the one collider body_shape[g]). A shape with a radius > 0 is a
     * sphere collider of that radius (collider_sphere_create, collider.cpp:12-18;
     * body_scale does not apply). Body g's colliders are
     * collider_shape[body_collider_offset[g] .. body_collider_offset[g+1]). */
    const float* shape_sphere_radius;        /* [num_shapes] or NULL */
    const uint32_t* body_collider_offset;    /* [nb+1] or NULL */
    const uint32_t* collider_shape;          /* [colliders] */
} orc_world_desc;

/* Per-body state, read and written as flat arrays (any pointer may be NULL). */
typedef struct {
    double* position;          /* [3nb] */
    double* rotation;          /* [4nb] */
    double* linear_velocity;   /* [3nb] */
    double* angular_velocity;  /* [3nb] */
    double* previous_position; /* [3nb] */
    double* previous_rotation; /* [4nb] */
    double* previous_linear_velocity;  /* [3nb] */
    double* previous_angular_velocity; /* [3nb] */
    int32_t* active;           /* [nb] */
    double* deactivation_time; /* [nb] */
} orc_state;

typedef struct orc_world orc_world;

/* Error bits: the conditions where the reference would abort (assert) or
 * read uninitialised memory. */
enum {
    ORC_ERR_GJK_SIMPLEX = 1u << 0,    /* add_to_simplex with num==4 (gjk.cpp:24-26) */
    ORC_ERR_EPA_DEGENERATE = 1u << 1, /* zero face normal / flat polytope (epa.cpp:41,72) */
    ORC_ERR_SUPPORT_NAN = 1u << 2,    /* no support index selected (support.cpp:6-16) */
    ORC_ERR_SKEW_LINES = 1u << 3,     /* parallel edges in edge contact (clipping.cpp:279) */
    ORC_ERR_ACTIVITY = 1u << 4,       /* pair of non-fixed bodies with different activity (pbd.cpp:589-591) */
    ORC_ERR_HULL = 1u << 5,           /* hull construction failed */
    ORC_ERR_ZERO_MASS = 1u << 6,      /* w1 + w2 == 0 (pbd_base_constraints.cpp:70,184) */
    ORC_ERR_SELECT = 1u << 7          /* face/edge selection left uninitialised (clipping.cpp:136-200) */
};

orc_world* orc_world_create(const orc_world_desc* desc);
void orc_world_destroy(orc_world* w);
/* Contact log (see rp_oracle.c): 1 = record from now, 2 = replay from the
 * start of the log, 0 = off. The error flag is set when a replayed visit
 * does not match the logged pair. */
void orc_world_contact_log(orc_world* w, int mode);
int orc_world_contact_log_error(const orc_world* w);
void orc_world_contact_log_copy(orc_world* dst, const orc_world* src);
/* Solve order for the next frame of one scene (a check of a parallel solve
 * that groups the frame's items into classes): item t < external
 * constraints is external constraint t, item (external constraints) + k the
 * k-th body pair the narrow phase visits, with all of its contacts. The
 * frame solves the items by ascending class and in item order within a
 * class, in both position iterations and the velocity solve; otherwise the
 * step is unchanged. Status bits (cumulative): 1 the frame's item count
 * differs from n, 2 two items of one class share a non-fixed body, 4 a
 * class outside [0, 64); with bit 1 or 4 the frame runs in list order. */
void orc_world_set_solve_order(orc_world* w, uint32_t scene, const int32_t* item_class, uint32_t n);
uint32_t orc_world_solve_order_status(const orc_world* w);
uint32_t orc_world_num_bodies(const orc_world* w);

/* Runs `frames` frames of the reference's caller pattern on every scene
 * (src/examples/stack.cpp:86-104): a force (0, -gravity*1.0/inverse_mass, 0) at
 * the body origin, then pbd_simulate_with_constraints(dt, scene, constraints,
 * num_substeps, num_pos_iters, enable_collisions). */
void orc_world_step(orc_world* w, double dt, uint32_t frames, uint32_t num_substeps,
                    uint32_t num_pos_iters, int32_t enable_collisions, double gravity);
void orc_world_get_state(const orc_world* w, const orc_state* out);
void orc_world_set_state(orc_world* w, const orc_state* in);
uint32_t orc_world_error(const orc_world* w, uint32_t scene);
/* Number of EPA runs that did not converge (epa.cpp:233-236), summed over the world. */
uint64_t orc_world_epa_failures(const orc_world* w);

/* Derived body data computed at creation (entity.cpp:24-65). */
void orc_world_body_info(const orc_world* w, double* inverse_mass, double* inertia, double* inverse_inertia,
                         double* bounding_radius);

/* Broad phase of the current poses for one scene (broad.cpp:6-29): pairs of
 * scene-local indices (i<j, i-major). Returns the pair count; writes at most
 * `cap` pairs. */
uint32_t orc_world_broad_pairs(const orc_world* w, uint32_t scene, uint32_t* pairs, uint32_t cap);

/* Contacts of two bodies at their current poses (colliders_get_contacts,
 * collider.cpp:560-572). Each contact is 9 doubles: point1, point2, normal.
 * Returns the count; writes at most `cap`. */
uint32_t orc_world_contacts(orc_world* w, uint32_t scene, uint32_t i, uint32_t j, double* out, uint32_t cap);

/* Statistics of the last step: collision constraints summed over scenes and
 * substeps, and the candidate pairs handed to the narrow phase. */
uint64_t orc_world_last_contacts(const orc_world* w);
uint64_t orc_world_last_narrow_pairs(const orc_world* w);

/* Hull topology of a body (collider_convex_hull_create, collider.cpp:194-364).
 * Fills the caller's arrays when non-NULL; the counts are always returned in
 * `counts`: nv, nf, face elements, vertex_to_faces, vertex_to_neighbors,
 * face_to_neighbors. CSR offsets have nv+1 / nf+1 entries. */
void orc_world_hull(const orc_world* w, uint32_t body, uint32_t counts[6], double* vertices, double* face_normals,
                    uint32_t* face_off, uint32_t* face_el, uint32_t* vf_off, uint32_t* vf, uint32_t* vn_off,
                    uint32_t* vn, uint32_t* ff_off, uint32_t* ff);

#ifdef __cplusplus
}
#endif
#endif
