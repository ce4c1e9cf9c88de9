/*
 * rawphys.h — C ABI of the MI355X-native raw-physics step (librawphys.so).
 *
 * Two ways in:
 *
 *  1. Drop-in for the reference's step (the "plain method"). The library
 *     exports, with the reference's C++ linkage and signatures
 *     (src/physics/pbd.h:93-94):
 *         void pbd_simulate(r64 dt, Entity** entities, u32 num_substeps,
 *                           u32 num_pos_iters, boolean enable_collisions);
 *         void pbd_simulate_with_constraints(r64 dt, Entity** entities,
 *                           Constraint* external_constraints, u32 num_substeps,
 *                           u32 num_pos_iters, boolean enable_collisions);
 *     so a program built against the reference headers links against this
 *     library instead of src/physics/pbd.cpp. The same entry points are
 *     exported with C linkage below (rp_pbd_simulate, physics_simulate) for
 *     FFI callers. Entities are read and written in place with the
 *     reference's struct layout (rp_entity mirrors Entity, src/entity.h:17-47).
 *
 *  2. Batched worlds (rp_world_*): thousands of independent scenes resident in
 *     HBM, stepped frame after frame on the GPU without host round trips.
 *
 * All functions are synchronous with respect to the host unless stated.
 * Errors that the reference turns into assert() aborts are reported as bits
 * of a per-scene error word (rp_world_error, rp_last_error).
 */
#ifndef RAWPHYS_H
#define RAWPHYS_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Reference-layout mirrors (include/gm.h:10-134, src/entity.h,
 *      src/physics/collider.h, src/physics/pbd.h) ------------------------- */
#pragma pack(push, 1)
typedef struct { double x, y, z; } rp_vec3;          /* vec3, packed */
typedef struct { double x, y, z, w; } rp_vec4;       /* vec4, packed */
typedef struct { double data[3][3]; } rp_mat3;       /* mat3, packed, row-major */
#pragma pack(pop)
typedef struct { double x, y, z, w; } rp_quaternion; /* Quaternion (src/quaternion.h:5-7) */
typedef struct { uint32_t VAO, VBO, EBO, num_indices; } rp_mesh; /* Mesh (src/render/mesh.h:13-16) */

/* Physics_Force (src/entity.h:11-15), 52 bytes */
typedef struct { rp_vec3 position; rp_vec3 force; int32_t local_coords; } rp_physics_force;

/* Collider_Convex_Hull_Face / Collider_Convex_Hull / Collider_Sphere / Collider
 * (src/physics/collider.h:12-44). Arrays are light_arrays: the element count
 * is ((size_t*)p)[-1] (include/light_array.h:73-89). */
typedef struct { uint32_t* elements; rp_vec3 normal; } rp_hull_face;
typedef struct {
    rp_vec3* vertices;
    rp_vec3* transformed_vertices;
    rp_hull_face* faces;
    rp_hull_face* transformed_faces;
    uint32_t** vertex_to_faces;
    uint32_t** vertex_to_neighbors;
    uint32_t** face_to_neighbors;
} rp_convex_hull;
typedef struct { float radius; rp_vec3 center; } rp_sphere;
enum { RP_COLLIDER_SPHERE = 0, RP_COLLIDER_CONVEX_HULL = 1 };
typedef struct {
    int32_t type;
    union { rp_convex_hull convex_hull; rp_sphere sphere; } u;
} rp_collider;

/* Entity (src/entity.h:17-47), 504 bytes */
typedef struct {
    uint64_t id;
    rp_mesh mesh;
    rp_vec3 world_position;
    rp_quaternion world_rotation;
    rp_vec3 world_scale;
    rp_vec4 color;
    rp_collider* colliders;
    double bounding_sphere_radius;
    rp_physics_force* forces;
    double inverse_mass;
    rp_mat3 inertia_tensor;
    rp_mat3 inverse_inertia_tensor;
    rp_vec3 angular_velocity;
    rp_vec3 linear_velocity;
    int32_t fixed;
    int32_t active;
    double deactivation_time;
    double static_friction_coefficient;
    double dynamic_friction_coefficient;
    double restitution_coefficient;
    rp_vec3 previous_world_position;
    rp_quaternion previous_world_rotation;
    rp_vec3 previous_linear_velocity;
    rp_vec3 previous_angular_velocity;
} rp_entity;

/* Constraint (src/physics/pbd.h:5-91), 144 bytes */
enum {
    RP_POSITIONAL_CONSTRAINT = 0,
    RP_COLLISION_CONSTRAINT = 1,
    RP_MUTUAL_ORIENTATION_CONSTRAINT = 2,
    RP_HINGE_JOINT_CONSTRAINT = 3,
    RP_SPHERICAL_JOINT_CONSTRAINT = 4
};
typedef struct { rp_vec3 r1_lc, r2_lc; double compliance, lambda; rp_vec3 distance; } rp_positional_constraint;
typedef struct { rp_vec3 r1_lc, r2_lc, normal; double lambda_t, lambda_n; } rp_collision_constraint;
typedef struct { double compliance, lambda; } rp_mutual_orientation_constraint;
typedef struct {
    rp_vec3 r1_lc, r2_lc;
    double compliance, lambda_pos;
    int32_t e1_aligned_axis, e2_aligned_axis;
    double lambda_aligned_axes;
    int32_t limited;
    double upper_limit, lower_limit;
    int32_t e1_limit_axis, e2_limit_axis;
    double lambda_limit_axes;
} rp_hinge_joint_constraint;
typedef struct {
    rp_vec3 r1_lc, r2_lc;
    double lambda_pos, lambda_swing, swing_upper_limit, swing_lower_limit;
    int32_t e1_swing_axis, e2_swing_axis;
    double lambda_twist, twist_upper_limit, twist_lower_limit;
    int32_t e1_twist_axis, e2_twist_axis;
} rp_spherical_joint_constraint;
typedef struct {
    int32_t type;
    uint64_t e1_id, e2_id;
    union {
        rp_positional_constraint positional_constraint;
        rp_collision_constraint collision_constraint;
        rp_mutual_orientation_constraint mutual_orientation_constraint;
        rp_hinge_joint_constraint hinge_joint_constraint;
        rp_spherical_joint_constraint spherical_joint_constraint;
    } u;
} rp_constraint;

/* ---- Drop-in step (plain method) -------------------------------------- */

/* pbd_simulate_with_constraints (src/physics/pbd.h:94, pbd.cpp:468-747) with C
 * linkage. `entities` is the reference's light_array of Entity*, and
 * `external_constraints` a light_array of Constraint or NULL. Entities are
 * updated in place, including the reference's side effect on their
 * colliders: with collisions on, every entity of a pair the collision pass
 * visits has its hulls' transformed_vertices / transformed_faces normals and
 * its spheres' centers set for its pose after the last substep's
 * integration (pbd.cpp:598-599, collider.cpp:409-438). Returns the error
 * word (0 on success): bits of RP_ERR_* where the reference would abort.
 * dt <= 0 is a no-op. */
uint32_t rp_pbd_simulate(double dt, rp_entity** entities, rp_constraint* external_constraints,
                         uint32_t num_substeps, uint32_t num_pos_iters, int32_t enable_collisions);

/* physics_simulate(entities, dt): pbd_simulate(dt, entities, 20, 1, true), the
 * reference's default substep configuration (src/examples/stack.cpp:98). */
void physics_simulate(rp_entity** entities, double dt);

/* ---- Batched worlds ---------------------------------------------------- */
typedef struct {
    int32_t type;               /* RP_*_CONSTRAINT except COLLISION */
    uint32_t body1, body2;      /* scene-local body indices */
    int32_t axis[4];            /* hinge: e1/e2 aligned, e1/e2 limit; spherical: e1/e2 swing, e1/e2 twist (PBD_Axis_Type) */
    int32_t limited;            /* hinge */
    int32_t pad;
    double r1_lc[3], r2_lc[3];
    double compliance;          /* positional, mutual orientation, hinge */
    double distance[3];         /* positional */
    double limits[4];           /* hinge: lower, upper; spherical: swing lower, swing upper, twist lower, twist upper */
} rp_constraint_desc;

/* A batch of independent scenes. Shapes are triangle corner streams as the
 * reference's obj_parse returns them (float xyz per corner, 3 corners per
 * triangle); a body's convex hull is built from its shape widened to double
 * and multiplied by the body's scale (src/examples/examples_util.cpp:5-24),
 * with the reference's hull construction (src/physics/collider.cpp:194-364). */
typedef struct {
    uint32_t num_shapes;
    const uint32_t* shape_corner_offset;  /* [num_shapes+1] */
    const float* shape_corners;           /* [3 * corners] */
    uint32_t num_scenes;
    const uint32_t* scene_body_offset;    /* [num_scenes+1] */
    const uint32_t* body_shape;           /* [nb] */
    const double* body_scale;             /* [3nb] */
    const double* body_position;          /* [3nb] */
    const double* body_rotation;          /* [4nb] x,y,z,w */
    const double* body_linear_velocity;   /* [3nb] */
    const double* body_angular_velocity;  /* [3nb] */
    const double* body_mass;              /* [nb], ignored for fixed bodies */
    const int32_t* body_fixed;            /* [nb] */
    const double* body_material;          /* [3nb] static friction, dynamic friction, restitution */
    const uint32_t* scene_constraint_offset; /* [num_scenes+1] or NULL */
    const rp_constraint_desc* constraints;   /* scene-grouped external constraints or NULL */
    /* Scenes with more bodies than this (0: 4096) run the large-scene path:
     * uniform-grid broad phase (same pair list as the all-pairs test),
     * parallel islands, and a graph-coloured solve instead of the reference's
     * sequential constraint order (results then differ from the reference
     * within a tolerance, and do not depend on thread timing). */
    uint32_t large_scene_bodies;
    /* Spheres and compound bodies (NULL: every shape is a hull and every body
     * has the one collider body_shape[g]). A shape with radius > 0 is a sphere
     * collider of that radius (collider_sphere_create, collider.cpp:12-18;
     * body_scale does not apply). Body g's colliders, in order, are
     * collider_shape[body_collider_offset[g] .. body_collider_offset[g+1]). */
    const float* shape_sphere_radius;     /* [num_shapes] or NULL */
    const uint32_t* body_collider_offset; /* [nb+1] or NULL */
    const uint32_t* collider_shape;       /* [colliders] */
} rp_world_desc;

/* Per-body state as flat arrays; NULL members are skipped. */
typedef struct {
    double* position;                  /* [3nb] */
    double* rotation;                  /* [4nb] */
    double* linear_velocity;           /* [3nb] */
    double* angular_velocity;          /* [3nb] */
    double* previous_position;         /* [3nb] */
    double* previous_rotation;         /* [4nb] */
    double* previous_linear_velocity;  /* [3nb] */
    double* previous_angular_velocity; /* [3nb] */
    int32_t* active;                   /* [nb] */
    double* deactivation_time;         /* [nb] */
} rp_state;

typedef struct rp_world rp_world;

/* Error bits (per scene). */
enum {
    RP_ERR_GJK_SIMPLEX = 1u << 0,     /* add_to_simplex with 4 points (gjk.cpp:24-26) */
    RP_ERR_EPA_DEGENERATE = 1u << 1,  /* zero EPA face normal / flat polytope (epa.cpp:41,72) */
    RP_ERR_SUPPORT_NAN = 1u << 2,     /* no support vertex selected (support.cpp:6-16) */
    RP_ERR_SKEW_LINES = 1u << 3,      /* parallel edges in an edge contact (clipping.cpp:279) */
    RP_ERR_ACTIVITY = 1u << 4,        /* colliding non-fixed bodies of different activity (pbd.cpp:589-591) */
    RP_ERR_HULL = 1u << 5,            /* hull construction / inertia inversion failed */
    RP_ERR_ZERO_MASS = 1u << 6,       /* w1 + w2 == 0 (pbd_base_constraints.cpp:70,184) */
    RP_ERR_SELECT = 1u << 7,          /* face / edge / axis selection failed */
    RP_ERR_EPA_CAPACITY = 1u << 8,    /* EPA polytope exceeded its fixed capacity */
    RP_ERR_CLIP_CAPACITY = 1u << 9,   /* clipped polygon exceeded its fixed capacity */
    RP_ERR_CONTACT_CAPACITY = 1u << 10, /* more than 16 contacts for one body pair */
    RP_ERR_SOLVER_CAPACITY = 1u << 11,  /* constraint list exceeded its capacity */
    RP_ERR_PAIR_CAPACITY = 1u << 12,    /* broad-phase pairs of a large scene exceeded its capacity */
    RP_ERR_ADJOINT_SCOPE = 1u << 13,    /* adjoint on a scene outside its scope (one dynamic body) */
    RP_ERR_OBJ = 1u << 14,              /* OBJ file unreadable or malformed (rp_obj_parse) */
    RP_ERR_RUNTIME = 1u << 31         /* HIP runtime failure (see rp_last_error_string) */
};

/* Creates a world on the current HIP device: builds hulls and derived body
 * data (entity.cpp:24-65) on the host and uploads everything to HBM.
 * Returns NULL on failure (rp_last_error_string says why). */
rp_world* rp_world_create(const rp_world_desc* desc);
void rp_world_destroy(rp_world* w);
uint32_t rp_world_num_bodies(const rp_world* w);
uint32_t rp_world_num_scenes(const rp_world* w);

/* Runs `frames` frames of the reference's caller pattern on every scene
 * (src/examples/stack.cpp:86-104): each frame a force (0, -gravity*1.0/m^-1, 0)
 * at every body's origin, then pbd_simulate_with_constraints(dt, scene,
 * constraints, num_substeps, num_pos_iters, enable_collisions).
 * Returns 0 or RP_ERR_RUNTIME. Blocks until the GPU is done. */
uint32_t rp_world_step(rp_world* w, double dt, uint32_t frames, uint32_t num_substeps, uint32_t num_pos_iters,
                       int32_t enable_collisions, double gravity);
/* Same, without waiting for completion (call rp_world_synchronize). */
uint32_t rp_world_step_async(rp_world* w, double dt, uint32_t frames, uint32_t num_substeps, uint32_t num_pos_iters,
                             int32_t enable_collisions, double gravity);
uint32_t rp_world_synchronize(rp_world* w);
/* rp_world_step without allocating: no hipMalloc, hipFree, graph capture or
 * other allocating call, for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation). No
 * counterpart in the reference. Same results as rp_world_step. The kernels
 * are launched one by one on the world's stream (no graph replay, one view)
 * and the narrow phase's lane-group working memory comes from a static
 * device array of 64 MiB (4519 lane-group slots) instead of the world's own
 * buffer; the narrow-phase grids are capped to the slots it holds. Every
 * call on a device shares that device's array: calls that overlap on one
 * device (worlds on different streams, other threads) race; calls on the
 * library's one shared stream are ordered. The world must have been created
 * on the current device. Blocks until done; returns 0 or RP_ERR_RUNTIME
 * (rp_last_error_string says why). */
uint32_t rp_world_step_noalloc(rp_world* w, double dt, uint32_t frames, uint32_t num_substeps, uint32_t num_pos_iters,
                               int32_t enable_collisions, double gravity);

void rp_world_get_state(rp_world* w, const rp_state* out);
void rp_world_set_state(rp_world* w, const rp_state* in);
uint32_t rp_world_error(rp_world* w, uint32_t scene);
/* Every scene's error word, in one copy, into out[num_scenes]. Returns 0, or
 * RP_ERR_RUNTIME if the device copy failed. */
uint32_t rp_world_errors(rp_world* w, uint32_t* out);
/* EPA runs that did not converge since creation (epa.cpp:233-236). */
uint64_t rp_world_epa_failures(rp_world* w);
/* GJK runs (gjk.cpp:350-379) that were ended early on a repeated loop state
 * since creation; each returns what the full 100 iterations would. */
uint64_t rp_world_gjk_cycle_exits(rp_world* w);
/* The HIP stream the world's kernels run on (a hipStream_t). */
void* rp_world_stream(rp_world* w);
/* Every body's model matrix as entity_get_model_matrix computes it
 * (src/entity.cpp:123-145: translation * rotation * scale, with the body's
 * world scale), row-major [nb][4][4]: doubles, or with as_float the float
 * array util_matrix_to_r32_array (src/util.cpp:77-94) makes for the renderer.
 * out_on_device: out is device memory and the kernel is queued on the
 * world's stream without waiting; otherwise out is host memory and the call
 * blocks. Returns 0 or RP_ERR_RUNTIME. */
uint32_t rp_world_model_matrices(rp_world* w, void* out, int32_t as_float, int32_t out_on_device);

/* ---- Gradients (reverse mode over a recorded run) --------------------------
 * Scope: every scene has one dynamic body (any number of fixed bodies), no
 * external constraints, and is not a large scene; otherwise
 * RP_ERR_ADJOINT_SCOPE. The contact set is frozen: the contacts the forward
 * pass generated (r1_lc, r2_lc, n per substep) are constants of the
 * gradient. No counterpart in the reference, which has no gradients. */

/* Mass (1/m, inertia and its inverse as at creation), static and dynamic
 * friction of every non-fixed body; NULL arrays are left unchanged. */
uint32_t rp_world_set_body_params(rp_world* w, const double* mass, const double* mu_s, const double* mu_d);
/* Steps `frames` frames (collisions on) from the current state, keeping a
 * device checkpoint of every frame's start state. */
uint32_t rp_world_adjoint_forward(rp_world* w, double dt, uint32_t frames, uint32_t num_substeps,
                                  uint32_t num_pos_iters, double gravity);
/* dl_dstate [nb][13]: dL/d(x, q(x,y,z,w), v, w) of the state at the end of the
 * recorded run. Out: grad_state [nb][13] (the same at its start) and
 * grad_param [nb][3] (dL/d mass, static friction, dynamic friction). The
 * world is left at the end state. Each frame is recomputed from its
 * checkpoint, then its substeps are reversed. */
uint32_t rp_world_adjoint_backward(rp_world* w, const double* dl_dstate, double* grad_state, double* grad_param);
/* Substeps whose recomputation in the backward pass did not reproduce the
 * recorded forward state bit for bit (0 when the backward pass is exact). */
uint64_t rp_world_adjoint_mismatches(rp_world* w);
/* Per body, after rp_world_adjoint_backward: the largest component of the
 * state adjoint dL/d(x, q, v, w) at any frame start of the recorded run,
 * over the largest component of dl_dstate (0 for bodies with dl_dstate 0).
 * It measures how much the substep Jacobians amplify the gradient. Impacts
 * of a tumbling body multiply it by 1e2-1e4 each; once it is large the
 * piecewise-smooth step flips a branch within any perturbation that finite
 * differences can resolve, and the gradient, although the exact derivative
 * of the frozen-contact map, no longer predicts a perturbed run. For bodies
 * with an amplification of at most 10, gradient entries agree with FP64
 * central differences (h = 1e-6) where those differences are reproducible
 * (h and h/10 agree to 1e-6): over 99 % to 1e-6, all to 1e-5 (the
 * frozen-contact approximation). */
uint32_t rp_world_adjoint_amplification(rp_world* w, double* amplification);

/* Narrow-phase schedules: lanes cooperating on one body pair in the GJK + EPA
 * kernel (narrow group) and in the contact-manifold kernel (manifold group):
 * 1 = one lane per pair over pre-transformed hulls; 4, 8 or 16 = a lane group
 * per pair with hull vertices staged in LDS and split support searches and
 * face / edge selections. Manifold group 0: no manifold kernel; the GJK + EPA
 * kernel builds each converged pair's manifold itself on the hulls it staged.
 * All schedules give identical results. Return 0, or -1 if the schedule does
 * not fit this world (LDS). rp_world_create picks a narrow group of 8 (else
 * 16, 4) where it fits with manifold group 0, else 1 and 1. */
int32_t rp_world_set_narrow_group(rp_world* w, int32_t lanes_per_pair);
int32_t rp_world_narrow_group(const rp_world* w);
int32_t rp_world_set_manifold_group(rp_world* w, int32_t lanes_per_pair);
int32_t rp_world_manifold_group(const rp_world* w);

/* Most blocks (one wave each) of a narrow-phase launch, 1..8192, or 0 for the
 * default 8192. Lane groups take the frame's live pairs in grid-stride
 * order, so fewer blocks give each group several pairs in turn; results are
 * identical. Returns 0 or -1. */
int32_t rp_world_set_narrow_block_cap(rp_world* w, int32_t blocks);

/* Dynamic LDS bytes one narrow-phase block may use in this world. A
 * lane-group schedule of G lanes per pair fits where 64 / G pairs' slices of
 * 2 x (the world's largest hull's vertices) x 24 bytes do. The limit is
 * 64 KiB where the 16-lane schedule fits in it; else it is the device's limit
 * per block (160 KiB on gfx950), and the world then takes lane groups only
 * in the fused schedule (manifold group 0), besides the one-lane schedule
 * (1, 1). */
int32_t rp_world_lds_limit(const rp_world* w);

/* Concurrent views: graph replay cuts the world into up to 4 scene ranges
 * of about equal body counts and replays each on a stream of its own, so
 * that one range's latency-bound kernels overlap another's (scenes never
 * interact, so results are identical). Used with the grouped narrow-phase
 * schedules (narrow group > 1, manifold group 0, 4, 8 or 16) in worlds
 * without sphere or compound pairs. views: 0 = automatic (3 from 64
 * scenes: with the runtime's default of four hardware queues per process,
 * a fourth view's stream shares a queue), 1 = none, 2..4. rp_world_views:
 * the number the last captured frame used. Returns 0 or -1. */
int32_t rp_world_set_views(rp_world* w, int32_t views);
int32_t rp_world_views(const rp_world* w);

/* Solve schedule: lanes serving one scene in the solve kernel (64, 32 or 16,
 * i.e. 1, 2 or 4 scenes per wave). Identical results. Returns 0 or -1. */
int32_t rp_world_set_solve_lanes(rp_world* w, int32_t lanes_per_scene);

/* Large scenes (coloured solve): the first `classes` colour classes are
 * solved one launch each over the whole device, the remaining colours and
 * the serial class by one 1024-thread block with a barrier between classes
 * (0..64; default 64, every class). Items of one colour are independent, so results are
 * identical. Returns 0 or -1. */
int32_t rp_world_set_colour_launches(rp_world* w, int32_t classes);
int32_t rp_world_solve_lanes(const rp_world* w);

/* Frame launches as a HIP graph (default on): one frame's kernels are
 * captured once per step parameters and schedules and replayed per frame.
 * Off: kernels are launched one by one. Identical results. */
void rp_world_set_graphs(rp_world* w, int32_t on);

/* Profiling mode: HIP events around every kernel launch of the step, read
 * after each frame (the frame loop then synchronizes once per frame).
 * rp_world_kernel_times fills ms[8] / launches[8] indexed by kernel class:
 * 0 frame kernels (forces, broad phase, islands, live-pair lists),
 * 1 integrate, 2 hull transform, 3 narrow phase, 4 solve. Enabling resets
 * the totals. */
void rp_world_set_profiling(rp_world* w, int32_t on);
void rp_world_kernel_times(rp_world* w, double* ms, uint64_t* launches);

/* Inspection (used by parity tests): derived body data, broad-phase pairs
 * and contacts at the current poses, hull topology of a body. Same
 * semantics as the reference functions cited in the oracle's header. */
void rp_world_body_info(rp_world* w, double* inverse_mass, double* inertia, double* inverse_inertia,
                        double* bounding_radius);
uint32_t rp_world_broad_pairs(rp_world* w, uint32_t scene, uint32_t* pairs, uint32_t cap);
uint32_t rp_world_contacts(rp_world* w, uint32_t scene, uint32_t i, uint32_t j, double* out, uint32_t cap);
/* Solve classes of a large scene's items in the last frame run: item t is
 * the scene's external constraint t, then (with collisions on) its live pair
 * t - (external constraints) in pair order; class 0..62 is a colour (items
 * of one colour share no non-fixed body and are solved in parallel), 63 the
 * class solved last, in item order. Position and velocity passes take the
 * classes in ascending order. Returns the item count (0 for a scene that is
 * not large) and writes min(count, cap) classes. No counterpart in the
 * reference, whose solve is sequential (pbd.cpp:615-620, 646-711). */
uint32_t rp_world_item_classes(rp_world* w, uint32_t scene, int32_t* classes, uint32_t cap);
void rp_world_hull(rp_world* w, uint32_t body, uint32_t counts[6], double* vertices, double* face_normals,
                   uint32_t* face_off, uint32_t* face_el, uint32_t* vf_off, uint32_t* vf, uint32_t* vn_off,
                   uint32_t* vn, uint32_t* ff_off, uint32_t* ff);

/* Device FP64 primitives on host arrays of n inputs: out[5i..5i+4] =
 * sqrt(a), a/b, asin(b/(|b|+1)), sin(b), cos(b). The step is bit-exact with
 * the reference only if the device sqrt and division are correctly rounded;
 * this entry point lets a test check that on the installed device. */
uint32_t rp_fp64_selftest(const double* a, const double* b, double* out, uint64_t n);

/* ---- OBJ assets --------------------------------------------------------------
 * obj_parse (src/render/obj.h:7, obj.cpp:7-151): the triangles of a Wavefront
 * OBJ file as one vertex per triangle corner, in face order, with the values
 * the reference's bundled tinyobjloader produces (its own decimal parser,
 * then float; polygons ear-clipped as it does). Normals come from the file
 * or, when it has none, are each triangle's unnormalised cross product.
 * obj_parse's index array is 0, 1, 2, ... and is not returned. The hull of a
 * shape is built from the corner positions (rp_world_desc.shape_corners, or
 * collider_convex_hull_create in the reference). */
#pragma pack(push, 1)
typedef struct {
    float position[3];
    float normal[3];
    float texture_coordinates[2]; /* 0 when the file has none */
} rp_vertex;                      /* Vertex (src/render/mesh.h:5-10), 32 B */
#pragma pack(pop)
/* Writes min(capacity, n) corners and sets *num_vertices = n (call with
 * vertices = NULL to size). Returns 0, or RP_ERR_OBJ where obj_parse would
 * exit (unreadable file, index 0) or read out of bounds (an index past the
 * file's arrays); rp_last_error_string says which. */
uint32_t rp_obj_parse(const char* obj_path, rp_vertex* vertices, uint32_t capacity, uint32_t* num_vertices);
/* The same over OBJ text in memory. */
uint32_t rp_obj_parse_memory(const char* text, size_t length, rp_vertex* vertices, uint32_t capacity,
                             uint32_t* num_vertices);

/* Last error of the calling thread: the error word of the last drop-in call
 * and a message for runtime failures. */
uint32_t rp_last_error(void);
const char* rp_last_error_string(void);

#ifdef __cplusplus
}  /* extern "C" */

#ifdef RP_DECLARE_REFERENCE_SYMBOLS
/* The reference's own C++ signatures (src/physics/pbd.h:93-94), exported by
 * librawphys.so. Declared only on request: a program that includes the
 * reference headers already declares them. */
struct Entity;
struct Constraint;
void pbd_simulate(double dt, Entity** entities, unsigned int num_substeps, unsigned int num_pos_iters,
                  int enable_collisions);
void pbd_simulate_with_constraints(double dt, Entity** entities, Constraint* external_constraints,
                                   unsigned int num_substeps, unsigned int num_pos_iters, int enable_collisions);
#endif
#endif
#endif
