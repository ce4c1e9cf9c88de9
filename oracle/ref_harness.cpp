/*
 * ref_harness.cpp — drives the reference's own step, compiled from its
 * sources under /root/reference by oracle/Makefile into oracle/_ref/libref.so.
 *
 * TEST INFRASTRUCTURE ONLY: it exists to pin the restatement in rp_oracle.c
 * (and to time the reference on the CPU). It builds entities through the
 * reference's own entity_create / collider_convex_hull_create and runs the
 * reference's caller pattern (src/examples/stack.cpp:86-104), exposing the
 * same flat world API as rp_oracle.h under the ref_ prefix.
 */
#define GRAPHICS_MATH_IMPLEMENT
#include <gm.h>
#define C_FEK_HASH_MAP_IMPLEMENT
#include <hash_map.h>
#include <light_array.h>
#include <string.h>
#include <vector>

#ifdef RP_LINK_DROPIN
/* Built against librawphys.so instead of pbd.o: the step's declarations keep
 * default visibility so that the calls bind to the library's exports. */
#pragma GCC visibility push(default)
#include "physics/pbd.h"
extern "C" void physics_simulate(Entity** entities, r64 dt);
#pragma GCC visibility pop
#else
#include "physics/pbd.h"
#endif
#include "physics/collider.h"
#include "physics/broad.h"
#include "examples/examples_util.h"
#include "render/obj.h"
#include "entity.h"
#include "rp_oracle.h"

#define API extern "C" __attribute__((visibility("default")))

struct ref_world {
    std::vector<Entity**> scenes;       /* light_array per scene */
    std::vector<Constraint*> constraints; /* light_array per scene or NULL */
};

static int g_module_ready = 0;

API void* ref_world_create(const orc_world_desc* d) {
    if (!g_module_ready) {
        entity_module_init();
        g_module_ready = 1;
    }
    ref_world* w = new ref_world();
    for (uint32_t si = 0; si < d->num_scenes; ++si) {
        Entity** ents = array_new(Entity*);
        std::vector<eid> ids;
        for (uint32_t g = d->scene_body_offset[si]; g < d->scene_body_offset[si + 1]; ++g) {
            vec3 scale = (vec3){d->body_scale[3 * g], d->body_scale[3 * g + 1], d->body_scale[3 * g + 2]};
            /* The body's colliders in order: spheres by collider_sphere_create,
             * hulls from the corner stream obj_parse would return
             * (src/render/obj.cpp:45-107) as examples_util does (examples_util.cpp:5-49). */
            uint32_t k0 = d->body_collider_offset ? d->body_collider_offset[g] : 0;
            uint32_t nc = d->body_collider_offset ? d->body_collider_offset[g + 1] - k0 : 1;
            Collider* cols = array_new(Collider);
            for (uint32_t ci = 0; ci < nc; ++ci) {
                uint32_t shp = d->body_collider_offset ? d->collider_shape[k0 + ci] : d->body_shape[g];
                if (d->shape_sphere_radius && d->shape_sphere_radius[shp] > 0.0f) {
                    Collider sph = collider_sphere_create(d->shape_sphere_radius[shp]);
                    array_push(cols, sph);
                    continue;
                }
                uint32_t c0 = d->shape_corner_offset[shp], c1 = d->shape_corner_offset[shp + 1];
                Vertex* verts = array_new(Vertex);
                u32* idx = array_new(u32);
                for (uint32_t c = c0; c < c1; ++c) {
                    Vertex v;
                    memset(&v, 0, sizeof(v));
                    v.position.x = d->shape_corners[3 * c];
                    v.position.y = d->shape_corners[3 * c + 1];
                    v.position.z = d->shape_corners[3 * c + 2];
                    array_push(verts, v);
                    u32 k = c - c0;
                    array_push(idx, k);
                }
                Collider hull = examples_util_create_convex_hull_collider(verts, idx, scale);
                array_push(cols, hull);
                array_free(verts);
                array_free(idx);
            }
            Mesh mesh;
            memset(&mesh, 0, sizeof(mesh));
            vec3 pos = (vec3){d->body_position[3 * g], d->body_position[3 * g + 1], d->body_position[3 * g + 2]};
            Quaternion rot = (Quaternion){d->body_rotation[4 * g], d->body_rotation[4 * g + 1],
                                          d->body_rotation[4 * g + 2], d->body_rotation[4 * g + 3]};
            vec4 color = (vec4){1.0, 1.0, 1.0, 1.0};
            const double* m = d->body_material + 3 * g;
            eid id = d->body_fixed[g] ? entity_create_fixed(mesh, pos, rot, scale, color, cols, m[0], m[1], m[2])
                                      : entity_create(mesh, pos, rot, scale, color, d->body_mass[g], cols, m[0], m[1], m[2]);
            Entity* e = entity_get_by_id(id);
            e->linear_velocity = (vec3){d->body_linear_velocity[3 * g], d->body_linear_velocity[3 * g + 1],
                                        d->body_linear_velocity[3 * g + 2]};
            e->angular_velocity = (vec3){d->body_angular_velocity[3 * g], d->body_angular_velocity[3 * g + 1],
                                         d->body_angular_velocity[3 * g + 2]};
            /* previous pose is never read before the first substep writes it */
            e->previous_world_position = (vec3){0.0, 0.0, 0.0};
            e->previous_world_rotation = (Quaternion){0.0, 0.0, 0.0, 0.0};
            array_push(ents, e);
            ids.push_back(id);
        }
        Constraint* cons = NULL;
        if (d->scene_constraint_offset && d->constraints &&
            d->scene_constraint_offset[si + 1] > d->scene_constraint_offset[si]) {
            cons = array_new(Constraint);
            for (uint32_t k = d->scene_constraint_offset[si]; k < d->scene_constraint_offset[si + 1]; ++k) {
                const orc_constraint_desc* cd = &d->constraints[k];
                Constraint c;
                memset(&c, 0, sizeof(c));
                eid a = ids[cd->body1], b = ids[cd->body2];
                vec3 r1 = (vec3){cd->r1_lc[0], cd->r1_lc[1], cd->r1_lc[2]};
                vec3 r2 = (vec3){cd->r2_lc[0], cd->r2_lc[1], cd->r2_lc[2]};
                switch (cd->type) {
                case POSITIONAL_CONSTRAINT:
                    pbd_positional_constraint_init(&c, a, b, r1, r2, cd->compliance,
                                                   (vec3){cd->distance[0], cd->distance[1], cd->distance[2]});
                    break;
                case MUTUAL_ORIENTATION_CONSTRAINT:
                    pbd_mutual_orientation_constraint_init(&c, a, b, cd->compliance);
                    break;
                case HINGE_JOINT_CONSTRAINT:
                    if (cd->limited)
                        pbd_hinge_joint_constraint_limited_init(&c, a, b, r1, r2, cd->compliance,
                                                                (PBD_Axis_Type)cd->axis[0], (PBD_Axis_Type)cd->axis[1],
                                                                (PBD_Axis_Type)cd->axis[2], (PBD_Axis_Type)cd->axis[3],
                                                                cd->limits[0], cd->limits[1]);
                    else
                        pbd_hinge_joint_constraint_unlimited_init(&c, a, b, r1, r2, cd->compliance,
                                                                  (PBD_Axis_Type)cd->axis[0], (PBD_Axis_Type)cd->axis[1]);
                    break;
                case SPHERICAL_JOINT_CONSTRAINT:
                    pbd_spherical_joint_constraint_init(&c, a, b, r1, r2, (PBD_Axis_Type)cd->axis[0],
                                                        (PBD_Axis_Type)cd->axis[1], (PBD_Axis_Type)cd->axis[2],
                                                        (PBD_Axis_Type)cd->axis[3], cd->limits[0], cd->limits[1],
                                                        cd->limits[2], cd->limits[3]);
                    break;
                }
                array_push(cons, c);
            }
        }
        w->scenes.push_back(ents);
        w->constraints.push_back(cons);
    }
    return w;
}

API void ref_world_destroy(void* wp) {
    ref_world* w = (ref_world*)wp;
    for (size_t s = 0; s < w->scenes.size(); ++s) {
        Entity** ents = w->scenes[s];
        for (u32 i = 0; i < array_length(ents); ++i) {
            colliders_destroy(ents[i]->colliders);
            array_free(ents[i]->colliders);
            entity_destroy(ents[i]);
        }
        array_free(ents);
        if (w->constraints[s]) array_free(w->constraints[s]);
    }
    delete w;
}

/* The frame driver of src/examples/stack.cpp:86-104 (arm.cpp:141-160 when
 * the scene carries external constraints). */
API void ref_world_step(void* wp, double dt, uint32_t frames, uint32_t substeps, uint32_t iters, int32_t collisions,
                        double gravity) {
    ref_world* w = (ref_world*)wp;
    for (uint32_t f = 0; f < frames; ++f) {
        for (size_t s = 0; s < w->scenes.size(); ++s) {
            Entity** ents = w->scenes[s];
            for (u32 i = 0; i < array_length(ents); ++i)
                colliders_update(ents[i]->colliders, ents[i]->world_position, &ents[i]->world_rotation);
            const r64 GRAVITY = gravity;
            for (u32 i = 0; i < array_length(ents); ++i)
                entity_add_force(ents[i], (vec3){0.0, 0.0, 0.0}, (vec3){0.0, -GRAVITY * 1.0 / ents[i]->inverse_mass, 0.0},
                                 false);
            pbd_simulate_with_constraints(dt, ents, w->constraints[s], substeps, iters, collisions);
            for (u32 i = 0; i < array_length(ents); ++i) entity_clear_forces(ents[i]);
        }
    }
}

/* The frame driver with a choice of entry point: 0 pbd_simulate_with_constraints
 * (as ref_world_step), 1 pbd_simulate (pbd.h:93; scenes without external
 * constraints), 2 physics_simulate(entities, dt) (20 substeps, 1 iteration,
 * collisions on; built against librawphys.so only). Returns 0, or 1 for a
 * mode this build or scene does not have. */
API uint32_t ref_world_step_mode(void* wp, int32_t mode, double dt, uint32_t frames, uint32_t substeps, uint32_t iters,
                                 int32_t collisions, double gravity) {
    ref_world* w = (ref_world*)wp;
#ifndef RP_LINK_DROPIN
    if (mode == 2) return 1;
#endif
    if (mode < 0 || mode > 2) return 1;
    for (size_t s = 0; s < w->scenes.size(); ++s)
        if (mode != 0 && w->constraints[s] && array_length(w->constraints[s])) return 1;
    for (uint32_t f = 0; f < frames; ++f) {
        for (size_t s = 0; s < w->scenes.size(); ++s) {
            Entity** ents = w->scenes[s];
            for (u32 i = 0; i < array_length(ents); ++i)
                colliders_update(ents[i]->colliders, ents[i]->world_position, &ents[i]->world_rotation);
            const r64 GRAVITY = gravity;
            for (u32 i = 0; i < array_length(ents); ++i)
                entity_add_force(ents[i], (vec3){0.0, 0.0, 0.0}, (vec3){0.0, -GRAVITY * 1.0 / ents[i]->inverse_mass, 0.0},
                                 false);
            if (mode == 0) pbd_simulate_with_constraints(dt, ents, w->constraints[s], substeps, iters, collisions);
            else if (mode == 1) pbd_simulate(dt, ents, substeps, iters, collisions);
#ifdef RP_LINK_DROPIN
            else physics_simulate(ents, dt);
#endif
            for (u32 i = 0; i < array_length(ents); ++i) entity_clear_forces(ents[i]);
        }
    }
    return 0;
}

static Entity* body_at(ref_world* w, uint32_t g) {
    for (size_t s = 0; s < w->scenes.size(); ++s) {
        u32 n = array_length(w->scenes[s]);
        if (g < n) return w->scenes[s][g];
        g -= n;
    }
    return NULL;
}

API uint32_t ref_world_num_bodies(void* wp) {
    ref_world* w = (ref_world*)wp;
    uint32_t n = 0;
    for (size_t s = 0; s < w->scenes.size(); ++s) n += array_length(w->scenes[s]);
    return n;
}

API void ref_world_get_state(void* wp, const orc_state* o) {
    ref_world* w = (ref_world*)wp;
    uint32_t n = ref_world_num_bodies(wp);
    for (uint32_t g = 0; g < n; ++g) {
        Entity* e = body_at(w, g);
        if (o->position) memcpy(o->position + 3 * g, &e->world_position, 24);
        if (o->rotation) memcpy(o->rotation + 4 * g, &e->world_rotation, 32);
        if (o->linear_velocity) memcpy(o->linear_velocity + 3 * g, &e->linear_velocity, 24);
        if (o->angular_velocity) memcpy(o->angular_velocity + 3 * g, &e->angular_velocity, 24);
        if (o->previous_position) memcpy(o->previous_position + 3 * g, &e->previous_world_position, 24);
        if (o->previous_rotation) memcpy(o->previous_rotation + 4 * g, &e->previous_world_rotation, 32);
        if (o->previous_linear_velocity) memcpy(o->previous_linear_velocity + 3 * g, &e->previous_linear_velocity, 24);
        if (o->previous_angular_velocity) memcpy(o->previous_angular_velocity + 3 * g, &e->previous_angular_velocity, 24);
        if (o->active) o->active[g] = e->active;
        if (o->deactivation_time) o->deactivation_time[g] = e->deactivation_time;
    }
}

API void ref_world_set_state(void* wp, const orc_state* in) {
    ref_world* w = (ref_world*)wp;
    uint32_t n = ref_world_num_bodies(wp);
    for (uint32_t g = 0; g < n; ++g) {
        Entity* e = body_at(w, g);
        if (in->position) memcpy(&e->world_position, in->position + 3 * g, 24);
        if (in->rotation) memcpy(&e->world_rotation, in->rotation + 4 * g, 32);
        if (in->linear_velocity) memcpy(&e->linear_velocity, in->linear_velocity + 3 * g, 24);
        if (in->angular_velocity) memcpy(&e->angular_velocity, in->angular_velocity + 3 * g, 24);
        if (in->previous_position) memcpy(&e->previous_world_position, in->previous_position + 3 * g, 24);
        if (in->previous_rotation) memcpy(&e->previous_world_rotation, in->previous_rotation + 4 * g, 32);
        if (in->previous_linear_velocity) memcpy(&e->previous_linear_velocity, in->previous_linear_velocity + 3 * g, 24);
        if (in->previous_angular_velocity)
            memcpy(&e->previous_angular_velocity, in->previous_angular_velocity + 3 * g, 24);
        if (in->active) e->active = in->active[g];
        if (in->deactivation_time) e->deactivation_time = in->deactivation_time[g];
    }
}

API void ref_world_body_info(void* wp, double* inv_mass, double* inertia, double* inv_inertia, double* radius) {
    ref_world* w = (ref_world*)wp;
    uint32_t n = ref_world_num_bodies(wp);
    for (uint32_t g = 0; g < n; ++g) {
        Entity* e = body_at(w, g);
        if (inv_mass) inv_mass[g] = e->inverse_mass;
        if (inertia) memcpy(inertia + 9 * g, e->inertia_tensor.data, 72);
        if (inv_inertia) memcpy(inv_inertia + 9 * g, e->inverse_inertia_tensor.data, 72);
        if (radius) radius[g] = e->bounding_sphere_radius;
    }
}

API uint32_t ref_world_broad_pairs(void* wp, uint32_t si, uint32_t* out, uint32_t cap) {
    ref_world* w = (ref_world*)wp;
    Entity** ents = w->scenes[si];
    Broad_Collision_Pair* pairs = broad_get_collision_pairs(ents);
    uint32_t n = array_length(pairs);
    for (uint32_t k = 0; k < n && k < cap; ++k) {
        uint32_t a = 0, b = 0;
        for (u32 i = 0; i < array_length(ents); ++i) {
            if (ents[i]->id == pairs[k].e1_id) a = i;
            if (ents[i]->id == pairs[k].e2_id) b = i;
        }
        out[2 * k] = a;
        out[2 * k + 1] = b;
    }
    array_free(pairs);
    return n;
}

API uint32_t ref_world_contacts(void* wp, uint32_t si, uint32_t i, uint32_t j, double* out, uint32_t cap) {
    ref_world* w = (ref_world*)wp;
    Entity* e1 = w->scenes[si][i];
    Entity* e2 = w->scenes[si][j];
    colliders_update(e1->colliders, e1->world_position, &e1->world_rotation);
    colliders_update(e2->colliders, e2->world_position, &e2->world_rotation);
    Collider_Contact* cs = colliders_get_contacts(e1->colliders, e2->colliders);
    uint32_t n = array_length(cs);
    for (uint32_t k = 0; k < n && k < cap; ++k) memcpy(out + 9 * k, &cs[k], 72);
    array_free(cs);
    return n;
}

API void ref_world_hull(void* wp, uint32_t g, uint32_t counts[6], double* vertices, double* face_normals,
                        uint32_t* face_off, uint32_t* face_el, uint32_t* vf_off, uint32_t* vf, uint32_t* vn_off,
                        uint32_t* vn, uint32_t* ff_off, uint32_t* ff) {
    ref_world* w = (ref_world*)wp;
    Entity* e = body_at(w, g);
    Collider_Convex_Hull* h = &e->colliders[0].convex_hull;
    uint32_t nv = array_length(h->vertices), nf = array_length(h->faces);
    uint32_t nfe = 0, nvf = 0, nvn = 0, nff = 0;
    for (uint32_t i = 0; i < nf; ++i) {
        if (face_off) face_off[i] = nfe;
        if (ff_off) ff_off[i] = nff;
        for (uint32_t k = 0; k < array_length(h->faces[i].elements); ++k, ++nfe)
            if (face_el) face_el[nfe] = h->faces[i].elements[k];
        for (uint32_t k = 0; k < array_length(h->face_to_neighbors[i]); ++k, ++nff)
            if (ff) ff[nff] = h->face_to_neighbors[i][k];
        if (face_normals) memcpy(face_normals + 3 * i, &h->faces[i].normal, 24);
    }
    if (face_off) face_off[nf] = nfe;
    if (ff_off) ff_off[nf] = nff;
    for (uint32_t i = 0; i < nv; ++i) {
        if (vf_off) vf_off[i] = nvf;
        if (vn_off) vn_off[i] = nvn;
        for (uint32_t k = 0; k < array_length(h->vertex_to_faces[i]); ++k, ++nvf)
            if (vf) vf[nvf] = h->vertex_to_faces[i][k];
        for (uint32_t k = 0; k < array_length(h->vertex_to_neighbors[i]); ++k, ++nvn)
            if (vn) vn[nvn] = h->vertex_to_neighbors[i][k];
        if (vertices) memcpy(vertices + 3 * i, &h->vertices[i], 24);
    }
    if (vf_off) vf_off[nv] = nvf;
    if (vn_off) vn_off[nv] = nvn;
    counts[0] = nv; counts[1] = nf; counts[2] = nfe; counts[3] = nvf; counts[4] = nvn; counts[5] = nff;
}

/* The same frame driver with another implementation of the step in place of
 * the reference's pbd_simulate_with_constraints: `fn` has the signature of
 * rp_pbd_simulate (include/rawphys.h). Used to check the drop-in on entities
 * and colliders built by the reference's own code. */
typedef uint32_t (*step_fn)(double, Entity**, Constraint*, uint32_t, uint32_t, int32_t);
API uint32_t ref_world_step_with(void* wp, void* fn, double dt, uint32_t frames, uint32_t substeps, uint32_t iters,
                                 int32_t collisions, double gravity) {
    ref_world* w = (ref_world*)wp;
    step_fn step = (step_fn)fn;
    uint32_t err = 0;
    for (uint32_t f = 0; f < frames; ++f) {
        for (size_t s = 0; s < w->scenes.size(); ++s) {
            Entity** ents = w->scenes[s];
            for (u32 i = 0; i < array_length(ents); ++i)
                colliders_update(ents[i]->colliders, ents[i]->world_position, &ents[i]->world_rotation);
            const r64 GRAVITY = gravity;
            for (u32 i = 0; i < array_length(ents); ++i)
                entity_add_force(ents[i], (vec3){0.0, 0.0, 0.0}, (vec3){0.0, -GRAVITY * 1.0 / ents[i]->inverse_mass, 0.0},
                                 false);
            err |= step(dt, ents, w->constraints[s], substeps, iters, collisions);
            for (u32 i = 0; i < array_length(ents); ++i) entity_clear_forces(ents[i]);
        }
    }
    return err;
}

/* The transformed geometry entity g's colliders hold (what collider_update,
 * collider.cpp:409-438, last wrote): per collider in order, a hull's
 * transformed vertices then its transformed face normals, or a sphere's
 * center, as doubles. Returns the number of doubles (written up to cap). */
API uint32_t ref_world_collider_geometry(void* wp, uint32_t g, double* out, uint32_t cap) {
    Entity* e = body_at((ref_world*)wp, g);
    uint32_t n = 0;
    auto put = [&](vec3 v) {
        if (n + 3 <= cap) memcpy(out + n, &v, 24);
        n += 3;
    };
    for (u32 ci = 0; ci < array_length(e->colliders); ++ci) {
        Collider* c = &e->colliders[ci];
        if (c->type == COLLIDER_TYPE_SPHERE) {
            put(c->sphere.center);
            continue;
        }
        for (u32 k = 0; k < array_length(c->convex_hull.transformed_vertices); ++k)
            put(c->convex_hull.transformed_vertices[k]);
        for (u32 k = 0; k < array_length(c->convex_hull.transformed_faces); ++k)
            put(c->convex_hull.transformed_faces[k].normal);
    }
    return n;
}

/* entity_get_model_matrix (entity.cpp:123-145) of entity g, row-major. */
API void ref_world_model_matrix(void* wp, uint32_t g, double out[16]) {
    mat4 m = entity_get_model_matrix(body_at((ref_world*)wp, g));
    memcpy(out, m.data, 128);
}

/* obj_parse (src/render/obj.cpp:7-151) of a file: copies min(cap, n) Vertex
 * records (32 bytes each) to out and returns n. obj_parse ends the process on
 * a file it cannot load, so only call this with valid files. */
API uint32_t ref_obj_parse(const char* path, void* out, uint32_t cap) {
    Vertex* verts = NULL;
    u32* idx = NULL;
    obj_parse(path, &verts, &idx);
    uint32_t n = array_length(verts);
    for (uint32_t k = 0; k < n; ++k)
        if (idx[k] != k) return UINT32_MAX; /* the index array is the identity */
    if (out) memcpy(out, verts, sizeof(Vertex) * (n < cap ? n : cap));
    array_free(verts);
    array_free(idx);
    return n;
}
