/*
 * rp_oracle.c — CPU restatement of raw-physics' rigid-body step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): see rp_oracle.h.
 *
 * Every function cites the reference file:line it restates. Floating-point
 * expressions are written in the reference's order and associativity, and
 * the file is compiled with -O2 -fno-fast-math -ffp-contract=off so that the
 * results are bit-identical to the reference built with the same flags.
 */
#include "rp_oracle.h"
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* The arithmetic type. It is double; the op-counting build (opcount.cpp)
 * compiles this file as C++ with `real` a double wrapper that counts every
 * add, subtract, multiply, divide and square root, attributed to the phase
 * named by ORC_PHASE. RD() reads the double value of a real. */
#if !defined(ORC_COUNT_OPS) && !defined(ORC_DUAL)
typedef double real;
#define RD(x) (x)
#define ORC_PHASE(p) ((void)0)
#endif
/* The dual-number build (dual.cpp) strips derivatives of the contact geometry
 * here when it differentiates with the contact set frozen. */
#ifndef ORC_FREEZE_V3
#define ORC_FREEZE_V3(v) ((void)0)
#endif
enum { PH_FRAME = 0, PH_INTEGRATE, PH_TRANSFORM, PH_NARROW, PH_POSITION, PH_VELOCITY_UPDATE, PH_VELOCITY_SOLVE, PH_SETUP,
       PH_MANIFOLD, PH_COUNT };

/* ------------------------------------------------------------------ */
/* Growable arrays with the reference's swap-remove semantics          */
/* (include/light_array.h:116-151).                                    */
/* ------------------------------------------------------------------ */
#define ARR(T) struct { T* d; uint32_t n, cap; }
#define ARR_PUSH(a, v)                                                          \
    do {                                                                        \
        if ((a).n == (a).cap) {                                                 \
            (a).cap = (a).cap ? 2 * (a).cap : 16;                               \
            (a).d = realloc((a).d, sizeof(*(a).d) * (a).cap);                   \
        }                                                                       \
        (a).d[(a).n++] = (v);                                                   \
    } while (0)
/* array_remove: length--, A[i] = A[length] (light_array.h:146) */
#define ARR_REMOVE(a, i) ((a).n--, (a).d[(i)] = (a).d[(a).n])
#define ARR_FREE(a) free((a).d)

/* ------------------------------------------------------------------ */
/* Vector / quaternion / matrix math (include/gm.h, src/quaternion.cpp) */
/* ------------------------------------------------------------------ */
typedef struct { real x, y, z; } v3;
typedef struct { real x, y, z, w; } qt;
typedef struct { real m[3][3]; } m3;
typedef struct { real m[4][4]; } m4;
typedef struct { int32_t x, y, z; } i3;
typedef struct { int32_t x, y; } i2;
typedef ARR(i3) i3arr;
typedef ARR(i2) i2arr;
typedef ARR(v3) v3arr;
typedef ARR(real) f64arr;

static v3 V(real x, real y, real z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
/* gm.h:667-670 */
static v3 add(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
/* gm.h:679-682 */
static v3 sub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
/* gm.h:615-618 */
static v3 scl(real s, v3 v) { return V(s * v.x, s * v.y, s * v.z); }
/* gm.h:691-694 */
static real dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* gm.h:722-729 */
static v3 cross(v3 a, v3 b) {
    return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* gm.h:649-652 */
static real len(v3 v) { return sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
/* gm.h:627-634 */
static v3 nrm(v3 v) {
    if (!(v.x != 0.0 || v.y != 0.0 || v.z != 0.0)) return V(0.0, 0.0, 0.0);
    real l = len(v);
    return V(v.x / l, v.y / l, v.z / l);
}
/* gm_vec3_invert is 0 - v, not -v (gm.h:588-590) */
static v3 inv3(v3 v) { return sub(V(0.0, 0.0, 0.0), v); }
/* gm.h:600-605 */
static int veq(v3 a, v3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

/* gm.h:370-381 */
static m3 m3_mul(const m3* a, const m3* b) {
    m3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r.m[i][j] = a->m[i][0] * b->m[0][j] + a->m[i][1] * b->m[1][j] + a->m[i][2] * b->m[2][j];
    return r;
}
/* gm.h:420-426 */
static v3 m3_mulv(const m3* m, v3 v) {
    return V(m->m[0][0] * v.x + m->m[0][1] * v.y + m->m[0][2] * v.z,
             m->m[1][0] * v.x + m->m[1][1] * v.y + m->m[1][2] * v.z,
             m->m[2][0] * v.x + m->m[2][1] * v.y + m->m[2][2] * v.z);
}
/* gm.h:445-452 */
static m3 m3_t(const m3* m) {
    m3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = m->m[j][i];
    return r;
}
/* gm.h:347-368 */
static int m3_inverse(const m3* m, m3* out) {
    const real (*a)[3] = m->m;
    real det = a[0][0] * (a[1][1] * a[2][2] - a[2][1] * a[1][2]) -
                 a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                 a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    if (det == 0.0) return 0;
    real invdet = 1 / det;
    out->m[0][0] = (a[1][1] * a[2][2] - a[2][1] * a[1][2]) * invdet;
    out->m[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * invdet;
    out->m[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * invdet;
    out->m[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) * invdet;
    out->m[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * invdet;
    out->m[1][2] = (a[1][0] * a[0][2] - a[0][0] * a[1][2]) * invdet;
    out->m[2][0] = (a[1][0] * a[2][1] - a[2][0] * a[1][1]) * invdet;
    out->m[2][1] = (a[2][0] * a[0][1] - a[0][0] * a[2][1]) * invdet;
    out->m[2][2] = (a[0][0] * a[1][1] - a[1][0] * a[0][1]) * invdet;
    return 1;
}
/* gm.h:327-345 */
static m4 m4_mul(const m4* a, const m4* b) {
    m4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            r.m[i][j] = a->m[i][0] * b->m[0][j] + a->m[i][1] * b->m[1][j] + a->m[i][2] * b->m[2][j] +
                        a->m[i][3] * b->m[3][j];
    return r;
}

/* quaternion.cpp:133-142 */
static qt q_mul(qt a, qt b) {
    qt r;
    r.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
    r.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
    r.y = a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z;
    r.z = a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x;
    return r;
}
/* quaternion.cpp:144-147 */
static qt q_nrm(qt q) {
    real l = sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    qt r;
    r.x = q.x / l; r.y = q.y / l; r.z = q.z / l; r.w = q.w / l;
    return r;
}
/* quaternion.cpp:80-87 */
static qt q_inv(qt q) {
    qt r;
    r.x = -q.x; r.y = -q.y; r.z = -q.z; r.w = q.w;
    return r;
}
/* quaternion.cpp:257-268 */
static v3 q_rot(qt q, v3 v) {
    real ix = q.w * v.x + q.y * v.z - q.z * v.y;
    real iy = q.w * v.y + q.z * v.x - q.x * v.z;
    real iz = q.w * v.z + q.x * v.y - q.y * v.x;
    real iw = -q.x * v.x - q.y * v.y - q.z * v.z;
    return V((ix * q.w) + (iw * -q.x) + (iy * -q.z) - (iz * -q.y),
             (iy * q.w) + (iw * -q.y) + (iz * -q.x) - (ix * -q.z),
             (iz * q.w) + (iw * -q.z) + (ix * -q.y) - (iy * -q.x));
}
/* quaternion.cpp:89-105 */
static m3 q_mat3(qt q) {
    m3 r;
    r.m[0][0] = 1.0 - 2.0 * q.y * q.y - 2.0 * q.z * q.z;
    r.m[1][0] = 2.0 * q.x * q.y + 2.0 * q.w * q.z;
    r.m[2][0] = 2.0 * q.x * q.z - 2.0 * q.w * q.y;
    r.m[0][1] = 2.0 * q.x * q.y - 2.0 * q.w * q.z;
    r.m[1][1] = 1.0 - (2.0 * q.x * q.x) - (2.0 * q.z * q.z);
    r.m[2][1] = 2.0 * q.y * q.z + 2.0 * q.w * q.x;
    r.m[0][2] = 2.0 * q.x * q.z + 2.0 * q.w * q.y;
    r.m[1][2] = 2.0 * q.y * q.z - 2.0 * q.w * q.x;
    r.m[2][2] = 1.0 - (2.0 * q.x * q.x) - (2.0 * q.y * q.y);
    return r;
}
/* quaternion.cpp:107-131 */
static m4 q_mat4(qt q) {
    m3 r3 = q_mat3(q);
    m4 r;
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) r.m[i][j] = r3.m[i][j];
        r.m[i][3] = 0.0;
        r.m[3][i] = 0.0;
    }
    r.m[3][3] = 1.0;
    return r;
}
/* quaternion.cpp:3-16 */
static qt q_from_axis_radians(v3 axis, real angle) {
    if (len(axis) != 0.0) axis = nrm(axis);
    real sang = sin(angle / 2.0);
    qt q;
    q.w = cos(angle / 2.0);
    q.x = axis.x * sang;
    q.y = axis.y * sang;
    q.z = axis.z * sang;
    return q;
}
/* Basis axes of a rotation (quaternion.cpp:33-78), indexed by PBD_Axis_Type
 * (pbd.cpp:219-243). */
static v3 q_axis(qt q, int axis, int* bad) {
    switch (axis) {
    case 0: /* quaternion_get_right */
        return V(1.0 - 2.0 * q.y * q.y - 2.0 * q.z * q.z, 2.0 * q.x * q.y - 2.0 * -q.w * q.z,
                 2.0 * q.x * q.z + 2.0 * -q.w * q.y);
    case 1: /* quaternion_get_right_inverted */
        return V(1.0 - 2.0 * q.y * q.y - 2.0 * q.z * q.z, 2.0 * q.x * q.y - 2.0 * q.w * q.z,
                 2.0 * q.x * q.z + 2.0 * q.w * q.y);
    case 2: /* quaternion_get_up */
        return V(2.0 * q.x * q.y + 2.0 * -q.w * q.z, 1.0 - (2.0 * q.x * q.x) - (2.0 * q.z * q.z),
                 2.0 * q.y * q.z - 2.0 * -q.w * q.x);
    case 3: /* quaternion_get_up_inverted */
        return V(2.0 * q.x * q.y + 2.0 * q.w * q.z, 1.0 - (2.0 * q.x * q.x) - (2.0 * q.z * q.z),
                 2.0 * q.y * q.z - 2.0 * q.w * q.x);
    case 4: /* quaternion_get_forward */
        return V(2.0 * q.x * q.z - 2.0 * -q.w * q.y, 2.0 * q.y * q.z + 2.0 * -q.w * q.x,
                 1.0 - (2.0 * q.x * q.x) - (2.0 * q.y * q.y));
    case 5: /* quaternion_get_forward_inverted */
        return V(2.0 * q.x * q.z - 2.0 * q.w * q.y, 2.0 * q.y * q.z + 2.0 * q.w * q.x,
                 1.0 - (2.0 * q.x * q.x) - (2.0 * q.y * q.y));
    }
    *bad = 1;
    return V(0.0, 0.0, 0.0);
}

#define PI_F 3.14159265358979 /* gm.h:8 */
#define RP_MIN(a, b) (((a) < (b)) ? (a) : (b))  /* common.h:17-19 */
#define RP_MAX(a, b) (((a) > (b)) ? (a) : (b))
#define RP_CLAMP(x, low, high) (((x) > (high)) ? (high) : (((x) < (low)) ? (low) : (x)))

/* ------------------------------------------------------------------ */
/* World data                                                          */
/* ------------------------------------------------------------------ */
typedef struct {
    uint32_t nv, nf;
    v3* v;   /* local vertices (after dedupe) */
    v3* fn;  /* local face normals */
    uint32_t *fe_off, *fe; /* face elements */
    uint32_t *vf_off, *vf; /* vertex_to_faces */
    uint32_t *vn_off, *vn; /* vertex_to_neighbors */
    uint32_t *ff_off, *ff; /* face_to_neighbors */
    v3* tv;  /* transformed vertices */
    v3* tn;  /* transformed face normals */
    int stamp; /* marked for transformation this substep */
} hull;

/* Collider (collider.h:27-44): a convex hull, or a sphere whose center
 * follows the entity position (collider.cpp:431-433). */
enum { COL_HULL = 0, COL_SPHERE = 1 };
typedef struct collider {
    int32_t type;
    hull h;
    float radius;  /* r32 (collider.h:29) */
    v3 center;
} collider;

typedef struct {
    v3 x; qt q; v3 v; v3 w;      /* world_position, world_rotation, linear/angular velocity */
    v3 px; qt pq; v3 pv; v3 pw;  /* previous_* */
    int32_t fixed, active;
    real deact;
    real inv_mass;
    m3 inertia, inv_inertia;
    real radius;
    real mu_s, mu_d, e;
    v3 force, torque;            /* external force / torque of the current frame */
    struct collider* col;        /* the entity's colliders, in order */
    uint32_t ncol;
    int stamp;                   /* colliders to be updated this substep */
} body;

typedef struct {
    int32_t type;
    uint32_t b1, b2;
    /* collision payload (pbd.h:30-36) */
    v3 r1, r2, n;
    real lambda_t, lambda_n;
    /* joint payloads */
    real compliance, lambda, lambda2, lambda3;
    v3 distance;
    int32_t axis[4];
    int32_t limited;
    real limits[4];
} constraint;

typedef struct {
    uint32_t nb;
    body* b;
    uint32_t nc_ext;
    constraint* ext;
    uint32_t err;
} scene;

/* Contact log: in record mode every narrow-phase visit appends its pair and
 * collision constraints (values); in replay mode the narrow phase is skipped
 * and the logged constraints are used in the same order. Replaying runs the
 * step with the contact set (and contact geometry) of the recorded run, the
 * map whose derivatives the device adjoint computes. */
typedef struct { uint32_t i1, i2, n; } log_pair;
typedef struct { double r1[3], r2[3], n[3]; } log_contact;
typedef ARR(log_pair) logpairarr;
typedef ARR(log_contact) logcontactarr;

struct orc_world {
    uint32_t ns;
    scene* s;
    uint32_t nb_total;
    uint64_t epa_failures;
    uint64_t last_contacts;
    uint64_t last_narrow_pairs;
    int log_mode;                 /* 0 off, 1 record, 2 replay */
    logpairarr log_pairs;
    logcontactarr log_contacts;
    uint32_t log_pair_cur, log_contact_cur;
    int log_error;
    /* Solve order given for one frame of one scene (orc_world_set_solve_order):
     * the class of every item of that frame, items solved by ascending class
     * and in item order within a class. */
    int64_t order_scene;          /* -1: none */
    int32_t* order_class;
    uint32_t order_n;
    uint32_t order_status;
};

/* ------------------------------------------------------------------ */
/* Hull construction (collider.cpp:194-364)                            */
/* ------------------------------------------------------------------ */
typedef ARR(uint32_t) u32arr;

/* do_triangles_share_same_vertex (collider.cpp:27-31) */
static int tri_share(i3 a, i3 b) {
    return a.x == b.x || a.x == b.y || a.x == b.z || a.y == b.x || a.y == b.y || a.y == b.z || a.z == b.x ||
           a.z == b.y || a.z == b.z;
}

/* collect_faces_planar_to (collider.cpp:47-77) */
static void collect_planar(const v3* hv, const i3* tris, u32arr* tri_nb, int* processed, uint32_t f, v3 target,
                           i3arr* out) {
    const double EPSILON = 0.000001;
    i3 t = tris[f];
    v3 v1 = hv[t.x], v2 = hv[t.y], v3_ = hv[t.z];
    v3 n = nrm(cross(sub(v2, v1), sub(v3_, v1)));
    if (processed[f]) return;
    real projection = dot(n, target);
    if ((projection - 1.0) > -EPSILON && (projection - 1.0) < EPSILON) {
        ARR_PUSH(*out, t);
        processed[f] = 1;
        for (uint32_t i = 0; i < tri_nb[f].n; ++i) collect_planar(hv, tris, tri_nb, processed, tri_nb[f].d[i], target, out);
    }
}

/* get_edge_index (collider.cpp:79-91) */
static int32_t edge_index(const i2* e, uint32_t n, i2 q) {
    for (uint32_t i = 0; i < n; ++i) {
        if (e[i].x == q.x && e[i].y == q.y) return (int32_t)i;
        if (e[i].x == q.y && e[i].y == q.x) return (int32_t)i;
    }
    return -1;
}

/* create_convex_hull_face (collider.cpp:93-166): boundary edges of a set of
 * coplanar triangles, chained into a loop. Returns 0 if the loop check fails. */
static int face_loop(const i3* tris, uint32_t nt, u32arr* elements) {
    i2arr edges = {0};
    for (uint32_t i = 0; i < nt; ++i) {
        i2 es[3] = {{tris[i].x, tris[i].y}, {tris[i].y, tris[i].z}, {tris[i].z, tris[i].x}};
        for (int k = 0; k < 3; ++k) {
            int32_t idx = edge_index(edges.d, edges.n, es[k]);
            if (idx >= 0) ARR_REMOVE(edges, idx);
            else ARR_PUSH(edges, es[k]);
        }
    }
    for (uint32_t i = 0; i < edges.n; ++i) {
        i2 cur = edges.d[i];
        for (uint32_t j = i + 1; j < edges.n; ++j) {
            i2 cand = edges.d[j];
            if (cur.y != cand.x && cur.y != cand.y) continue;
            if (cur.y == cand.y) { int32_t t = cand.x; cand.x = cand.y; cand.y = t; }
            i2 tmp = edges.d[i + 1];
            edges.d[i + 1] = cand;
            edges.d[j] = tmp;
        }
    }
    int ok = edges.n > 0 && edges.d[0].x == edges.d[edges.n - 1].y;
    for (uint32_t i = 0; i < edges.n; ++i) ARR_PUSH(*elements, (uint32_t)edges.d[i].x);
    ARR_FREE(edges);
    return ok;
}

static int contains(const u32arr* a, uint32_t v) {
    for (uint32_t i = 0; i < a->n; ++i)
        if (a->d[i] == v) return 1;
    return 0;
}

static uint32_t* csr_pack(u32arr* lists, uint32_t n, uint32_t** off_out) {
    uint32_t* off = malloc(sizeof(uint32_t) * (n + 1));
    off[0] = 0;
    for (uint32_t i = 0; i < n; ++i) off[i + 1] = off[i] + lists[i].n;
    uint32_t* d = malloc(sizeof(uint32_t) * (off[n] ? off[n] : 1));
    for (uint32_t i = 0; i < n; ++i) {
        memcpy(d + off[i], lists[i].d, sizeof(uint32_t) * lists[i].n);
        ARR_FREE(lists[i]);
    }
    *off_out = off;
    return d;
}

/* collider_convex_hull_create (collider.cpp:194-364) on the corner stream
 * `corners` (nc corners) with indices 0..nc-1, as obj_parse provides them. */
static int hull_build(hull* h, const v3* corners, uint32_t nc) {
    v3arr hv = {0};
    uint32_t* corner_idx = malloc(sizeof(uint32_t) * (nc ? nc : 1));
    /* Vertex dedupe in first-appearance order (collider.cpp:200-209). */
    for (uint32_t i = 0; i < nc; ++i) {
        uint32_t k;
        for (k = 0; k < hv.n; ++k)
            if (veq(hv.d[k], corners[i])) break;
        if (k == hv.n) ARR_PUSH(hv, corners[i]);
        corner_idx[i] = k;
    }
    uint32_t nt = nc / 3;
    i3* tris = malloc(sizeof(i3) * (nt ? nt : 1));
    for (uint32_t i = 0; i < nt; ++i) {
        tris[i].x = (int32_t)corner_idx[3 * i];
        tris[i].y = (int32_t)corner_idx[3 * i + 1];
        tris[i].z = (int32_t)corner_idx[3 * i + 2];
    }
    free(corner_idx);
    uint32_t nv = hv.n;
    u32arr* vf = calloc(nv, sizeof(u32arr));
    u32arr* vn = calloc(nv, sizeof(u32arr));
    u32arr* tn = calloc(nt ? nt : 1, sizeof(u32arr));
    /* vertex_to_neighbors and triangle neighbours (collider.cpp:248-282) */
    for (uint32_t i = 0; i < nt; ++i) {
        i3 t = tris[i];
        for (uint32_t j = 0; j < nt; ++j) {
            if (i == j) continue;
            if (tri_share(t, tris[j])) ARR_PUSH(tn[i], j);
        }
        if (!contains(&vn[t.x], t.y)) ARR_PUSH(vn[t.x], (uint32_t)t.y);
        if (!contains(&vn[t.x], t.z)) ARR_PUSH(vn[t.x], (uint32_t)t.z);
        if (!contains(&vn[t.y], t.x)) ARR_PUSH(vn[t.y], (uint32_t)t.x);
        if (!contains(&vn[t.y], t.z)) ARR_PUSH(vn[t.y], (uint32_t)t.z);
        if (!contains(&vn[t.z], t.x)) ARR_PUSH(vn[t.z], (uint32_t)t.x);
        if (!contains(&vn[t.z], t.y)) ARR_PUSH(vn[t.z], (uint32_t)t.y);
    }
    /* Faces: flood-merge coplanar triangles (collider.cpp:284-319). */
    int ok = 1;
    int* processed = calloc(nt ? nt : 1, sizeof(int));
    v3arr fnorm = {0};
    ARR(u32arr) felems = {0};
    for (uint32_t i = 0; i < nt; ++i) {
        if (processed[i]) continue;
        i3 t = tris[i];
        v3 n = nrm(cross(sub(hv.d[t.y], hv.d[t.x]), sub(hv.d[t.z], hv.d[t.x])));
        i3arr planar = {0};
        collect_planar(hv.d, tris, tn, processed, i, n, &planar);
        u32arr el = {0};
        ok &= face_loop(planar.d, planar.n, &el);
        uint32_t fidx = fnorm.n;
        ARR_PUSH(fnorm, n);
        ARR_PUSH(felems, el);
        for (uint32_t j = 0; j < planar.n; ++j) {
            ARR_PUSH(vf[planar.d[j].x], fidx);
            ARR_PUSH(vf[planar.d[j].y], fidx);
            ARR_PUSH(vf[planar.d[j].z], fidx);
        }
        ARR_FREE(planar);
    }
    uint32_t nf = fnorm.n;
    /* face_to_neighbors: faces sharing any vertex (collider.cpp:321-341) */
    u32arr* ff = calloc(nf ? nf : 1, sizeof(u32arr));
    for (uint32_t i = 0; i < nf; ++i) {
        for (uint32_t j = 0; j < nf; ++j) {
            if (i == j) continue;
            int share = 0;
            for (uint32_t a = 0; a < felems.d[i].n && !share; ++a)
                for (uint32_t b = 0; b < felems.d[j].n; ++b)
                    if (felems.d[i].d[a] == felems.d[j].d[b]) { share = 1; break; }
            if (share) ARR_PUSH(ff[i], j);
        }
    }
    h->nv = nv;
    h->nf = nf;
    h->v = hv.d;
    h->fn = fnorm.d;
    h->fe = csr_pack(felems.d, nf, &h->fe_off);
    ARR_FREE(felems);
    h->vf = csr_pack(vf, nv, &h->vf_off);
    h->vn = csr_pack(vn, nv, &h->vn_off);
    h->ff = csr_pack(ff, nf, &h->ff_off);
    free(vf); free(vn); free(ff);
    for (uint32_t i = 0; i < nt; ++i) ARR_FREE(tn[i]);
    free(tn); free(processed); free(tris);
    h->tv = malloc(sizeof(v3) * (nv ? nv : 1));
    h->tn = malloc(sizeof(v3) * (nf ? nf : 1));
    memcpy(h->tv, h->v, sizeof(v3) * nv);
    memcpy(h->tn, h->fn, sizeof(v3) * nf);
    return ok && nv > 0 && nf > 0;
}

static void hull_free(hull* h) {
    free(h->v); free(h->fn); free(h->fe_off); free(h->fe); free(h->vf_off); free(h->vf);
    free(h->vn_off); free(h->vn); free(h->ff_off); free(h->ff); free(h->tv); free(h->tn);
}

/* collider_update for a convex hull (collider.cpp:409-430) with
 * util_get_model_matrix_no_scale (util.cpp:60-74): M = T * R. */
static void hull_transform(hull* h, v3 t, qt q) {
    m4 R = q_mat4(q);
    m4 T = {{{1.0, 0.0, 0.0, t.x}, {0.0, 1.0, 0.0, t.y}, {0.0, 0.0, 1.0, t.z}, {0.0, 0.0, 0.0, 1.0}}};
    m4 M = m4_mul(&T, &R);
    for (uint32_t i = 0; i < h->nv; ++i) {
        v3 v = h->v[i];
        /* gm_mat4_multiply_vec4 (gm.h:405-413) with w = 1 */
        real x = v.x * M.m[0][0] + v.y * M.m[0][1] + v.z * M.m[0][2] + 1.0 * M.m[0][3];
        real y = v.x * M.m[1][0] + v.y * M.m[1][1] + v.z * M.m[1][2] + 1.0 * M.m[1][3];
        real z = v.x * M.m[2][0] + v.y * M.m[2][1] + v.z * M.m[2][2] + 1.0 * M.m[2][3];
        real w = v.x * M.m[3][0] + v.y * M.m[3][1] + v.z * M.m[3][2] + 1.0 * M.m[3][3];
        real s = 1.0 / w;
        h->tv[i] = V(s * x, s * y, s * z);
    }
    for (uint32_t i = 0; i < h->nf; ++i) {
        v3 n = h->fn[i];
        /* gm_mat4_multiply_vec3(..., false) (gm.h:414-419) */
        v3 r = V(M.m[0][0] * n.x + M.m[0][1] * n.y + M.m[0][2] * n.z,
                 M.m[1][0] * n.x + M.m[1][1] * n.y + M.m[1][2] * n.z,
                 M.m[2][0] * n.x + M.m[2][1] * n.y + M.m[2][2] * n.z);
        h->tn[i] = nrm(r);
    }
}

/* ------------------------------------------------------------------ */
/* Support mapping (support.cpp:5-39)                                  */
/* ------------------------------------------------------------------ */
static uint32_t support_index(const hull* h, v3 d, uint32_t* err) {
    uint32_t sel = UINT32_MAX;
    real max_dot = -DBL_MAX;
    for (uint32_t i = 0; i < h->nv; ++i) {
        real dd = dot(h->tv[i], d);
        if (dd > max_dot) { sel = i; max_dot = dd; }
    }
    if (sel == UINT32_MAX) { *err |= ORC_ERR_SUPPORT_NAN; sel = 0; }
    return sel;
}
/* support_point (support.cpp:19-31) */
static v3 support_point(const collider* c, v3 d, uint32_t* err) {
    if (c->type == COL_SPHERE) return add(c->center, scl((real)c->radius, nrm(d)));
    return c->h.tv[support_index(&c->h, d, err)];
}
/* support_point_of_minkowski_difference (support.cpp:34-39) */
static v3 support_md(const collider* a, const collider* b, v3 d, uint32_t* err) {
    v3 s1 = support_point(a, d, err);
    v3 s2 = support_point(b, scl(-1.0, d), err);
    return sub(s1, s2);
}

/* ------------------------------------------------------------------ */
/* GJK (gjk.cpp:7-379)                                                  */
/* ------------------------------------------------------------------ */
typedef struct { v3 a, b, c, d; uint32_t num; } simplex;

static v3 triple_cross(v3 a, v3 b, v3 c) { return cross(cross(a, b), c); }

static int add_to_simplex(simplex* s, v3 p) {
    switch (s->num) {
    case 1: s->b = s->a; s->a = p; break;
    case 2: s->c = s->b; s->b = s->a; s->a = p; break;
    case 3: s->d = s->c; s->c = s->b; s->b = s->a; s->a = p; break;
    default: return 0; /* assert(0) in the reference */
    }
    ++s->num;
    return 1;
}

static void do_simplex_2(simplex* s, v3* dir) {
    v3 a = s->a, b = s->b;
    v3 ao = inv3(a), ab = sub(b, a);
    if (dot(ab, ao) >= 0.0) {
        s->a = a; s->b = b; s->num = 2; *dir = triple_cross(ab, ao, ab);
    } else {
        s->a = a; s->num = 1; *dir = ao;
    }
}

static void do_simplex_3(simplex* s, v3* dir) {
    v3 a = s->a, b = s->b, c = s->c;
    v3 ao = inv3(a), ab = sub(b, a), ac = sub(c, a), abc = cross(ab, ac);
    if (dot(cross(abc, ac), ao) >= 0.0) {
        if (dot(ac, ao) >= 0.0) {
            s->a = a; s->b = c; s->num = 2; *dir = triple_cross(ac, ao, ac);
        } else if (dot(ab, ao) >= 0.0) {
            s->a = a; s->b = b; s->num = 2; *dir = triple_cross(ab, ao, ab);
        } else {
            s->a = a; *dir = ao; /* num unchanged (gjk.cpp:81-85) */
        }
    } else {
        if (dot(cross(ab, abc), ao) >= 0.0) {
            if (dot(ab, ao) >= 0.0) {
                s->a = a; s->b = b; s->num = 2; *dir = triple_cross(ab, ao, ab);
            } else {
                s->a = a; *dir = ao; /* num unchanged (gjk.cpp:95-98) */
            }
        } else if (dot(abc, ao) >= 0.0) {
            s->a = a; s->b = b; s->c = c; s->num = 3; *dir = abc;
        } else {
            s->a = a; s->b = c; s->c = b; s->num = 3; *dir = inv3(abc);
        }
    }
}

static int do_simplex_4(simplex* s, v3* dir) {
    v3 a = s->a, b = s->b, c = s->c, d = s->d;
    v3 ao = inv3(a), ab = sub(b, a), ac = sub(c, a), ad = sub(d, a);
    v3 abc = cross(ab, ac), acd = cross(ac, ad), adb = cross(ad, ab);
    unsigned pi = 0;
    if (dot(abc, ao) >= 0.0) pi |= 1;
    if (dot(acd, ao) >= 0.0) pi |= 2;
    if (dot(adb, ao) >= 0.0) pi |= 4;
    switch (pi) {
    case 0: return 1;
    case 1: /* triangle ABC (gjk.cpp:153-197) */
        if (dot(cross(abc, ac), ao) >= 0.0) {
            if (dot(ac, ao) >= 0.0) { s->a = a; s->b = c; s->num = 2; *dir = triple_cross(ac, ao, ac); }
            else if (dot(ab, ao) >= 0.0) { s->a = a; s->b = b; s->num = 2; *dir = triple_cross(ab, ao, ab); }
            else { s->a = a; *dir = ao; }
        } else {
            if (dot(cross(ab, abc), ao) >= 0.0) {
                if (dot(ab, ao) >= 0.0) { s->a = a; s->b = b; s->num = 2; *dir = triple_cross(ab, ao, ab); }
                else { s->a = a; *dir = ao; }
            } else { s->a = a; s->b = b; s->c = c; s->num = 3; *dir = abc; }
        }
        break;
    case 2: /* triangle ACD (gjk.cpp:198-242); the AC branch uses ab (gjk.cpp:213) */
        if (dot(cross(acd, ad), ao) >= 0.0) {
            if (dot(ad, ao) >= 0.0) { s->a = a; s->b = d; s->num = 2; *dir = triple_cross(ad, ao, ad); }
            else if (dot(ac, ao) >= 0.0) { s->a = a; s->b = c; s->num = 2; *dir = triple_cross(ab, ao, ab); }
            else { s->a = a; *dir = ao; }
        } else {
            if (dot(cross(ac, acd), ao) >= 0.0) {
                if (dot(ac, ao) >= 0.0) { s->a = a; s->b = c; s->num = 2; *dir = triple_cross(ac, ao, ac); }
                else { s->a = a; *dir = ao; }
            } else { s->a = a; s->b = c; s->c = d; s->num = 3; *dir = acd; }
        }
        break;
    case 3: /* line AC (gjk.cpp:243-256) */
        if (dot(ac, ao) >= 0.0) { s->a = a; s->b = c; s->num = 2; *dir = triple_cross(ac, ao, ac); }
        else { s->a = a; s->num = 1; *dir = ao; }
        break;
    case 4: /* triangle ADB (gjk.cpp:257-301) */
        if (dot(cross(adb, ab), ao) >= 0.0) {
            if (dot(ab, ao) >= 0.0) { s->a = a; s->b = b; s->num = 2; *dir = triple_cross(ab, ao, ab); }
            else if (dot(ad, ao) >= 0.0) { s->a = a; s->b = d; s->num = 2; *dir = triple_cross(ad, ao, ad); }
            else { s->a = a; *dir = ao; }
        } else {
            if (dot(cross(ad, adb), ao) >= 0.0) {
                if (dot(ad, ao) >= 0.0) { s->a = a; s->b = d; s->num = 2; *dir = triple_cross(ad, ao, ad); }
                else { s->a = a; *dir = ao; }
            } else { s->a = a; s->b = d; s->c = b; s->num = 3; *dir = adb; }
        }
        break;
    case 5: /* line AB (gjk.cpp:302-314) */
        if (dot(ab, ao) >= 0.0) { s->a = a; s->b = b; s->num = 2; *dir = triple_cross(ab, ao, ab); }
        else { s->a = a; s->num = 1; *dir = ao; }
        break;
    case 6: /* line AD (gjk.cpp:315-327) */
        if (dot(ad, ao) >= 0.0) { s->a = a; s->b = d; s->num = 2; *dir = triple_cross(ad, ao, ad); }
        else { s->a = a; s->num = 1; *dir = ao; }
        break;
    case 7: s->a = a; s->num = 1; *dir = ao; break;
    }
    return 0;
}

/* gjk_collides (gjk.cpp:350-379) */
static int gjk(const collider* h1, const collider* h2, simplex* out, uint32_t* err) {
    simplex s;
    memset(&s, 0, sizeof(s));
    s.a = support_md(h1, h2, V(0.0, 0.0, 1.0), err);
    s.num = 1;
    v3 dir = scl(-1.0, s.a);
    for (uint32_t i = 0; i < 100; ++i) {
        v3 p = support_md(h1, h2, dir, err);
        if (dot(p, dir) < 0.0) return 0;
        if (!add_to_simplex(&s, p)) { *err |= ORC_ERR_GJK_SIMPLEX; return 0; }
        int hit = 0;
        switch (s.num) {
        case 2: do_simplex_2(&s, &dir); break;
        case 3: do_simplex_3(&s, &dir); break;
        case 4: hit = do_simplex_4(&s, &dir); break;
        }
        if (hit) { *out = s; return 1; }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* EPA (epa.cpp:8-237)                                                  */
/* ------------------------------------------------------------------ */
/* get_face_normal_and_distance_to_origin (epa.cpp:32-77) */
static int face_normal(i3 f, const v3* poly, uint32_t np, v3* n_out, real* d_out) {
    v3 a = poly[f.x], b = poly[f.y], c = poly[f.z];
    v3 n = nrm(cross(sub(b, a), sub(c, a)));
    if (!(n.x != 0.0 || n.y != 0.0 || n.z != 0.0)) return 0;
    const double TOL = 0.0000000000000;
    real d = dot(n, a);
    if (d < -TOL) {
        n = inv3(n);
        d = -d;
    } else if (d >= -TOL && d <= TOL) {
        int found = 0;
        for (uint32_t i = 0; i < np; ++i) {
            real ad = dot(n, poly[i]);
            if (ad < -TOL || ad > TOL) {
                n = ad < -TOL ? n : inv3(n);
                found = 1;
                break;
            }
        }
        if (!found) return 0;
    }
    *n_out = n;
    *d_out = d;
    return 1;
}

/* add_edge (epa.cpp:79-110) */
static void add_edge(i2arr* edges, i2 e, const v3* poly) {
    for (uint32_t i = 0; i < edges->n; ++i) {
        i2 cur = edges->d[i];
        if (e.x == cur.x && e.y == cur.y) { ARR_REMOVE(*edges, i); return; }
        if (e.x == cur.y && e.y == cur.x) { ARR_REMOVE(*edges, i); return; }
        v3 c1 = poly[cur.x], c2 = poly[cur.y], e1 = poly[e.x], e2 = poly[e.y];
        if (veq(c1, e1) && veq(c2, e2)) { ARR_REMOVE(*edges, i); return; }
        if (veq(c1, e2) && veq(c2, e1)) { ARR_REMOVE(*edges, i); return; }
    }
    ARR_PUSH(*edges, e);
}

/* triangle_centroid (epa.cpp:112-116) */
static v3 centroid(v3 p1, v3 p2, v3 p3) { return scl(1.0 / 3.0, add(add(p2, p3), p1)); }

/* epa (epa.cpp:118-237). Returns 1 on convergence. */
static int epa(const collider* h1, const collider* h2, const simplex* s, v3* n_out, real* pen_out, uint32_t* err) {
    v3arr poly = {0};
    i3arr faces = {0};
    v3arr normals = {0};
    f64arr dist = {0};
    i2arr edges = {0};
    ARR_PUSH(poly, s->a); ARR_PUSH(poly, s->b); ARR_PUSH(poly, s->c); ARR_PUSH(poly, s->d);
    i3 f0 = {0, 1, 2}, f1 = {0, 2, 3}, f2 = {0, 3, 1}, f3 = {1, 2, 3};
    ARR_PUSH(faces, f0); ARR_PUSH(faces, f1); ARR_PUSH(faces, f2); ARR_PUSH(faces, f3);
    v3 min_n = V(0.0, 0.0, 0.0);
    real min_d = DBL_MAX;
    int converged = 0, bad = 0;
    for (uint32_t i = 0; i < faces.n; ++i) {
        v3 n; real d;
        if (!face_normal(faces.d[i], poly.d, poly.n, &n, &d)) { bad = 1; goto done; }
        ARR_PUSH(normals, n);
        ARR_PUSH(dist, d);
        if (d < min_d) { min_d = d; min_n = n; }
    }
    for (uint32_t it = 0; it < 100; ++it) {
        v3 sp = support_md(h1, h2, min_n, err);
        real d = dot(min_n, sp);
        if (fabs(d - min_d) < 0.0001) {
            *n_out = min_n;
            *pen_out = min_d;
            converged = 1;
            break;
        }
        uint32_t new_idx = poly.n;
        ARR_PUSH(poly, sp);
        for (uint32_t i = 0; i < normals.n; ++i) {
            v3 n = normals.d[i];
            i3 f = faces.d[i];
            v3 c = centroid(poly.d[f.x], poly.d[f.y], poly.d[f.z]);
            if (dot(n, sub(sp, c)) > 0.0) {
                i2 e1 = {f.x, f.y}, e2 = {f.y, f.z}, e3 = {f.z, f.x};
                add_edge(&edges, e1, poly.d);
                add_edge(&edges, e2, poly.d);
                add_edge(&edges, e3, poly.d);
                ARR_REMOVE(faces, i);
                ARR_REMOVE(dist, i);
                ARR_REMOVE(normals, i);
                --i;
            }
        }
        for (uint32_t i = 0; i < edges.n; ++i) {
            i3 nf = {edges.d[i].x, edges.d[i].y, (int32_t)new_idx};
            ARR_PUSH(faces, nf);
            v3 n; real fd;
            if (!face_normal(nf, poly.d, poly.n, &n, &fd)) { bad = 1; goto done; }
            ARR_PUSH(normals, n);
            ARR_PUSH(dist, fd);
        }
        min_d = DBL_MAX;
        for (uint32_t i = 0; i < dist.n; ++i) {
            if (dist.d[i] < min_d) { min_d = dist.d[i]; min_n = normals.d[i]; }
        }
        edges.n = 0;
    }
done:
    if (bad) *err |= ORC_ERR_EPA_DEGENERATE;
    ARR_FREE(poly); ARR_FREE(faces); ARR_FREE(normals); ARR_FREE(dist); ARR_FREE(edges);
    return converged && !bad;
}

/* ------------------------------------------------------------------ */
/* Contact manifold (clipping.cpp:7-371)                               */
/* ------------------------------------------------------------------ */
typedef struct { v3 normal, point; } plane;

/* is_point_in_plane (clipping.cpp:12-19): note the float truncation */
static int point_in_plane(const plane* p, v3 pos) {
    float distance = (float)RD(-dot(p->normal, p->point));
    if (dot(pos, p->normal) + distance < 0.0) return 0;
    return 1;
}

/* plane_edge_intersection (clipping.cpp:21-47) */
static int plane_edge(const plane* p, v3 start, v3 end, v3* out) {
    const double EPSILON = 0.000001;
    v3 ab = sub(end, start);
    float ab_p = (float)RD(dot(p->normal, ab));
    if (fabs(ab_p) > EPSILON) {
        float distance = (float)RD(-dot(p->normal, p->point));
        v3 p_co = scl(-distance, p->normal);
        float fac = (float)RD(-dot(p->normal, sub(start, p_co)) / ab_p);
        fac = RP_MIN(RP_MAX(fac, 0.0), 1.0);
        *out = add(start, scl(fac, ab));
        return 1;
    }
    return 0;
}

/* sutherland_hodgman (clipping.cpp:52-113) */
static void clip_polygon(const v3arr* in_poly, uint32_t nplanes, const plane* planes, v3arr* out_poly, int remove) {
    v3arr input = {0}, output = {0};
    for (uint32_t i = 0; i < in_poly->n; ++i) ARR_PUSH(input, in_poly->d[i]);
    for (uint32_t i = 0; i < nplanes; ++i) {
        if (input.n == 0) break;
        const plane* p = &planes[i];
        v3 tmp, start = input.d[input.n - 1];
        for (uint32_t j = 0; j < input.n; ++j) {
            v3 end = input.d[j];
            int s_in = point_in_plane(p, start);
            int e_in = point_in_plane(p, end);
            if (remove) {
                if (e_in) ARR_PUSH(output, end);
            } else {
                if (s_in && e_in) {
                    ARR_PUSH(output, end);
                } else if (s_in && !e_in) {
                    if (plane_edge(p, start, end, &tmp)) ARR_PUSH(output, tmp);
                } else if (!s_in && e_in) {
                    if (plane_edge(p, start, end, &tmp)) ARR_PUSH(output, tmp);
                    ARR_PUSH(output, end);
                }
            }
            start = end;
        }
        v3arr t = input;
        input = output;
        output = t;
        output.n = 0;
    }
    *out_poly = input;
    ARR_FREE(output);
}

/* get_closest_point_polygon (clipping.cpp:115-119) */
static v3 closest_on_plane(v3 pos, const plane* rp) {
    real d = dot(scl(-1.0, rp->normal), rp->point);
    return sub(pos, scl(dot(rp->normal, pos) + d, rp->normal));
}

/* get_face_with_most_fitting_normal (clipping.cpp:136-152) */
static uint32_t best_face(uint32_t sidx, const hull* h, v3 n, uint32_t* err) {
    real max_proj = -DBL_MAX;
    uint32_t sel = UINT32_MAX;
    for (uint32_t i = h->vf_off[sidx]; i < h->vf_off[sidx + 1]; ++i) {
        real proj = dot(h->tn[h->vf[i]], n);
        if (proj > max_proj) { max_proj = proj; sel = h->vf[i]; }
    }
    if (sel == UINT32_MAX) { *err |= ORC_ERR_SELECT; sel = 0; }
    return sel;
}

/* get_edge_with_most_fitting_normal (clipping.cpp:154-200) */
static void best_edge(uint32_t s1, uint32_t s2, const hull* h1, const hull* h2, v3 n, v3* edge_n, uint32_t sel[4],
                      uint32_t* err) {
    v3 sp1 = h1->tv[s1], sp2 = h2->tv[s2];
    real max_dot = -DBL_MAX;
    int found = 0;
    for (uint32_t i = h1->vn_off[s1]; i < h1->vn_off[s1 + 1]; ++i) {
        v3 e1 = sub(sp1, h1->tv[h1->vn[i]]);
        for (uint32_t j = h2->vn_off[s2]; j < h2->vn_off[s2 + 1]; ++j) {
            v3 e2 = sub(sp2, h2->tv[h2->vn[j]]);
            v3 cn = nrm(cross(e1, e2));
            v3 cni = inv3(cn);
            real dd = dot(cn, n);
            if (dd > max_dot) {
                max_dot = dd; sel[0] = s1; sel[1] = h1->vn[i]; sel[2] = s2; sel[3] = h2->vn[j];
                *edge_n = cn; found = 1;
            }
            dd = dot(cni, n);
            if (dd > max_dot) {
                max_dot = dd; sel[0] = s1; sel[1] = h1->vn[i]; sel[2] = s2; sel[3] = h2->vn[j];
                *edge_n = cni; found = 1;
            }
        }
    }
    if (!found) *err |= ORC_ERR_SELECT;
}

/* collision_distance_between_skew_lines (clipping.cpp:210-239) */
static int skew_lines(v3 p1, v3 d1, v3 p2, v3 d2, v3* l1, v3* l2) {
    real n1 = d1.x * d2.x + d1.y * d2.y + d1.z * d2.z;
    real n2 = d2.x * d2.x + d2.y * d2.y + d2.z * d2.z;
    real m1 = -d1.x * d1.x - d1.y * d1.y - d1.z * d1.z;
    real m2 = -d2.x * d1.x - d2.y * d1.y - d2.z * d1.z;
    real r1 = -d1.x * p2.x + d1.x * p1.x - d1.y * p2.y + d1.y * p1.y - d1.z * p2.z + d1.z * p1.z;
    real r2 = -d2.x * p2.x + d2.x * p1.x - d2.y * p2.y + d2.y * p1.y - d2.z * p2.z + d2.z * p1.z;
    if ((n1 * m2) - (n2 * m1) == 0) return 0;
    real n = ((r1 * m2) - (r2 * m1)) / ((n1 * m2) - (n2 * m1));
    real m = ((n1 * r2) - (n2 * r1)) / ((n1 * m2) - (n2 * m1));
    *l1 = add(p1, scl(m, d1));
    *l2 = add(p2, scl(n, d2));
    return 1;
}

typedef struct { v3 p1, p2, n; } contact;
typedef ARR(contact) contactarr;

/* convex_convex_contact_manifold (clipping.cpp:249-341) */
static void manifold(const hull* h1, const hull* h2, v3 normal, contactarr* out, uint32_t* err) {
    const double EPSILON = 0.0001;
    v3 inormal = inv3(normal);
    v3 edge_n = V(0.0, 0.0, 0.0);
    uint32_t s1 = support_index(h1, normal, err);
    uint32_t s2 = support_index(h2, inormal, err);
    uint32_t f1 = best_face(s1, h1, normal, err);
    uint32_t f2 = best_face(s2, h2, inormal, err);
    uint32_t edges[4] = {0, 0, 0, 0};
    best_edge(s1, s2, h1, h2, normal, &edge_n, edges, err);
    real d1 = dot(h1->tn[f1], normal);
    real d2 = dot(h2->tn[f2], inormal);
    real de = dot(edge_n, normal);
    if (de > d1 + EPSILON && de > d2 + EPSILON) {
        v3 l1, l2;
        v3 p1 = h1->tv[edges[0]];
        v3 dd1 = sub(h1->tv[edges[1]], p1);
        v3 p2 = h2->tv[edges[2]];
        v3 dd2 = sub(h2->tv[edges[3]], p2);
        if (!skew_lines(p1, dd1, p2, dd2, &l1, &l2)) { *err |= ORC_ERR_SKEW_LINES; return; }
        contact c = {l1, l2, normal};
        ARR_PUSH(*out, c);
    } else {
        int ref1 = d1 > d2;
        const hull* rh = ref1 ? h1 : h2;
        const hull* ih = ref1 ? h2 : h1;
        uint32_t rf = ref1 ? f1 : f2, inf = ref1 ? f2 : f1;
        v3arr inc = {0};
        for (uint32_t i = ih->fe_off[inf]; i < ih->fe_off[inf + 1]; ++i) ARR_PUSH(inc, ih->tv[ih->fe[i]]);
        /* build_boundary_planes (clipping.cpp:121-134) */
        uint32_t nbp = rh->ff_off[rf + 1] - rh->ff_off[rf];
        plane* bp = malloc(sizeof(plane) * (nbp ? nbp : 1));
        for (uint32_t i = 0; i < nbp; ++i) {
            uint32_t nf = rh->ff[rh->ff_off[rf] + i];
            bp[i].point = rh->tv[rh->fe[rh->fe_off[nf]]];
            bp[i].normal = inv3(rh->tn[nf]);
        }
        v3arr clipped = {0}, final_pts = {0};
        if (nbp > 0) clip_polygon(&inc, nbp, bp, &clipped, 0);
        else for (uint32_t i = 0; i < inc.n; ++i) ARR_PUSH(clipped, inc.d[i]);
        plane rp;
        rp.normal = inv3(rh->tn[rf]);
        rp.point = rh->tv[rh->fe[rh->fe_off[rf]]];
        clip_polygon(&clipped, 1, &rp, &final_pts, 1);
        for (uint32_t i = 0; i < final_pts.n; ++i) {
            v3 pt = final_pts.d[i];
            v3 cp = closest_on_plane(pt, &rp);
            v3 diff = sub(pt, cp);
            real pen;
            contact c;
            if (ref1) {
                pen = dot(diff, normal);
                c.p1 = sub(pt, scl(pen, normal));
                c.p2 = pt;
            } else {
                pen = -dot(diff, normal);
                c.p1 = pt;
                c.p2 = add(pt, scl(pen, normal));
            }
            c.n = normal;
            if (pen < 0.0) ARR_PUSH(*out, c);
        }
        free(bp);
        ARR_FREE(inc); ARR_FREE(clipped); ARR_FREE(final_pts);
    }
}

/* clipping_get_contact_manifold (clipping.cpp:343-371): a sphere gives one
 * contact at its support point along the normal. */
static void contact_manifold(const collider* c1, const collider* c2, v3 normal, real penetration, contactarr* out,
                             uint32_t* err) {
    if (c1->type == COL_SPHERE) {
        v3 p = support_point(c1, normal, err);
        contact c;
        c.p1 = p;
        c.p2 = sub(p, scl(penetration, normal));
        c.n = normal;
        ARR_PUSH(*out, c);
    } else if (c2->type == COL_SPHERE) {
        v3 inormal = inv3(normal);
        v3 p = support_point(c2, inormal, err);
        contact c;
        c.p1 = add(p, scl(penetration, normal));
        c.p2 = p;
        c.n = normal;
        ARR_PUSH(*out, c);
    } else {
        manifold(&c1->h, &c2->h, normal, out, err);
    }
}

/* collider_get_contacts (collider.cpp:523-558). Two spheres are resolved
 * analytically with the reference's float arithmetic; otherwise GJK, EPA and
 * the contact manifold. */
static void collider_contacts(const collider* h1, const collider* h2, contactarr* out, uint32_t* err,
                              uint64_t* epa_fail) {
    if (h1->type == COL_SPHERE && h2->type == COL_SPHERE) {
        v3 dv = sub(h1->center, h2->center);
        float dsq = (float)RD(dot(dv, dv));
        float mind = h1->radius + h2->radius;
        if (dsq < (mind * mind)) {
            v3 n = nrm(sub(h2->center, h1->center));
            real pen = (real)(double)(mind - sqrtf(dsq));
            ORC_PHASE(PH_MANIFOLD);
            contact_manifold(h1, h2, n, pen, out, err);
            ORC_PHASE(PH_NARROW);
        }
        return;
    }
    simplex s;
    if (gjk(h1, h2, &s, err)) {
        v3 n; real pen;
        if (!epa(h1, h2, &s, &n, &pen, err)) { ++*epa_fail; return; }
        ORC_PHASE(PH_MANIFOLD);
        contact_manifold(h1, h2, n, pen, out, err);
        ORC_PHASE(PH_NARROW);
    }
}

/* colliders_get_contacts (collider.cpp:560-572): every collider pair, in order. */
static void pair_contacts(const body* e1, const body* e2, contactarr* out, uint32_t* err, uint64_t* epa_fail) {
    for (uint32_t i = 0; i < e1->ncol; ++i)
        for (uint32_t j = 0; j < e2->ncol; ++j) collider_contacts(&e1->col[i], &e2->col[j], out, err, epa_fail);
}

/* colliders_update (collider.cpp:409-445) */
static void colliders_update(body* e) {
    for (uint32_t i = 0; i < e->ncol; ++i) {
        if (e->col[i].type == COL_SPHERE) e->col[i].center = e->x;
        else hull_transform(&e->col[i].h, e->x, e->q);
    }
}

/* ------------------------------------------------------------------ */
/* Body dynamics helpers (physics_util.cpp:26-61)                      */
/* ------------------------------------------------------------------ */
static m3 dyn_inv_inertia(const body* e) {
    m3 R = q_mat3(e->q);
    m3 Rt = m3_t(&R);
    m3 aux = m3_mul(&R, &e->inv_inertia);
    return m3_mul(&aux, &Rt);
}
static m3 dyn_inertia(const body* e) {
    m3 R = q_mat3(e->q);
    m3 Rt = m3_t(&R);
    m3 aux = m3_mul(&R, &e->inertia);
    return m3_mul(&aux, &Rt);
}

/* ------------------------------------------------------------------ */
/* Base XPBD constraints (pbd_base_constraints.cpp:6-227)              */
/* ------------------------------------------------------------------ */
typedef struct { body *e1, *e2; v3 r1, r2; m3 I1, I2; } pcpd;
typedef struct { body *e1, *e2; m3 I1, I2; } acpd;

static void pos_prep(body* e1, body* e2, v3 r1_lc, v3 r2_lc, pcpd* p) {
    p->e1 = e1; p->e2 = e2;
    p->r1 = q_rot(e1->q, r1_lc);
    p->r2 = q_rot(e2->q, r2_lc);
    p->I1 = dyn_inv_inertia(e1);
    p->I2 = dyn_inv_inertia(e2);
}

static real pos_dlambda(const pcpd* p, real h, real compliance, real lambda, v3 dx, uint32_t* err) {
    real c = len(dx);
    if (c <= 1e-50) return 0.0;
    v3 n = V(dx.x / c, dx.y / c, dx.z / c);
    real w1 = p->e1->inv_mass + dot(cross(p->r1, n), m3_mulv(&p->I1, cross(p->r1, n)));
    real w2 = p->e2->inv_mass + dot(cross(p->r2, n), m3_mulv(&p->I2, cross(p->r2, n)));
    if (w1 + w2 == 0.0) *err |= ORC_ERR_ZERO_MASS;
    real tc = compliance / (h * h);
    return (-c - tc * lambda) / (w1 + w2 + tc);
}

static void pos_apply(const pcpd* p, real dl, v3 dx) {
    real c = len(dx);
    if (c <= 1e-50) return;
    body *e1 = p->e1, *e2 = p->e2;
    v3 n = V(dx.x / c, dx.y / c, dx.z / c);
    v3 imp = scl(dl, n);
    if (!e1->fixed) e1->x = add(e1->x, scl(e1->inv_mass, imp));
    if (!e2->fixed) e2->x = add(e2->x, scl(-e2->inv_mass, imp));
    v3 a1 = m3_mulv(&p->I1, cross(p->r1, imp));
    v3 a2 = m3_mulv(&p->I2, cross(p->r2, imp));
    qt aq1 = {a1.x, a1.y, a1.z, 0.0}, aq2 = {a2.x, a2.y, a2.z, 0.0};
    qt q1 = q_mul(aq1, e1->q);
    qt q2 = q_mul(aq2, e2->q);
    if (!e1->fixed) {
        e1->q.x = e1->q.x + 0.5 * q1.x; e1->q.y = e1->q.y + 0.5 * q1.y;
        e1->q.z = e1->q.z + 0.5 * q1.z; e1->q.w = e1->q.w + 0.5 * q1.w;
        e1->q = q_nrm(e1->q);
    }
    if (!e2->fixed) {
        e2->q.x = e2->q.x - 0.5 * q2.x; e2->q.y = e2->q.y - 0.5 * q2.y;
        e2->q.z = e2->q.z - 0.5 * q2.z; e2->q.w = e2->q.w - 0.5 * q2.w;
        e2->q = q_nrm(e2->q);
    }
}

static void ang_prep(body* e1, body* e2, acpd* p) {
    p->e1 = e1; p->e2 = e2;
    p->I1 = dyn_inv_inertia(e1);
    p->I2 = dyn_inv_inertia(e2);
}

static real ang_dlambda(const acpd* p, real h, real compliance, real lambda, v3 dq, uint32_t* err) {
    real theta = len(dq);
    if (theta <= 1e-50) return 0.0;
    v3 n = V(dq.x / theta, dq.y / theta, dq.z / theta);
    real w1 = dot(n, m3_mulv(&p->I1, n));
    real w2 = dot(n, m3_mulv(&p->I2, n));
    if (w1 + w2 == 0.0) *err |= ORC_ERR_ZERO_MASS;
    real tc = compliance / (h * h);
    return (-theta - tc * lambda) / (w1 + w2 + tc);
}

static void ang_apply(const acpd* p, real dl, v3 dq) {
    real theta = len(dq);
    if (theta <= 1e-50) return;
    body *e1 = p->e1, *e2 = p->e2;
    v3 n = V(dq.x / theta, dq.y / theta, dq.z / theta);
    v3 imp = scl(-dl, n);
    v3 a1 = m3_mulv(&p->I1, imp);
    v3 a2 = m3_mulv(&p->I2, imp);
    qt aq1 = {a1.x, a1.y, a1.z, 0.0}, aq2 = {a2.x, a2.y, a2.z, 0.0};
    qt q1 = q_mul(aq1, e1->q);
    qt q2 = q_mul(aq2, e2->q);
    if (!e1->fixed) {
        e1->q.x = e1->q.x + 0.5 * q1.x; e1->q.y = e1->q.y + 0.5 * q1.y;
        e1->q.z = e1->q.z + 0.5 * q1.z; e1->q.w = e1->q.w + 0.5 * q1.w;
        e1->q = q_nrm(e1->q);
    }
    if (!e2->fixed) {
        e2->q.x = e2->q.x - 0.5 * q2.x; e2->q.y = e2->q.y - 0.5 * q2.y;
        e2->q.z = e2->q.z - 0.5 * q2.z; e2->q.w = e2->q.w - 0.5 * q2.w;
        e2->q = q_nrm(e2->q);
    }
}

/* ------------------------------------------------------------------ */
/* Constraint solves (pbd.cpp:81-406)                                  */
/* ------------------------------------------------------------------ */
/* positional_constraint_solve (pbd.cpp:81-97) */
static void solve_positional(constraint* c, body* e1, body* e2, real h, uint32_t* err) {
    v3 dx = sub(sub(e1->x, e2->x), c->distance);
    pcpd p;
    pos_prep(e1, e2, c->r1, c->r2, &p);
    real dl = pos_dlambda(&p, h, c->compliance, c->lambda, dx, err);
    pos_apply(&p, dl, dx);
    c->lambda += dl;
}

/* collision_constraint_solve (pbd.cpp:107-154) */
static void solve_collision(constraint* c, body* e1, body* e2, real h, uint32_t* err) {
    pcpd p;
    pos_prep(e1, e2, c->r1, c->r2, &p);
    v3 p1 = add(e1->x, p.r1);
    v3 p2 = add(e2->x, p.r2);
    real d = dot(sub(p1, p2), c->n);
    if (d > 0.0) {
        v3 dx = scl(d, c->n);
        real dl = pos_dlambda(&p, h, 0.0, c->lambda_n, dx, err);
        pos_apply(&p, dl, dx);
        c->lambda_n += dl;
        pos_prep(e1, e2, c->r1, c->r2, &p);
        p1 = add(e1->x, p.r1);
        p2 = add(e2->x, p.r2);
        dl = pos_dlambda(&p, h, 0.0, c->lambda_t, dx, err);
        const real mu_s = (e1->mu_s + e2->mu_s) / 2.0f;
        real lambda_n = c->lambda_n;
        real lambda_t = c->lambda_t + dl;
        if (lambda_t > mu_s * lambda_n) {
            v3 p1_til = add(e1->px, q_rot(e1->pq, c->r1));
            v3 p2_til = add(e2->px, q_rot(e2->pq, c->r2));
            v3 dp = sub(sub(p1, p1_til), sub(p2, p2_til));
            v3 dp_t = sub(dp, scl(dot(dp, c->n), c->n));
            pos_apply(&p, dl, dp_t);
            c->lambda_t += dl;
        }
    }
}

/* mutual_orientation_constraint_solve (pbd.cpp:156-173) */
static void solve_mutual(constraint* c, body* e1, body* e2, real h, uint32_t* err) {
    acpd a;
    ang_prep(e1, e2, &a);
    qt q2i = q_inv(e2->q);
    qt aux = q_mul(e1->q, q2i);
    v3 dq = V(2.0 * aux.x, 2.0 * aux.y, 2.0 * aux.z);
    real dl = ang_dlambda(&a, h, c->compliance, c->lambda, dq, err);
    ang_apply(&a, dl, dq);
    c->lambda += dl;
}

/* limit_angle (pbd.cpp:175-217) */
static int limit_angle(v3 n, v3 n1, v3 n2, real alpha, real beta, v3* dq) {
    real phi = asin(dot(cross(n1, n2), n));
    if (dot(n1, n2) < 0.0) phi = PI_F - phi;
    if (phi > PI_F) phi = phi - 2.0 * PI_F;
    if (phi < -PI_F) phi = phi + 2.0 * PI_F;
    if (phi < alpha || phi > beta) {
        phi = RP_CLAMP(phi, alpha, beta);
        qt rot = q_from_axis_radians(n, phi);
        n1 = q_rot(rot, n1);
        *dq = cross(n1, n2);
        return 1;
    }
    return 0;
}

/* hinge_joint_constraint_solve (pbd.cpp:245-299); lambda = aligned axes,
 * lambda2 = position, lambda3 = limit */
static void solve_hinge(constraint* c, body* e1, body* e2, real h, uint32_t* err) {
    int bad = 0;
    acpd a;
    ang_prep(e1, e2, &a);
    v3 a1 = q_axis(e1->q, c->axis[0], &bad);
    v3 a2 = q_axis(e2->q, c->axis[1], &bad);
    v3 dq = cross(a1, a2);
    real dl = ang_dlambda(&a, h, c->compliance, c->lambda, dq, err);
    ang_apply(&a, dl, dq);
    c->lambda += dl;
    pcpd p;
    pos_prep(e1, e2, c->r1, c->r2, &p);
    v3 p1 = add(e1->x, p.r1);
    v3 p2 = add(e2->x, p.r2);
    v3 dx = sub(p1, p2);
    dl = pos_dlambda(&p, h, 0.0, c->lambda2, dx, err);
    pos_apply(&p, dl, dx);
    c->lambda2 += dl;
    if (c->limited) {
        v3 n1 = q_axis(e1->q, c->axis[2], &bad);
        v3 n2 = q_axis(e2->q, c->axis[3], &bad);
        v3 n = q_axis(e1->q, c->axis[0], &bad);
        if (limit_angle(n, n1, n2, c->limits[0], c->limits[1], &dq)) {
            acpd a2p;
            ang_prep(e1, e2, &a2p);
            real dl2 = ang_dlambda(&a2p, h, 0.0, c->lambda3, dq, err);
            ang_apply(&a2p, dl2, dq);
            c->lambda3 += dl2;
        }
    }
    if (bad) *err |= ORC_ERR_SELECT;
}

/* spherical_joint_constraint_solve (pbd.cpp:301-379); lambda = position,
 * lambda2 = swing, lambda3 = twist */
static void solve_spherical(constraint* c, body* e1, body* e2, real h, uint32_t* err) {
    const double EPSILON = 1e-50;
    int bad = 0;
    pcpd p;
    pos_prep(e1, e2, c->r1, c->r2, &p);
    v3 p1 = add(e1->x, p.r1);
    v3 p2 = add(e2->x, p.r2);
    v3 dx = sub(p1, p2);
    real dl = pos_dlambda(&p, h, 0.0, c->lambda, dx, err);
    pos_apply(&p, dl, dx);
    c->lambda += dl;

    v3 n1 = q_axis(e1->q, c->axis[0], &bad);
    v3 n2 = q_axis(e2->q, c->axis[1], &bad);
    v3 n = cross(n1, n2);
    real n_len = len(n);
    if (n_len > EPSILON) {
        n = V(n.x / n_len, n.y / n_len, n.z / n_len);
        v3 dq;
        if (limit_angle(n, n1, n2, c->limits[0], c->limits[1], &dq)) {
            acpd a;
            ang_prep(e1, e2, &a);
            real d2 = ang_dlambda(&a, h, 0.0, c->lambda2, dq, err);
            ang_apply(&a, d2, dq);
            c->lambda2 += d2;
        }
    }
    v3 a1 = q_axis(e1->q, c->axis[0], &bad);
    v3 b1 = q_axis(e1->q, c->axis[2], &bad);
    v3 a2 = q_axis(e2->q, c->axis[1], &bad);
    v3 b2 = q_axis(e2->q, c->axis[3], &bad);
    n = add(a1, a2);
    n_len = len(n);
    if (n_len > EPSILON) {
        n = V(n.x / n_len, n.y / n_len, n.z / n_len);
        n1 = sub(b1, scl(dot(n, b1), n));
        n2 = sub(b2, scl(dot(n, b2), n));
        real n1_len = len(n1);
        real n2_len = len(n2);
        if (n1_len > EPSILON && n2_len > EPSILON) {
            n1 = V(n1.x / n1_len, n1.y / n1_len, n1.z / n1_len);
            n2 = V(n2.x / n2_len, n2.y / n2_len, n2.z / n2_len);
            v3 dq;
            if (limit_angle(n, n1, n2, c->limits[2], c->limits[3], &dq)) {
                acpd a;
                ang_prep(e1, e2, &a);
                real d3 = ang_dlambda(&a, h, 0.0, c->lambda3, dq, err);
                ang_apply(&a, d3, dq);
                c->lambda3 += d3;
            }
        }
    }
    if (bad) *err |= ORC_ERR_SELECT;
}

static void solve_constraint(scene* sc, constraint* c, real h) {
    body* e1 = &sc->b[c->b1];
    body* e2 = &sc->b[c->b2];
    switch (c->type) {
    case ORC_POSITIONAL_CONSTRAINT: solve_positional(c, e1, e2, h, &sc->err); break;
    case ORC_COLLISION_CONSTRAINT: solve_collision(c, e1, e2, h, &sc->err); break;
    case ORC_MUTUAL_ORIENTATION_CONSTRAINT: solve_mutual(c, e1, e2, h, &sc->err); break;
    case ORC_HINGE_JOINT_CONSTRAINT: solve_hinge(c, e1, e2, h, &sc->err); break;
    case ORC_SPHERICAL_JOINT_CONSTRAINT: solve_spherical(c, e1, e2, h, &sc->err); break;
    default: sc->err |= ORC_ERR_SELECT; break;
    }
}

/* Velocity solve of one collision constraint (pbd.cpp:646-711) */
static void velocity_solve(constraint* c, body* e1, body* e2, real h) {
    v3 n = c->n;
    real lambda_n = c->lambda_n;
    pcpd p;
    pos_prep(e1, e2, c->r1, c->r2, &p);
    v3 v1 = e1->v, w1 = e1->w, v2 = e2->v, w2 = e2->w;
    v3 v = sub(add(v1, cross(w1, p.r1)), add(v2, cross(w2, p.r2)));
    real vn = dot(n, v);
    v3 vt = sub(v, scl(vn, n));
    v3 dv = V(0.0, 0.0, 0.0);
    const real mu_d = (e1->mu_d + e2->mu_d) / 2.0f;
    real fn = lambda_n / h;
    real fact = RP_MIN(mu_d * fabs(fn), len(vt));
    dv = add(dv, scl(-fact, nrm(vt)));
    v3 ov1 = e1->pv, ow1 = e1->pw, ov2 = e2->pv, ow2 = e2->pw;
    v3 vtil = sub(add(ov1, cross(ow1, p.r1)), add(ov2, cross(ow2, p.r2)));
    real vn_til = dot(n, vtil);
    real e = e1->e * e2->e;
    fact = -vn + RP_MIN(-e * vn_til, 0.0);
    dv = add(dv, scl(fact, n));
    real _w1 = e1->inv_mass + dot(cross(p.r1, n), m3_mulv(&p.I1, cross(p.r1, n)));
    real _w2 = e2->inv_mass + dot(cross(p.r2, n), m3_mulv(&p.I2, cross(p.r2, n)));
    v3 pimp = scl(1.0 / (_w1 + _w2), dv);
    if (!e1->fixed) {
        e1->v = add(e1->v, scl(e1->inv_mass, pimp));
        e1->w = add(e1->w, m3_mulv(&p.I1, cross(p.r1, pimp)));
    }
    if (!e2->fixed) {
        e2->v = add(e2->v, inv3(scl(e2->inv_mass, pimp)));
        e2->w = add(e2->w, inv3(m3_mulv(&p.I2, cross(p.r2, pimp))));
    }
}

/* clipping_contact_to_collision_constraint (pbd.cpp:408-424) */
static constraint contact_to_constraint(const body* e1, const body* e2, uint32_t i1, uint32_t i2, const contact* ct) {
    constraint c;
    memset(&c, 0, sizeof(c));
    c.type = ORC_COLLISION_CONSTRAINT;
    c.b1 = i1; c.b2 = i2;
    c.n = ct->n;
    c.lambda_n = 0.0; c.lambda_t = 0.0;
    v3 r1w = sub(ct->p1, e1->x);
    v3 r2w = sub(ct->p2, e2->x);
    qt q1i = q_inv(e1->q);
    c.r1 = q_rot(q1i, r1w);
    qt q2i = q_inv(e2->q);
    c.r2 = q_rot(q2i, r2w);
    ORC_FREEZE_V3(c.n);
    ORC_FREEZE_V3(c.r1);
    ORC_FREEZE_V3(c.r2);
    return c;
}

/* ------------------------------------------------------------------ */
/* Frame driver                                                         */
/* ------------------------------------------------------------------ */
typedef ARR(constraint) constraintarr;

/* broad_get_collision_pairs (broad.cpp:6-29). A pair whose x offset alone
 * exceeds the bound is skipped before the test: sqrt(fl(x*x)) == |x| for
 * correctly rounded IEEE arithmetic and the other two terms only add, so
 * the reference's length is at least |x| and its test fails too. The skip
 * is plain double arithmetic (not counted as algorithmic work). */
static uint32_t broad(const scene* sc, uint32_t** out) {
    u32arr pairs = {0};
    double* xr = malloc(sizeof(double) * 2 * (sc->nb ? sc->nb : 1));
    for (uint32_t i = 0; i < sc->nb; ++i) {
        xr[2 * i] = RD(sc->b[i].x.x);
        xr[2 * i + 1] = RD(sc->b[i].radius);
    }
    for (uint32_t i = 0; i < sc->nb; ++i) {
        const body* e1 = &sc->b[i];
        double x1 = xr[2 * i], r1 = xr[2 * i + 1];
        for (uint32_t j = i + 1; j < sc->nb; ++j) {
            if (fabs(x1 - xr[2 * j]) > (r1 + xr[2 * j + 1]) + 0.1) continue;
            const body* e2 = &sc->b[j];
            real dd = len(sub(e1->x, e2->x));
            real maxd = e1->radius + e2->radius + 0.1;
            if (dd <= maxd) { ARR_PUSH(pairs, i); ARR_PUSH(pairs, j); }
        }
    }
    free(xr);
    *out = pairs.d;
    return pairs.n / 2;
}

static uint32_t uf_find(uint32_t* parent, uint32_t x) {
    while (parent[x] != x) x = parent[x];
    return x;
}
/* uf_union (broad.cpp:55-59): parent[find(y)] = find(x) */
static void uf_union(uint32_t* parent, uint32_t x, uint32_t y) {
    uint32_t key = uf_find(parent, y);
    uint32_t value = uf_find(parent, x);
    parent[key] = value;
}

/* Islands + sleeping (broad.cpp:31-116, pbd.cpp:476-506). The reference
 * groups bodies by union-find root; the active flags depend only on the
 * partition, so the union-find here keeps indices instead of a hash map. */
static void islands_and_sleep(scene* sc, const uint32_t* pairs, uint32_t np, real dt) {
    uint32_t* parent = malloc(sizeof(uint32_t) * (sc->nb ? sc->nb : 1));
    for (uint32_t i = 0; i < sc->nb; ++i) parent[i] = i;
    for (uint32_t k = 0; k < np; ++k) {
        uint32_t i = pairs[2 * k], j = pairs[2 * k + 1];
        if (!sc->b[i].fixed && !sc->b[j].fixed) uf_union(parent, i, j);
    }
    for (uint32_t k = 0; k < sc->nc_ext; ++k) {
        uint32_t i = sc->ext[k].b1, j = sc->ext[k].b2;
        if (!sc->b[i].fixed && !sc->b[j].fixed) uf_union(parent, i, j);
    }
    int* all_inactive = malloc(sizeof(int) * (sc->nb ? sc->nb : 1));
    for (uint32_t i = 0; i < sc->nb; ++i) all_inactive[i] = 1;
    for (uint32_t i = 0; i < sc->nb; ++i) {
        body* e = &sc->b[i];
        if (e->fixed) continue;
        real lv = len(e->v);
        real av = len(e->w);
        if (lv < 0.10 && av < 0.10) e->deact += dt;
        else e->deact = 0.0;
        if (e->deact < 1.0) all_inactive[uf_find(parent, i)] = 0;
    }
    for (uint32_t i = 0; i < sc->nb; ++i) {
        body* e = &sc->b[i];
        if (e->fixed) continue;
        e->active = !all_inactive[uf_find(parent, i)];
    }
    free(all_inactive);
    free(parent);
}

/* A frame's solve items in a given class order (orc_world_set_solve_order):
 * item t < nc_ext is external constraint t, item nc_ext + k the k-th pair the
 * narrow phase visits (all of its contacts, in order). perm lists the items
 * by ascending class, in item order within a class. The frame's items must
 * match the given count, no class may hold two items that share a non-fixed
 * body (such items would not commute), and classes must lie in [0, 64). */
typedef struct {
    int on;
    uint32_t n;
    uint32_t* perm;
    uint32_t* first; /* [n + 1] constraint range of each item in the substep's list */
} solve_order;

static void order_begin(orc_world* w, scene* sc, const uint32_t* pairs, uint32_t np, int32_t collisions,
                        solve_order* so) {
    memset(so, 0, sizeof(*so));
    if (w->order_scene != (int64_t)(sc - w->s)) return;
    w->order_scene = -1;
    uint32_t n = sc->nc_ext;
    uint32_t* ib = malloc(sizeof(uint32_t) * 2 * (sc->nc_ext + np + 1));
    for (uint32_t k = 0; k < sc->nc_ext; ++k) { ib[2 * k] = sc->ext[k].b1; ib[2 * k + 1] = sc->ext[k].b2; }
    for (uint32_t k = 0; collisions && k < np; ++k) {
        const body *e1 = &sc->b[pairs[2 * k]], *e2 = &sc->b[pairs[2 * k + 1]];
        if ((e1->fixed || !e1->active) && (e2->fixed || !e2->active)) continue;
        ib[2 * n] = pairs[2 * k];
        ib[2 * n + 1] = pairs[2 * k + 1];
        ++n;
    }
    if (n != w->order_n) {
        w->order_status |= 1u;
        free(ib);
        return;
    }
    for (uint32_t t = 0; t < n; ++t)
        if (w->order_class[t] < 0 || w->order_class[t] >= 64) {
            w->order_status |= 4u;
            free(ib);
            return;
        }
    /* counting sort by class, stable */
    uint32_t cnt[65] = {0};
    for (uint32_t t = 0; t < n; ++t) ++cnt[w->order_class[t] + 1];
    for (int c = 0; c < 64; ++c) cnt[c + 1] += cnt[c];
    so->perm = malloc(sizeof(uint32_t) * (n ? n : 1));
    for (uint32_t t = 0; t < n; ++t) so->perm[cnt[w->order_class[t]]++] = t;
    /* no two items of one class share a non-fixed body */
    int32_t* seen = malloc(sizeof(int32_t) * (sc->nb ? sc->nb : 1));
    for (uint32_t i = 0; i < sc->nb; ++i) seen[i] = -1;
    for (uint32_t k = 0; k < n; ++k) {
        uint32_t t = so->perm[k];
        int32_t c = w->order_class[t];
        for (int side = 0; side < 2; ++side) {
            uint32_t b = ib[2 * t + side];
            if (sc->b[b].fixed || (side == 1 && b == ib[2 * t])) continue;
            if (seen[b] == c) w->order_status |= 2u;
            seen[b] = c;
        }
    }
    free(seen);
    free(ib);
    so->n = n;
    so->first = malloc(sizeof(uint32_t) * (n + 1));
    so->on = 1;
}

static void order_end(solve_order* so) {
    free(so->perm);
    free(so->first);
}

/* pbd_simulate_with_constraints (pbd.cpp:468-747) for one scene */
static void simulate(orc_world* w, scene* sc, real dt, uint32_t substeps, uint32_t iters, int32_t collisions) {
    if (dt <= 0.0) return;
    ORC_PHASE(PH_FRAME);
    real h = dt / substeps;
    uint32_t* pairs = NULL;
    uint32_t np = broad(sc, &pairs);
    islands_and_sleep(sc, pairs, np, dt);
    constraintarr cons = {0};
    contactarr cts = {0};
    solve_order so;
    order_begin(w, sc, pairs, np, collisions, &so);
    for (uint32_t s = 0; s < substeps; ++s) {
        /* integrate (pbd.cpp:537-577) */
        ORC_PHASE(PH_INTEGRATE);
        for (uint32_t j = 0; j < sc->nb; ++j) {
            body* e = &sc->b[j];
            e->px = e->x;
            e->pq = e->q;
            if (e->fixed) continue;
            if (!e->active) continue;
            v3 F = e->force, T = e->torque;
            e->v = add(e->v, scl(h * e->inv_mass, F));
            e->x = add(e->x, scl(h, e->v));
            m3 Iinv = dyn_inv_inertia(e);
            m3 I = dyn_inertia(e);
            e->w = add(e->w, scl(h, m3_mulv(&Iinv, sub(T, cross(e->w, m3_mulv(&I, e->w))))));
            qt aux = {e->w.x, e->w.y, e->w.z, 0.0};
            qt q = q_mul(aux, e->q);
            e->q.x = e->q.x + h * 0.5 * q.x;
            e->q.y = e->q.y + h * 0.5 * q.y;
            e->q.z = e->q.z + h * 0.5 * q.z;
            e->q.w = e->q.w + h * 0.5 * q.w;
            e->q = q_nrm(e->q);
        }
        /* constraints := copy of the external ones with zeroed lambdas (pbd.cpp:426-462) */
        cons.n = 0;
        for (uint32_t k = 0; k < sc->nc_ext; ++k) {
            constraint c = sc->ext[k];
            c.lambda = c.lambda2 = c.lambda3 = 0.0;
            c.lambda_t = c.lambda_n = 0.0;
            if (so.on) so.first[k] = k;
            ARR_PUSH(cons, c);
        }
        uint32_t visited = sc->nc_ext;
        if (collisions) {
            /* Transform each hull in a visited pair once for this substep:
             * the reference transforms both bodies of each pair right before
             * use (pbd.cpp:598-599); poses do not change between pairs, so
             * the values are identical. */
            ORC_PHASE(PH_TRANSFORM);
            for (uint32_t k = 0; k < np; ++k) {  /* bodies of the pairs visited below, once each */
                body *e1 = &sc->b[pairs[2 * k]], *e2 = &sc->b[pairs[2 * k + 1]];
                if ((e1->fixed || !e1->active) && (e2->fixed || !e2->active)) continue;
                e1->stamp = e2->stamp = 1;
            }
            for (uint32_t j = 0; j < sc->nb; ++j) {
                if (!sc->b[j].stamp) continue;
                sc->b[j].stamp = 0;
                colliders_update(&sc->b[j]);
            }
            ORC_PHASE(PH_NARROW);
            for (uint32_t k = 0; k < np; ++k) {
                uint32_t i1 = pairs[2 * k], i2 = pairs[2 * k + 1];
                body *e1 = &sc->b[i1], *e2 = &sc->b[i2];
                if (!e1->fixed && !e2->fixed && !((e1->active && e2->active) || (!e1->active && !e2->active)))
                    sc->err |= ORC_ERR_ACTIVITY;
                if ((e1->fixed || !e1->active) && (e2->fixed || !e2->active)) continue;
                ++w->last_narrow_pairs;
                if (so.on) so.first[visited] = cons.n;
                ++visited;
                if (w->log_mode == 2) {  /* replay the logged constraints of this visit */
                    if (w->log_pair_cur >= w->log_pairs.n || w->log_pairs.d[w->log_pair_cur].i1 != i1 ||
                        w->log_pairs.d[w->log_pair_cur].i2 != i2) {
                        w->log_error = 1;
                        continue;
                    }
                    uint32_t nl = w->log_pairs.d[w->log_pair_cur++].n;
                    for (uint32_t l = 0; l < nl; ++l) {
                        const log_contact* lc = &w->log_contacts.d[w->log_contact_cur++];
                        constraint c;
                        memset(&c, 0, sizeof(c));
                        c.type = ORC_COLLISION_CONSTRAINT;
                        c.b1 = i1; c.b2 = i2;
                        c.r1 = V(lc->r1[0], lc->r1[1], lc->r1[2]);
                        c.r2 = V(lc->r2[0], lc->r2[1], lc->r2[2]);
                        c.n = V(lc->n[0], lc->n[1], lc->n[2]);
                        c.lambda_n = 0.0; c.lambda_t = 0.0;
                        ARR_PUSH(cons, c);
                    }
                    w->last_contacts += nl;
                    continue;
                }
                cts.n = 0;
                pair_contacts(e1, e2, &cts, &sc->err, &w->epa_failures);
                for (uint32_t l = 0; l < cts.n; ++l) {
                    constraint c = contact_to_constraint(e1, e2, i1, i2, &cts.d[l]);
                    ARR_PUSH(cons, c);
                    if (w->log_mode == 1) {
                        log_contact lc = {{RD(c.r1.x), RD(c.r1.y), RD(c.r1.z)}, {RD(c.r2.x), RD(c.r2.y), RD(c.r2.z)},
                                          {RD(c.n.x), RD(c.n.y), RD(c.n.z)}};
                        ARR_PUSH(w->log_contacts, lc);
                    }
                }
                if (w->log_mode == 1) {
                    log_pair lp = {i1, i2, cts.n};
                    ARR_PUSH(w->log_pairs, lp);
                }
                w->last_contacts += cts.n;
            }
        }
        if (so.on) so.first[so.n] = cons.n;
        /* position solve (pbd.cpp:615-620) */
        ORC_PHASE(PH_POSITION);
        for (uint32_t it = 0; it < iters; ++it) {
            if (!so.on) {
                for (uint32_t k = 0; k < cons.n; ++k) solve_constraint(sc, &cons.d[k], h);
                continue;
            }
            for (uint32_t m = 0; m < so.n; ++m) {
                uint32_t t = so.perm[m];
                for (uint32_t k = so.first[t]; k < so.first[t + 1]; ++k) solve_constraint(sc, &cons.d[k], h);
            }
        }
        /* velocity update (pbd.cpp:623-643) */
        ORC_PHASE(PH_VELOCITY_UPDATE);
        for (uint32_t j = 0; j < sc->nb; ++j) {
            body* e = &sc->b[j];
            if (e->fixed) continue;
            if (!e->active) continue;
            e->pv = e->v;
            e->pw = e->w;
            e->v = scl(1.0 / h, sub(e->x, e->px));
            qt pinv = q_inv(e->pq);
            qt dq = q_mul(e->q, pinv);
            if (dq.w >= 0.0) e->w = scl(2.0 / h, V(dq.x, dq.y, dq.z));
            else e->w = scl(-2.0 / h, V(dq.x, dq.y, dq.z));
        }
        /* velocity solve (pbd.cpp:646-711) */
        ORC_PHASE(PH_VELOCITY_SOLVE);
        for (uint32_t m = 0; m < (so.on ? so.n : 1); ++m) {
            uint32_t k0 = so.on ? so.first[so.perm[m]] : 0, k1 = so.on ? so.first[so.perm[m] + 1] : cons.n;
            for (uint32_t k = k0; k < k1; ++k) {
                constraint* c = &cons.d[k];
                if (c->type == ORC_COLLISION_CONSTRAINT) velocity_solve(c, &sc->b[c->b1], &sc->b[c->b2], h);
            }
        }
    }
    order_end(&so);
    ARR_FREE(cons);
    ARR_FREE(cts);
    free(pairs);
}

/* ------------------------------------------------------------------ */
/* Public API                                                           */
/* ------------------------------------------------------------------ */
static v3 ld3(const double* p, uint32_t i) { return V(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
static qt ld4(const double* p, uint32_t i) {
    qt q; q.x = p[4 * i]; q.y = p[4 * i + 1]; q.z = p[4 * i + 2]; q.w = p[4 * i + 3];
    return q;
}
static void st3(double* p, uint32_t i, v3 v) { p[3 * i] = RD(v.x); p[3 * i + 1] = RD(v.y); p[3 * i + 2] = RD(v.z); }
static void st4(double* p, uint32_t i, qt q) { p[4 * i] = RD(q.x); p[4 * i + 1] = RD(q.y); p[4 * i + 2] = RD(q.z); p[4 * i + 3] = RD(q.w); }

orc_world* orc_world_create(const orc_world_desc* d) {
    ORC_PHASE(PH_SETUP);
    orc_world* w = calloc(1, sizeof(orc_world));
    w->order_scene = -1;
    w->ns = d->num_scenes;
    w->s = calloc(w->ns ? w->ns : 1, sizeof(scene));
    w->nb_total = d->scene_body_offset[d->num_scenes];
    for (uint32_t si = 0; si < w->ns; ++si) {
        scene* sc = &w->s[si];
        uint32_t b0 = d->scene_body_offset[si], b1 = d->scene_body_offset[si + 1];
        sc->nb = b1 - b0;
        sc->b = calloc(sc->nb ? sc->nb : 1, sizeof(body));
        for (uint32_t k = 0; k < sc->nb; ++k) {
            uint32_t g = b0 + k;
            body* e = &sc->b[k];
            /* colliders: the body's shape list, hull corners widened and scaled
             * (examples_util.cpp:7-16), spheres of their shape's radius */
            uint32_t k0 = d->body_collider_offset ? d->body_collider_offset[g] : 0;
            uint32_t nc = d->body_collider_offset ? d->body_collider_offset[g + 1] - k0 : 1;
            e->ncol = nc;
            e->col = calloc(nc ? nc : 1, sizeof(collider));
            for (uint32_t ci = 0; ci < nc; ++ci) {
                uint32_t shp = d->body_collider_offset ? d->collider_shape[k0 + ci] : d->body_shape[g];
                collider* col = &e->col[ci];
                if (d->shape_sphere_radius && d->shape_sphere_radius[shp] > 0.0f) {
                    col->type = COL_SPHERE;
                    col->radius = d->shape_sphere_radius[shp];
                    col->center = V(0.0, 0.0, 0.0);
                    continue;
                }
                col->type = COL_HULL;
                uint32_t c0 = d->shape_corner_offset[shp], c1 = d->shape_corner_offset[shp + 1];
                v3* corners = malloc(sizeof(v3) * ((c1 - c0) ? (c1 - c0) : 1));
                v3 sc3 = ld3(d->body_scale, g);
                for (uint32_t c = c0; c < c1; ++c) {
                    v3 p = V((double)d->shape_corners[3 * c], (double)d->shape_corners[3 * c + 1],
                             (double)d->shape_corners[3 * c + 2]);
                    p.x *= sc3.x; p.y *= sc3.y; p.z *= sc3.z;
                    corners[c - c0] = p;
                }
                if (!hull_build(&col->h, corners, c1 - c0)) sc->err |= ORC_ERR_HULL;
                free(corners);
            }
            /* entity_create_ex (entity.cpp:24-65) */
            e->x = ld3(d->body_position, g);
            e->q = ld4(d->body_rotation, g);
            e->v = ld3(d->body_linear_velocity, g);
            e->w = ld3(d->body_angular_velocity, g);
            e->pv = V(0.0, 0.0, 0.0);
            e->pw = V(0.0, 0.0, 0.0);
            e->px = V(0.0, 0.0, 0.0);
            e->pq.x = e->pq.y = e->pq.z = e->pq.w = 0.0;
            /* colliders_get_bounding_sphere_radius (collider.cpp:178-189, 510-521) */
            real maxr = -DBL_MAX;
            for (uint32_t ci = 0; ci < e->ncol; ++ci) {
                const collider* col = &e->col[ci];
                real r = 0.0;
                if (col->type == COL_SPHERE) {
                    r = col->radius;
                } else {
                    for (uint32_t i = 0; i < col->h.nv; ++i) {
                        real l = len(col->h.v[i]);
                        if (l > r) r = l;
                    }
                }
                if (r > maxr) maxr = r;
            }
            e->radius = maxr;
            e->fixed = d->body_fixed[g] != 0;
            if (e->fixed) {
                e->inv_mass = 0.0;
                memset(&e->inertia, 0, sizeof(m3));
                memset(&e->inv_inertia, 0, sizeof(m3));
            } else {
                real mass = d->body_mass[g];
                e->inv_mass = 1.0 / mass;
                /* colliders_get_default_inertia_tensor (collider.cpp:448-494): a
                 * lone sphere is a solid ball; otherwise the mass is spread
                 * evenly over the vertices of all hull colliders (a sphere
                 * among several colliders is an assert in the reference) */
                m3 I;
                memset(&I, 0, sizeof(I));
                if (e->ncol == 1 && e->col[0].type == COL_SPHERE) {
                    real r = e->col[0].radius;
                    real Id = (2.0 / 5.0) * mass * r * r;
                    I.m[0][0] = Id; I.m[1][1] = Id; I.m[2][2] = Id;
                } else {
                    uint32_t total = 0;
                    for (uint32_t ci = 0; ci < e->ncol; ++ci) {
                        if (e->col[ci].type == COL_SPHERE) sc->err |= ORC_ERR_HULL;
                        else total += e->col[ci].h.nv;
                    }
                    real mpv = mass / (real)(double)total;
                    for (uint32_t ci = 0; ci < e->ncol; ++ci) {
                        if (e->col[ci].type == COL_SPHERE) continue;
                        const hull* h = &e->col[ci].h;
                        for (uint32_t i = 0; i < h->nv; ++i) {
                            v3 v = h->v[i];
                            I.m[0][0] += mpv * (v.y * v.y + v.z * v.z);
                            I.m[0][1] += mpv * v.x * v.y;
                            I.m[0][2] += mpv * v.x * v.z;
                            I.m[1][0] += mpv * v.x * v.y;
                            I.m[1][1] += mpv * (v.x * v.x + v.z * v.z);
                            I.m[1][2] += mpv * v.y * v.z;
                            I.m[2][0] += mpv * v.x * v.z;
                            I.m[2][1] += mpv * v.y * v.z;
                            I.m[2][2] += mpv * (v.x * v.x + v.y * v.y);
                        }
                    }
                }
                e->inertia = I;
                if (!m3_inverse(&e->inertia, &e->inv_inertia)) sc->err |= ORC_ERR_HULL;
            }
            e->active = 1;
            e->deact = 0.0;
            e->mu_s = d->body_material[3 * g];
            e->mu_d = d->body_material[3 * g + 1];
            e->e = d->body_material[3 * g + 2];
        }
        if (d->scene_constraint_offset && d->constraints) {
            uint32_t k0 = d->scene_constraint_offset[si], k1 = d->scene_constraint_offset[si + 1];
            sc->nc_ext = k1 - k0;
            sc->ext = calloc(sc->nc_ext ? sc->nc_ext : 1, sizeof(constraint));
            for (uint32_t k = 0; k < sc->nc_ext; ++k) {
                const orc_constraint_desc* cd = &d->constraints[k0 + k];
                constraint* c = &sc->ext[k];
                c->type = cd->type;
                c->b1 = cd->body1; c->b2 = cd->body2;
                c->r1 = V(cd->r1_lc[0], cd->r1_lc[1], cd->r1_lc[2]);
                c->r2 = V(cd->r2_lc[0], cd->r2_lc[1], cd->r2_lc[2]);
                c->compliance = cd->compliance;
                c->distance = V(cd->distance[0], cd->distance[1], cd->distance[2]);
                for (int a = 0; a < 4; ++a) { c->axis[a] = cd->axis[a]; c->limits[a] = cd->limits[a]; }
                c->limited = cd->limited;
                if (c->b1 >= sc->nb || c->b2 >= sc->nb) { sc->err |= ORC_ERR_SELECT; c->b1 = c->b2 = 0; }
            }
        }
    }
    return w;
}

void orc_world_contact_log(orc_world* w, int mode) {
    if (mode == 1) {
        w->log_pairs.n = 0;
        w->log_contacts.n = 0;
    }
    w->log_mode = mode;
    w->log_pair_cur = w->log_contact_cur = 0;
    w->log_error = 0;
}

int orc_world_contact_log_error(const orc_world* w) { return w->log_error; }

void orc_world_set_solve_order(orc_world* w, uint32_t scene_idx, const int32_t* item_class, uint32_t n) {
    free(w->order_class);
    w->order_class = malloc(sizeof(int32_t) * (n ? n : 1));
    if (n) memcpy(w->order_class, item_class, sizeof(int32_t) * n);
    w->order_n = n;
    w->order_scene = scene_idx < w->ns ? (int64_t)scene_idx : -1;
}

uint32_t orc_world_solve_order_status(const orc_world* w) { return w->order_status; }

/* Copies src's contact log into dst (for replay in a world built with other
 * parameters). */
void orc_world_contact_log_copy(orc_world* dst, const orc_world* src) {
    dst->log_pairs.n = 0;
    dst->log_contacts.n = 0;
    for (uint32_t k = 0; k < src->log_pairs.n; ++k) ARR_PUSH(dst->log_pairs, src->log_pairs.d[k]);
    for (uint32_t k = 0; k < src->log_contacts.n; ++k) ARR_PUSH(dst->log_contacts, src->log_contacts.d[k]);
}

void orc_world_destroy(orc_world* w) {
    if (!w) return;
    free(w->log_pairs.d);
    free(w->log_contacts.d);
    free(w->order_class);
    for (uint32_t si = 0; si < w->ns; ++si) {
        for (uint32_t k = 0; k < w->s[si].nb; ++k) {
            body* e = &w->s[si].b[k];
            for (uint32_t ci = 0; ci < e->ncol; ++ci)
                if (e->col[ci].type == COL_HULL) hull_free(&e->col[ci].h);
            free(e->col);
        }
        free(w->s[si].b);
        free(w->s[si].ext);
    }
    free(w->s);
    free(w);
}

uint32_t orc_world_num_bodies(const orc_world* w) { return w->nb_total; }

void orc_world_step(orc_world* w, double dt, uint32_t frames, uint32_t substeps, uint32_t iters, int32_t collisions,
                    double gravity) {
    w->last_contacts = 0;
    w->last_narrow_pairs = 0;
    for (uint32_t f = 0; f < frames; ++f) {
        for (uint32_t si = 0; si < w->ns; ++si) {
            scene* sc = &w->s[si];
            ORC_PHASE(PH_FRAME);
            /* Caller pattern (stack.cpp:93-96) and the force/torque sums
             * (physics_util.cpp:5-23) over the single gravity force. */
            for (uint32_t k = 0; k < sc->nb; ++k) {
                body* e = &sc->b[k];
                v3 pos = V(0.0, 0.0, 0.0);
                v3 force = V(0.0, -gravity * 1.0 / e->inv_mass, 0.0);
                e->force = add(V(0.0, 0.0, 0.0), force);
                e->torque = add(V(0.0, 0.0, 0.0), cross(sub(pos, V(0.0, 0.0, 0.0)), force));
            }
            simulate(w, sc, dt, substeps, iters, collisions);
        }
    }
}

void orc_world_get_state(const orc_world* w, const orc_state* o) {
    uint32_t g = 0;
    for (uint32_t si = 0; si < w->ns; ++si) {
        const scene* sc = &w->s[si];
        for (uint32_t k = 0; k < sc->nb; ++k, ++g) {
            const body* e = &sc->b[k];
            if (o->position) st3(o->position, g, e->x);
            if (o->rotation) st4(o->rotation, g, e->q);
            if (o->linear_velocity) st3(o->linear_velocity, g, e->v);
            if (o->angular_velocity) st3(o->angular_velocity, g, e->w);
            if (o->previous_position) st3(o->previous_position, g, e->px);
            if (o->previous_rotation) st4(o->previous_rotation, g, e->pq);
            if (o->previous_linear_velocity) st3(o->previous_linear_velocity, g, e->pv);
            if (o->previous_angular_velocity) st3(o->previous_angular_velocity, g, e->pw);
            if (o->active) o->active[g] = e->active;
            if (o->deactivation_time) o->deactivation_time[g] = RD(e->deact);
        }
    }
}

void orc_world_set_state(orc_world* w, const orc_state* in) {
    uint32_t g = 0;
    for (uint32_t si = 0; si < w->ns; ++si) {
        scene* sc = &w->s[si];
        for (uint32_t k = 0; k < sc->nb; ++k, ++g) {
            body* e = &sc->b[k];
            if (in->position) e->x = ld3(in->position, g);
            if (in->rotation) e->q = ld4(in->rotation, g);
            if (in->linear_velocity) e->v = ld3(in->linear_velocity, g);
            if (in->angular_velocity) e->w = ld3(in->angular_velocity, g);
            if (in->previous_position) e->px = ld3(in->previous_position, g);
            if (in->previous_rotation) e->pq = ld4(in->previous_rotation, g);
            if (in->previous_linear_velocity) e->pv = ld3(in->previous_linear_velocity, g);
            if (in->previous_angular_velocity) e->pw = ld3(in->previous_angular_velocity, g);
            if (in->active) e->active = in->active[g];
            if (in->deactivation_time) e->deact = in->deactivation_time[g];
        }
    }
}

uint32_t orc_world_error(const orc_world* w, uint32_t scene_idx) {
    return scene_idx < w->ns ? w->s[scene_idx].err : 0;
}
uint64_t orc_world_epa_failures(const orc_world* w) { return w->epa_failures; }
uint64_t orc_world_last_contacts(const orc_world* w) { return w->last_contacts; }
uint64_t orc_world_last_narrow_pairs(const orc_world* w) { return w->last_narrow_pairs; }

void orc_world_body_info(const orc_world* w, double* inv_mass, double* inertia, double* inv_inertia, double* radius) {
    uint32_t g = 0;
    for (uint32_t si = 0; si < w->ns; ++si) {
        for (uint32_t k = 0; k < w->s[si].nb; ++k, ++g) {
            const body* e = &w->s[si].b[k];
            if (inv_mass) inv_mass[g] = RD(e->inv_mass);
            if (inertia) memcpy(inertia + 9 * g, e->inertia.m, sizeof(double) * 9);
            if (inv_inertia) memcpy(inv_inertia + 9 * g, e->inv_inertia.m, sizeof(double) * 9);
            if (radius) radius[g] = RD(e->radius);
        }
    }
}

uint32_t orc_world_broad_pairs(const orc_world* w, uint32_t si, uint32_t* out, uint32_t cap) {
    if (si >= w->ns) return 0;
    uint32_t* pairs = NULL;
    uint32_t np = broad(&w->s[si], &pairs);
    for (uint32_t k = 0; k < np && k < cap; ++k) { out[2 * k] = pairs[2 * k]; out[2 * k + 1] = pairs[2 * k + 1]; }
    free(pairs);
    return np;
}

uint32_t orc_world_contacts(orc_world* w, uint32_t si, uint32_t i, uint32_t j, double* out, uint32_t cap) {
    if (si >= w->ns) return 0;
    scene* sc = &w->s[si];
    if (i >= sc->nb || j >= sc->nb) return 0;
    body *e1 = &sc->b[i], *e2 = &sc->b[j];
    colliders_update(e1);
    colliders_update(e2);
    contactarr cts = {0};
    uint64_t fails = 0;
    pair_contacts(e1, e2, &cts, &sc->err, &fails);
    for (uint32_t k = 0; k < cts.n && k < cap; ++k) {
        const contact* c = &cts.d[k];
        double* o = out + 9 * k;
        o[0] = RD(c->p1.x); o[1] = RD(c->p1.y); o[2] = RD(c->p1.z);
        o[3] = RD(c->p2.x); o[4] = RD(c->p2.y); o[5] = RD(c->p2.z);
        o[6] = RD(c->n.x); o[7] = RD(c->n.y); o[8] = RD(c->n.z);
    }
    uint32_t n = cts.n;
    ARR_FREE(cts);
    return n;
}

void orc_world_hull(const orc_world* w, uint32_t body_idx, uint32_t counts[6], double* vertices, double* face_normals,
                    uint32_t* face_off, uint32_t* face_el, uint32_t* vf_off, uint32_t* vf, uint32_t* vn_off,
                    uint32_t* vn, uint32_t* ff_off, uint32_t* ff) {
    uint32_t g = 0;
    for (uint32_t si = 0; si < w->ns; ++si) {
        if (body_idx < g + w->s[si].nb) {
            const body* e = &w->s[si].b[body_idx - g];
            if (e->col[0].type != COL_HULL) break;
            const hull* h = &e->col[0].h;  /* the body's first collider */
            counts[0] = h->nv; counts[1] = h->nf;
            counts[2] = h->fe_off[h->nf]; counts[3] = h->vf_off[h->nv];
            counts[4] = h->vn_off[h->nv]; counts[5] = h->ff_off[h->nf];
            if (vertices) memcpy(vertices, h->v, sizeof(v3) * h->nv);
            if (face_normals) memcpy(face_normals, h->fn, sizeof(v3) * h->nf);
            if (face_off) memcpy(face_off, h->fe_off, sizeof(uint32_t) * (h->nf + 1));
            if (face_el) memcpy(face_el, h->fe, sizeof(uint32_t) * counts[2]);
            if (vf_off) memcpy(vf_off, h->vf_off, sizeof(uint32_t) * (h->nv + 1));
            if (vf) memcpy(vf, h->vf, sizeof(uint32_t) * counts[3]);
            if (vn_off) memcpy(vn_off, h->vn_off, sizeof(uint32_t) * (h->nv + 1));
            if (vn) memcpy(vn, h->vn, sizeof(uint32_t) * counts[4]);
            if (ff_off) memcpy(ff_off, h->ff_off, sizeof(uint32_t) * (h->nf + 1));
            if (ff) memcpy(ff, h->ff, sizeof(uint32_t) * counts[5]);
            return;
        }
        g += w->s[si].nb;
    }
    memset(counts, 0, sizeof(uint32_t) * 6);
}
