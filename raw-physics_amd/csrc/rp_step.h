// rp_step.h — per-pair and per-constraint device functions of the step.
//
// Each function restates one reference routine (file:line given) with the
// same floating-point evaluation order, so that results are bit-identical
// to the reference's CPU step. Memory is bounded: where the reference grows
// heap arrays, these functions use fixed-capacity per-thread arrays and set
// an error bit if a capacity would be exceeded.
#pragma once
#include <float.h>
#include "rp_math.h"

namespace rp {

// ---------------------------------------------------------------------------
// Device data layout
// ---------------------------------------------------------------------------
template <class S> struct BodyT {   // one record per body (464 bytes for S = double)
    V3T<S> x; Q4T<S> q; V3T<S> v; V3T<S> w;      // world_position, world_rotation, velocities
    V3T<S> px; Q4T<S> pq; V3T<S> pv; V3T<S> pw;  // previous_* (entity.h:43-46)
    S deact;                                     // deactivation_time
    S inv_mass;
    M3T<S> I, Iinv;                              // local inertia tensor and its inverse
    S mu_s, mu_d, e, radius;
    V3T<S> F, T;                                 // external force / torque of the frame
    int32_t fixed, active;
};
using Body = BodyT<double>;

// A collision constraint as built by clipping_contact_to_collision_constraint
// (pbd.cpp:408-424) plus its multipliers.
template <class S> struct ContactT {
    V3T<S> r1, r2, n;
    S lambda_n, lambda_t;
};
using Contact = ContactT<double>;

// External constraint (Constraint, pbd.h:22-91); b1/b2 are global body indices.
struct ExtC {
    int32_t type, b1, b2, limited;
    int32_t axis[4];
    V3 r1, r2, distance;
    double compliance;
    double limits[4];
};

// Hull topology and geometry of one body at the current substep. Vertices
// are pre-transformed; face normals are transformed when first needed, with
// the reference's arithmetic (collider.cpp:426-428), so that only the few
// faces the manifold visits are computed.
struct HullView {
    const V3* tv;  // transformed vertices
    const V3* fn;  // local face normals
    M3 R;          // upper-left 3x3 of the model matrix T*R (util.cpp:60-74)
    int nv, nf;
    const int32_t *fe_off, *fe, *vf_off, *vf, *vn_off, *vn, *ff_off, *ff;
};

// A collider of a body with more than one collider or a sphere (the general
// narrow-phase path): a hull view, or a sphere whose center is the body
// position (collider.cpp:431-433) with an r32 radius (collider.h:29).
enum { COL_HULL = 0, COL_SPHERE = 1 };
struct ColView {
    int type;
    HullView h;
    V3 center;
    float radius;
};

// gm_mat4_multiply_vec3(M, n, false) then normalize (collider.cpp:426-428)
RP_HD V3 face_normal_world(const HullView& h, int f) {
    V3 n = h.fn[f];
    const M3& M = h.R;
    V3 r = v3(M.m[0][0] * n.x + M.m[0][1] * n.y + M.m[0][2] * n.z, M.m[1][0] * n.x + M.m[1][1] * n.y + M.m[1][2] * n.z,
              M.m[2][0] * n.x + M.m[2][1] * n.y + M.m[2][2] * n.z);
    return nrm(r);
}

enum : uint32_t {
    ERR_GJK_SIMPLEX = 1u << 0,
    ERR_EPA_DEGENERATE = 1u << 1,
    ERR_SUPPORT_NAN = 1u << 2,
    ERR_SKEW_LINES = 1u << 3,
    ERR_ACTIVITY = 1u << 4,
    ERR_HULL = 1u << 5,
    ERR_ZERO_MASS = 1u << 6,
    ERR_SELECT = 1u << 7,
    ERR_EPA_CAPACITY = 1u << 8,
    ERR_CLIP_CAPACITY = 1u << 9,
    ERR_CONTACT_CAPACITY = 1u << 10,
    ERR_SOLVER_CAPACITY = 1u << 11,
    ERR_PAIR_CAPACITY = 1u << 12,
    ERR_ADJOINT_SCOPE = 1u << 13,
};

constexpr int EPA_MAXV = 104;  // 4 simplex points + at most 100 expansions (epa.cpp:149)
constexpr int EPA_MAXF = 256;
constexpr int EPA_MAXE = 128;
constexpr int CLIP_MAX = 64;
constexpr int MAXC = 16;       // contacts kept per body pair
// Index of a first-maximum / first-minimum search that has found nothing:
// larger than every real index, so "ties to the lower index" also prefers
// a lane that found something (its value is then strictly better, since
// the search starts from the most extreme finite value).
constexpr int NO_SEL = 0x7fffffff;

// ---------------------------------------------------------------------------
// Support mapping (support.cpp:5-39): first index wins on ties
// ---------------------------------------------------------------------------
// G = 1: one lane searches all vertices; they are loaded eight at a time so
// that their memory latencies overlap, and compared in index order.
// G > 1: the G lanes of an aligned group (one body pair per group) each scan
// every G-th vertex, then reduce (larger dot wins, ties to the lower index),
// which selects exactly the first maximum of the sequential scan; every lane
// of the group returns the same index.

// Lane exchanges inside a row of 16 lanes (DPP): partner = lane ^ 1, lane ^ 2
// (quad permutations), 7 - lane within 8 (half mirror), 15 - lane (mirror).
enum : int { DPP_XOR1 = 0xB1, DPP_XOR2 = 0x4E, DPP_HALF_MIRROR = 0x141, DPP_MIRROR = 0x140 };
template <int CTRL> __device__ __forceinline__ int dpp_int(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, true);
}
template <int CTRL> __device__ __forceinline__ double dpp_double(double v) {
    long long b = __double_as_longlong(v);
    int lo = dpp_int<CTRL>((int)b), hi = dpp_int<CTRL>((int)(b >> 32));
    return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}
// (dot, index) of the partner lane, kept if its dot is larger, or equal
// with a lower index. A lane that has found nothing holds (-DBL_MAX,
// NO_SEL); a lane that has found something holds a dot above -DBL_MAX (its
// scan took only values above it), so one comparison of each kind decides
// every case.
// Written as selects over a non-short-circuit predicate, so that it compiles
// to compares and v_cndmask instead of a branch per term.
template <int CTRL> __device__ __forceinline__ void argmax_step(double& m, int& sel) {
    double om = dpp_double<CTRL>(m);
    int os = dpp_int<CTRL>(sel);
    bool take = (om > m) | ((om == m) & (os < sel));
    m = take ? om : m;
    sel = take ? os : sel;
}
// First maximum over the G lanes of the group (G = 4, 8 or 16): after each
// step every lane holds the best of a block of 2, 4, 8, 16 lanes.
template <int G> __device__ __forceinline__ void group_argmax(double& m, int& sel) {
    static_assert(G == 4 || G == 8 || G == 16, "lane groups of 4, 8 or 16");
    argmax_step<DPP_XOR1>(m, sel);
    argmax_step<DPP_XOR2>(m, sel);
    if constexpr (G >= 8) argmax_step<DPP_HALF_MIRROR>(m, sel);
    if constexpr (G >= 16) argmax_step<DPP_MIRROR>(m, sel);
}
// Two independent first maxima, their steps interleaved so that the two
// dependency chains overlap.
template <int G> __device__ __forceinline__ void group_argmax2(double& m1, int& s1, double& m2, int& s2) {
    static_assert(G == 4 || G == 8 || G == 16, "lane groups of 4, 8 or 16");
    argmax_step<DPP_XOR1>(m1, s1);
    argmax_step<DPP_XOR1>(m2, s2);
    argmax_step<DPP_XOR2>(m1, s1);
    argmax_step<DPP_XOR2>(m2, s2);
    if constexpr (G >= 8) {
        argmax_step<DPP_HALF_MIRROR>(m1, s1);
        argmax_step<DPP_HALF_MIRROR>(m2, s2);
    }
    if constexpr (G >= 16) {
        argmax_step<DPP_MIRROR>(m1, s1);
        argmax_step<DPP_MIRROR>(m2, s2);
    }
}

// (value, index) of the partner lane, kept if its value is smaller, or equal
// with a lower index; a lane that has found nothing holds (DBL_MAX, NO_SEL).
template <int CTRL> __device__ __forceinline__ void argmin_step(double& m, int& sel) {
    double om = dpp_double<CTRL>(m);
    int os = dpp_int<CTRL>(sel);
    bool take = (om < m) | ((om == m) & (os < sel));
    m = take ? om : m;
    sel = take ? os : sel;
}
// First minimum over the G lanes of the group.
template <int G> __device__ __forceinline__ void group_argmin(double& m, int& sel) {
    static_assert(G == 4 || G == 8 || G == 16, "lane groups of 4, 8 or 16");
    argmin_step<DPP_XOR1>(m, sel);
    argmin_step<DPP_XOR2>(m, sel);
    if constexpr (G >= 8) argmin_step<DPP_HALF_MIRROR>(m, sel);
    if constexpr (G >= 16) argmin_step<DPP_MIRROR>(m, sel);
}

template <int G = 1>
RP_HD int support_index(const HullView& h, V3 d, uint32_t& err) {
    int sel = NO_SEL;
    double max_dot = -DBL_MAX;
    if constexpr (G == 1) {
        int i = 0;
        for (; i + 8 <= h.nv; i += 8) {
            V3 v[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) v[k] = h.tv[i + k];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                double dd = dot(v[k], d);
                if (dd > max_dot) { sel = i + k; max_dot = dd; }
            }
        }
        for (; i < h.nv; ++i) {
            double dd = dot(h.tv[i], d);
            if (dd > max_dot) { sel = i; max_dot = dd; }
        }
    } else {
        int gl = (int)(__lane_id() & (G - 1));
        for (int i = gl; i < h.nv; i += G) {
            double dd = dot(h.tv[i], d);
            bool t = dd > max_dot;
            sel = t ? i : sel;
            max_dot = t ? dd : max_dot;
        }
        group_argmax<G>(max_dot, sel);
    }
    if (sel == NO_SEL) { err |= ERR_SUPPORT_NAN; sel = 0; }
    return sel;
}
// Support indices of two hulls in directions d1 and d2 at once: the two
// scans advance together (eight vertices of each hull per batch for G = 1,
// one vertex of each per lane and step for G > 1) so that their memory
// latencies overlap; each scan keeps its own first-maximum order, and for
// G > 1 the two group reductions run side by side.
template <int G = 1>
RP_HD void support_index2(const HullView& a, V3 d1, const HullView& b, V3 d2, int& i1, int& i2, uint32_t& err) {
    int s1 = NO_SEL, s2 = NO_SEL;
    double m1 = -DBL_MAX, m2 = -DBL_MAX;
    int n = a.nv > b.nv ? a.nv : b.nv;
    if constexpr (G == 1) {
        for (int i = 0; i < n; i += 8) {
            V3 va[8], vb[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if (i + k < a.nv) va[k] = a.tv[i + k];
                if (i + k < b.nv) vb[k] = b.tv[i + k];
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if (i + k < a.nv) {
                    double dd = dot(va[k], d1);
                    if (dd > m1) { s1 = i + k; m1 = dd; }
                }
                if (i + k < b.nv) {
                    double dd = dot(vb[k], d2);
                    if (dd > m2) { s2 = i + k; m2 = dd; }
                }
            }
        }
    } else {
        // a lane visits its vertices gl, gl + G, ... in order, four of each
        // hull per batch, all loaded before any is compared; positions past a
        // hull's end load its last vertex (in bounds) and are masked out of
        // the comparison, so the batch has no branches
        int gl = (int)(__lane_id() & (G - 1));
        for (int i = gl; i < n; i += 4 * G) {
            V3 va[4], vb[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                int j = i + k * G;
                va[k] = a.tv[j < a.nv ? j : a.nv - 1];
                vb[k] = b.tv[j < b.nv ? j : b.nv - 1];
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                int j = i + k * G;
                double da = dot(va[k], d1);
                bool ta = (j < a.nv) & (da > m1);
                s1 = ta ? j : s1;
                m1 = ta ? da : m1;
                double db = dot(vb[k], d2);
                bool tb = (j < b.nv) & (db > m2);
                s2 = tb ? j : s2;
                m2 = tb ? db : m2;
            }
        }
        group_argmax2<G>(m1, s1, m2, s2);
    }
    if (s1 == NO_SEL) { err |= ERR_SUPPORT_NAN; s1 = 0; }
    if (s2 == NO_SEL) { err |= ERR_SUPPORT_NAN; s2 = 0; }
    i1 = s1;
    i2 = s2;
}

template <int G = 1>
RP_HD V3 support_md(const HullView& a, const HullView& b, V3 d, uint32_t& err) {
    int i1, i2;
    support_index2<G>(a, d, b, scl(-1.0, d), i1, i2, err);
    return sub(a.tv[i1], b.tv[i2]);
}

// support_point of one collider (support.cpp:19-31); the general path runs
// one lane per pair.
RP_HD V3 support_point(const ColView& c, V3 d, uint32_t& err) {
    if (c.type == COL_SPHERE) return add(c.center, scl((double)c.radius, nrm(d)));
    return c.h.tv[support_index<1>(c.h, d, err)];
}
// support_point_of_minkowski_difference (support.cpp:34-39)
template <int G = 1>
RP_HD V3 support_md(const ColView& a, const ColView& b, V3 d, uint32_t& err) {
    V3 s1 = support_point(a, d, err);
    V3 s2 = support_point(b, scl(-1.0, d), err);
    return sub(s1, s2);
}

// ---------------------------------------------------------------------------
// GJK (gjk.cpp:7-379)
// ---------------------------------------------------------------------------
struct Simplex { V3 a, b, c, d; int num; };

RP_HD V3 triple_cross(V3 a, V3 b, V3 c) { return cross(cross(a, b), c); }

// gjk.cpp:36-55
RP_HD void do_simplex_2(Simplex& s, V3& dir) {
    V3 a = s.a, b = s.b;
    V3 ao = inv3(a), ab = sub(b, a);
    if (dot(ab, ao) >= 0.0) { s.a = a; s.b = b; s.num = 2; dir = triple_cross(ab, ao, ab); }
    else { s.a = a; s.num = 1; dir = ao; }
}
// gjk.cpp:57-120 ("A region" keeps num, gjk.cpp:81-85,95-98)
RP_HD void do_simplex_3(Simplex& s, V3& dir) {
    V3 a = s.a, b = s.b, c = s.c;
    V3 ao = inv3(a), ab = sub(b, a), ac = sub(c, a), abc = cross(ab, ac);
    if (dot(cross(abc, ac), ao) >= 0.0) {
        if (dot(ac, ao) >= 0.0) { s.a = a; s.b = c; s.num = 2; dir = triple_cross(ac, ao, ac); }
        else if (dot(ab, ao) >= 0.0) { s.a = a; s.b = b; s.num = 2; dir = triple_cross(ab, ao, ab); }
        else { s.a = a; dir = ao; }
    } else {
        if (dot(cross(ab, abc), ao) >= 0.0) {
            if (dot(ab, ao) >= 0.0) { s.a = a; s.b = b; s.num = 2; dir = triple_cross(ab, ao, ab); }
            else { s.a = a; dir = ao; }
        } else if (dot(abc, ao) >= 0.0) { s.a = a; s.b = b; s.c = c; s.num = 3; dir = abc; }
        else { s.a = a; s.b = c; s.c = b; s.num = 3; dir = inv3(abc); }
    }
}
// gjk.cpp:122-337 (the ACD/AC branch uses ab, gjk.cpp:213)
RP_HD bool do_simplex_4(Simplex& s, V3& dir) {
    V3 a = s.a, b = s.b, c = s.c, d = s.d;
    V3 ao = inv3(a), ab = sub(b, a), ac = sub(c, a), ad = sub(d, a);
    V3 abc = cross(ab, ac), acd = cross(ac, ad), adb = cross(ad, ab);
    unsigned pi = 0;
    if (dot(abc, ao) >= 0.0) pi |= 1;
    if (dot(acd, ao) >= 0.0) pi |= 2;
    if (dot(adb, ao) >= 0.0) pi |= 4;
    switch (pi) {
    case 0: return true;
    case 1:
        if (dot(cross(abc, ac), ao) >= 0.0) {
            if (dot(ac, ao) >= 0.0) { s.a = a; s.b = c; s.num = 2; dir = triple_cross(ac, ao, ac); }
            else if (dot(ab, ao) >= 0.0) { s.a = a; s.b = b; s.num = 2; dir = triple_cross(ab, ao, ab); }
            else { s.a = a; dir = ao; }
        } else {
            if (dot(cross(ab, abc), ao) >= 0.0) {
                if (dot(ab, ao) >= 0.0) { s.a = a; s.b = b; s.num = 2; dir = triple_cross(ab, ao, ab); }
                else { s.a = a; dir = ao; }
            } else { s.a = a; s.b = b; s.c = c; s.num = 3; dir = abc; }
        }
        break;
    case 2:
        if (dot(cross(acd, ad), ao) >= 0.0) {
            if (dot(ad, ao) >= 0.0) { s.a = a; s.b = d; s.num = 2; dir = triple_cross(ad, ao, ad); }
            else if (dot(ac, ao) >= 0.0) { s.a = a; s.b = c; s.num = 2; dir = triple_cross(ab, ao, ab); }
            else { s.a = a; dir = ao; }
        } else {
            if (dot(cross(ac, acd), ao) >= 0.0) {
                if (dot(ac, ao) >= 0.0) { s.a = a; s.b = c; s.num = 2; dir = triple_cross(ac, ao, ac); }
                else { s.a = a; dir = ao; }
            } else { s.a = a; s.b = c; s.c = d; s.num = 3; dir = acd; }
        }
        break;
    case 3:
        if (dot(ac, ao) >= 0.0) { s.a = a; s.b = c; s.num = 2; dir = triple_cross(ac, ao, ac); }
        else { s.a = a; s.num = 1; dir = ao; }
        break;
    case 4:
        if (dot(cross(adb, ab), ao) >= 0.0) {
            if (dot(ab, ao) >= 0.0) { s.a = a; s.b = b; s.num = 2; dir = triple_cross(ab, ao, ab); }
            else if (dot(ad, ao) >= 0.0) { s.a = a; s.b = d; s.num = 2; dir = triple_cross(ad, ao, ad); }
            else { s.a = a; dir = ao; }
        } else {
            if (dot(cross(ad, adb), ao) >= 0.0) {
                if (dot(ad, ao) >= 0.0) { s.a = a; s.b = d; s.num = 2; dir = triple_cross(ad, ao, ad); }
                else { s.a = a; dir = ao; }
            } else { s.a = a; s.b = d; s.c = b; s.num = 3; dir = adb; }
        }
        break;
    case 5:
        if (dot(ab, ao) >= 0.0) { s.a = a; s.b = b; s.num = 2; dir = triple_cross(ab, ao, ab); }
        else { s.a = a; s.num = 1; dir = ao; }
        break;
    case 6:
        if (dot(ad, ao) >= 0.0) { s.a = a; s.b = d; s.num = 2; dir = triple_cross(ad, ao, ad); }
        else { s.a = a; s.num = 1; dir = ao; }
        break;
    case 7: s.a = a; s.num = 1; dir = ao; break;
    }
    return false;
}

RP_HD bool same_bits(V3 a, V3 b) {
    return __double_as_longlong(a.x) == __double_as_longlong(b.x) &&
           __double_as_longlong(a.y) == __double_as_longlong(b.y) &&
           __double_as_longlong(a.z) == __double_as_longlong(b.z);
}
RP_HD bool same_state(const Simplex& s, V3 dir, const Simplex& t, V3 tdir) {
    return s.num == t.num && same_bits(dir, tdir) && same_bits(s.a, t.a) && same_bits(s.b, t.b) &&
           same_bits(s.c, t.c) && same_bits(s.d, t.d);
}

// gjk_collides (gjk.cpp:350-379).
// Early exit on a repeated state, with the same result as the full loop: the
// loop state (simplex, its count and the search direction) determines the
// rest of the loop, and it takes values in a finite set (support points are
// differences of hull vertices, the direction a function of the simplex). A
// run that neither separates nor encloses the origin therefore cycles until
// its 100th iteration and returns false. Brent's method finds the cycle: the
// state is saved at iterations 1, 2, 4, 8, ... and compared bit for bit with
// every later state; a match means the iterations since the save returned
// nothing and would repeat until the limit (their error bits are already
// set), so false is returned at once. `cycled` is set when this happens.
template <int G = 1, class View = HullView>
RP_HD bool gjk(const View& h1, const View& h2, Simplex& s, uint32_t& err, bool& cycled) {
    s.a = support_md<G>(h1, h2, v3(0.0, 0.0, 1.0), err);
    s.b = s.c = s.d = v3(0.0, 0.0, 0.0);
    s.num = 1;
    V3 dir = scl(-1.0, s.a);
    Simplex saved = s;
    V3 saved_dir = dir;
    int power = 1, lam = 0;
    for (int i = 0; i < 100; ++i) {
        if (lam == power) {
            saved = s;
            saved_dir = dir;
            power <<= 1;
            lam = 0;
        }
        ++lam;
        if (i > 0 && same_state(s, dir, saved, saved_dir) && lam > 1) {
            cycled = true;
            return false;
        }
        V3 p = support_md<G>(h1, h2, dir, err);
        if (dot(p, dir) < 0.0) return false;
        // add_to_simplex (gjk.cpp:7-30); num == 4 is an assert in the reference
        switch (s.num) {
        case 1: s.b = s.a; s.a = p; break;
        case 2: s.c = s.b; s.b = s.a; s.a = p; break;
        case 3: s.d = s.c; s.c = s.b; s.b = s.a; s.a = p; break;
        default: err |= ERR_GJK_SIMPLEX; return false;
        }
        ++s.num;
        bool hit = false;
        if (s.num == 2) do_simplex_2(s, dir);
        else if (s.num == 3) do_simplex_3(s, dir);
        else hit = do_simplex_4(s, dir);
        if (hit) return true;
    }
    return false;
}

// ---------------------------------------------------------------------------
// EPA (epa.cpp:8-237)
// ---------------------------------------------------------------------------
struct I3 { int32_t x, y, z; };

// get_face_normal_and_distance_to_origin (epa.cpp:32-77)
RP_HD bool epa_face(I3 f, const V3* poly, int np, V3& n_out, double& d_out) {
    V3 a = poly[f.x], b = poly[f.y], c = poly[f.z];
    V3 n = nrm(cross(sub(b, a), sub(c, a)));
    if (is_zero(n)) return false;
    const double TOL = 0.0;
    double d = dot(n, a);
    if (d < -TOL) {
        n = inv3(n);
        d = -d;
    } else if (d >= -TOL && d <= TOL) {
        bool found = false;
        for (int i = 0; i < np; ++i) {
            double ad = dot(n, poly[i]);
            if (ad < -TOL || ad > TOL) {
                n = ad < -TOL ? n : inv3(n);
                found = true;
                break;
            }
        }
        if (!found) return false;
    }
    n_out = n;
    d_out = d;
    return true;
}

// The EPA polytope (epa.cpp:118-237) over caller-owned arrays (separate
// locals, so that the compiler can overlap their scratch with later phases).
struct EpaPoly {
    V3* poly;      // [EPA_MAXV]
    I3* faces;     // [EPA_MAXF]
    V3* normals;   // [EPA_MAXF]
    double* dist;  // [EPA_MAXF]
    int2* edges;   // [EPA_MAXE]
    int np, nf, ne;
    V3 min_n;
    double min_d;

    // The simplex's four faces and the closest one; false on a degenerate face.
    RP_HD bool init(const Simplex& s, uint32_t& err) {
        np = 4; nf = 4; ne = 0;
        poly[0] = s.a; poly[1] = s.b; poly[2] = s.c; poly[3] = s.d;
        faces[0] = {0, 1, 2}; faces[1] = {0, 2, 3}; faces[2] = {0, 3, 1}; faces[3] = {1, 2, 3};
        min_n = v3(0.0, 0.0, 0.0);
        min_d = DBL_MAX;
        for (int i = 0; i < nf; ++i) {
            V3 n; double d;
            if (!epa_face(faces[i], poly, np, n, d)) { err |= ERR_EPA_DEGENERATE; return false; }
            normals[i] = n;
            dist[i] = d;
            if (d < min_d) { min_d = d; min_n = n; }
        }
        return true;
    }
    // add_edge (epa.cpp:79-110): an edge equal to a listed one, by index or
    // by position, cancels it (swap-remove), else it is appended; false when
    // the list is full.
    RP_HD bool add_edge(int2 e, uint32_t& err) {
        for (int j = 0; j < ne; ++j) {
            int2 cur = edges[j];
            bool hit = (e.x == cur.x && e.y == cur.y) || (e.x == cur.y && e.y == cur.x);
            if (!hit) {
                V3 c1 = poly[cur.x], c2 = poly[cur.y], e1 = poly[e.x], e2 = poly[e.y];
                hit = (veq(c1, e1) && veq(c2, e2)) || (veq(c1, e2) && veq(c2, e1));
            }
            if (hit) {
                --ne;
                edges[j] = edges[ne];
                return true;
            }
        }
        if (ne == EPA_MAXE) { err |= ERR_EPA_CAPACITY; return false; }
        edges[ne++] = e;
        return true;
    }
    // triangle_centroid (epa.cpp:112-116) test: does face i see sp?
    RP_HD bool sees(int i, V3 sp) const {
        I3 f = faces[i];
        V3 c = scl(1.0 / 3.0, add(add(poly[f.y], poly[f.z]), poly[f.x]));
        return dot(normals[i], sub(sp, c)) > 0.0;
    }
    // One iteration with the support point sp in direction min_n:
    // 1 = converged (n_out, pen_out set), 2 = expanded, -1 = error.
    RP_HD int step(V3 sp, V3& n_out, double& pen_out, uint32_t& err) {
        double d = dot(min_n, sp);
        if (fabs(d - min_d) < 0.0001) {
            n_out = min_n;
            pen_out = min_d;
            return 1;
        }
        int new_idx = np;
        poly[np++] = sp;
        for (int i = 0; i < nf; ++i) {
            V3 n = normals[i];
            I3 f = faces[i];
            // triangle_centroid (epa.cpp:112-116)
            V3 c = scl(1.0 / 3.0, add(add(poly[f.y], poly[f.z]), poly[f.x]));
            if (dot(n, sub(sp, c)) > 0.0) {
                if (!add_edge(make_int2(f.x, f.y), err) || !add_edge(make_int2(f.y, f.z), err) ||
                    !add_edge(make_int2(f.z, f.x), err))
                    return -1;
                // swap-remove faces / distances / normals together (epa.cpp:191-195)
                --nf;
                faces[i] = faces[nf];
                dist[i] = dist[nf];
                normals[i] = normals[nf];
                --i;
            }
        }
        for (int i = 0; i < ne; ++i) {
            if (nf == EPA_MAXF) { err |= ERR_EPA_CAPACITY; return -1; }
            I3 f = {edges[i].x, edges[i].y, new_idx};
            faces[nf] = f;
            V3 n; double fd;
            if (!epa_face(f, poly, np, n, fd)) { err |= ERR_EPA_DEGENERATE; return -1; }
            normals[nf] = n;
            dist[nf] = fd;
            ++nf;
        }
        // first minimum: selects over the distances, then one normal load
        min_d = DBL_MAX;
        int mi = -1;
        for (int i = 0; i < nf; ++i) {
            double di = dist[i];
            bool t = di < min_d;
            min_d = t ? di : min_d;
            mi = t ? i : mi;
        }
        if (mi >= 0) min_n = normals[mi];
        ne = 0;
        return 2;
    }
};

// Orders the memory accesses of a lane group (one wave): writes before the
// call are seen by the group's reads after it.
__device__ __forceinline__ void group_mem_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Lane exchanges of a lane group by DPP moves (register to register, no LDS
// round trip): the group's lanes all take part (control flow is uniform per
// group), and a group of 4, 8 or 16 lanes never spans two 16-lane rows.
enum : int { DPP_QUAD_BCAST0 = 0x00, DPP_QUAD_BCAST3 = 0xFF, DPP_ROW_SHR0 = 0x110 };
// Value of lane 0 of the caller's aligned group of G lanes: each quad takes
// its first lane's value, then the upper quad (and the upper half of a
// 16-lane group) takes the mirrored lane's, which holds lane 0's by then.
template <int G> __device__ __forceinline__ int group_bcast(int v) {
    static_assert(G == 4 || G == 8 || G == 16, "lane groups of 4, 8 or 16");
    int gl = (int)(__lane_id() & (G - 1));
    v = dpp_int<DPP_QUAD_BCAST0>(v);
    if constexpr (G >= 8) { int m = dpp_int<DPP_HALF_MIRROR>(v); v = (gl & 4) ? m : v; }
    if constexpr (G >= 16) { int m = dpp_int<DPP_MIRROR>(v); v = (gl & 8) ? m : v; }
    return v;
}
// Value of the last lane (G - 1) of the caller's group, the same way.
template <int G> __device__ __forceinline__ int group_bcast_last(int v) {
    static_assert(G == 4 || G == 8 || G == 16, "lane groups of 4, 8 or 16");
    int gl = (int)(__lane_id() & (G - 1));
    v = dpp_int<DPP_QUAD_BCAST3>(v);
    if constexpr (G >= 8) { int m = dpp_int<DPP_HALF_MIRROR>(v); v = (gl & 4) ? v : m; }
    if constexpr (G >= 16) { int m = dpp_int<DPP_MIRROR>(v); v = (gl & 8) ? v : m; }
    return v;
}
template <int G> __device__ __forceinline__ double group_bcast(double v) {
    long long b = __double_as_longlong(v);
    int lo = group_bcast<G>((int)b), hi = group_bcast<G>((int)(b >> 32));
    return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}
// Minimum over the group's lanes (every lane gets it).
template <int G> __device__ __forceinline__ int group_min(int v) {
    static_assert(G == 4 || G == 8 || G == 16, "lane groups of 4, 8 or 16");
    v = min(v, dpp_int<DPP_XOR1>(v));
    v = min(v, dpp_int<DPP_XOR2>(v));
    if constexpr (G >= 8) v = min(v, dpp_int<DPP_HALF_MIRROR>(v));
    if constexpr (G >= 16) v = min(v, dpp_int<DPP_MIRROR>(v));
    return v;
}
template <int G> __device__ __forceinline__ V3 group_bcast(V3 v) {
    return v3(group_bcast<G>(v.x), group_bcast<G>(v.y), group_bcast<G>(v.z));
}

// Working memory of the first lane of a lane group (G > 1): the EPA polytope,
// then (reusing it) the clipping polygons and the contact points. A slice of
// a device buffer, one per group, so that one lane's arrays are contiguous
// (private arrays are interleaved across the 64 lanes of a wave, and a lane
// that works alone would touch a cache line per element).
constexpr size_t EPA_BYTES = sizeof(V3) * EPA_MAXV + sizeof(V3) * EPA_MAXF + sizeof(double) * EPA_MAXF +
                             sizeof(int2) * EPA_MAXE + sizeof(I3) * EPA_MAXF;
constexpr size_t CLIP_BYTES = 2 * sizeof(V3) * CLIP_MAX;
constexpr size_t GROUP_WORK_BYTES = ((EPA_BYTES > CLIP_BYTES + MAXC * sizeof(V3) * 3 ? EPA_BYTES
                                      : CLIP_BYTES + MAXC * sizeof(V3) * 3) + 255) / 256 * 256;
RP_HD void epa_bind(EpaPoly& P, char* m) {
    P.poly = (V3*)m;
    m += sizeof(V3) * EPA_MAXV;
    P.normals = (V3*)m;
    m += sizeof(V3) * EPA_MAXF;
    P.dist = (double*)m;
    m += sizeof(double) * EPA_MAXF;
    P.edges = (int2*)m;
    m += sizeof(int2) * EPA_MAXE;
    P.faces = (I3*)m;
}

// Visibility flags of up to EPA_MAXF = 256 faces in four words.
struct FaceBits {
    unsigned long long w0 = 0, w1 = 0, w2 = 0, w3 = 0;
    __device__ bool get(int i) const {
        unsigned long long w = i < 64 ? w0 : i < 128 ? w1 : i < 192 ? w2 : w3;
        return (w >> (i & 63)) & 1ull;
    }
    __device__ void put(int i, bool v) {
        unsigned long long b = 1ull << (i & 63);
        int k = i >> 6;
        if (k == 0) w0 = v ? (w0 | b) : (w0 & ~b);
        else if (k == 1) w1 = v ? (w1 | b) : (w1 & ~b);
        else if (k == 2) w2 = v ? (w2 | b) : (w2 & ~b);
        else w3 = v ? (w3 | b) : (w3 & ~b);
    }
    __device__ void or_bits(int c0, unsigned long long m) {  // c0 a multiple of the group size
        unsigned long long b = m << (c0 & 63);
        int k = c0 >> 6;
        w0 |= k == 0 ? b : 0ull;
        w1 |= k == 1 ? b : 0ull;
        w2 |= k == 2 ? b : 0ull;
        w3 |= k == 3 ? b : 0ull;
    }
};
// epa for a lane group (G > 1), with the results of EpaPoly::step: the
// polytope lives in the group's working memory; the lanes test which faces
// see the support point, compute the new faces' normals and distances, and
// search the closest face (first minimum), while lane 0 keeps the edge list
// and the face list's swap-remove order (epa.cpp:165-195). Every lane
// returns the status and the result; the EPA error bits are lane 0's.
template <int G, class View>
__device__ int epa_group(const View& h1, const View& h2, const Simplex& s, V3& n_out, double& pen_out, uint32_t& err,
                         char* work) {
    EpaPoly P;
    epa_bind(P, work);
    int gl = (int)(__lane_id() & (G - 1));
    bool leader = gl == 0;
    int gbase = (int)__lane_id() & ~(G - 1);
    int st = 2;
    if (leader && !P.init(s, err)) st = -1;
    st = group_bcast<G>(st);
    if (st != 2) return st;
    V3 min_n = group_bcast<G>(P.min_n);
    double min_d = group_bcast<G>(P.min_d);
    int np = 4, nf = 4;
    group_mem_sync();  // lane 0's initial faces
    for (int it = 0; it < 100; ++it) {
        V3 sp = support_md<G>(h1, h2, min_n, err);
        double d = dot(min_n, sp);
        if (fabs(d - min_d) < 0.0001) {
            n_out = min_n;
            pen_out = min_d;
            return 1;
        }
        int new_idx = np;
        if (leader) P.poly[np] = sp;
        ++np;
        // which faces see sp (a property of each face: computed up front)
        FaceBits vis;
        for (int c0 = 0; c0 < nf; c0 += G) {
            int i = c0 + gl;
            bool v = i < nf && P.sees(i, sp);
            unsigned long long m = (__ballot(v) >> gbase) & ((1ull << G) - 1ull);
            vis.or_bits(c0, m);
        }
        // lane 0: edges of the visible faces and the swap-removes, in order
        int ne = 0;
        if (leader) {
            P.nf = nf;
            P.ne = 0;
            for (int i = 0; i < P.nf; ++i) {
                if (vis.get(i)) {
                    I3 f = P.faces[i];
                    if (!P.add_edge(make_int2(f.x, f.y), err) || !P.add_edge(make_int2(f.y, f.z), err) ||
                        !P.add_edge(make_int2(f.z, f.x), err)) {
                        st = -1;
                        break;
                    }
                    --P.nf;
                    P.faces[i] = P.faces[P.nf];
                    P.dist[i] = P.dist[P.nf];
                    P.normals[i] = P.normals[P.nf];
                    vis.put(i, vis.get(P.nf));
                    --i;
                }
            }
            nf = P.nf;
            ne = P.ne;
        }
        st = group_bcast<G>(st);
        if (st != 2) return st;
        nf = group_bcast<G>(nf);
        ne = group_bcast<G>(ne);
        group_mem_sync();
        // the new faces (edge i, new point) at nf + i; the sequential loop
        // stops at the first full list or degenerate face, in that order
        int room = EPA_MAXF - nf;
        int nnew = ne < room ? ne : room;
        int first_bad = nnew;
        for (int i = gl; i < nnew; i += G) {
            int2 e = P.edges[i];
            I3 f = {e.x, e.y, new_idx};
            V3 n;
            double fd;
            bool ok = epa_face(f, P.poly, np, n, fd);
            P.faces[nf + i] = f;
            P.normals[nf + i] = n;
            P.dist[nf + i] = fd;
            if (!ok && i < first_bad) first_bad = i;
        }
        first_bad = group_min<G>(first_bad);
        if (first_bad < nnew) {
            if (leader) err |= ERR_EPA_DEGENERATE;
            return -1;
        }
        if (ne > room) {
            if (leader) err |= ERR_EPA_CAPACITY;
            return -1;
        }
        nf += ne;
        group_mem_sync();
        // the closest face: first minimum of the distances
        double md = DBL_MAX;
        int mi = NO_SEL;
        for (int i = gl; i < nf; i += G) {
            double di = P.dist[i];
            bool t = di < md;
            md = t ? di : md;
            mi = t ? i : mi;
        }
        group_argmin<G>(md, mi);
        min_d = md;
        if (mi != NO_SEL) min_n = P.normals[mi];
    }
    return 0;
}

// epa (epa.cpp:118-237): 0 = no convergence, 1 = converged, -1 = error.
// G > 1: the group's lanes compute each support point together; the
// polytope lives in lane 0 alone (in the group's working memory `work`),
// which expands it and hands the next direction and the status to the group. Every lane returns
// the status and, on convergence, lane 0's n_out / pen_out; the EPA error
// bits are set in lane 0's err only.
template <int G = 1, class View = HullView>
RP_HD int epa(const View& h1, const View& h2, const Simplex& s, V3& n_out, double& pen_out, uint32_t& err,
              char* work = nullptr) {
    if constexpr (G == 1) {
        V3 poly[EPA_MAXV];
        I3 faces[EPA_MAXF];
        V3 normals[EPA_MAXF];
        double dist[EPA_MAXF];
        int2 edges[EPA_MAXE];
        EpaPoly P;
        P.poly = poly; P.faces = faces; P.normals = normals; P.dist = dist; P.edges = edges;
        if (!P.init(s, err)) return -1;
        for (int it = 0; it < 100; ++it) {
            V3 sp = support_md<G>(h1, h2, P.min_n, err);
            int st = P.step(sp, n_out, pen_out, err);
            if (st != 2) return st;
        }
        return 0;
    } else {
        return epa_group<G>(h1, h2, s, n_out, pen_out, err, work);
    }
}

// ---------------------------------------------------------------------------
// Contact manifold (clipping.cpp:7-371)// ---------------------------------------------------------------------------
// Contact manifold (clipping.cpp:7-371)
// ---------------------------------------------------------------------------
struct Plane { V3 normal, point; };

// is_point_in_plane (clipping.cpp:12-19): the plane offset is truncated to float
RP_HD bool point_in_plane(const Plane& p, V3 pos) {
    float distance = -dot(p.normal, p.point);
    if (dot(pos, p.normal) + distance < 0.0) return false;
    return true;
}
// plane_edge_intersection (clipping.cpp:21-47), with its float truncations
RP_HD bool plane_edge(const Plane& p, V3 start, V3 end, V3& out) {
    const double EPSILON = 0.000001;
    V3 ab = sub(end, start);
    float ab_p = dot(p.normal, ab);
    if (fabs((double)ab_p) > EPSILON) {
        float distance = -dot(p.normal, p.point);
        V3 p_co = scl(-distance, p.normal);
        float fac = -dot(p.normal, sub(start, p_co)) / ab_p;
        fac = RP_MIN(RP_MAX(fac, 0.0), 1.0);
        out = add(start, scl(fac, ab));
        return true;
    }
    return false;
}

// One Sutherland-Hodgman pass against one plane (clipping.cpp:62-108).
RP_HD int clip_pass(const V3* in, int n, const Plane& p, V3* out, bool remove, uint32_t& err) {
    int m = 0;
    if (n == 0) return 0;
    V3 tmp, start = in[n - 1];
    for (int j = 0; j < n; ++j) {
        V3 end = in[j];
        bool s_in = point_in_plane(p, start);
        bool e_in = point_in_plane(p, end);
        if (remove) {
            if (e_in) { if (m < CLIP_MAX) out[m++] = end; else err |= ERR_CLIP_CAPACITY; }
        } else if (s_in && e_in) {
            if (m < CLIP_MAX) out[m++] = end; else err |= ERR_CLIP_CAPACITY;
        } else if (s_in && !e_in) {
            if (plane_edge(p, start, end, tmp)) { if (m < CLIP_MAX) out[m++] = tmp; else err |= ERR_CLIP_CAPACITY; }
        } else if (!s_in && e_in) {
            if (plane_edge(p, start, end, tmp)) { if (m < CLIP_MAX) out[m++] = tmp; else err |= ERR_CLIP_CAPACITY; }
            if (m < CLIP_MAX) out[m++] = end; else err |= ERR_CLIP_CAPACITY;
        }
        start = end;
    }
    return m;
}

// Lane-group clipping (G > 1): the polygons live in LDS, and each pass gives
// every lane one edge (start = in[j - 1], end = in[j]); a lane's 0-2 output
// points go to their place in the sequential output order by a prefix sum
// over the group, so the polygon equals clip_pass's point for point.
template <int G> __device__ __forceinline__ int group_inclusive_sum(int v) {
    int gl = (int)(__lane_id() & (G - 1));
    // row_shr:d reads lane - d of the same 16-lane row; every lane moves its
    // value before the group's first d lanes drop what they received
    int o = dpp_int<DPP_ROW_SHR0 + 1>(v);
    v += gl >= 1 ? o : 0;
    o = dpp_int<DPP_ROW_SHR0 + 2>(v);
    v += gl >= 2 ? o : 0;
    if constexpr (G >= 8) {
        o = dpp_int<DPP_ROW_SHR0 + 4>(v);
        v += gl >= 4 ? o : 0;
    }
    if constexpr (G >= 16) {
        o = dpp_int<DPP_ROW_SHR0 + 8>(v);
        v += gl >= 8 ? o : 0;
    }
    return v;
}
// One pass (clip_pass's rules) over `in` (n points) into `out` (cap points).
// Returns the output count; more than cap sets `ovf` (no point is lost:
// the caller redoes the clipping in CLIP_MAX buffers).
template <int G>
__device__ int clip_pass_group(const V3* in, int n, const Plane& p, V3* out, bool remove, int cap, bool& ovf) {
    int gl = (int)(__lane_id() & (G - 1));
    int m = 0;
    for (int c0 = 0; c0 < n; c0 += G) {
        int j = c0 + gl;
        int cnt = 0;
        V3 o0 = v3(0.0, 0.0, 0.0), o1 = o0;
        if (j < n) {
            V3 start = in[j == 0 ? n - 1 : j - 1], end = in[j];
            bool s_in = point_in_plane(p, start);
            bool e_in = point_in_plane(p, end);
            V3 t;
            if (remove) {
                if (e_in) { o0 = end; cnt = 1; }
            } else if (s_in && e_in) {
                o0 = end; cnt = 1;
            } else if (s_in && !e_in) {
                if (plane_edge(p, start, end, t)) { o0 = t; cnt = 1; }
            } else if (!s_in && e_in) {
                if (plane_edge(p, start, end, t)) { o0 = t; o1 = end; cnt = 2; }
                else { o0 = end; cnt = 1; }
            }
        }
        int incl = group_inclusive_sum<G>(cnt);
        int total = group_bcast_last<G>(incl);
        int pos = m + incl - cnt;
        if (cnt >= 1 && pos < cap) out[pos] = o0;
        if (cnt == 2 && pos + 1 < cap) out[pos + 1] = o1;
        m += total;
    }
    if (m > cap) ovf = true;
    group_mem_sync();
    return m;
}
// clip_incident on the group's lanes with LDS buffers of `cap` points;
// -1 when they overflow. Every lane of the group calls it.
template <int G>
__device__ int clip_incident_group(const HullView& rh, int rf, const HullView& ih, int inf, const Plane& rp, V3* bufA,
                                   V3* bufB, int cap, const V3*& res) {
    int gl = (int)(__lane_id() & (G - 1));
    int f0 = ih.fe_off[inf], n = ih.fe_off[inf + 1] - f0;
    if (n > cap) return -1;
    for (int i = gl; i < n; i += G) bufA[i] = ih.tv[ih.fe[f0 + i]];
    group_mem_sync();
    bool ovf = false;
    V3* in = bufA;
    V3* outp = bufB;
    for (int k = rh.ff_off[rf]; k < rh.ff_off[rf + 1]; ++k) {
        if (n == 0) break;
        int nf = rh.ff[k];
        Plane p;
        p.point = rh.tv[rh.fe[rh.fe_off[nf]]];
        p.normal = inv3(face_normal_world(rh, nf));
        n = clip_pass_group<G>(in, n, p, outp, false, cap, ovf);
        if (ovf) return -1;
        V3* t = in; in = outp; outp = t;
    }
    n = clip_pass_group<G>(in, n, rp, outp, true, cap, ovf);
    if (ovf) return -1;
    res = outp;
    return n;
}

// get_face_with_most_fitting_normal (clipping.cpp:136-152)
RP_HD int best_face(int sidx, const HullView& h, V3 n, uint32_t& err) {
    double max_proj = -DBL_MAX;
    int sel = -1;
    for (int i = h.vf_off[sidx]; i < h.vf_off[sidx + 1]; ++i) {
        double proj = dot(face_normal_world(h, h.vf[i]), n);
        if (proj > max_proj) { max_proj = proj; sel = h.vf[i]; }
    }
    if (sel < 0) { err |= ERR_SELECT; sel = 0; }
    return sel;
}

// collision_distance_between_skew_lines (clipping.cpp:210-239)
RP_HD bool skew_lines(V3 p1, V3 d1, V3 p2, V3 d2, V3& l1, V3& l2) {
    double n1 = d1.x * d2.x + d1.y * d2.y + d1.z * d2.z;
    double n2 = d2.x * d2.x + d2.y * d2.y + d2.z * d2.z;
    double m1 = -d1.x * d1.x - d1.y * d1.y - d1.z * d1.z;
    double m2 = -d2.x * d1.x - d2.y * d1.y - d2.z * d1.z;
    double r1 = -d1.x * p2.x + d1.x * p1.x - d1.y * p2.y + d1.y * p1.y - d1.z * p2.z + d1.z * p1.z;
    double r2 = -d2.x * p2.x + d2.x * p1.x - d2.y * p2.y + d2.y * p1.y - d2.z * p2.z + d2.z * p1.z;
    if ((n1 * m2) - (n2 * m1) == 0) return false;
    double n = ((r1 * m2) - (r2 * m1)) / ((n1 * m2) - (n2 * m1));
    double m = ((n1 * r2) - (n2 * r1)) / ((n1 * m2) - (n2 * m1));
    l1 = add(p1, scl(m, d1));
    l2 = add(p2, scl(n, d2));
    return true;
}

struct ContactPt { V3 p1, p2, n; };

// convex_convex_contact_manifold (clipping.cpp:249-341). Returns the number
// of contacts written to `out` (at most MAXC). G > 1: the group's lanes find
// the support points, the best faces and the best edge pair together (split
// scans, first-maximum reductions); clipping and the contact points run on
// lane 0 of the group alone, on the group's working memory `work` (its first
// CLIP_BYTES), whose `out`, return value and error bits are the results (the
// other lanes return 0).
// get_face_with_most_fitting_normal over the group: lane k scores the
// entries k, k + G, ... of the vertex's face list, then the group keeps the
// first maximum (larger projection, ties to the earlier entry).
template <int G>
__device__ __forceinline__ int best_face_group(int sidx, const HullView& h, V3 n, uint32_t& err) {
    int gl = (int)(__lane_id() & (G - 1));
    int b = h.vf_off[sidx], cnt = h.vf_off[sidx + 1] - b;
    double max_proj = -DBL_MAX;
    int pos = NO_SEL;
    for (int k = gl; k < cnt; k += G) {
        double proj = dot(face_normal_world(h, h.vf[b + k]), n);
        bool t = proj > max_proj;
        max_proj = t ? proj : max_proj;
        pos = t ? k : pos;
    }
    group_argmax<G>(max_proj, pos);
    if (pos == NO_SEL) { err |= ERR_SELECT; return 0; }
    return h.vf[b + pos];
}

// The incident face clipped against the reference face's boundary planes
// (clipping.cpp:121-134, 294) and the points above the reference plane
// removed (clipping.cpp:296-302), by one lane, in two caller-owned buffers
// of CLIP_MAX points; `res` is set to the one holding the result.
RP_HD int clip_serial(const HullView& rh, int rf, const HullView& ih, int inf, const Plane& rp, V3* bufA, V3* bufB,
                      const V3*& res, uint32_t& err) {
    int n = 0;
    for (int i = ih.fe_off[inf]; i < ih.fe_off[inf + 1]; ++i) {
        if (n < CLIP_MAX) bufA[n++] = ih.tv[ih.fe[i]]; else err |= ERR_CLIP_CAPACITY;
    }
    // clip against the reference face's boundary planes (clipping.cpp:121-134, 294)
    V3* in = bufA;
    V3* outp = bufB;
    for (int k = rh.ff_off[rf]; k < rh.ff_off[rf + 1]; ++k) {
        if (n == 0) break;
        int nf = rh.ff[k];
        Plane p;
        p.point = rh.tv[rh.fe[rh.fe_off[nf]]];
        p.normal = inv3(face_normal_world(rh, nf));
        n = clip_pass(in, n, p, outp, false, err);
        V3* t = in; in = outp; outp = t;
    }
    n = clip_pass(in, n, rp, outp, true, err);
    res = outp;
    return n;
}

// get_closest_point_polygon (clipping.cpp:115-119) and the contact points
// (clipping.cpp:304-329) of the clipped polygon.
RP_HD int contacts_from_clip(const V3* pts, int n, const Plane& rp, V3 normal, bool ref1, ContactPt* out,
                             uint32_t& err) {
    int nc = 0;
    double pd = dot(scl(-1.0, rp.normal), rp.point);
    for (int i = 0; i < n; ++i) {
        V3 pt = pts[i];
        V3 cp = sub(pt, scl(dot(rp.normal, pt) + pd, rp.normal));
        V3 diff = sub(pt, cp);
        double pen;
        ContactPt c;
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
        if (pen < 0.0) {
            if (nc < MAXC) out[nc++] = c; else err |= ERR_CONTACT_CAPACITY;
        }
    }
    return nc;
}

// contacts_from_clip on the group's lanes: lane j takes point j; the points
// kept (negative penetration) go to `out` in point order by a prefix sum,
// the first MAXC of them (more set ERR_CONTACT_CAPACITY, as in order).
// Returns the count on every lane of the group.
template <int G>
__device__ int contacts_from_clip_group(const V3* pts, int n, const Plane& rp, V3 normal, bool ref1, ContactPt* out,
                                        uint32_t& err) {
    int gl = (int)(__lane_id() & (G - 1));
    int nc = 0;
    double pd = dot(scl(-1.0, rp.normal), rp.point);
    for (int c0 = 0; c0 < n; c0 += G) {
        int i = c0 + gl;
        bool keep = false;
        ContactPt c;
        if (i < n) {
            V3 pt = pts[i];
            V3 cp = sub(pt, scl(dot(rp.normal, pt) + pd, rp.normal));
            V3 diff = sub(pt, cp);
            double pen;
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
            keep = pen < 0.0;
        }
        int k = keep ? 1 : 0;
        int incl = group_inclusive_sum<G>(k);
        int total = group_bcast_last<G>(incl);
        int pos = nc + incl - k;
        if (keep && pos < MAXC) out[pos] = c;
        nc += total;
    }
    if (nc > MAXC) {
        if (gl == 0) err |= ERR_CONTACT_CAPACITY;
        nc = MAXC;
    }
    return nc;
}

// G > 1: the group clips in `spare` (spare_n points: the unused end of the
// pair's LDS slice) when both polygons fit there, else lane 0 clips in
// `work`. The contact points go to `out`; every lane of the group returns
// their count.
template <int G = 1>
RP_HD int manifold(const HullView& h1, const HullView& h2, V3 normal, ContactPt* out, uint32_t& err,
                   char* work = nullptr, V3* spare = nullptr, int spare_n = 0) {
    const double EPSILON = 0.0001;
    V3 inormal = inv3(normal);
    int s1, s2;
    support_index2<G>(h1, normal, h2, inormal, s1, s2, err);
    int f1, f2;
    if constexpr (G > 1) {
        f1 = best_face_group<G>(s1, h1, normal, err);
        f2 = best_face_group<G>(s2, h2, inormal, err);
    } else {
        f1 = best_face(s1, h1, normal, err);
        f2 = best_face(s2, h2, inormal, err);
    }
    // get_edge_with_most_fitting_normal (clipping.cpp:154-200)
    V3 edge_n = v3(0.0, 0.0, 0.0);
    int e1a = 0, e1b = 0, e2a = 0, e2b = 0;
    if constexpr (G > 1) {
        // candidates in the sequential order (neighbour i of s1, neighbour j
        // of s2, then n before -n) are numbered 2 (i n2 + j) + variant; lane k
        // scores the pairs k, k + G, ... and the group keeps the first maximum
        int gl = (int)(__lane_id() & (G - 1));
        V3 sp1 = h1.tv[s1], sp2 = h2.tv[s2];
        int i0 = h1.vn_off[s1], n1 = h1.vn_off[s1 + 1] - i0;
        int j0 = h2.vn_off[s2], n2 = h2.vn_off[s2 + 1] - j0;
        double max_dot = -DBL_MAX;
        int best = NO_SEL;
        for (int k = gl; k < n1 * n2; k += G) {
            int a = k / n2, b = k - a * n2;
            V3 ed1 = sub(sp1, h1.tv[h1.vn[i0 + a]]);
            V3 ed2 = sub(sp2, h2.tv[h2.vn[j0 + b]]);
            V3 cn = nrm(cross(ed1, ed2));
            V3 cni = inv3(cn);
            double dd = dot(cn, normal);
            bool t = dd > max_dot;
            max_dot = t ? dd : max_dot;
            best = t ? 2 * k : best;
            dd = dot(cni, normal);
            t = dd > max_dot;
            max_dot = t ? dd : max_dot;
            best = t ? 2 * k + 1 : best;
        }
        group_argmax<G>(max_dot, best);
        if (best == NO_SEL) {
            err |= ERR_SELECT;
        } else {
            int k = best >> 1, a = k / n2, b = k - a * n2;
            e1a = s1; e1b = h1.vn[i0 + a]; e2a = s2; e2b = h2.vn[j0 + b];
            V3 cn = nrm(cross(sub(sp1, h1.tv[e1b]), sub(sp2, h2.tv[e2b])));
            edge_n = (best & 1) ? inv3(cn) : cn;
        }
    } else {
        V3 sp1 = h1.tv[s1], sp2 = h2.tv[s2];
        double max_dot = -DBL_MAX;
        bool found = false;
        for (int i = h1.vn_off[s1]; i < h1.vn_off[s1 + 1]; ++i) {
            V3 ed1 = sub(sp1, h1.tv[h1.vn[i]]);
            for (int j = h2.vn_off[s2]; j < h2.vn_off[s2 + 1]; ++j) {
                V3 ed2 = sub(sp2, h2.tv[h2.vn[j]]);
                V3 cn = nrm(cross(ed1, ed2));
                V3 cni = inv3(cn);
                double dd = dot(cn, normal);
                if (dd > max_dot) { max_dot = dd; e1a = s1; e1b = h1.vn[i]; e2a = s2; e2b = h2.vn[j]; edge_n = cn; found = true; }
                dd = dot(cni, normal);
                if (dd > max_dot) { max_dot = dd; e1a = s1; e1b = h1.vn[i]; e2a = s2; e2b = h2.vn[j]; edge_n = cni; found = true; }
            }
        }
        if (!found) err |= ERR_SELECT;
    }
    double d1 = dot(face_normal_world(h1, f1), normal);
    double d2 = dot(face_normal_world(h2, f2), inormal);
    double de = dot(edge_n, normal);
    if (de > d1 + EPSILON && de > d2 + EPSILON) {
        V3 l1, l2;
        V3 p1 = h1.tv[e1a];
        V3 dd1 = sub(h1.tv[e1b], p1);
        V3 p2 = h2.tv[e2a];
        V3 dd2 = sub(h2.tv[e2b], p2);
        if (!skew_lines(p1, dd1, p2, dd2, l1, l2)) { err |= ERR_SKEW_LINES; return 0; }
        if constexpr (G > 1) {
            if ((__lane_id() & (G - 1)) != 0) return 1;  // lane 0 writes the contact
        }
        out[0].p1 = l1; out[0].p2 = l2; out[0].n = normal;
        return 1;
    }
    bool ref1 = d1 > d2;
    const HullView& rh = ref1 ? h1 : h2;
    const HullView& ih = ref1 ? h2 : h1;
    int rf = ref1 ? f1 : f2, inf = ref1 ? f2 : f1;
    // remove the points above the reference plane (clipping.cpp:296-302)
    Plane rp;
    rp.normal = inv3(face_normal_world(rh, rf));
    rp.point = rh.tv[rh.fe[rh.fe_off[rf]]];
    const V3* outp = nullptr;
    if constexpr (G > 1) {
        int cap = spare_n / 2 < CLIP_MAX ? spare_n / 2 : CLIP_MAX;
        int n = -1;
        if (cap >= 8) n = clip_incident_group<G>(rh, rf, ih, inf, rp, spare, spare + cap, cap, outp);
        if (n >= 0) return contacts_from_clip_group<G>(outp, n, rp, normal, ref1, out, err);
        // lane 0 clips in the group's working memory
        int nc = 0;
        if ((__lane_id() & (G - 1)) == 0) {
            int m = clip_serial(rh, rf, ih, inf, rp, (V3*)work, (V3*)work + CLIP_MAX, outp, err);
            nc = contacts_from_clip(outp, m, rp, normal, ref1, out, err);
        }
        return group_bcast<G>(nc);
    } else {
        V3 localA[CLIP_MAX], localB[CLIP_MAX];  // here, in the frame that reads the result
        int n = clip_serial(rh, rf, ih, inf, rp, localA, localB, outp, err);
        return contacts_from_clip(outp, n, rp, normal, ref1, out, err);
    }
}

// collider_get_contacts for two hulls (collider.cpp:523-558): GJK, EPA, manifold.
// `epa_fail` counts EPA runs that did not converge (the reference prints and
// drops the pair's contacts, epa.cpp:233-236, collider.cpp:549-551).
// `flags` gets PAIR_EPA_FAIL and PAIR_GJK_CYCLE (a cycle exit of gjk).
enum { PAIR_EPA_FAIL = 1, PAIR_GJK_CYCLE = 2 };
// G > 1: lane-group working memory `work`; the results are lane 0's.
template <int G = 1>
RP_HD int pair_contacts(const HullView& h1, const HullView& h2, ContactPt* out, uint32_t& err, int& flags,
                        char* work = nullptr) {
    Simplex s;
    bool cycled = false;
    if (!gjk<G>(h1, h2, s, err, cycled)) {
        if (cycled) flags |= PAIR_GJK_CYCLE;
        return 0;
    }
    V3 n; double pen;
    int r = epa<G>(h1, h2, s, n, pen, err, work);
    if (r != 1) { flags |= PAIR_EPA_FAIL; return 0; }
    return manifold<G>(h1, h2, n, out, err, work);
}

// clipping_get_contact_manifold (clipping.cpp:343-371) for colliders: a
// sphere gives one contact at its support point along the normal.
RP_HD int collider_manifold(const ColView& c1, const ColView& c2, V3 normal, double penetration, ContactPt* out,
                            uint32_t& err) {
    if (c1.type == COL_SPHERE) {
        V3 p = support_point(c1, normal, err);
        out[0].p1 = p;
        out[0].p2 = sub(p, scl(penetration, normal));
        out[0].n = normal;
        return 1;
    }
    if (c2.type == COL_SPHERE) {
        V3 inormal = inv3(normal);
        V3 p = support_point(c2, inormal, err);
        out[0].p1 = add(p, scl(penetration, normal));
        out[0].p2 = p;
        out[0].n = normal;
        return 1;
    }
    return manifold<1>(c1.h, c2.h, normal, out, err);
}

// collider_get_contacts (collider.cpp:523-558) for any two colliders. Two
// spheres are resolved with the reference's float arithmetic; the float
// square root is taken in double and rounded once, which is the correctly
// rounded result (a 53-bit square root of a 24-bit value). `cap` is the room
// left in `out`.
RP_HD int collider_contacts(const ColView& c1, const ColView& c2, ContactPt* out, int cap, uint32_t& err, int& flags) {
    ContactPt tmp[MAXC];
    int n = 0;
    if (c1.type == COL_SPHERE && c2.type == COL_SPHERE) {
        V3 dv = sub(c1.center, c2.center);
        float dsq = (float)dot(dv, dv);
        float mind = c1.radius + c2.radius;
        if (dsq < (mind * mind)) {
            V3 nrm_ = nrm(sub(c2.center, c1.center));
            float root = (float)sqrt((double)dsq);
            double pen = (double)(mind - root);
            n = collider_manifold(c1, c2, nrm_, pen, tmp, err);
        }
    } else {
        Simplex s;
        bool cycled = false;
        if (gjk<1>(c1, c2, s, err, cycled)) {
            V3 nn;
            double pen;
            int r = epa<1>(c1, c2, s, nn, pen, err);
            if (r != 1) flags |= PAIR_EPA_FAIL;
            else n = collider_manifold(c1, c2, nn, pen, tmp, err);
        } else if (cycled) {
            flags |= PAIR_GJK_CYCLE;
        }
    }
    if (n > cap) {
        err |= ERR_CONTACT_CAPACITY;
        n = cap;
    }
    for (int k = 0; k < n; ++k) out[k] = tmp[k];
    return n;
}

// ---------------------------------------------------------------------------
// XPBD base constraints (pbd_base_constraints.cpp:6-227)
// ---------------------------------------------------------------------------
template <class S> struct PosPrepT { V3T<S> r1, r2; M3T<S> I1, I2; };
using PosPrep = PosPrepT<double>;

template <class S>
RP_HD void pos_prep(const BodyT<S>& e1, const BodyT<S>& e2, V3T<S> r1_lc, V3T<S> r2_lc, PosPrepT<S>& p) {
    p.r1 = q_rot(e1.q, r1_lc);
    p.r2 = q_rot(e2.q, r2_lc);
    p.I1 = world_tensor(e1.q, e1.Iinv);
    p.I2 = world_tensor(e2.q, e2.Iinv);
}
template <class S>
RP_HD S pos_dlambda(const BodyT<S>& e1, const BodyT<S>& e2, const PosPrepT<S>& p, double h, double compliance, S lambda,
                    V3T<S> dx, uint32_t& err) {
    S c = len(dx);
    if (c <= 1e-50) return S(0.0);
    V3T<S> n = v3<S>(dx.x / c, dx.y / c, dx.z / c);
    S w1 = e1.inv_mass + dot(cross(p.r1, n), m3_mulv(p.I1, cross(p.r1, n)));
    S w2 = e2.inv_mass + dot(cross(p.r2, n), m3_mulv(p.I2, cross(p.r2, n)));
    if (w1 + w2 == 0.0) err |= ERR_ZERO_MASS;
    double tc = compliance / (h * h);
    return (-c - tc * lambda) / (w1 + w2 + tc);
}
template <class S>
RP_HD void rot_update(BodyT<S>& e, V3T<S> a, bool plus) {
    Q4T<S> aq = {a.x, a.y, a.z, S(0.0)};
    Q4T<S> qq = q_mul(aq, e.q);
    if (plus) {
        e.q.x = e.q.x + 0.5 * qq.x; e.q.y = e.q.y + 0.5 * qq.y; e.q.z = e.q.z + 0.5 * qq.z; e.q.w = e.q.w + 0.5 * qq.w;
    } else {
        e.q.x = e.q.x - 0.5 * qq.x; e.q.y = e.q.y - 0.5 * qq.y; e.q.z = e.q.z - 0.5 * qq.z; e.q.w = e.q.w - 0.5 * qq.w;
    }
    e.q = q_nrm(e.q);
}
template <class S>
RP_HD void pos_apply(BodyT<S>& e1, BodyT<S>& e2, const PosPrepT<S>& p, S dl, V3T<S> dx) {
    S c = len(dx);
    if (c <= 1e-50) return;
    V3T<S> n = v3<S>(dx.x / c, dx.y / c, dx.z / c);
    V3T<S> imp = scl(dl, n);
    if (!e1.fixed) e1.x = add(e1.x, scl(e1.inv_mass, imp));
    if (!e2.fixed) e2.x = add(e2.x, scl(-e2.inv_mass, imp));
    V3T<S> a1 = m3_mulv(p.I1, cross(p.r1, imp));
    V3T<S> a2 = m3_mulv(p.I2, cross(p.r2, imp));
    // quaternion products use the rotation before either body is updated
    // (pbd_base_constraints.cpp:114-117); e1 and e2 are distinct bodies.
    if (!e1.fixed) rot_update(e1, a1, true);
    if (!e2.fixed) rot_update(e2, a2, false);
}
RP_HD double ang_dlambda(const M3& I1, const M3& I2, double h, double compliance, double lambda, V3 dq, uint32_t& err) {
    double theta = len(dq);
    if (theta <= 1e-50) return 0.0;
    V3 n = v3(dq.x / theta, dq.y / theta, dq.z / theta);
    double w1 = dot(n, m3_mulv(I1, n));
    double w2 = dot(n, m3_mulv(I2, n));
    if (w1 + w2 == 0.0) err |= ERR_ZERO_MASS;
    double tc = compliance / (h * h);
    return (-theta - tc * lambda) / (w1 + w2 + tc);
}
RP_HD void ang_apply(Body& e1, Body& e2, const M3& I1, const M3& I2, double dl, V3 dq) {
    double theta = len(dq);
    if (theta <= 1e-50) return;
    V3 n = v3(dq.x / theta, dq.y / theta, dq.z / theta);
    V3 imp = scl(-dl, n);
    V3 a1 = m3_mulv(I1, imp);
    V3 a2 = m3_mulv(I2, imp);
    if (!e1.fixed) rot_update(e1, a1, true);
    if (!e2.fixed) rot_update(e2, a2, false);
}

// collision_constraint_solve (pbd.cpp:107-154)
template <class S>
RP_HD void solve_collision(ContactT<S>& c, BodyT<S>& e1, BodyT<S>& e2, double h, uint32_t& err) {
    PosPrepT<S> p;
    pos_prep(e1, e2, c.r1, c.r2, p);
    V3T<S> p1 = add(e1.x, p.r1);
    V3T<S> p2 = add(e2.x, p.r2);
    S d = dot(sub(p1, p2), c.n);
    if (d > 0.0) {
        V3T<S> dx = scl(d, c.n);
        S dl = pos_dlambda(e1, e2, p, h, 0.0, c.lambda_n, dx, err);
        pos_apply(e1, e2, p, dl, dx);
        c.lambda_n += dl;
        pos_prep(e1, e2, c.r1, c.r2, p);
        p1 = add(e1.x, p.r1);
        p2 = add(e2.x, p.r2);
        dl = pos_dlambda(e1, e2, p, h, 0.0, c.lambda_t, dx, err);
        const S mu_s = (e1.mu_s + e2.mu_s) / 2.0f;
        S lambda_n = c.lambda_n;
        S lambda_t = c.lambda_t + dl;
        if (lambda_t > mu_s * lambda_n) {
            V3T<S> p1_til = add(e1.px, q_rot(e1.pq, c.r1));
            V3T<S> p2_til = add(e2.px, q_rot(e2.pq, c.r2));
            V3T<S> dp = sub(sub(p1, p1_til), sub(p2, p2_til));
            V3T<S> dp_t = sub(dp, scl(dot(dp, c.n), c.n));
            pos_apply(e1, e2, p, dl, dp_t);
            c.lambda_t += dl;
        }
    }
}

// limit_angle (pbd.cpp:175-217)
RP_HD bool limit_angle(V3 n, V3 n1, V3 n2, double alpha, double beta, V3& dq) {
    double phi = asin(dot(cross(n1, n2), n));
    if (dot(n1, n2) < 0.0) phi = RP_PI_F - phi;
    if (phi > RP_PI_F) phi = phi - 2.0 * RP_PI_F;
    if (phi < -RP_PI_F) phi = phi + 2.0 * RP_PI_F;
    if (phi < alpha || phi > beta) {
        phi = RP_CLAMP(phi, alpha, beta);
        Q4 rot = q_from_axis_radians(n, phi);
        n1 = q_rot(rot, n1);
        dq = cross(n1, n2);
        return true;
    }
    return false;
}

// Joint and positional constraints (pbd.cpp:81-97, 156-173, 245-379).
// lam[0..2] are the constraint's multipliers for this substep:
//   positional: lambda; mutual orientation: lambda;
//   hinge: aligned axes, position, limit; spherical: position, swing, twist.
RP_HD void solve_external(const ExtC& c, double* lam, Body& e1, Body& e2, double h, uint32_t& err) {
    bool bad = false;
    switch (c.type) {
    case 0: {  // POSITIONAL_CONSTRAINT
        V3 dx = sub(sub(e1.x, e2.x), c.distance);
        PosPrep p;
        pos_prep(e1, e2, c.r1, c.r2, p);
        double dl = pos_dlambda(e1, e2, p, h, c.compliance, lam[0], dx, err);
        pos_apply(e1, e2, p, dl, dx);
        lam[0] += dl;
    } break;
    case 2: {  // MUTUAL_ORIENTATION_CONSTRAINT
        M3 I1 = world_tensor(e1.q, e1.Iinv), I2 = world_tensor(e2.q, e2.Iinv);
        Q4 aux = q_mul(e1.q, q_inv(e2.q));
        V3 dq = v3(2.0 * aux.x, 2.0 * aux.y, 2.0 * aux.z);
        double dl = ang_dlambda(I1, I2, h, c.compliance, lam[0], dq, err);
        ang_apply(e1, e2, I1, I2, dl, dq);
        lam[0] += dl;
    } break;
    case 3: {  // HINGE_JOINT_CONSTRAINT
        M3 I1 = world_tensor(e1.q, e1.Iinv), I2 = world_tensor(e2.q, e2.Iinv);
        V3 a1 = q_axis(e1.q, c.axis[0], bad);
        V3 a2 = q_axis(e2.q, c.axis[1], bad);
        V3 dq = cross(a1, a2);
        double dl = ang_dlambda(I1, I2, h, c.compliance, lam[0], dq, err);
        ang_apply(e1, e2, I1, I2, dl, dq);
        lam[0] += dl;
        PosPrep p;
        pos_prep(e1, e2, c.r1, c.r2, p);
        V3 dx = sub(add(e1.x, p.r1), add(e2.x, p.r2));
        dl = pos_dlambda(e1, e2, p, h, 0.0, lam[1], dx, err);
        pos_apply(e1, e2, p, dl, dx);
        lam[1] += dl;
        if (c.limited) {
            V3 n1 = q_axis(e1.q, c.axis[2], bad);
            V3 n2 = q_axis(e2.q, c.axis[3], bad);
            V3 n = q_axis(e1.q, c.axis[0], bad);
            if (limit_angle(n, n1, n2, c.limits[0], c.limits[1], dq)) {
                M3 J1 = world_tensor(e1.q, e1.Iinv), J2 = world_tensor(e2.q, e2.Iinv);
                double d2 = ang_dlambda(J1, J2, h, 0.0, lam[2], dq, err);
                ang_apply(e1, e2, J1, J2, d2, dq);
                lam[2] += d2;
            }
        }
    } break;
    case 4: {  // SPHERICAL_JOINT_CONSTRAINT
        const double EPSILON = 1e-50;
        PosPrep p;
        pos_prep(e1, e2, c.r1, c.r2, p);
        V3 dx = sub(add(e1.x, p.r1), add(e2.x, p.r2));
        double dl = pos_dlambda(e1, e2, p, h, 0.0, lam[0], dx, err);
        pos_apply(e1, e2, p, dl, dx);
        lam[0] += dl;
        V3 n1 = q_axis(e1.q, c.axis[0], bad);
        V3 n2 = q_axis(e2.q, c.axis[1], bad);
        V3 n = cross(n1, n2);
        double n_len = len(n);
        if (n_len > EPSILON) {
            n = v3(n.x / n_len, n.y / n_len, n.z / n_len);
            V3 dq;
            if (limit_angle(n, n1, n2, c.limits[0], c.limits[1], dq)) {
                M3 I1 = world_tensor(e1.q, e1.Iinv), I2 = world_tensor(e2.q, e2.Iinv);
                double d2 = ang_dlambda(I1, I2, h, 0.0, lam[1], dq, err);
                ang_apply(e1, e2, I1, I2, d2, dq);
                lam[1] += d2;
            }
        }
        V3 a1 = q_axis(e1.q, c.axis[0], bad);
        V3 b1 = q_axis(e1.q, c.axis[2], bad);
        V3 a2 = q_axis(e2.q, c.axis[1], bad);
        V3 b2 = q_axis(e2.q, c.axis[3], bad);
        n = add(a1, a2);
        n_len = len(n);
        if (n_len > EPSILON) {
            n = v3(n.x / n_len, n.y / n_len, n.z / n_len);
            n1 = sub(b1, scl(dot(n, b1), n));
            n2 = sub(b2, scl(dot(n, b2), n));
            double n1_len = len(n1);
            double n2_len = len(n2);
            if (n1_len > EPSILON && n2_len > EPSILON) {
                n1 = v3(n1.x / n1_len, n1.y / n1_len, n1.z / n1_len);
                n2 = v3(n2.x / n2_len, n2.y / n2_len, n2.z / n2_len);
                V3 dq;
                if (limit_angle(n, n1, n2, c.limits[2], c.limits[3], dq)) {
                    M3 I1 = world_tensor(e1.q, e1.Iinv), I2 = world_tensor(e2.q, e2.Iinv);
                    double d3 = ang_dlambda(I1, I2, h, 0.0, lam[2], dq, err);
                    ang_apply(e1, e2, I1, I2, d3, dq);
                    lam[2] += d3;
                }
            }
        }
    } break;
    default: bad = true; break;
    }
    if (bad) err |= ERR_SELECT;
}

// Velocity solve of one collision constraint (pbd.cpp:646-711)
template <class S>
RP_HD void velocity_solve(const ContactT<S>& c, BodyT<S>& e1, BodyT<S>& e2, double h) {
    V3T<S> n = c.n;
    S lambda_n = c.lambda_n;
    PosPrepT<S> p;
    pos_prep(e1, e2, c.r1, c.r2, p);
    V3T<S> v = sub(add(e1.v, cross(e1.w, p.r1)), add(e2.v, cross(e2.w, p.r2)));
    S vn = dot(n, v);
    V3T<S> vt = sub(v, scl(vn, n));
    V3T<S> dv = v3<S>(S(0.0), S(0.0), S(0.0));
    const S mu_d = (e1.mu_d + e2.mu_d) / 2.0f;
    S fn = lambda_n / h;
    S fact = RP_MIN(mu_d * fabs(fn), len(vt));
    dv = add(dv, scl(-fact, nrm(vt)));
    V3T<S> vtil = sub(add(e1.pv, cross(e1.pw, p.r1)), add(e2.pv, cross(e2.pw, p.r2)));
    S vn_til = dot(n, vtil);
    S e = e1.e * e2.e;
    fact = -vn + RP_MIN(-e * vn_til, S(0.0));
    dv = add(dv, scl(fact, n));
    S w1 = e1.inv_mass + dot(cross(p.r1, n), m3_mulv(p.I1, cross(p.r1, n)));
    S w2 = e2.inv_mass + dot(cross(p.r2, n), m3_mulv(p.I2, cross(p.r2, n)));
    V3T<S> pimp = scl(1.0 / (w1 + w2), dv);
    if (!e1.fixed) {
        e1.v = add(e1.v, scl(e1.inv_mass, pimp));
        e1.w = add(e1.w, m3_mulv(p.I1, cross(p.r1, pimp)));
    }
    if (!e2.fixed) {
        e2.v = add(e2.v, inv3(scl(e2.inv_mass, pimp)));
        e2.w = add(e2.w, inv3(m3_mulv(p.I2, cross(p.r2, pimp))));
    }
}

// Integration of one body (pbd.cpp:537-577)
template <class S>
RP_HD void integrate(BodyT<S>& e, double h) {
    e.px = e.x;
    e.pq = e.q;
    if (e.fixed || !e.active) return;
    e.v = add(e.v, scl(h * e.inv_mass, e.F));
    e.x = add(e.x, scl(h, e.v));
    M3T<S> Iinv = world_tensor(e.q, e.Iinv);
    M3T<S> I = world_tensor(e.q, e.I);
    e.w = add(e.w, scl(h, m3_mulv(Iinv, sub(e.T, cross(e.w, m3_mulv(I, e.w))))));
    Q4T<S> aux = {e.w.x, e.w.y, e.w.z, S(0.0)};
    Q4T<S> q = q_mul(aux, e.q);
    e.q.x = e.q.x + h * 0.5 * q.x;
    e.q.y = e.q.y + h * 0.5 * q.y;
    e.q.z = e.q.z + h * 0.5 * q.z;
    e.q.w = e.q.w + h * 0.5 * q.w;
    e.q = q_nrm(e.q);
}

// Velocity update of one body (pbd.cpp:623-643)
template <class S>
RP_HD void velocity_update(BodyT<S>& e, double h) {
    if (e.fixed || !e.active) return;
    e.pv = e.v;
    e.pw = e.w;
    e.v = scl(1.0 / h, sub(e.x, e.px));
    Q4T<S> dq = q_mul(e.q, q_inv(e.pq));
    if (dq.w >= 0.0) e.w = scl(2.0 / h, v3<S>(dq.x, dq.y, dq.z));
    else e.w = scl(-2.0 / h, v3<S>(dq.x, dq.y, dq.z));
}

// clipping_contact_to_collision_constraint (pbd.cpp:408-424)
RP_HD void contact_to_constraint(const Body& e1, const Body& e2, const ContactPt& ct, Contact& c) {
    c.n = ct.n;
    c.lambda_n = 0.0;
    c.lambda_t = 0.0;
    V3 r1w = sub(ct.p1, e1.x);
    V3 r2w = sub(ct.p2, e2.x);
    c.r1 = q_rot(q_inv(e1.q), r1w);
    c.r2 = q_rot(q_inv(e2.q), r2w);
}

}  // namespace rp
