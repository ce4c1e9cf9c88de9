// rp_math.h — FP64 vector / quaternion / matrix primitives for the device step.
//
// Bit-exactness with the reference depends on evaluating every expression in
// the reference's order and association (include/gm.h, src/quaternion.cpp)
// with no FMA contraction: all sources are compiled with -ffp-contract=off,
// and no expression here may be "simplified" (e.g. gm_vec3_invert is 0 - v,
// which differs from -v on signed zeros).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define RP_HD __host__ __device__ __forceinline__

namespace rp {

// Value types over a scalar T: double for the step; the adjoint evaluates
// the solver functions with a dual-number scalar (rp_dual.h). The double
// instantiations perform exactly the operations written here.
template <class T> struct V3T { T x, y, z; };
template <class T> struct Q4T { T x, y, z, w; };
template <class T> struct M3T { T m[3][3]; };
using V3 = V3T<double>;
using Q4 = Q4T<double>;
using M3 = M3T<double>;
struct M4 { double m[4][4]; };

template <class T> RP_HD V3T<T> v3(T x, T y, T z) { V3T<T> r; r.x = x; r.y = y; r.z = z; return r; }
// gm.h:667-694, 722-729
template <class T> RP_HD V3T<T> add(V3T<T> a, V3T<T> b) { return v3<T>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T> RP_HD V3T<T> sub(V3T<T> a, V3T<T> b) { return v3<T>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class S, class T> RP_HD V3T<T> scl(S s, V3T<T> v) { return v3<T>(s * v.x, s * v.y, s * v.z); }
template <class T> RP_HD T dot(V3T<T> a, V3T<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> RP_HD V3T<T> cross(V3T<T> a, V3T<T> b) {
    return v3<T>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// gm.h:649-652
template <class T> RP_HD T len(V3T<T> v) { return sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
// gm.h:627-634: zero stays zero
template <class T> RP_HD V3T<T> nrm(V3T<T> v) {
    if (!(v.x != 0.0 || v.y != 0.0 || v.z != 0.0)) return v3<T>(T(0.0), T(0.0), T(0.0));
    T l = len(v);
    return v3<T>(v.x / l, v.y / l, v.z / l);
}
// gm.h:588-590: 0 - v
template <class T> RP_HD V3T<T> inv3(V3T<T> v) { return sub(v3<T>(T(0.0), T(0.0), T(0.0)), v); }
RP_HD bool veq(V3 a, V3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
RP_HD bool is_zero(V3 v) { return !(v.x != 0.0 || v.y != 0.0 || v.z != 0.0); }

// gm.h:370-381
template <class T> RP_HD M3T<T> m3_mul(const M3T<T>& a, const M3T<T>& b) {
    M3T<T> r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j];
    return r;
}
// gm.h:420-426
template <class T> RP_HD V3T<T> m3_mulv(const M3T<T>& m, V3T<T> v) {
    return v3<T>(m.m[0][0] * v.x + m.m[0][1] * v.y + m.m[0][2] * v.z, m.m[1][0] * v.x + m.m[1][1] * v.y + m.m[1][2] * v.z,
                 m.m[2][0] * v.x + m.m[2][1] * v.y + m.m[2][2] * v.z);
}
// gm.h:445-452
template <class T> RP_HD M3T<T> m3_t(const M3T<T>& m) {
    M3T<T> r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.m[i][j] = m.m[j][i];
    return r;
}
// gm.h:327-345
RP_HD M4 m4_mul(const M4& a, const M4& b) {
    M4 r;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
            r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j] + a.m[i][3] * b.m[3][j];
    return r;
}

// quaternion.cpp:133-142
template <class T> RP_HD Q4T<T> q_mul(Q4T<T> a, Q4T<T> b) {
    Q4T<T> r;
    r.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
    r.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
    r.y = a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z;
    r.z = a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x;
    return r;
}
// quaternion.cpp:144-147
template <class T> RP_HD Q4T<T> q_nrm(Q4T<T> q) {
    T l = sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    Q4T<T> r;
    r.x = q.x / l; r.y = q.y / l; r.z = q.z / l; r.w = q.w / l;
    return r;
}
// quaternion.cpp:80-87
template <class T> RP_HD Q4T<T> q_inv(Q4T<T> q) {
    Q4T<T> r;
    r.x = -q.x; r.y = -q.y; r.z = -q.z; r.w = q.w;
    return r;
}
// quaternion.cpp:257-268
template <class T> RP_HD V3T<T> q_rot(Q4T<T> q, V3T<T> v) {
    T ix = q.w * v.x + q.y * v.z - q.z * v.y;
    T iy = q.w * v.y + q.z * v.x - q.x * v.z;
    T iz = q.w * v.z + q.x * v.y - q.y * v.x;
    T iw = -q.x * v.x - q.y * v.y - q.z * v.z;
    return v3<T>((ix * q.w) + (iw * -q.x) + (iy * -q.z) - (iz * -q.y), (iy * q.w) + (iw * -q.y) + (iz * -q.x) - (ix * -q.z),
              (iz * q.w) + (iw * -q.z) + (ix * -q.y) - (iy * -q.x));
}
// quaternion.cpp:89-105
template <class T> RP_HD M3T<T> q_mat3(Q4T<T> q) {
    M3T<T> r;
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
// Model matrix without scale, T * R (util.cpp:60-74, quaternion.cpp:107-131)
RP_HD M4 model_matrix(V3 t, Q4 q) {
    M3 r3 = q_mat3(q);
    M4 R;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j) R.m[i][j] = r3.m[i][j];
        R.m[i][3] = 0.0;
        R.m[3][i] = 0.0;
    }
    R.m[3][3] = 1.0;
    M4 T = {{{1.0, 0.0, 0.0, t.x}, {0.0, 1.0, 0.0, t.y}, {0.0, 0.0, 1.0, t.z}, {0.0, 0.0, 0.0, 1.0}}};
    return m4_mul(T, R);
}
// entity_get_model_matrix (entity.cpp:123-145): T * (R * S), the matrix a
// renderer draws the body's mesh with.
RP_HD M4 entity_model_matrix(V3 t, Q4 q, V3 s) {
    M3 r3 = q_mat3(q);
    M4 Rq;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j) Rq.m[i][j] = r3.m[i][j];
        Rq.m[i][3] = 0.0;
        Rq.m[3][i] = 0.0;
    }
    Rq.m[3][3] = 1.0;
    M4 S = {{{s.x, 0.0, 0.0, 0.0}, {0.0, s.y, 0.0, 0.0}, {0.0, 0.0, s.z, 0.0}, {0.0, 0.0, 0.0, 1.0}}};
    M4 T = {{{1.0, 0.0, 0.0, t.x}, {0.0, 1.0, 0.0, t.y}, {0.0, 0.0, 1.0, t.z}, {0.0, 0.0, 0.0, 1.0}}};
    return m4_mul(T, m4_mul(Rq, S));
}
// collider_update for one hull vertex (collider.cpp:413-422): gm_mat4_multiply_vec4
// of M with [v, 1] (gm.h:405-413), then scaled by 1/w.
RP_HD V3 transform_vertex(const M4& M, V3 v) {
    double x = v.x * M.m[0][0] + v.y * M.m[0][1] + v.z * M.m[0][2] + 1.0 * M.m[0][3];
    double y = v.x * M.m[1][0] + v.y * M.m[1][1] + v.z * M.m[1][2] + 1.0 * M.m[1][3];
    double z = v.x * M.m[2][0] + v.y * M.m[2][1] + v.z * M.m[2][2] + 1.0 * M.m[2][3];
    double ww = v.x * M.m[3][0] + v.y * M.m[3][1] + v.z * M.m[3][2] + 1.0 * M.m[3][3];
    // A model matrix's last row is (0, 0, 0, 1), so ww is exactly 1 for a
    // finite vertex; then 1/ww = 1 and each sc * t = t bit for bit.
    if (ww == 1.0) return v3(x, y, z);
    double sc = 1.0 / ww;
    return v3(sc * x, sc * y, sc * z);
}
// collider_update for one face normal (collider.cpp:425-429):
// gm_mat4_multiply_vec3(M, n, false) (gm.h:448-460), then normalized.
RP_HD V3 transform_normal(const M4& M, V3 n) {
    V3 r = v3(M.m[0][0] * n.x + M.m[0][1] * n.y + M.m[0][2] * n.z, M.m[1][0] * n.x + M.m[1][1] * n.y + M.m[1][2] * n.z,
              M.m[2][0] * n.x + M.m[2][1] * n.y + M.m[2][2] * n.z);
    return nrm(r);
}
// quaternion.cpp:3-16
RP_HD Q4 q_from_axis_radians(V3 axis, double angle) {
    if (len(axis) != 0.0) axis = nrm(axis);
    double sang = sin(angle / 2.0);
    Q4 q;
    q.w = cos(angle / 2.0);
    q.x = axis.x * sang;
    q.y = axis.y * sang;
    q.z = axis.z * sang;
    return q;
}
// Basis axes (quaternion.cpp:33-78) indexed by PBD_Axis_Type (pbd.cpp:219-243)
RP_HD V3 q_axis(Q4 q, int axis, bool& bad) {
    switch (axis) {
    case 0: return v3(1.0 - 2.0 * q.y * q.y - 2.0 * q.z * q.z, 2.0 * q.x * q.y - 2.0 * -q.w * q.z, 2.0 * q.x * q.z + 2.0 * -q.w * q.y);
    case 1: return v3(1.0 - 2.0 * q.y * q.y - 2.0 * q.z * q.z, 2.0 * q.x * q.y - 2.0 * q.w * q.z, 2.0 * q.x * q.z + 2.0 * q.w * q.y);
    case 2: return v3(2.0 * q.x * q.y + 2.0 * -q.w * q.z, 1.0 - (2.0 * q.x * q.x) - (2.0 * q.z * q.z), 2.0 * q.y * q.z - 2.0 * -q.w * q.x);
    case 3: return v3(2.0 * q.x * q.y + 2.0 * q.w * q.z, 1.0 - (2.0 * q.x * q.x) - (2.0 * q.z * q.z), 2.0 * q.y * q.z - 2.0 * q.w * q.x);
    case 4: return v3(2.0 * q.x * q.z - 2.0 * -q.w * q.y, 2.0 * q.y * q.z + 2.0 * -q.w * q.x, 1.0 - (2.0 * q.x * q.x) - (2.0 * q.y * q.y));
    case 5: return v3(2.0 * q.x * q.z - 2.0 * q.w * q.y, 2.0 * q.y * q.z + 2.0 * q.w * q.x, 1.0 - (2.0 * q.x * q.x) - (2.0 * q.y * q.y));
    }
    bad = true;
    return v3(0.0, 0.0, 0.0);
}

// Dynamic (world) inverse inertia R * I^-1 * R^T (physics_util.cpp:44-61)
template <class T> RP_HD M3T<T> world_tensor(Q4T<T> q, const M3T<T>& local) {
    M3T<T> R = q_mat3(q);
    M3T<T> Rt = m3_t(R);
    M3T<T> aux = m3_mul(R, local);
    return m3_mul(aux, Rt);
}

#define RP_PI_F 3.14159265358979 /* gm.h:8 */
#define RP_MIN(a, b) (((a) < (b)) ? (a) : (b))
#define RP_MAX(a, b) (((a) > (b)) ? (a) : (b))
#define RP_CLAMP(x, low, high) (((x) > (high)) ? (high) : (((x) < (low)) ? (low) : (x)))

}  // namespace rp
