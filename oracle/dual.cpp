// dual.cpp — the oracle compiled with dual numbers (forward-mode derivatives).
//
// TEST INFRASTRUCTURE ONLY. `real` carries a value and ORC_ND directional
// derivatives; the value parts are computed exactly as in the double build,
// comparisons look at values (the branches are the primal run's), and the
// float truncations of the clipping code (clipping.cpp:13,27-39) drop the
// derivative. Seeds set the derivative slots of a body's initial state
// (x, q, v, w), of its mass (1/m, the inertia tensor and its inverse) or of
// its friction coefficients; the derivatives of any later state are read
// back per slot. With the freeze switch on, the contact geometry that
// contact_to_constraint hands to the solver (r1_lc, r2_lc, n) carries no
// derivative: the frozen-contact map the device adjoint differentiates.
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#define ORC_ND 8

struct real {
    double v;
    double d[ORC_ND];
    real() : v(0.0) { for (double& x : d) x = 0.0; }
    real(double x) : v(x) { for (double& y : d) y = 0.0; }
    explicit operator double() const { return v; }
};
static inline real operator+(const real& a, const real& b) {
    real r(a.v + b.v);
    for (int k = 0; k < ORC_ND; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
static inline real operator-(const real& a, const real& b) {
    real r(a.v - b.v);
    for (int k = 0; k < ORC_ND; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
static inline real operator*(const real& a, const real& b) {
    real r(a.v * b.v);
    for (int k = 0; k < ORC_ND; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
static inline real operator/(const real& a, const real& b) {
    real r(a.v / b.v);
    for (int k = 0; k < ORC_ND; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
    return r;
}
static inline real operator-(const real& a) {
    real r(-a.v);
    for (int k = 0; k < ORC_ND; ++k) r.d[k] = -a.d[k];
    return r;
}
static inline real& operator+=(real& a, const real& b) { return a = a + b; }
static inline real& operator-=(real& a, const real& b) { return a = a - b; }
static inline real& operator*=(real& a, const real& b) { return a = a * b; }
static inline real& operator/=(real& a, const real& b) { return a = a / b; }
static inline bool operator==(const real& a, const real& b) { return a.v == b.v; }
static inline bool operator!=(const real& a, const real& b) { return a.v != b.v; }
static inline bool operator<(const real& a, const real& b) { return a.v < b.v; }
static inline bool operator>(const real& a, const real& b) { return a.v > b.v; }
static inline bool operator<=(const real& a, const real& b) { return a.v <= b.v; }
static inline bool operator>=(const real& a, const real& b) { return a.v >= b.v; }
static inline real sqrt(const real& a) {
    real r(std::sqrt(a.v));
    double g = r.v > 0.0 ? 0.5 / r.v : 0.0;
    for (int k = 0; k < ORC_ND; ++k) r.d[k] = g * a.d[k];
    return r;
}
static inline real fabs(const real& a) {
    real r(std::fabs(a.v));
    double g = a.v < 0.0 ? -1.0 : 1.0;
    for (int k = 0; k < ORC_ND; ++k) r.d[k] = g * a.d[k];
    return r;
}
static inline real sin(const real& a) {
    real r(std::sin(a.v));
    double g = std::cos(a.v);
    for (int k = 0; k < ORC_ND; ++k) r.d[k] = g * a.d[k];
    return r;
}
static inline real cos(const real& a) {
    real r(std::cos(a.v));
    double g = -std::sin(a.v);
    for (int k = 0; k < ORC_ND; ++k) r.d[k] = g * a.d[k];
    return r;
}
static inline real asin(const real& a) {
    real r(std::asin(a.v));
    double g = 1.0 / std::sqrt(1.0 - a.v * a.v);
    for (int k = 0; k < ORC_ND; ++k) r.d[k] = g * a.d[k];
    return r;
}

namespace {
int g_freeze = 0;
inline void strip(real& x) {
    for (double& y : x.d) y = 0.0;
}
}  // namespace

#define ORC_DUAL 1
#define RD(x) ((x).v)
#define ORC_PHASE(p) ((void)0)
#define ORC_FREEZE_V3(vec) \
    do {                   \
        if (g_freeze) {    \
            strip((vec).x);  \
            strip((vec).y);  \
            strip((vec).z);  \
        }                  \
    } while (0)
#include "rp_oracle.c"

static body* body_at(orc_world* w, uint32_t g) {
    for (uint32_t si = 0; si < w->ns; ++si) {
        if (g < w->s[si].nb) return &w->s[si].b[g];
        g -= w->s[si].nb;
    }
    return nullptr;
}
static real* state_comp(body* e, int k) {
    real* c[13] = {&e->x.x, &e->x.y, &e->x.z, &e->q.x, &e->q.y, &e->q.z, &e->q.w,
                   &e->v.x, &e->v.y, &e->v.z, &e->w.x, &e->w.y, &e->w.z};
    return c[k];
}

extern "C" {
__attribute__((visibility("default"))) void orc_dual_set_freeze(int on) { g_freeze = on; }

/* Sets derivative slot `slot` of body g to the direction of: kind 0 = state
 * component comp of (x, q, v, w); 1 = mass; 2 = static friction; 3 = dynamic
 * friction. Returns 0, or -1 for a bad argument. */
__attribute__((visibility("default"))) int orc_dual_seed(orc_world* w, uint32_t g, int slot, int kind, int comp) {
    body* e = body_at(w, g);
    if (!e || slot < 0 || slot >= ORC_ND) return -1;
    switch (kind) {
    case 0:
        if (comp < 0 || comp >= 13) return -1;
        state_comp(e, comp)->d[slot] = 1.0;
        break;
    case 1: {  // m: 1/m, I = m C, I^-1 = C^-1 / m
        double im = e->inv_mass.v;
        e->inv_mass.d[slot] = -im * im;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                e->inertia.m[i][j].d[slot] = e->inertia.m[i][j].v * im;
                e->inv_inertia.m[i][j].d[slot] = -e->inv_inertia.m[i][j].v * im;
            }
    } break;
    case 2: e->mu_s.d[slot] = 1.0; break;
    case 3: e->mu_d.d[slot] = 1.0; break;
    default: return -1;
    }
    return 0;
}

/* d(x, q, v, w)/d(seed of slot) of body g: out[13]. */
__attribute__((visibility("default"))) int orc_dual_state(orc_world* w, uint32_t g, int slot, double out[13]) {
    body* e = body_at(w, g);
    if (!e || slot < 0 || slot >= ORC_ND) return -1;
    for (int k = 0; k < 13; ++k) out[k] = state_comp(e, k)->d[slot];
    return 0;
}
}
