// opcount.cpp — the oracle compiled with an op-counting arithmetic type.
//
// TEST / MEASUREMENT INFRASTRUCTURE ONLY. It counts the algorithmic FP64 work
// of the restatement: every add, subtract, multiply, divide and square root
// (sin/cos/asin count 1 each; negation, comparison, fabs and conversions
// count 0), attributed to the step's phases. The counts are exact for the
// given inputs because the control flow is the oracle's own. bench.py uses
// them as the "algorithmic flops" of each kernel class.
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace {
uint64_t g_ops[16];
int g_phase = 7;
}  // namespace

struct real {
    double v;
    real() = default;
    real(double x) : v(x) {}
    explicit operator double() const { return v; }
};
static inline real operator+(real a, real b) { ++g_ops[g_phase]; return real(a.v + b.v); }
static inline real operator-(real a, real b) { ++g_ops[g_phase]; return real(a.v - b.v); }
static inline real operator*(real a, real b) { ++g_ops[g_phase]; return real(a.v * b.v); }
static inline real operator/(real a, real b) { ++g_ops[g_phase]; return real(a.v / b.v); }
static inline real operator-(real a) { return real(-a.v); }
static inline real& operator+=(real& a, real b) { a = a + b; return a; }
static inline real& operator-=(real& a, real b) { a = a - b; return a; }
static inline real& operator*=(real& a, real b) { a = a * b; return a; }
static inline real& operator/=(real& a, real b) { a = a / b; return a; }
static inline bool operator==(real a, real b) { return a.v == b.v; }
static inline bool operator!=(real a, real b) { return a.v != b.v; }
static inline bool operator<(real a, real b) { return a.v < b.v; }
static inline bool operator>(real a, real b) { return a.v > b.v; }
static inline bool operator<=(real a, real b) { return a.v <= b.v; }
static inline bool operator>=(real a, real b) { return a.v >= b.v; }
static inline real sqrt(real a) { ++g_ops[g_phase]; return real(std::sqrt(a.v)); }
static inline real fabs(real a) { return real(std::fabs(a.v)); }
static inline real sin(real a) { ++g_ops[g_phase]; return real(std::sin(a.v)); }
static inline real cos(real a) { ++g_ops[g_phase]; return real(std::cos(a.v)); }
static inline real asin(real a) { ++g_ops[g_phase]; return real(std::asin(a.v)); }

#define ORC_COUNT_OPS 1
#define RD(x) ((x).v)
#define ORC_PHASE(p) (g_phase = (p))
#include "rp_oracle.c"

extern "C" {
__attribute__((visibility("default"))) void orc_opcount_get(uint64_t out[16]) {
    for (int i = 0; i < 16; ++i) out[i] = g_ops[i];
}
__attribute__((visibility("default"))) void orc_opcount_reset(void) {
    for (int i = 0; i < 16; ++i) g_ops[i] = 0;
}
}
