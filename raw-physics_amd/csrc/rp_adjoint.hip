// rp_adjoint.hip — reverse-mode gradients of the step over a recorded run.
//
// Scope: scenes with one dynamic body (any number of fixed bodies), as in a
// body thrown onto static geometry. The contact set is frozen: the contacts
// the forward pass generated in each substep (r1_lc, r2_lc, n) are constants
// of the backward pass, so the gradient is that of the step with those
// contacts (the discrete collision decisions and the contact geometry carry
// no derivative).
//
// Per substep the state is (x, q, v, w) of the dynamic body: integration
// overwrites the previous pose at the substep start and the velocity update
// overwrites the previous velocities before the velocity solve reads them.
// The parameters are the body's mass (through 1/m, the inertia tensor and
// its inverse, and the gravity force m g), static and dynamic friction.
//
// The backward sweep runs frame by frame from the last: the host restores
// the frame's checkpoint and replays the frame with recording (substep start
// states and contacts), then this kernel walks the substeps in reverse. For
// each substep, lane k of the scene's 16-lane group re-evaluates the substep
// from its recorded start state with a dual number seeded on input k, which
// gives column k of the substep Jacobian J; the lane then forms
// (J^T abar)_k. Lanes 0-12 produce the new state adjoint, lanes 13-15
// accumulate the parameter gradients. The dual value parts repeat the primal
// operations, so each re-evaluation reproduces the recorded next state bit
// for bit; a mismatch is counted (STAT_ADJ_MISMATCH).
#include "rp_device.h"
#include "rp_dual.h"

namespace rp {

using D1 = ad::Dual<1>;
constexpr int ADJ_LANES = 16;       // 13 state inputs + mass, static and dynamic friction
constexpr int ADJ_MAX_CONTACTS = 32;  // contacts of one scene in one substep

template <class S> RP_HD V3T<S> lift(V3 a) { return v3<S>(S(a.x), S(a.y), S(a.z)); }
template <class S> RP_HD Q4T<S> lift(Q4 a) { Q4T<S> r = {S(a.x), S(a.y), S(a.z), S(a.w)}; return r; }
template <class S> RP_HD M3T<S> lift(const M3& a) {
    M3T<S> r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = S(a.m[i][j]);
    return r;
}
template <class S> RP_HD BodyT<S> lift(const Body& b) {
    BodyT<S> r;
    r.x = lift<S>(b.x); r.q = lift<S>(b.q); r.v = lift<S>(b.v); r.w = lift<S>(b.w);
    r.px = lift<S>(b.px); r.pq = lift<S>(b.pq); r.pv = lift<S>(b.pv); r.pw = lift<S>(b.pw);
    r.deact = S(b.deact); r.inv_mass = S(b.inv_mass);
    r.I = lift<S>(b.I); r.Iinv = lift<S>(b.Iinv);
    r.mu_s = S(b.mu_s); r.mu_d = S(b.mu_d); r.e = S(b.e); r.radius = S(b.radius);
    r.F = lift<S>(b.F); r.T = lift<S>(b.T);
    r.fixed = b.fixed; r.active = b.active;
    return r;
}

// State component k of (x, q, v, w).
__device__ __forceinline__ D1& state_comp(BodyT<D1>& e, int k) {
    switch (k) {
    case 0: return e.x.x; case 1: return e.x.y; case 2: return e.x.z;
    case 3: return e.q.x; case 4: return e.q.y; case 5: return e.q.z; case 6: return e.q.w;
    case 7: return e.v.x; case 8: return e.v.y; case 9: return e.v.z;
    case 10: return e.w.x; case 11: return e.w.y; default: return e.w.z;
    }
}
__device__ __forceinline__ double state_val(const Body& e, int k) {
    switch (k) {
    case 0: return e.x.x; case 1: return e.x.y; case 2: return e.x.z;
    case 3: return e.q.x; case 4: return e.q.y; case 5: return e.q.z; case 6: return e.q.w;
    case 7: return e.v.x; case 8: return e.v.y; case 9: return e.v.z;
    case 10: return e.w.x; case 11: return e.w.y; default: return e.w.z;
    }
}

// One substep of a scene with dynamic body `bd` (global index), with the
// recorded contacts (pbd.cpp:537-711 for this scene: integration, contact
// position solve in pair order, velocity update, velocity solve).
__device__ void adj_substep(const DevWorld& w, const Recorder& r, int step, int s, int bd, BodyT<D1>& e,
                            const FrameParams& fp, uint32_t& err) {
    int b0 = w.scene_body_off[s];
    // the frame's gravity force (stack.cpp:93-96, physics_util.cpp:5-23)
    V3T<D1> force = v3<D1>(D1(0.0), -fp.gravity * 1.0 / e.inv_mass, D1(0.0));
    e.F = add(v3<D1>(D1(0.0), D1(0.0), D1(0.0)), force);
    e.T = add(v3<D1>(D1(0.0), D1(0.0), D1(0.0)),
              cross(sub(v3<D1>(D1(0.0), D1(0.0), D1(0.0)), v3<D1>(D1(0.0), D1(0.0), D1(0.0))), force));
    integrate(e, fp.h);
    // contacts of the scene's live pairs, in pair order
    ContactT<D1> cs[ADJ_MAX_CONTACTS];
    int other[ADJ_MAX_CONTACTS];
    bool dyn_first[ADJ_MAX_CONTACTS];
    int ncs = 0;
    int g0 = w.live_prefix[s], g1 = g0 + w.live_count[s];
    for (int g = g0; g < g1; ++g) {
        int2 p = w.pairs[w.live_global[g].y];
        int i1 = b0 + p.x, i2 = b0 + p.y;
        int nc = r.ncontact[(size_t)step * r.pair_cap + g];
        for (int c = 0; c < nc; ++c) {
            if (ncs == ADJ_MAX_CONTACTS) { err |= ERR_SOLVER_CAPACITY; break; }
            const Contact& k = r.contacts[((size_t)step * r.pair_cap + g) * MAXC + c];
            cs[ncs].r1 = lift<D1>(k.r1);
            cs[ncs].r2 = lift<D1>(k.r2);
            cs[ncs].n = lift<D1>(k.n);
            cs[ncs].lambda_n = D1(0.0);
            cs[ncs].lambda_t = D1(0.0);
            dyn_first[ncs] = (i1 == bd);
            other[ncs] = (i1 == bd) ? i2 : i1;
            ++ncs;
        }
    }
    for (unsigned it = 0; it < fp.iters; ++it)
        for (int c = 0; c < ncs; ++c) {
            BodyT<D1> f = lift<D1>(w.body[other[c]]);  // fixed: never written
            if (dyn_first[c]) solve_collision(cs[c], e, f, fp.h, err);
            else solve_collision(cs[c], f, e, fp.h, err);
        }
    velocity_update(e, fp.h);
    for (int c = 0; c < ncs; ++c) {
        BodyT<D1> f = lift<D1>(w.body[other[c]]);
        if (dyn_first[c]) velocity_solve(cs[c], e, f, fp.h);
        else velocity_solve(cs[c], f, e, fp.h);
    }
}

__global__ void __launch_bounds__(64) k_adjoint_frame(DevWorld w, Recorder r, const int32_t* scene_dyn,
                                                      const Body* end_state, double* abar, double* gparam,
                                                      double* amax, FrameParams fp) {
    constexpr int SPB = 64 / ADJ_LANES;
    int s = blockIdx.x * SPB + (int)threadIdx.x / ADJ_LANES;
    int k = (int)threadIdx.x % ADJ_LANES;
    int slot = (int)threadIdx.x / ADJ_LANES;
    __shared__ double a_out[SPB][13];
    __shared__ double a_new[SPB][ADJ_LANES];
    int bd = (s < w.ns) ? scene_dyn[s] : -1;
    if (bd >= 0 && k < 13) a_out[slot][k] = abar[13 * bd + k];
    __syncthreads();
    uint32_t err = 0;
    bool mismatch = false;
    for (int sub = (int)fp.substeps - 1; sub >= 0; --sub) {
        double acc = 0.0;
        if (bd >= 0) {
            const Body& b = r.body[(size_t)sub * w.nb + bd];
            BodyT<D1> e = lift<D1>(b);
            if (k < 13) {
                state_comp(e, k).d[0] = 1.0;
            } else if (k == 13) {  // mass m: 1/m, I = m C, I^-1 = C^-1 / m
                double im = b.inv_mass;
                e.inv_mass.d[0] = -im * im;
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) {
                        e.I.m[i][j].d[0] = b.I.m[i][j] * im;
                        e.Iinv.m[i][j].d[0] = -b.Iinv.m[i][j] * im;
                    }
            } else if (k == 14) {
                e.mu_s.d[0] = 1.0;
            } else {
                e.mu_d.d[0] = 1.0;
            }
            adj_substep(w, r, sub, s, bd, e, fp, err);
            const Body& nx = (sub + 1 < (int)fp.substeps) ? r.body[(size_t)(sub + 1) * w.nb + bd] : end_state[bd];
            for (int j = 0; j < 13; ++j) {
                const D1& o = state_comp(e, j);
                acc += a_out[slot][j] * o.d[0];
                if (__double_as_longlong(o.v) != __double_as_longlong(state_val(nx, j))) mismatch = true;
            }
        }
        a_new[slot][k] = acc;
        __syncthreads();
        if (bd >= 0) {
            if (k < 13) a_out[slot][k] = a_new[slot][k];
            else gparam[3 * bd + (k - 13)] += a_new[slot][k];
        }
        __syncthreads();
    }
    if (bd >= 0 && k < 13) abar[13 * bd + k] = a_out[slot][k];
    if (bd >= 0 && k == 0) {  // largest state adjoint at a frame start so far
        double m = amax[bd];
        for (int j = 0; j < 13; ++j) m = fmax(m, fabs(a_out[slot][j]));
        amax[bd] = m;
    }
    if (mismatch) atomicAdd(&w.stats[STAT_ADJ_MISMATCH], 1ull);
    if (err && bd >= 0 && k == 0) atomicOr(&w.scene_err[s], err);
}

hipError_t launch_adjoint_frame(const DevWorld& w, const Recorder& rec, const int32_t* scene_dyn, const Body* end_state,
                                double* abar, double* gparam, double* amax, const FrameParams& fp, hipStream_t st) {
    constexpr int SPB = 64 / ADJ_LANES;
    hipLaunchKernelGGL(k_adjoint_frame, dim3((w.ns + SPB - 1) / SPB), dim3(64), 0, st, w, rec, scene_dyn, end_state,
                       abar, gparam, amax, fp);
    return hipGetLastError();
}

}  // namespace rp
