// rp_device.h — device helpers shared by the small-scene and large-scene kernels.
#pragma once
#include "rp_kernels.h"

namespace rp {

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Ordered wave compaction: returns this lane's output slot among lanes with
// pred set (in lane order) and the wave total in *total.
__device__ __forceinline__ int wave_compact(bool pred, int* total) {
    unsigned long long m = __ballot(pred);
    *total = __popcll(m);
    unsigned long long below = (lane_id() == 0) ? 0ull : (m & ((~0ull) >> (64 - lane_id())));
    return __popcll(below);
}

// Loads of the body fields each solve reads (the rest of Body is unused).
__device__ __forceinline__ void load_pos_body(const Body& g, Body& b) {
    b.x = g.x; b.q = g.q; b.px = g.px; b.pq = g.pq;
    b.inv_mass = g.inv_mass; b.Iinv = g.Iinv; b.mu_s = g.mu_s; b.fixed = g.fixed;
}
__device__ __forceinline__ void load_vel_body(const Body& g, Body& b) {
    b.q = g.q; b.v = g.v; b.w = g.w; b.pv = g.pv; b.pw = g.pw;
    b.inv_mass = g.inv_mass; b.Iinv = g.Iinv; b.mu_d = g.mu_d; b.e = g.e; b.fixed = g.fixed;
}

// The broad-phase test of two bodies (broad.cpp:17-21), body i first.
__device__ __forceinline__ bool broad_hit(V3 x1, double r1, V3 x2, double r2) {
    double dd = len(sub(x1, x2));
    double maxd = r1 + r2 + 0.1;
    return dd <= maxd;
}
__device__ __forceinline__ bool broad_hit(const Body& e1, const Body& e2) {
    return broad_hit(e1.x, e1.radius, e2.x, e2.radius);
}

// The frame's records of live pair g between global bodies b1 and b2: the
// bodies and their hull instances, and both hulls' local-vertex offsets and
// counts (so that the narrow phase loads the vertices right after the
// records). A hull index < 0 (sphere or compound body) gives a zero hull.
__device__ __forceinline__ void write_live_records(const DevWorld& w, int g, int b1, int b2) {
    int h1 = w.body_hull[b1], h2 = w.body_hull[b2];
    w.live_rec[g] = make_int4(b1, b2, h1, h2);
    w.live_hull[g] = make_int4(h1 >= 0 ? w.hull_v_off[h1] : 0, h2 >= 0 ? w.hull_v_off[h2] : 0,
                               h1 >= 0 ? w.hull_nv[h1] : 0, h2 >= 0 ? w.hull_nv[h2] : 0);
}

// Lock-free union-find over scene-local parent indices (islands,
// broad.cpp:31-116): any order of the unions gives the same partition.
__device__ __forceinline__ int uf_load(const int32_t* p, int x) {
    return __hip_atomic_load(p + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Root of x, halving the path on the way (x's parent becomes its
// grandparent). Only non-roots are written, and hooks only change roots, so
// the stores never race with a hook; every parent stays an ancestor with a
// smaller or equal index, so the partition and the walks' termination are
// unchanged. A settled pile is one island of 10^5 bodies: without the
// halving its walks took the union pass to 2.7 ms per frame.
__device__ __forceinline__ int uf_root(int32_t* p, int x) {
    while (true) {
        int y = uf_load(p, x);
        if (y == x) return x;
        int z = uf_load(p, y);
        if (z == y) return y;
        __hip_atomic_store(p + x, z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = z;
    }
}
// Links the sets of a and b: the larger root is hooked under the smaller, so
// parent indices only decrease and every root walk ends.
__device__ __forceinline__ void uf_link(int32_t* p, int a, int b) {
    while (true) {
        a = uf_root(p, a);
        b = uf_root(p, b);
        if (a == b) return;
        if (a > b) { int t = a; a = b; b = t; }
        int old = atomicCAS(p + b, b, a);
        if (old == b) return;
    }
}

// Global body indices of a solve item: an external constraint (item < 0,
// index ~item) or live pair `item` of the scene whose first body is b0.
// Computed once per frame into item_bodies beside item_order, so that the
// solve reaches an item's bodies in one load instead of a chain of three.
__device__ __forceinline__ int2 item_body_pair(const DevWorld& w, int item, int b0) {
    if (item < 0) return make_int2(w.ext[~item].b1, w.ext[~item].b2);
    int2 p = w.pairs[w.live_global[item].y];
    return make_int2(b0 + p.x, b0 + p.y);
}

// One solve item: an external constraint (item < 0, index ~item) or the
// contacts of live pair `item` in order (pbd.cpp:381-406, 615-620), on
// bodies bb (item_body_pair).
template <bool kExt>
__device__ __forceinline__ void solve_item_pos(const DevWorld& w, int item, int2 bb, double h, uint32_t& err) {
    int bi1 = bb.x, bi2 = bb.y;
    Body B1, B2;
    load_pos_body(w.body[bi1], B1);
    load_pos_body(w.body[bi2], B2);
    if (item < 0) {
        if constexpr (kExt) solve_external(w.ext[~item], w.ext_lambda + 3 * (~item), B1, B2, h, err);
    } else {
        int nc = w.ncontact[item];
        Contact* cs = w.contacts + (size_t)item * MAXC;
        for (int c = 0; c < nc; ++c) {
            Contact ct = cs[c];
            solve_collision(ct, B1, B2, h, err);
            cs[c].lambda_n = ct.lambda_n;
            cs[c].lambda_t = ct.lambda_t;
        }
    }
    if (!B1.fixed) { w.body[bi1].x = B1.x; w.body[bi1].q = B1.q; }
    if (!B2.fixed) { w.body[bi2].x = B2.x; w.body[bi2].q = B2.q; }
}

// Velocity solve of the contacts of one item (pbd.cpp:646-711); external
// constraints take no part.
__device__ __forceinline__ void solve_item_vel(const DevWorld& w, int item, int2 bb, double h) {
    if (item < 0) return;
    int nc = w.ncontact[item];
    if (nc == 0) return;
    int bi1 = bb.x, bi2 = bb.y;
    Body B1, B2;
    load_vel_body(w.body[bi1], B1);
    load_vel_body(w.body[bi2], B2);
    const Contact* cs = w.contacts + (size_t)item * MAXC;
    for (int c = 0; c < nc; ++c) {
        Contact ct = cs[c];
        velocity_solve(ct, B1, B2, h);
    }
    if (!B1.fixed) { w.body[bi1].v = B1.v; w.body[bi1].w = B1.w; }
    if (!B2.fixed) { w.body[bi2].v = B2.v; w.body[bi2].w = B2.w; }
}

// Velocity update of one body (pbd.cpp:623-643).
__device__ __forceinline__ void velocity_update_body(Body& g, double h) {
    if (g.fixed || !g.active) return;
    Body e;
    e.x = g.x; e.px = g.px; e.q = g.q; e.pq = g.pq; e.v = g.v; e.w = g.w;
    e.fixed = 0; e.active = 1;
    velocity_update(e, h);
    g.pv = e.pv; g.pw = e.pw; g.v = e.v; g.w = e.w;
}

// Appends live pair g to the small-hull or the large-hull list (one atomic
// per wave and list; the lists' order does not matter). All lanes of the
// wave call it; `valid` marks the lanes holding a pair.
__device__ __forceinline__ void classify_pair(const DevWorld& w, bool valid, int g, int b1, int b2) {
    bool big = false;
    if (valid) {
        int h1 = w.body_hull[b1], h2 = w.body_hull[b2];
        big = (h1 >= 0 && w.hull_nv[h1] > w.small_nv) || (h2 >= 0 && w.hull_nv[h2] > w.small_nv);
    }
    unsigned long long mb = __ballot(valid && big), ms = __ballot(valid && !big);
    int lane = lane_id();
    unsigned long long below = lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
    int base_b = 0, base_s = 0;
    int leader = __ffsll((long long)(mb | ms)) - 1;
    if (lane == leader) {
        if (mb) base_b = atomicAdd(&w.counters[CNT_BIG], __popcll(mb));
        if (ms) base_s = atomicAdd(&w.counters[CNT_SMALL], __popcll(ms));
    }
    if (leader < 0) return;
    base_b = __shfl(base_b, leader);
    base_s = __shfl(base_s, leader);
    if (!valid) return;
    if (big) w.pair_lists[w.list_stride + base_b + __popcll(mb & below)] = g;
    else w.pair_lists[base_s + __popcll(ms & below)] = g;
}

}  // namespace rp
