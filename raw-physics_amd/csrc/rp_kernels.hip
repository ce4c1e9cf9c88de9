// rp_kernels.hip — CDNA4 kernels of the batched step.
//
// One frame of every scene (pbd_simulate_with_constraints, pbd.cpp:468-747):
//   k_frame_begin   body lanes:  external force / torque of the frame
//   k_broad         wave/scene:  all-pairs bounding-sphere test, ordered compaction
//   k_islands       wave/scene:  union-find islands, sleeping, live-pair list
//   k_scan_live     one block:   prefix over scenes' live-pair counts
//   k_scatter_live  wave/scene:  global live-pair list
//   then num_substeps times:
//   k_integrate     body lanes:  integration + model matrix
//   k_transform     element lanes: hull vertices / face normals to world
//   k_narrow        pair lanes:  GJK -> EPA -> manifold -> collision constraints
//   k_solve         wave/scene:  dependency-levelled Gauss-Seidel position solve,
//                                velocity update, velocity solve
//
// The Gauss-Seidel order of the reference is kept exactly: constraints are
// assigned levels so that two constraints on the same non-fixed body keep
// their relative order, and constraints of one level (which touch disjoint
// writable bodies) run on different lanes. Fixed bodies are never written by
// any solve (pbd_base_constraints.cpp:103-133, pbd.cpp:702-711), so sharing
// one does not order two constraints. Every body therefore sees the same
// sequence of updates as in the sequential loop and results are bit-identical.
#include <algorithm>
#include <mutex>
#include <vector>

#include "rp_device.h"

namespace rp {

// ---------------------------------------------------------------------------
__global__ void k_frame_begin(DevWorld w, double gravity) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= w.nb) return;
    b += w.body_base;
    Body& e = w.body[b];
    // Caller pattern (stack.cpp:93-96) and force / torque sums over its one
    // force (physics_util.cpp:5-23).
    V3 pos = v3(0.0, 0.0, 0.0);
    V3 force = v3(0.0, -gravity * 1.0 / e.inv_mass, 0.0);
    e.F = add(v3(0.0, 0.0, 0.0), force);
    e.T = add(v3(0.0, 0.0, 0.0), cross(sub(pos, v3(0.0, 0.0, 0.0)), force));
}

// broad_get_collision_pairs (broad.cpp:6-29): pairs (i<j) in i-major order.
__global__ void k_broad(DevWorld w) {
    int s = blockIdx.x;
    if (w.scene_big[s]) return;
    int lane = threadIdx.x;
    int b0 = w.scene_body_off[s], n = w.scene_body_off[s + 1] - b0;
    int base = w.scene_pair_base[s];
    int count = 0;
    for (int i = lane; i < n; i += 64) w.need[b0 + i] = 0;
    for (int i = 0; i < n; ++i) {
        const Body& e1 = w.body[b0 + i];
        for (int j0 = i + 1; j0 < n; j0 += 64) {
            int j = j0 + lane;
            bool hit = false;
            if (j < n) hit = broad_hit(e1, w.body[b0 + j]);
            int total;
            int slot = wave_compact(hit, &total);
            if (hit) w.pairs[base + count + slot] = make_int2(i, j);
            count += total;
        }
    }
    if (lane == 0) w.pair_count[s] = count;
}

// Islands and sleeping (broad.cpp:31-116, pbd.cpp:476-506), then the list of
// pairs the narrow phase visits (pbd.cpp:593-596). The reference joins the
// islands one pair after another (uf_union, broad.cpp:55-59); the wave's
// lanes join them at once with the lock-free union-find, which gives the same
// partition, and a body is active unless every body of its island has slept
// for a second, which depends on the partition only.
__global__ void k_islands(DevWorld w, double dt) {
    int s = blockIdx.x;
    if (w.scene_big[s]) return;
    int lane = threadIdx.x;
    int b0 = w.scene_body_off[s], n = w.scene_body_off[s + 1] - b0;
    int base = w.scene_pair_base[s];
    int np = w.pair_count[s];
    int* parent = w.uf_parent + b0;  // scene-local indices
    int* alli = w.uf_flag + b0;
    for (int i = lane; i < n; i += 64) { parent[i] = i; alli[i] = 1; }
    __syncthreads();
    for (int k = lane; k < np; k += 64) {
        int2 p = w.pairs[base + k];
        if (!w.body[b0 + p.x].fixed && !w.body[b0 + p.y].fixed) uf_link(parent, p.x, p.y);
    }
    for (int k = w.scene_ext_off[s] + lane; k < w.scene_ext_off[s + 1]; k += 64) {
        int i = w.ext[k].b1 - b0, j = w.ext[k].b2 - b0;
        if (!w.body[b0 + i].fixed && !w.body[b0 + j].fixed) uf_link(parent, i, j);
    }
    __syncthreads();
    for (int i = lane; i < n; i += 64) {
        Body& e = w.body[b0 + i];
        if (e.fixed) continue;
        double lv = len(e.v), av = len(e.w);
        if (lv < 0.10 && av < 0.10) e.deact += dt;
        else e.deact = 0.0;
        if (e.deact < 1.0) alli[uf_root(parent, i)] = 0;
    }
    __syncthreads();
    for (int i = lane; i < n; i += 64) {
        Body& e = w.body[b0 + i];
        if (e.fixed) continue;
        e.active = !alli[uf_root(parent, i)];
    }
    __syncthreads();
    // live pairs, in pair order
    int count = 0;
    uint32_t err = 0;
    for (int k0 = 0; k0 < np; k0 += 64) {
        int k = k0 + lane;
        bool live = false;
        if (k < np) {
            int2 p = w.pairs[base + k];
            const Body& e1 = w.body[b0 + p.x];
            const Body& e2 = w.body[b0 + p.y];
            if (!e1.fixed && !e2.fixed && !((e1.active && e2.active) || (!e1.active && !e2.active))) err |= ERR_ACTIVITY;
            live = !((e1.fixed || !e1.active) && (e2.fixed || !e2.active));
        }
        int total;
        int slot = wave_compact(live, &total);
        if (live) w.live_local[base + count + slot] = k;
        count += total;
    }
    if (err) atomicOr(&w.scene_err[s], err);
    if (lane == 0) w.live_count[s] = count;
}

// Exclusive prefix sum of live_count over scenes -> live_prefix[0..ns].
__global__ void k_scan_live(DevWorld w) {
    __shared__ int part[1024];
    int t = threadIdx.x, nt = blockDim.x;
    if (t == 0) {  // the size-class lists of this frame start empty
        w.counters[CNT_SMALL] = 0;
        w.counters[CNT_BIG] = 0;
    }
    int ns = w.ns;
    int chunk = (ns + nt - 1) / nt;
    int lo = min(ns, t * chunk), hi = min(ns, lo + chunk);
    int sum = 0;
    for (int i = lo; i < hi; ++i) sum += w.live_count[i];
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < nt; off <<= 1) {
        int v = (t >= off) ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - sum;
    for (int i = lo; i < hi; ++i) {
        w.live_prefix[i] = run;
        run += w.live_count[i];
    }
    if (t == nt - 1) w.live_prefix[ns] = part[t];
}

__global__ void k_scatter_live(DevWorld w) {
    int s = blockIdx.x;
    if (w.scene_big[s]) return;
    int base = w.scene_pair_base[s];
    int g0 = w.live_prefix[s], n = w.live_count[s];
    int b0 = w.scene_body_off[s];
    for (int m0 = 0; m0 < n; m0 += blockDim.x) {
        int m = m0 + threadIdx.x;
        bool valid = m < n;
        int b1 = 0, b2 = 0;
        if (valid) {
            int slot = base + w.live_local[base + m];
            w.live_global[g0 + m] = make_int2(s, slot);
            int2 p = w.pairs[slot];
            b1 = b0 + p.x;
            b2 = b0 + p.y;
            write_live_records(w, g0 + m, b1, b2);
            w.need[b1] = 1;  // same value from every writer
            w.need[b2] = 1;
        }
        if (w.small_nv > 0) classify_pair(w, valid, g0 + m, b1, b2);
    }
}

// ---------------------------------------------------------------------------
// Integration of body b (pbd.cpp:565-596) and its model matrix.
__device__ __forceinline__ void integrate_body(const DevWorld& w, int b, double h) {
    Body e = w.body[b];
    integrate(e, h);
    w.body[b].x = e.x;
    w.body[b].q = e.q;
    w.body[b].v = e.v;
    w.body[b].w = e.w;
    w.body[b].px = e.px;
    w.body[b].pq = e.pq;
    w.model[b] = model_matrix(e.x, e.q);
}

__global__ void k_integrate(DevWorld w, double h) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b == 0) w.counters[CNT_MANIFOLD] = 0;  // the substep's manifold queue starts empty
    if (b >= w.nb) return;
    integrate_body(w, b + w.body_base, h);
}

// The pose every body has when the substep's collision pass reads it
// (colliders_update at pbd.cpp:598-599 runs after integration): launched
// after the last substep's integration when the caller wants the colliders'
// transformed geometry of that pose written back.
__global__ void k_keep_pose(DevWorld w) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= w.nb) return;
    b += w.body_base;
    const Body& e = w.body[b];
    double* p = w.pose + 7 * (size_t)b;
    p[0] = e.x.x; p[1] = e.x.y; p[2] = e.x.z;
    p[3] = e.q.x; p[4] = e.q.y; p[5] = e.q.z; p[6] = e.q.w;
}

// entity_get_model_matrix of every body, row-major 4x4, as double or as the
// float array util_matrix_to_r32_array (util.cpp:77-94) hands to the renderer.
template <class F>
__global__ void k_model_out(DevWorld w, F* out) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= w.nb) return;
    const Body& e = w.body[b];
    M4 M = entity_model_matrix(e.x, e.q, w.body_scale[b]);
    F* o = out + 16 * (size_t)b;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) o[4 * i + j] = (F)M.m[i][j];
}

// Transformed vertices of the bodies that take part in a live pair this
// frame (lane-per-pair narrow phase and the inspection path).
__global__ void k_transform(DevWorld w, int all) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= w.n_xf) return;
    int4 el = w.xf[t];  // (body, hull instance, vertex, tv slot)
    int b = el.x;
    if (!all && !w.need[b]) return;
    w.tv[el.w] = transform_vertex(w.model[b], w.hull_v[w.hull_v_off[el.y] + el.z]);
}

// View of hull instance `hid` of body b whose transformed vertices start at tv_off.
__device__ __forceinline__ HullView hull_view_of(const DevWorld& w, int b, int hid, int tv_off) {
    HullView hv;
    int tp = w.hull_topo[hid];
    const int32_t* tb = w.topo_base + 8 * tp;
    hv.tv = w.tv + tv_off;
    hv.fn = w.hull_fn + w.hull_fn_off[hid];
    const M4& M = w.model[b];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) hv.R.m[i][j] = M.m[i][j];
    hv.nv = w.hull_nv[hid];
    hv.nf = w.hull_nf[hid];
    hv.fe_off = w.topo_pool + tb[0];
    hv.fe = w.topo_pool + tb[1];
    hv.vf_off = w.topo_pool + tb[2];
    hv.vf = w.topo_pool + tb[3];
    hv.vn_off = w.topo_pool + tb[4];
    hv.vn = w.topo_pool + tb[5];
    hv.ff_off = w.topo_pool + tb[6];
    hv.ff = w.topo_pool + tb[7];
    return hv;
}
// The hull of a body with one hull collider (the fast narrow path).
__device__ __forceinline__ HullView hull_view(const DevWorld& w, int b) {
    return hull_view_of(w, b, w.body_hull[b], w.body_tv_off[b]);
}
// Collider c (global index) of body b.
__device__ __forceinline__ ColView col_view(const DevWorld& w, int b, int c) {
    ColView v;
    v.type = w.col_type[c];
    v.radius = w.col_radius[c];
    v.center = w.body[b].x;  // colliders_update: the sphere follows the entity (collider.cpp:431-433)
    if (v.type == COL_HULL) v.h = hull_view_of(w, b, w.col_hull[c], w.col_tv_off[c]);
    else memset(&v.h, 0, sizeof(v.h));
    return v;
}

// Narrow phase over the live pairs of all scenes (pbd.cpp:583-612), in two
// kernels: GJK + EPA for every live pair, then the contact manifold for the
// pairs whose EPA converged (collider.cpp:523-558). Both take G lanes per
// pair: G = 1 is one lane per pair over the pre-transformed vertices
// (k_transform); G = 4, 8, 16 put 64/G pairs in a wave, transform both
// hulls' vertices into LDS and split every support search across the group
// (support_index<G>); the rest runs identically on the group's lanes and
// lane 0 of the group writes the results. Results go to the pair's slot g in
// the live list (scene, then pair order), so the queue order between the
// kernels does not matter.
__device__ __forceinline__ void group_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Scene and global body indices of live pair g (both loads depend on g
// only: the bodies come from the frame's live-pair record).
__device__ __forceinline__ void pair_bodies(const DevWorld& w, int g, int& s, int& b1, int& b2) {
    s = w.live_global[g].x;
    int4 r = w.live_rec[g];
    b1 = r.x;
    b2 = r.y;
}
// Pairs with a sphere or a multi-collider body take the general path.
__device__ __forceinline__ bool general_pair(const DevWorld& w, int b1, int b2) {
    return w.has_general && (w.body_general[b1] | w.body_general[b2]);
}

// Hull views of live pair g; for G > 1 the group's lanes transform both
// hulls' vertices into the group's LDS stage `st`.
template <int G>
__device__ __forceinline__ void pair_views(const DevWorld& w, int g, V3* st, int gl, int& s, int& b1, int& b2,
                                           HullView& h1, HullView& h2) {
    s = w.live_global[g].x;
    int4 r = w.live_rec[g];  // bodies and hull instances: no pair-slot and scene-offset loads first
    b1 = r.x;
    b2 = r.y;
    h1 = hull_view_of(w, b1, r.z, w.body_tv_off[b1]);
    h2 = hull_view_of(w, b2, r.w, w.body_tv_off[b2]);
    if constexpr (G > 1) {
        // vertex offsets and counts from the pair's record: the vertex loads
        // depend on that one load only
        int4 hv = w.live_hull[g];
        const V3* lv1 = w.hull_v + hv.x;
        const V3* lv2 = w.hull_v + hv.y;
        int nv1 = hv.z, nv2 = hv.w;
        const M4 M1 = w.model[b1], M2 = w.model[b2];
        // both hulls together, two vertices of each per lane and batch, all
        // loaded before any is transformed
        int n = nv1 > nv2 ? nv1 : nv2;
        for (int i = gl; i < n; i += 2 * G) {
            V3 va[2], vb[2];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                int j = i + k * G;
                if (j < nv1) va[k] = lv1[j];
                if (j < nv2) vb[k] = lv2[j];
            }
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                int j = i + k * G;
                if (j < nv1) st[j] = transform_vertex(M1, va[k]);
                if (j < nv2) st[nv1 + j] = transform_vertex(M2, vb[k]);
            }
        }
        group_lds_sync();
        h1.tv = st;
        h2.tv = st + nv1;
    }
}

// kManifold: the contact manifold of a converged pair is built here too, on
// the same staged hulls, and its contacts written (no manifold queue).
// list / count: visit the pairs of a size-class list (LDS slices of `stage`
// vertices per pair), or (list null) every live pair with slices of
// stage_cap vertices.
template <int G, bool kManifold>
__device__ __forceinline__ void gjk_epa_pairs(DevWorld& w, const int32_t* list, const int32_t* count, int stage) {
    extern __shared__ V3 rp_stage[];
    constexpr int NG = 64 / G;
    int grp = threadIdx.x / G, gl = threadIdx.x & (G - 1);
    V3* st = rp_stage + grp * stage;
    int total = list ? *count : w.live_prefix[w.ns];
    for (int k = blockIdx.x * NG + grp; k < total; k += gridDim.x * NG) {
        int g = list ? list[k] : k;
        int s, b1, b2;
        pair_bodies(w, g, s, b1, b2);
        if (general_pair(w, b1, b2)) continue;  // k_narrow_general (uniform per group)
        HullView h1, h2;
        pair_views<G>(w, g, st, gl, s, b1, b2, h1, h2);
        uint32_t err = 0;
        bool cycled = false;
        Simplex sx;
        V3 n = v3(0.0, 0.0, 0.0);
        double pen;
        int r = 0;
        char* work = G > 1 ? w.group_work + (size_t)(blockIdx.x * NG + grp) * GROUP_WORK_BYTES : nullptr;
        bool hit = gjk<G>(h1, h2, sx, err, cycled);
        if (hit) r = epa<G>(h1, h2, sx, n, pen, err, work);
        if constexpr (kManifold) {
            ContactPt local[G == 1 ? MAXC : 1];
            ContactPt* cts = G == 1 ? local : (ContactPt*)(work + CLIP_BYTES);
            int nc = 0;
            if (r == 1) {
                int used = h1.nv + h2.nv;  // the rest of the pair's LDS slice is free
                nc = manifold<G>(h1, h2, n, cts, err, work, st + used, G > 1 ? stage - used : 0);
            }
            // one contact per lane (the group's count is uniform)
            if constexpr (G > 1) group_lds_sync();
            const Body& e1 = w.body[b1];
            const Body& e2 = w.body[b2];
            Contact* out = w.contacts + (size_t)g * MAXC;
            for (int c = gl; c < nc; c += G) contact_to_constraint(e1, e2, cts[c], out[c]);
            if (gl == 0) w.ncontact[g] = nc;
        }
        if (gl == 0) {
            if constexpr (!kManifold) {
                if (r == 1) {
                    w.pair_normal[g] = n;
                    w.mlist[atomicAdd(&w.counters[CNT_MANIFOLD], 1)] = g;
                } else {
                    w.ncontact[g] = 0;
                }
            }
            if (err) atomicOr(&w.scene_err[s], err);
            if (hit && r != 1) atomicAdd(&w.stats[STAT_EPA_FAIL], 1ull);  // epa.cpp:233-236
            if (cycled) atomicAdd(&w.stats[STAT_GJK_CYCLE], 1ull);
        }
        if constexpr (G > 1) group_lds_sync();
    }
}

// Lane groups (G = 4, 8, 16) at two waves per SIMD; one lane per pair (G = 1).
template <int G, bool kManifold>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
k_gjk_epa(DevWorld w, const int32_t* list, const int32_t* count, int stage) {
    static_assert(G > 1, "lane groups");
    gjk_epa_pairs<G, kManifold>(w, list, count, stage);
}
template <bool kManifold>
__global__ void __launch_bounds__(64) k_gjk_epa_lane(DevWorld w) {
    gjk_epa_pairs<1, kManifold>(w, nullptr, nullptr, 0);
}

template <int G>
__global__ void __launch_bounds__(64) k_manifold(DevWorld w) {
    extern __shared__ V3 rp_stage[];
    constexpr int NG = 64 / G;
    int grp = threadIdx.x / G, gl = threadIdx.x & (G - 1);
    V3* st = rp_stage + grp * w.stage_cap;
    int total = w.counters[CNT_MANIFOLD];
    for (int k = blockIdx.x * NG + grp; k < total; k += gridDim.x * NG) {
        int g = w.mlist[k];
        int s, b1, b2;
        HullView h1, h2;
        pair_views<G>(w, g, st, gl, s, b1, b2, h1, h2);
        uint32_t err = 0;
        char* work = G > 1 ? w.group_work + (size_t)(blockIdx.x * NG + grp) * GROUP_WORK_BYTES : nullptr;
        ContactPt local[G == 1 ? MAXC : 1];
        ContactPt* cts = G == 1 ? local : (ContactPt*)(work + CLIP_BYTES);
        int used = h1.nv + h2.nv;
        int nc = manifold<G>(h1, h2, w.pair_normal[g], cts, err, work, st + used, G > 1 ? w.stage_cap - used : 0);
        if constexpr (G > 1) group_lds_sync();
        const Body& e1 = w.body[b1];
        const Body& e2 = w.body[b2];
        Contact* out = w.contacts + (size_t)g * MAXC;
        for (int c = gl; c < nc; c += G) contact_to_constraint(e1, e2, cts[c], out[c]);
        if (gl == 0) {
            w.ncontact[g] = nc;
            if (err) atomicOr(&w.scene_err[s], err);
        }
        if constexpr (G > 1) group_lds_sync();
    }
}

// Live pairs with a sphere collider or a multi-collider body, one lane per
// pair: colliders_get_contacts (collider.cpp:560-572) over every collider
// pair in order, each by collider_get_contacts (collider.cpp:523-558).
__global__ void __launch_bounds__(64) k_narrow_general(DevWorld w) {
    int total = w.live_prefix[w.ns];
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < total; g += gridDim.x * blockDim.x) {
        int s, b1, b2;
        pair_bodies(w, g, s, b1, b2);
        if (!general_pair(w, b1, b2)) continue;
        uint32_t err = 0;
        int flags = 0;
        ContactPt cts[MAXC];
        int nc = 0;
        for (int c1 = w.body_col_off[b1]; c1 < w.body_col_off[b1 + 1]; ++c1)
            for (int c2 = w.body_col_off[b2]; c2 < w.body_col_off[b2 + 1]; ++c2) {
                ColView v1 = col_view(w, b1, c1), v2 = col_view(w, b2, c2);
                nc += collider_contacts(v1, v2, cts + nc, MAXC - nc, err, flags);
            }
        const Body& e1 = w.body[b1];
        const Body& e2 = w.body[b2];
        Contact* out = w.contacts + (size_t)g * MAXC;
        for (int c = 0; c < nc; ++c) contact_to_constraint(e1, e2, cts[c], out[c]);
        w.ncontact[g] = nc;
        if (err) atomicOr(&w.scene_err[s], err);
        if (flags & PAIR_EPA_FAIL) atomicAdd(&w.stats[STAT_EPA_FAIL], 1ull);
        if (flags & PAIR_GJK_CYCLE) atomicAdd(&w.stats[STAT_GJK_CYCLE], 1ull);
    }
}

// Dependency levels of one frame's solve items (wave/scene). An item is one
// external constraint or the contacts of one live pair (solved in order by
// one lane). The reference's sequence is: external constraints, then the
// contacts of each pair in pair order (pbd.cpp:580-620). An item's level is
// one above the last level of either of its non-fixed bodies, so items that
// share a writable body keep their order and items of one level touch
// disjoint writable bodies. Pairs are fixed for the frame (the broad phase
// and activity run once per frame, pbd.cpp:474-506), so levels are too.
constexpr int LEV_CHUNK = 512;
constexpr int LEV_LDS_BODIES = 1024;

__global__ void __launch_bounds__(64) k_levels(DevWorld w, int collisions) {
    int s = blockIdx.x;
    if (w.scene_big[s]) return;
    int lane = threadIdx.x;
    int b0 = w.scene_body_off[s], n = w.scene_body_off[s + 1] - b0;
    int e0 = w.scene_ext_off[s], next = w.scene_ext_off[s + 1] - e0;
    int g0 = w.live_prefix[s];
    int np = collisions ? w.live_count[s] : 0;
    int K = next + np;
    int ib = w.scene_item_base[s];
    int* lev = w.item_level + ib;
    int* order = w.item_order + ib;
    int* ls = w.item_lstart + ib + s;
    __shared__ int sb1[LEV_CHUNK], sb2[LEV_CHUNK];
    __shared__ int slast[LEV_LDS_BODIES];
    __shared__ int sL;
    bool lds_last = n <= LEV_LDS_BODIES;
    int* last = lds_last ? slast : (w.body_last + b0);
    for (int i = lane; i < n; i += 64) last[i] = 0;
    if (lane == 0) sL = 0;
    __syncthreads();
    for (int c0 = 0; c0 < K; c0 += LEV_CHUNK) {
        int cn = min(LEV_CHUNK, K - c0);
        for (int k = lane; k < cn; k += 64) {
            int t = c0 + k, i1, i2;
            if (t < next) {
                i1 = w.ext[e0 + t].b1 - b0;
                i2 = w.ext[e0 + t].b2 - b0;
            } else {
                int2 p = w.pairs[w.live_global[g0 + t - next].y];
                i1 = p.x;
                i2 = p.y;
            }
            sb1[k] = w.body[b0 + i1].fixed ? -1 : i1;  // fixed bodies never order items
            sb2[k] = w.body[b0 + i2].fixed ? -1 : i2;
        }
        __syncthreads();
        if (lane == 0) {
            int L = sL;
            for (int k = 0; k < cn; ++k) {
                int i1 = sb1[k], i2 = sb2[k];
                int lv = 0;
                if (i1 >= 0) lv = last[i1];
                if (i2 >= 0) lv = max(lv, last[i2]);
                ++lv;
                if (i1 >= 0) last[i1] = lv;
                if (i2 >= 0) last[i2] = lv;
                lev[c0 + k] = lv - 1;
                L = max(L, lv);
            }
            sL = L;
        }
        __syncthreads();
    }
    int L = sL;
    // counting sort by level; the order inside a level does not matter (its
    // items touch disjoint writable bodies)
    int* cur = w.item_cursor + ib + s;
    for (int l = lane; l <= L; l += 64) ls[l] = 0;
    __syncthreads();
    for (int t = lane; t < K; t += 64) atomicAdd(&ls[lev[t] + 1], 1);
    __syncthreads();
    if (lane == 0) {
        for (int l = 0; l < L; ++l) ls[l + 1] += ls[l];
        w.scene_nlevels[s] = L;
    }
    __syncthreads();
    for (int l = lane; l <= L; l += 64) cur[l] = ls[l];
    __syncthreads();
    int2* bodies = w.item_bodies + ib;
    for (int t = lane; t < K; t += 64) {
        int pos = atomicAdd(&cur[lev[t]], 1);
        int code = (t < next) ? ~(e0 + t) : (g0 + t - next);
        order[pos] = code;
        bodies[pos] = item_body_pair(w, code, b0);
    }
}

// Solve of one scene and substep: position Gauss-Seidel (pbd.cpp:615-620),
// velocity update (pbd.cpp:623-643), velocity solve (pbd.cpp:646-711), level
// by level over the frame's items. LPS lanes serve one scene, so a wave
// solves 64/LPS scenes side by side (small scenes leave most of a 64-lane
// wave idle otherwise).
template <bool kExt, int LPS>
__global__ void __launch_bounds__(64) k_solve(DevWorld w, double h, int iters, int collisions, int integrate_next) {
    constexpr int SPW = 64 / LPS;
    int s = blockIdx.x * SPW + (int)threadIdx.x / LPS;
    int lane = (int)threadIdx.x % LPS;
    bool valid = s < w.ns && !w.scene_big[s];  // large scenes: coloured solve (rp_large.hip)
    int b0 = 0, n = 0, e0 = 0, e1x = 0, L = 0;
    const int* order = nullptr;
    const int2* bodies = nullptr;
    const int* ls = nullptr;
    if (valid) {
        b0 = w.scene_body_off[s];
        n = w.scene_body_off[s + 1] - b0;
        e0 = w.scene_ext_off[s];
        e1x = w.scene_ext_off[s + 1];
        int ib = w.scene_item_base[s];
        order = w.item_order + ib;
        bodies = w.item_bodies + ib;
        ls = w.item_lstart + ib + s;
        L = w.scene_nlevels[s];
    }
    int Lmax = L;  // levels of the wave's busiest scene: every lane joins every barrier
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) Lmax = max(Lmax, __shfl_xor(Lmax, off));
    // level starts ls[0..L] in the scene's lanes' registers (lane l holds
    // ls[l]) when every scene of the wave fits, read back by lane exchange
    // instead of a global load per level (the branch is uniform per wave)
    bool ls_in_lanes = Lmax + 1 <= LPS;
    int my_ls = (valid && lane <= L) ? ls[lane] : 0;
    auto level_start = [&](int l) -> int {
        if (ls_in_lanes) return __shfl(my_ls, l, LPS);
        return (l <= L) ? ls[l] : 0;
    };
    uint32_t err = 0;
    (void)collisions;

    // multipliers of the external constraints start at zero (pbd.cpp:426-462)
    if constexpr (kExt) {
        for (int k = e0 + lane; k < e1x; k += LPS) {
            w.ext_lambda[3 * k] = 0.0;
            w.ext_lambda[3 * k + 1] = 0.0;
            w.ext_lambda[3 * k + 2] = 0.0;
        }
    }
    __syncthreads();
    for (int it = 0; it < iters; ++it) {
        for (int l = 0; l < Lmax; ++l) {
            int a = level_start(l), b = level_start(l + 1);
            int i0 = (l < L) ? a : 0, i1 = (l < L) ? b : 0;
            for (int idx = i0 + lane; idx < i1; idx += LPS) solve_item_pos<kExt>(w, order[idx], bodies[idx], h, err);
            __syncthreads();
        }
    }
    // velocity update
    for (int i = lane; i < n; i += LPS) velocity_update_body(w.body[b0 + i], h);
    __syncthreads();
    // velocity solve over the collision items, in the same levels
    for (int l = 0; l < Lmax; ++l) {
        int a = level_start(l), b = level_start(l + 1);
        int i0 = (l < L) ? a : 0, i1 = (l < L) ? b : 0;
        for (int idx = i0 + lane; idx < i1; idx += LPS) solve_item_vel(w, order[idx], bodies[idx], h);
        __syncthreads();
    }
    if (err && valid) atomicOr(&w.scene_err[s], err);
    // the next substep's integration of the scene's bodies (worlds without
    // large scenes), in place of a k_integrate launch
    if (integrate_next) {
        __syncthreads();
        for (int i = lane; i < n; i += LPS) integrate_body(w, b0 + i, h);
        if (blockIdx.x == 0 && threadIdx.x == 0) w.counters[CNT_MANIFOLD] = 0;
    }
}

template <bool kExt>
static void launch_solve(const DevWorld& w, const FrameParams& fp, hipStream_t st, int integrate_next) {
    int spw = 64 / w.solve_lanes;
    dim3 grid((w.ns + spw - 1) / spw);
    int it = (int)fp.iters, co = fp.collisions ? 1 : 0;
    switch (w.solve_lanes) {
    case 16: hipLaunchKernelGGL((k_solve<kExt, 16>), grid, dim3(64), 0, st, w, fp.h, it, co, integrate_next); break;
    case 32: hipLaunchKernelGGL((k_solve<kExt, 32>), grid, dim3(64), 0, st, w, fp.h, it, co, integrate_next); break;
    default: hipLaunchKernelGGL((k_solve<kExt, 64>), grid, dim3(64), 0, st, w, fp.h, it, co, integrate_next); break;
    }
}

// Contacts of one body pair at the current transformed poses (debug / parity).
// One lane group of 8 over the pre-transformed hulls, its polytope and
// polygons in lane-group working memory slot 0 (no large private arrays).
__global__ void __launch_bounds__(64) k_pair_contacts(DevWorld w, int b1, int b2, double* out, int* count,
                                                     uint32_t* err_out) {
    if (blockIdx.x != 0 || threadIdx.x >= 8) return;
    HullView h1 = hull_view(w, b1), h2 = hull_view(w, b2);
    uint32_t err = 0;
    int flags = 0;
    char* work = w.group_work;
    ContactPt* cts = (ContactPt*)(work + CLIP_BYTES);
    int nc = pair_contacts<8>(h1, h2, cts, err, flags, work);
    if (threadIdx.x != 0) return;
    for (int k = 0; k < nc; ++k) {
        double* o = out + 9 * k;
        o[0] = cts[k].p1.x; o[1] = cts[k].p1.y; o[2] = cts[k].p1.z;
        o[3] = cts[k].p2.x; o[4] = cts[k].p2.y; o[5] = cts[k].p2.z;
        o[6] = cts[k].n.x; o[7] = cts[k].n.y; o[8] = cts[k].n.z;
    }
    *count = nc;
    *err_out = err;
}

// Model matrices of all bodies at their current pose (no integration).
__global__ void k_model(DevWorld w) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= w.nb) return;
    w.model[b] = model_matrix(w.body[b].x, w.body[b].q);
}

// Elementwise FP64 primitives the step relies on: sqrt and division must be
// correctly rounded for bit-exactness; asin/sin/cos feed the joint limits.
__global__ void k_fp64_selftest(const double* a, const double* b, double* out, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = a[i], y = b[i];
    out[5 * i + 0] = sqrt(x);
    out[5 * i + 1] = x / y;
    out[5 * i + 2] = asin(y / (fabs(y) + 1.0));
    out[5 * i + 3] = sin(y);
    out[5 * i + 4] = cos(y);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// Recording for the adjoint: body states at the start of substep k, and the
// substep's contacts after the narrow phase.
__global__ void k_rec_bodies(DevWorld w, Recorder r, int k) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < w.nb) r.body[(size_t)k * w.nb + b] = w.body[b];
}
__global__ void k_rec_contacts(DevWorld w, Recorder r, int k) {
    int total = w.live_prefix[w.ns];
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < total; g += gridDim.x * blockDim.x) {
        int nc = w.ncontact[g];
        r.ncontact[(size_t)k * r.pair_cap + g] = nc;
        for (int c = 0; c < nc; ++c)
            r.contacts[((size_t)k * r.pair_cap + g) * MAXC + c] = w.contacts[(size_t)g * MAXC + c];
    }
}

hipError_t launch_frame(const DevWorld& w, const FrameParams& fp, hipStream_t st, KernelTimer* kt,
                        const BigScene* bigs, int nbig, const Recorder* rec) {
    if (w.nb == 0) return hipSuccess;
    KernelTimer none;
    KernelTimer& t = kt ? *kt : none;
    t.begin(KC_FRAME, st);
    if (!fp.forces_given) hipLaunchKernelGGL(k_frame_begin, dim3(cdiv(w.nb, 256)), dim3(256), 0, st, w, fp.gravity);
    hipLaunchKernelGGL(k_broad, dim3(w.ns), dim3(64), 0, st, w);
    for (int k = 0; k < nbig; ++k) big_broad(w, bigs[k], st);
    hipLaunchKernelGGL(k_islands, dim3(w.ns), dim3(64), 0, st, w, fp.dt);
    for (int k = 0; k < nbig; ++k) big_islands(w, bigs[k], fp.dt, st);
    hipLaunchKernelGGL(k_scan_live, dim3(1), dim3(1024), 0, st, w);
    hipLaunchKernelGGL(k_scatter_live, dim3(w.ns), dim3(64), 0, st, w);
    for (int k = 0; k < nbig; ++k) big_scatter(w, bigs[k], st);
    hipLaunchKernelGGL(k_levels, dim3(w.ns), dim3(64), 0, st, w, fp.collisions ? 1 : 0);
    for (int k = 0; k < nbig; ++k) big_colour(w, bigs[k], fp.collisions, st);
    t.end(st);
    // without large scenes or recording, the solve kernel integrates its
    // scenes' bodies for the next substep
    bool fuse_integrate = nbig == 0 && !rec;
    for (unsigned sub = 0; sub < fp.substeps; ++sub) {
        if (rec) hipLaunchKernelGGL(k_rec_bodies, dim3(cdiv(w.nb, 256)), dim3(256), 0, st, w, *rec, (int)sub);
        if (sub == 0 || !fuse_integrate) {
            t.begin(KC_INTEGRATE, st);
            hipLaunchKernelGGL(k_integrate, dim3(cdiv(w.nb, 128)), dim3(128), 0, st, w, fp.h);
            t.end(st);
        }
        if (fp.keep_pose && sub + 1 == fp.substeps)
            hipLaunchKernelGGL(k_keep_pose, dim3(cdiv(w.nb, 256)), dim3(256), 0, st, w);
        if (fp.collisions) {
            bool fused = w.manifold_group == 0;
            if (w.narrow_group == 1 || w.manifold_group == 1 || w.has_general) {
                t.begin(KC_TRANSFORM, st);
                hipLaunchKernelGGL(k_transform, dim3(cdiv(w.n_xf, 256)), dim3(256), 0, st, w, 0);
                t.end(st);
            }
            t.begin(KC_NARROW, st);
            size_t lds = w.narrow_group > 1 ? (size_t)(64 / w.narrow_group) * w.stage_cap * sizeof(V3) : 0;
            dim3 nb(w.narrow_blocks), t64(64);
            if (fused && w.small_nv > 0 && w.narrow_group > 1) {
                // two size classes: small-hull pairs with small LDS slices on the
                // narrow schedule, the others on the schedule with the fewest
                // groups per block that fits
                const int32_t* small = w.pair_lists;
                const int32_t* big = w.pair_lists + w.list_stride;
                size_t lds_s = (size_t)(64 / w.narrow_group) * w.stage_small * sizeof(V3);
                switch (w.narrow_group) {
                case 4: hipLaunchKernelGGL((k_gjk_epa<4, true>), nb, t64, lds_s, st, w, small, w.counters + CNT_SMALL, w.stage_small); break;
                case 8: hipLaunchKernelGGL((k_gjk_epa<8, true>), nb, t64, lds_s, st, w, small, w.counters + CNT_SMALL, w.stage_small); break;
                default: hipLaunchKernelGGL((k_gjk_epa<16, true>), nb, t64, lds_s, st, w, small, w.counters + CNT_SMALL, w.stage_small); break;
                }
                size_t lds_b = (size_t)(64 / 16) * w.stage_cap * sizeof(V3);
                hipLaunchKernelGGL((k_gjk_epa<16, true>), nb, t64, lds_b, st, w, big, w.counters + CNT_BIG, w.stage_cap);
            } else if (fused) {
                switch (w.narrow_group) {
                case 4: hipLaunchKernelGGL((k_gjk_epa<4, true>), nb, t64, lds, st, w, nullptr, nullptr, w.stage_cap); break;
                case 8: hipLaunchKernelGGL((k_gjk_epa<8, true>), nb, t64, lds, st, w, nullptr, nullptr, w.stage_cap); break;
                case 16: hipLaunchKernelGGL((k_gjk_epa<16, true>), nb, t64, lds, st, w, nullptr, nullptr, w.stage_cap); break;
                default: hipLaunchKernelGGL(k_gjk_epa_lane<true>, nb, t64, 0, st, w); break;
                }
            } else {
                switch (w.narrow_group) {
                case 4: hipLaunchKernelGGL((k_gjk_epa<4, false>), nb, t64, lds, st, w, nullptr, nullptr, w.stage_cap); break;
                case 8: hipLaunchKernelGGL((k_gjk_epa<8, false>), nb, t64, lds, st, w, nullptr, nullptr, w.stage_cap); break;
                case 16: hipLaunchKernelGGL((k_gjk_epa<16, false>), nb, t64, lds, st, w, nullptr, nullptr, w.stage_cap); break;
                default: hipLaunchKernelGGL(k_gjk_epa_lane<false>, nb, t64, 0, st, w); break;
                }
            }
            if (w.has_general) hipLaunchKernelGGL(k_narrow_general, dim3(1024), t64, 0, st, w);
            t.end(st);
            if (!fused) {
                t.begin(KC_MANIFOLD, st);
                lds = w.manifold_group > 1 ? (size_t)(64 / w.manifold_group) * w.stage_cap * sizeof(V3) : 0;
                dim3 mb(w.manifold_blocks);
                switch (w.manifold_group) {
                case 4: hipLaunchKernelGGL(k_manifold<4>, mb, t64, lds, st, w); break;
                case 8: hipLaunchKernelGGL(k_manifold<8>, mb, t64, lds, st, w); break;
                case 16: hipLaunchKernelGGL(k_manifold<16>, mb, t64, lds, st, w); break;
                default: hipLaunchKernelGGL(k_manifold<1>, mb, t64, 0, st, w); break;
                }
                t.end(st);
            }
            if (rec) hipLaunchKernelGGL(k_rec_contacts, dim3(256), dim3(256), 0, st, w, *rec, (int)sub);
        }
        t.begin(KC_SOLVE, st);
        int next = fuse_integrate && sub + 1 < fp.substeps ? 1 : 0;
        if (w.n_ext > 0) launch_solve<true>(w, fp, st, next);
        else launch_solve<false>(w, fp, st, next);
        for (int k = 0; k < nbig; ++k) big_solve(w, bigs[k], fp, st);
        t.end(st);
    }
    return hipGetLastError();
}

hipError_t launch_pair_contacts(const DevWorld& w, int b1, int b2, double* out, int* count, uint32_t* err,
                                hipStream_t st) {
    hipLaunchKernelGGL(k_model, dim3(cdiv(w.nb, 128)), dim3(128), 0, st, w);
    hipLaunchKernelGGL(k_transform, dim3(cdiv(w.n_xf, 256)), dim3(256), 0, st, w, 1);
    hipLaunchKernelGGL(k_pair_contacts, dim3(1), dim3(64), 0, st, w, b1, b2, out, count, err);
    return hipGetLastError();
}

hipError_t launch_model_matrices(const DevWorld& w, void* out, bool f32, hipStream_t st) {
    if (w.nb == 0) return hipSuccess;
    if (f32) hipLaunchKernelGGL(k_model_out<float>, dim3(cdiv(w.nb, 256)), dim3(256), 0, st, w, (float*)out);
    else hipLaunchKernelGGL(k_model_out<double>, dim3(cdiv(w.nb, 256)), dim3(256), 0, st, w, (double*)out);
    return hipGetLastError();
}

hipError_t launch_fp64_selftest(const double* a, const double* b, double* out, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fp64_selftest, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, b, out, n);
    return hipGetLastError();
}

// Lane-group working memory of rp_world_step_noalloc: a file-scope device
// array instead of a hipMalloc'd buffer. Each device that runs the library
// holds its own copy until the process ends. Host code never passes the
// array's name to a kernel (it is a host-side stand-in there): it takes the
// device address from noalloc_scratch(), which looks it up once per device.
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

unsigned char* noalloc_scratch() {
    static std::mutex mu;
    static std::vector<unsigned char*> ptr;  // per device, null until looked up
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if ((size_t)dev >= ptr.size()) ptr.resize(dev + 1, nullptr);
    if (!ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        ptr[dev] = (unsigned char*)p;
    }
    return ptr[dev];
}

int64_t group_work_slots_needed(const DevWorld& w) {
    int64_t n = 1;  // k_pair_contacts
    if (w.narrow_group > 1) n = std::max<int64_t>(n, (int64_t)w.narrow_blocks * (64 / w.narrow_group));
    if (w.manifold_group > 1) n = std::max<int64_t>(n, (int64_t)w.manifold_blocks * (64 / w.manifold_group));
    return n;
}

bool narrow_group_supported(const DevWorld& w, int g) {
    if (g == 1) return true;
    if (g != 4 && g != 8 && g != 16) return false;
    return (size_t)(64 / g) * w.stage_cap * sizeof(V3) <= (size_t)w.lds_limit;
}

hipError_t allow_narrow_lds(int bytes) {
    const void* k[] = {(const void*)k_gjk_epa<4, true>, (const void*)k_gjk_epa<8, true>,
                       (const void*)k_gjk_epa<16, true>, (const void*)k_gjk_epa<4, false>,
                       (const void*)k_gjk_epa<8, false>, (const void*)k_gjk_epa<16, false>,
                       (const void*)k_manifold<4>, (const void*)k_manifold<8>, (const void*)k_manifold<16>};
    for (const void* f : k) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Blocks of the narrow-phase kernels: one lane group per pair slot, up to
// 8192 one-wave blocks (32 per CU). Live pairs beyond the groups are taken in
// grid-stride order; frames with more live pairs than 4096 blocks hold ran
// 5 % faster with the larger cap (batched 4096-scene workload, 600 frames).
int narrow_blocks_for(const DevWorld& w, int g, int64_t pcap) {
    int64_t per_block = 64 / g;
    int64_t blocks = (pcap + per_block - 1) / per_block;
    int64_t cap = w.narrow_block_cap > 0 ? w.narrow_block_cap : 256 * 32;
    return (int)std::max<int64_t>(1, std::min<int64_t>(blocks, cap));
}

hipError_t launch_broad(const DevWorld& w, hipStream_t st, const BigScene* bigs, int nbig) {
    hipLaunchKernelGGL(k_broad, dim3(w.ns), dim3(64), 0, st, w);
    for (int k = 0; k < nbig; ++k) big_broad(w, bigs[k], st);
    return hipGetLastError();
}

}  // namespace rp
