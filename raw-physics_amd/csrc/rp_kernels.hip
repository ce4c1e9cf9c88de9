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
#include "rp_kernels.h"

namespace rp {

static __device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Ordered wave compaction: returns this lane's output slot among lanes with
// pred set (in lane order) and the wave total in *total.
static __device__ __forceinline__ int wave_compact(bool pred, int* total) {
    unsigned long long m = __ballot(pred);
    *total = __popcll(m);
    unsigned long long below = (lane_id() == 0) ? 0ull : (m & ((~0ull) >> (64 - lane_id())));
    return __popcll(below);
}

// ---------------------------------------------------------------------------
__global__ void k_frame_begin(DevWorld w, double gravity) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= w.nb) return;
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
    int lane = threadIdx.x;
    int b0 = w.scene_body_off[s], n = w.scene_body_off[s + 1] - b0;
    int base = w.scene_pair_base[s];
    int count = 0;
    for (int i = lane; i < n; i += 64) w.need[b0 + i] = 0;
    for (int i = 0; i < n; ++i) {
        const Body& e1 = w.body[b0 + i];
        V3 x1 = e1.x;
        double r1 = e1.radius;
        for (int j0 = i + 1; j0 < n; j0 += 64) {
            int j = j0 + lane;
            bool hit = false;
            if (j < n) {
                const Body& e2 = w.body[b0 + j];
                double dd = len(sub(x1, e2.x));
                double maxd = r1 + e2.radius + 0.1;
                hit = dd <= maxd;
            }
            int total;
            int slot = wave_compact(hit, &total);
            if (hit) w.pairs[base + count + slot] = make_int2(i, j);
            count += total;
        }
    }
    if (lane == 0) w.pair_count[s] = count;
}

static __device__ int uf_find(int* parent, int x) {
    while (parent[x] != x) x = parent[x];
    return x;
}

// Islands and sleeping (broad.cpp:31-116, pbd.cpp:476-506), then the list of
// pairs the narrow phase visits (pbd.cpp:593-596).
__global__ void k_islands(DevWorld w, double dt) {
    int s = blockIdx.x;
    int lane = threadIdx.x;
    int b0 = w.scene_body_off[s], n = w.scene_body_off[s + 1] - b0;
    int base = w.scene_pair_base[s];
    int np = w.pair_count[s];
    int* parent = w.uf_parent + b0;  // scene-local indices
    int* alli = w.uf_flag + b0;
    if (lane == 0) {
        for (int i = 0; i < n; ++i) { parent[i] = i; alli[i] = 1; }
        // uf_union (broad.cpp:55-59): parent[find(y)] = find(x)
        for (int k = 0; k < np; ++k) {
            int2 p = w.pairs[base + k];
            if (!w.body[b0 + p.x].fixed && !w.body[b0 + p.y].fixed) {
                int key = uf_find(parent, p.y), value = uf_find(parent, p.x);
                parent[key] = value;
            }
        }
        for (int k = w.scene_ext_off[s]; k < w.scene_ext_off[s + 1]; ++k) {
            int i = w.ext[k].b1 - b0, j = w.ext[k].b2 - b0;
            if (!w.body[b0 + i].fixed && !w.body[b0 + j].fixed) {
                int key = uf_find(parent, j), value = uf_find(parent, i);
                parent[key] = value;
            }
        }
        for (int i = 0; i < n; ++i) {
            Body& e = w.body[b0 + i];
            if (e.fixed) continue;
            double lv = len(e.v), av = len(e.w);
            if (lv < 0.10 && av < 0.10) e.deact += dt;
            else e.deact = 0.0;
            if (e.deact < 1.0) alli[uf_find(parent, i)] = 0;
        }
        for (int i = 0; i < n; ++i) {
            Body& e = w.body[b0 + i];
            if (e.fixed) continue;
            e.active = !alli[uf_find(parent, i)];
        }
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
    int base = w.scene_pair_base[s];
    int g0 = w.live_prefix[s], n = w.live_count[s];
    int b0 = w.scene_body_off[s];
    for (int m = threadIdx.x; m < n; m += blockDim.x) {
        int slot = base + w.live_local[base + m];
        w.live_global[g0 + m] = make_int2(s, slot);
        int2 p = w.pairs[slot];
        w.need[b0 + p.x] = 1;  // same value from every writer
        w.need[b0 + p.y] = 1;
    }
}

// ---------------------------------------------------------------------------
__global__ void k_integrate(DevWorld w, double h) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= w.nb) return;
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

// collider_update for the vertices (collider.cpp:409-421): M*[v,1], then
// scaled by 1/w. Only bodies that take part in a live pair this frame.
__global__ void k_transform(DevWorld w, int all) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= w.n_xf) return;
    int2 el = w.xf[t];
    int b = el.x;
    if (!all && !w.need[b]) return;
    const M4& M = w.model[b];
    int hid = w.body_hull[b];
    int i = el.y;
    V3 v = w.hull_v[w.hull_v_off[hid] + i];
    double x = v.x * M.m[0][0] + v.y * M.m[0][1] + v.z * M.m[0][2] + 1.0 * M.m[0][3];
    double y = v.x * M.m[1][0] + v.y * M.m[1][1] + v.z * M.m[1][2] + 1.0 * M.m[1][3];
    double z = v.x * M.m[2][0] + v.y * M.m[2][1] + v.z * M.m[2][2] + 1.0 * M.m[2][3];
    double ww = v.x * M.m[3][0] + v.y * M.m[3][1] + v.z * M.m[3][2] + 1.0 * M.m[3][3];
    double sc = 1.0 / ww;
    w.tv[w.body_tv_off[b] + i] = v3(sc * x, sc * y, sc * z);
}

__device__ __forceinline__ HullView hull_view(const DevWorld& w, int b) {
    HullView hv;
    int hid = w.body_hull[b];
    int tp = w.hull_topo[hid];
    const int32_t* tb = w.topo_base + 8 * tp;
    hv.tv = w.tv + w.body_tv_off[b];
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

// Narrow phase over the live pairs of all scenes (pbd.cpp:583-612). Each
// lane owns one pair; results go to the pair's slot in the live list, which
// is in scene then pair order.
__global__ void __launch_bounds__(64) k_narrow(DevWorld w) {
    int total = w.live_prefix[w.ns];
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < total; g += gridDim.x * blockDim.x) {
        int2 sp = w.live_global[g];
        int s = sp.x;
        int2 p = w.pairs[sp.y];
        int b0 = w.scene_body_off[s];
        int b1 = b0 + p.x, b2 = b0 + p.y;
        HullView h1 = hull_view(w, b1), h2 = hull_view(w, b2);
        uint32_t err = 0;
        int epa_fail = 0;
        ContactPt cts[MAXC];
        int nc = pair_contacts(h1, h2, cts, err, epa_fail);
        const Body& e1 = w.body[b1];
        const Body& e2 = w.body[b2];
        Contact* out = w.contacts + (size_t)g * MAXC;
        for (int k = 0; k < nc; ++k) contact_to_constraint(e1, e2, cts[k], out[k]);
        w.ncontact[g] = nc;
        if (err) atomicOr(&w.scene_err[s], err);
        if (epa_fail) atomicAdd(&w.stats[STAT_EPA_FAIL], 1ull);
    }
}

// Solve of one scene: position Gauss-Seidel (pbd.cpp:615-620), velocity
// update (pbd.cpp:623-643), velocity solve (pbd.cpp:646-711).
__global__ void __launch_bounds__(64) k_solve(DevWorld w, double h, int iters, int collisions) {
    int s = blockIdx.x;
    int lane = threadIdx.x;
    int b0 = w.scene_body_off[s], n = w.scene_body_off[s + 1] - b0;
    int e0 = w.scene_ext_off[s], e1x = w.scene_ext_off[s + 1];
    int next = e1x - e0;
    int g0 = collisions ? w.live_prefix[s] : 0;
    int g1 = collisions ? w.live_prefix[s] + w.live_count[s] : 0;
    int cb = w.scene_con_base[s];
    int ccap = w.scene_con_base[s + 1] - cb;
    int* cref = w.con_ref + cb;     // constraint sequence
    int* clev = w.con_level + cb;   // level of each constraint
    int* order = w.con_order + cb;  // constraints sorted by level
    int* lstart = w.con_lstart + cb + s;  // level starts (ccap + 1 entries)
    int* last = w.body_last + b0;
    __shared__ int sh_K, sh_L;
    uint32_t err = 0;

    // multipliers of the external constraints start at zero (pbd.cpp:426-462)
    for (int k = e0 + lane; k < e1x; k += 64) {
        w.ext_lambda[3 * k] = 0.0;
        w.ext_lambda[3 * k + 1] = 0.0;
        w.ext_lambda[3 * k + 2] = 0.0;
    }
    for (int i = lane; i < n; i += 64) last[i] = 0;
    __syncthreads();
    if (lane == 0) {
        // constraint sequence: external constraints, then contacts in pair order
        int K = 0;
        bool overflow = false;
        for (int k = 0; k < next; ++k) {
            if (K < ccap) cref[K++] = ~(e0 + k); else overflow = true;
        }
        for (int g = g0; g < g1; ++g) {
            int nc = w.ncontact[g];
            for (int c = 0; c < nc; ++c) {
                if (K < ccap) cref[K++] = g * MAXC + c; else overflow = true;
            }
        }
        if (overflow) err |= ERR_SOLVER_CAPACITY;
        // levels: one more than the last level of either writable body
        int L = 0;
        for (int t = 0; t < K; ++t) {
            int r = cref[t];
            int i1, i2;
            if (r < 0) { i1 = w.ext[~r].b1 - b0; i2 = w.ext[~r].b2 - b0; }
            else { int2 p = w.pairs[w.live_global[r / MAXC].y]; i1 = p.x; i2 = p.y; }
            bool f1 = w.body[b0 + i1].fixed, f2 = w.body[b0 + i2].fixed;
            int lv = 0;
            if (!f1) lv = last[i1];
            if (!f2) lv = max(lv, last[i2]);
            ++lv;
            if (!f1) last[i1] = lv;
            if (!f2) last[i2] = lv;
            clev[t] = lv - 1;
            L = max(L, lv);
        }
        // counting sort by level (stable)
        for (int l = 0; l <= L; ++l) lstart[l] = 0;
        for (int t = 0; t < K; ++t) ++lstart[clev[t] + 1];
        for (int l = 0; l < L; ++l) lstart[l + 1] += lstart[l];
        for (int t = 0; t < K; ++t) order[lstart[clev[t]]++] = t;
        for (int l = L; l > 0; --l) lstart[l] = lstart[l - 1];
        lstart[0] = 0;
        sh_K = K;
        sh_L = L;
    }
    __syncthreads();
    int K = sh_K, L = sh_L;

    for (int it = 0; it < iters; ++it) {
        for (int l = 0; l < L; ++l) {
            for (int idx = lstart[l] + lane; idx < lstart[l + 1]; idx += 64) {
                int r = cref[order[idx]];
                int bi1, bi2;
                if (r < 0) { bi1 = w.ext[~r].b1; bi2 = w.ext[~r].b2; }
                else { int2 p = w.pairs[w.live_global[r / MAXC].y]; bi1 = b0 + p.x; bi2 = b0 + p.y; }
                Body B1 = w.body[bi1];
                Body B2 = w.body[bi2];
                if (r < 0) {
                    solve_external(w.ext[~r], w.ext_lambda + 3 * (~r), B1, B2, h, err);
                } else {
                    Contact c = w.contacts[r];
                    solve_collision(c, B1, B2, h, err);
                    w.contacts[r].lambda_n = c.lambda_n;
                    w.contacts[r].lambda_t = c.lambda_t;
                }
                if (!B1.fixed) { w.body[bi1].x = B1.x; w.body[bi1].q = B1.q; }
                if (!B2.fixed) { w.body[bi2].x = B2.x; w.body[bi2].q = B2.q; }
            }
            __syncthreads();
        }
    }
    // velocity update
    for (int i = lane; i < n; i += 64) {
        Body e = w.body[b0 + i];
        if (e.fixed || !e.active) continue;
        velocity_update(e, h);
        w.body[b0 + i].pv = e.pv;
        w.body[b0 + i].pw = e.pw;
        w.body[b0 + i].v = e.v;
        w.body[b0 + i].w = e.w;
    }
    __syncthreads();
    // velocity solve over collision constraints in the same levels
    for (int l = 0; l < L; ++l) {
        for (int idx = lstart[l] + lane; idx < lstart[l + 1]; idx += 64) {
            int r = cref[order[idx]];
            if (r < 0) continue;
            int2 p = w.pairs[w.live_global[r / MAXC].y];
            int bi1 = b0 + p.x, bi2 = b0 + p.y;
            Body B1 = w.body[bi1];
            Body B2 = w.body[bi2];
            Contact c = w.contacts[r];
            velocity_solve(c, B1, B2, h);
            if (!B1.fixed) { w.body[bi1].v = B1.v; w.body[bi1].w = B1.w; }
            if (!B2.fixed) { w.body[bi2].v = B2.v; w.body[bi2].w = B2.w; }
        }
        __syncthreads();
    }
    (void)K;
    if (err) atomicOr(&w.scene_err[s], err);
}

// Contacts of one body pair at the current transformed poses (debug / parity).
__global__ void k_pair_contacts(DevWorld w, int b1, int b2, double* out, int* count, uint32_t* err_out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    HullView h1 = hull_view(w, b1), h2 = hull_view(w, b2);
    uint32_t err = 0;
    int epa_fail = 0;
    ContactPt cts[MAXC];
    int nc = pair_contacts(h1, h2, cts, err, epa_fail);
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

hipError_t launch_frame(const DevWorld& w, const FrameParams& fp, hipStream_t st, KernelTimer* kt) {
    if (w.nb == 0) return hipSuccess;
    KernelTimer none;
    KernelTimer& t = kt ? *kt : none;
    t.begin(KC_FRAME, st);
    if (!fp.forces_given) hipLaunchKernelGGL(k_frame_begin, dim3(cdiv(w.nb, 256)), dim3(256), 0, st, w, fp.gravity);
    hipLaunchKernelGGL(k_broad, dim3(w.ns), dim3(64), 0, st, w);
    hipLaunchKernelGGL(k_islands, dim3(w.ns), dim3(64), 0, st, w, fp.dt);
    hipLaunchKernelGGL(k_scan_live, dim3(1), dim3(1024), 0, st, w);
    hipLaunchKernelGGL(k_scatter_live, dim3(w.ns), dim3(64), 0, st, w);
    t.end(st);
    for (unsigned sub = 0; sub < fp.substeps; ++sub) {
        t.begin(KC_INTEGRATE, st);
        hipLaunchKernelGGL(k_integrate, dim3(cdiv(w.nb, 128)), dim3(128), 0, st, w, fp.h);
        t.end(st);
        if (fp.collisions) {
            t.begin(KC_TRANSFORM, st);
            hipLaunchKernelGGL(k_transform, dim3(cdiv(w.n_xf, 256)), dim3(256), 0, st, w, 0);
            t.end(st);
            t.begin(KC_NARROW, st);
            hipLaunchKernelGGL(k_narrow, dim3(w.narrow_blocks), dim3(64), 0, st, w);
            t.end(st);
        }
        t.begin(KC_SOLVE, st);
        hipLaunchKernelGGL(k_solve, dim3(w.ns), dim3(64), 0, st, w, fp.h, (int)fp.iters, fp.collisions ? 1 : 0);
        t.end(st);
    }
    return hipGetLastError();
}

hipError_t launch_pair_contacts(const DevWorld& w, int b1, int b2, double* out, int* count, uint32_t* err,
                                hipStream_t st) {
    hipLaunchKernelGGL(k_model, dim3(cdiv(w.nb, 128)), dim3(128), 0, st, w);
    hipLaunchKernelGGL(k_transform, dim3(cdiv(w.n_xf, 256)), dim3(256), 0, st, w, 1);
    hipLaunchKernelGGL(k_pair_contacts, dim3(1), dim3(1), 0, st, w, b1, b2, out, count, err);
    return hipGetLastError();
}

hipError_t launch_fp64_selftest(const double* a, const double* b, double* out, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fp64_selftest, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, b, out, n);
    return hipGetLastError();
}

hipError_t launch_broad(const DevWorld& w, hipStream_t st) {
    hipLaunchKernelGGL(k_broad, dim3(w.ns), dim3(64), 0, st, w);
    return hipGetLastError();
}

}  // namespace rp
