// rp_large.hip — the frame of a large scene (thousands to 10^5+ bodies).
//
// The per-scene wave kernels (k_broad, k_islands, k_levels, k_solve) serialise
// a scene on one wave, which a 10^5-body scene cannot afford. A large scene
// instead runs:
//   broad phase   uniform hash grid over the small bodies, waves over chunks
//                 of the later bodies for each large body, counting sort and
//                 per-body sorted segments: the pair
//                 list equals the all-pairs test of broad.cpp:6-29 in set and
//                 order (i-major, j ascending)
//   islands       lock-free union-find over pairs and constraints, then the
//                 sleep rule of pbd.cpp:476-506 (identical partition)
//   live pairs    ordered compaction by a device-wide scan
//   solve         items (external constraints, live pairs) coloured once per
//                 frame (Jones-Plassmann on the item conflict graph: items
//                 sharing a non-fixed body), then solved colour by colour
//                 with one lane per item (the first colours one launch each,
//                 the small rest on one block). This replaces the reference's
//                 sequential Gauss-Seidel order, so results differ from it
//                 within a stated tolerance; they do not depend on thread
//                 timing (colours and the serial class are deterministic, and
//                 items of one colour touch disjoint writable bodies).
#include "rp_device.h"

namespace rp {

// ---------------------------------------------------------------------------
// Device-wide exclusive scan of int32 (in place allowed): out[n] = total.
// ---------------------------------------------------------------------------
constexpr int SCAN_T = 256;           // threads per block
constexpr int SCAN_PER = 4;           // elements per thread
constexpr int SCAN_B = SCAN_T * SCAN_PER;

int scan_block_count(int64_t n) { return (int)((n + SCAN_B - 1) / SCAN_B) + 1; }

__device__ __forceinline__ int block_exclusive_scan(int v, int* sh, int* total) {
    int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_T; off <<= 1) {
        int x = (t >= off) ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += x;
        __syncthreads();
    }
    int incl = sh[t];
    *total = sh[SCAN_T - 1];
    __syncthreads();
    return incl - v;
}

__global__ void __launch_bounds__(SCAN_T) k_scan_local(const int32_t* in, int32_t* out, int n, int32_t* sums) {
    __shared__ int sh[SCAN_T];
    int base = blockIdx.x * SCAN_B + threadIdx.x * SCAN_PER;
    int v[SCAN_PER];
    int s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        s += v[k];
    }
    int total;
    int run = block_exclusive_scan(s, sh, &total);
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Scans the block sums (one block, any count) and writes out[n] = total.
__global__ void __launch_bounds__(SCAN_T) k_scan_sums(int32_t* sums, int nb, int32_t* out, int n) {
    __shared__ int sh[SCAN_T];
    int chunk = (nb + SCAN_T - 1) / SCAN_T;
    int lo = min(nb, (int)threadIdx.x * chunk), hi = min(nb, lo + chunk);
    int s = 0;
    for (int i = lo; i < hi; ++i) s += sums[i];
    int total;
    int run = block_exclusive_scan(s, sh, &total);
    for (int i = lo; i < hi; ++i) {
        int x = sums[i];
        sums[i] = run;
        run += x;
    }
    if (threadIdx.x == 0) out[n] = total;
}

__global__ void __launch_bounds__(SCAN_T) k_scan_add(int32_t* out, int n, const int32_t* sums) {
    int base = blockIdx.x * SCAN_B + threadIdx.x * SCAN_PER;
    int add = sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k)
        if (base + k < n) out[base + k] += add;
}

// in[0..n) -> out[0..n] (exclusive prefix, out[n] = total); `sums` holds
// scan_block_count(n) entries.
static void scan(const int32_t* in, int32_t* out, int n, int32_t* sums, hipStream_t st) {
    int nb = (n + SCAN_B - 1) / SCAN_B;
    if (nb == 0) nb = 1;
    hipLaunchKernelGGL(k_scan_local, dim3(nb), dim3(SCAN_T), 0, st, in, out, n, sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_T), 0, st, sums, nb, out, n);
    hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(SCAN_T), 0, st, out, n, sums);
}

static inline int grid_for(int64_t n, int t) {
    int64_t b = (n + t - 1) / t;
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, 4096));
}

__global__ void k_fill(int32_t* p, int n, int v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_copy(const int32_t* a, int32_t* b, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b[i] = a[i];
}

// ---------------------------------------------------------------------------
// Broad phase
// ---------------------------------------------------------------------------
struct Cell { int x, y, z; };

__device__ __forceinline__ Cell cell_of(V3 p, double c) {
    // clamped so that far-away bodies still hash (they only test against
    // bodies in the same clamped cell range, and far bodies cannot pair)
    double lim = 1.0e9;
    Cell k;
    k.x = (int)floor(fmin(fmax(p.x / c, -lim), lim));
    k.y = (int)floor(fmin(fmax(p.y / c, -lim), lim));
    k.z = (int)floor(fmin(fmax(p.z / c, -lim), lim));
    return k;
}
__device__ __forceinline__ int cell_hash(int x, int y, int z, int T) {
    uint32_t h = (uint32_t)x * 73856093u ^ (uint32_t)y * 19349663u ^ (uint32_t)z * 83492791u;
    return (int)(h & (uint32_t)(T - 1));
}

// Per body: bucket counts of the small bodies; clears the `need` flags.
__global__ void k_big_count(DevWorld w, BigScene b) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < b.n; i += gridDim.x * blockDim.x) {
        w.need[b.b0 + i] = 0;
        if (b.is_large[i]) continue;
        Cell c = cell_of(w.body[b.b0 + i].x, b.cell);
        atomicAdd(&b.cell_cursor[cell_hash(c.x, c.y, c.z, b.T)], 1);
    }
}

__global__ void k_big_fill(DevWorld w, BigScene b) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < b.n; i += gridDim.x * blockDim.x) {
        if (b.is_large[i]) continue;
        const Body& e = w.body[b.b0 + i];
        Cell c = cell_of(e.x, b.cell);
        int pos = atomicAdd(&b.cell_cursor[cell_hash(c.x, c.y, c.z, b.T)], 1);
        b.sorted[pos] = i;
        b.spos[pos] = make_double4(e.x.x, e.x.y, e.x.z, e.radius);
    }
}

// Pairs (i, j > i) of small body i: small bodies in the 27 neighbouring
// cells (each bucket visited once) and every large body. pass 0 counts into
// pair_off[i]; pass 1 writes at pair_off[i] and sorts the segment by j.
// Bodies are visited in bucket order, so that neighbouring lanes search the
// same cells (each body's segment is its own, whatever the visiting order).
template <int PASS>
__global__ void k_big_pairs_small(DevWorld w, BigScene b) {
    int nsmall = b.cell_start[b.T];
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nsmall; t += gridDim.x * blockDim.x) {
        int i = b.sorted[t];
        double4 p1 = b.spos[t];
        const V3 x1 = v3(p1.x, p1.y, p1.z);
        const double r1 = p1.w;
        Cell c = cell_of(x1, b.cell);
        int keys[27];
        int nk = 0, cnt = 0;
        int off = 0;
        if (PASS == 1) off = b.pair_off[i];
        int2* out = w.pairs + b.pair_base;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int key = cell_hash(c.x + dx, c.y + dy, c.z + dz, b.T);
                    bool seen = false;
                    for (int k = 0; k < nk; ++k) seen |= keys[k] == key;
                    if (seen) continue;
                    keys[nk++] = key;
                    for (int m = b.cell_start[key]; m < b.cell_start[key + 1]; ++m) {
                        int j = b.sorted[m];
                        double4 q = b.spos[m];  // the bucket's bodies, contiguous
                        if (j <= i || !broad_hit(x1, r1, v3(q.x, q.y, q.z), q.w)) continue;
                        if (PASS == 1 && off + cnt < b.pair_cap) out[off + cnt] = make_int2(i, j);
                        ++cnt;
                    }
                }
        for (int k = 0; k < b.n_large; ++k) {
            int j = b.large[k];
            const Body& e2 = w.body[b.b0 + j];
            if (j <= i || !broad_hit(x1, r1, e2.x, e2.radius)) continue;
            if (PASS == 1 && off + cnt < b.pair_cap) out[off + cnt] = make_int2(i, j);
            ++cnt;
        }
        if (PASS == 0) {
            b.pair_off[i] = cnt;
        } else {
            // insertion sort of the segment by j (segments are short)
            int end = min(off + cnt, b.pair_cap);
            for (int a = off + 1; a < end; ++a) {
                int2 v = out[a];
                int q = a - 1;
                while (q >= off && out[q].y > v.y) {
                    out[q + 1] = out[q];
                    --q;
                }
                out[q + 1] = v;
            }
        }
    }
}

// Pairs (l, k > l) of large body l: the later bodies in chunks of
// LARGE_CHUNK, one wave per (large body, chunk) scanning its chunk in order.
// PASS 0 counts each chunk; the counts are scanned (per large body: the
// difference from its first chunk) and PASS 1 writes each chunk's hits at
// its offset, so the segment comes out sorted.
template <int PASS>
__global__ void __launch_bounds__(64) k_big_pairs_large(DevWorld w, BigScene b) {
    int li = blockIdx.x / b.large_chunks, c = blockIdx.x % b.large_chunks;
    int l = b.large[li];
    const Body& e1 = w.body[b.b0 + l];
    int cnt = 0;
    int off = 0;
    if (PASS == 1) {
        const int32_t* row = b.large_cnt + (size_t)li * b.large_chunks;
        off = b.pair_off[l] + row[c] - row[0];
    }
    int2* out = w.pairs + b.pair_base;
    int k0 = l + 1 + c * LARGE_CHUNK;
    int k1 = min(b.n, k0 + LARGE_CHUNK);
    for (int kb = k0; kb < k1; kb += 64) {
        int k = kb + (int)threadIdx.x;
        bool hit = k < k1 && broad_hit(e1, w.body[b.b0 + k]);
        int total;
        int slot = wave_compact(hit, &total);
        if (PASS == 1 && hit && off + cnt + slot < b.pair_cap) out[off + cnt + slot] = make_int2(l, k);
        cnt += total;
    }
    if (PASS == 0 && threadIdx.x == 0) b.large_cnt[blockIdx.x] = cnt;
}

// Segment lengths of the large bodies from their scanned chunk counts.
__global__ void k_big_large_total(DevWorld w, BigScene b) {
    for (int li = blockIdx.x * blockDim.x + threadIdx.x; li < b.n_large; li += gridDim.x * blockDim.x) {
        const int32_t* row = b.large_cnt + (size_t)li * b.large_chunks;
        b.pair_off[b.large[li]] = row[b.large_chunks] - row[0];
    }
}

__global__ void k_big_pair_count(DevWorld w, BigScene b) {
    int total = b.pair_off[b.n];
    w.pair_count[b.s] = min(total, b.pair_cap);
    if (total > b.pair_cap) atomicOr(&w.scene_err[b.s], (uint32_t)ERR_PAIR_CAPACITY);
}

void big_broad(const DevWorld& w, const BigScene& b, hipStream_t st) {
    int gb = grid_for(b.n, 256);
    hipLaunchKernelGGL(k_fill, dim3(grid_for(b.T, 256)), dim3(256), 0, st, b.cell_cursor, b.T, 0);
    hipLaunchKernelGGL(k_big_count, dim3(gb), dim3(256), 0, st, w, b);
    scan(b.cell_cursor, b.cell_start, b.T, b.scan_blocks, st);
    hipLaunchKernelGGL(k_copy, dim3(grid_for(b.T, 256)), dim3(256), 0, st, b.cell_start, b.cell_cursor, b.T);
    hipLaunchKernelGGL(k_big_fill, dim3(gb), dim3(256), 0, st, w, b);
    hipLaunchKernelGGL(k_big_pairs_small<0>, dim3(gb), dim3(256), 0, st, w, b);
    int nlc = b.n_large * b.large_chunks;
    if (b.n_large > 0) {
        hipLaunchKernelGGL(k_big_pairs_large<0>, dim3(nlc), dim3(64), 0, st, w, b);
        scan(b.large_cnt, b.large_cnt, nlc, b.scan_blocks, st);
        hipLaunchKernelGGL(k_big_large_total, dim3(grid_for(b.n_large, 256)), dim3(256), 0, st, w, b);
    }
    scan(b.pair_off, b.pair_off, b.n, b.scan_blocks, st);
    hipLaunchKernelGGL(k_big_pairs_small<1>, dim3(gb), dim3(256), 0, st, w, b);
    if (b.n_large > 0) hipLaunchKernelGGL(k_big_pairs_large<1>, dim3(nlc), dim3(64), 0, st, w, b);
    hipLaunchKernelGGL(k_big_pair_count, dim3(1), dim3(1), 0, st, w, b);
}

// ---------------------------------------------------------------------------
// Islands and sleeping
// ---------------------------------------------------------------------------
__global__ void k_big_uf_init(DevWorld w, BigScene b) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < b.n; i += gridDim.x * blockDim.x) {
        w.uf_parent[b.b0 + i] = i;
        w.uf_flag[b.b0 + i] = 1;
    }
}

// broad_collect_simulation_islands (broad.cpp:31-116): pairs and external
// constraints between non-fixed bodies join islands.
__global__ void k_big_uf_union(DevWorld w, BigScene b) {
    int np = w.pair_count[b.s];
    int32_t* parent = w.uf_parent + b.b0;
    const int2* pairs = w.pairs + b.pair_base;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < np; k += gridDim.x * blockDim.x) {
        int2 p = pairs[k];
        if (!w.body[b.b0 + p.x].fixed && !w.body[b.b0 + p.y].fixed) uf_link(parent, p.x, p.y);
    }
    int e0 = w.scene_ext_off[b.s], e1 = w.scene_ext_off[b.s + 1];
    for (int k = e0 + blockIdx.x * blockDim.x + threadIdx.x; k < e1; k += gridDim.x * blockDim.x) {
        int i = w.ext[k].b1 - b.b0, j = w.ext[k].b2 - b.b0;
        if (!w.body[b.b0 + i].fixed && !w.body[b.b0 + j].fixed) uf_link(parent, i, j);
    }
}

// Sleep timers (pbd.cpp:483-495) and the island's all-asleep flag.
__global__ void k_big_sleep(DevWorld w, BigScene b, double dt) {
    int32_t* parent = w.uf_parent + b.b0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < b.n; i += gridDim.x * blockDim.x) {
        Body& e = w.body[b.b0 + i];
        if (e.fixed) continue;
        double lv = len(e.v), av = len(e.w);
        if (lv < 0.10 && av < 0.10) e.deact += dt;
        else e.deact = 0.0;
        if (e.deact < 1.0) w.uf_flag[b.b0 + uf_root(parent, i)] = 0;
    }
}

__global__ void k_big_active(DevWorld w, BigScene b) {
    int32_t* parent = w.uf_parent + b.b0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < b.n; i += gridDim.x * blockDim.x) {
        Body& e = w.body[b.b0 + i];
        if (e.fixed) continue;
        e.active = !w.uf_flag[b.b0 + uf_root(parent, i)];
    }
}

// Live pairs (pbd.cpp:593-596), flags then ordered scatter.
__global__ void k_big_live_flags(DevWorld w, BigScene b) {
    int np = w.pair_count[b.s];
    const int2* pairs = w.pairs + b.pair_base;
    uint32_t err = 0;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < np; k += gridDim.x * blockDim.x) {
        int2 p = pairs[k];
        const Body& e1 = w.body[b.b0 + p.x];
        const Body& e2 = w.body[b.b0 + p.y];
        if (!e1.fixed && !e2.fixed && !((e1.active && e2.active) || (!e1.active && !e2.active))) err |= ERR_ACTIVITY;
        b.tmp[k] = !((e1.fixed || !e1.active) && (e2.fixed || !e2.active));
    }
    if (err) atomicOr(&w.scene_err[b.s], err);
}

__global__ void k_big_live_write(DevWorld w, BigScene b) {
    int np = w.pair_count[b.s];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < np; k += gridDim.x * blockDim.x)
        if (b.tmp[k + 1] != b.tmp[k]) w.live_local[b.pair_base + b.tmp[k]] = k;
    if (blockIdx.x == 0 && threadIdx.x == 0) w.live_count[b.s] = b.tmp[np];
}

void big_islands(const DevWorld& w, const BigScene& b, double dt, hipStream_t st) {
    int gb = grid_for(b.n, 256), gp = grid_for(b.pair_cap, 256);
    hipLaunchKernelGGL(k_big_uf_init, dim3(gb), dim3(256), 0, st, w, b);
    hipLaunchKernelGGL(k_big_uf_union, dim3(gp), dim3(256), 0, st, w, b);
    hipLaunchKernelGGL(k_big_sleep, dim3(gb), dim3(256), 0, st, w, b, dt);
    hipLaunchKernelGGL(k_big_active, dim3(gb), dim3(256), 0, st, w, b);
    // live flags over the capacity (entries past pair_count are zero), then scan
    hipLaunchKernelGGL(k_fill, dim3(gp), dim3(256), 0, st, b.tmp, b.pair_cap, 0);
    hipLaunchKernelGGL(k_big_live_flags, dim3(gp), dim3(256), 0, st, w, b);
    scan(b.tmp, b.tmp, b.pair_cap, b.scan_blocks, st);
    hipLaunchKernelGGL(k_big_live_write, dim3(gp), dim3(256), 0, st, w, b);
}

// Global live list and transform flags of the scene (after k_scan_live).
__global__ void k_big_scatter(DevWorld w, BigScene b) {
    int g0 = w.live_prefix[b.s], n = w.live_count[b.s];
    // whole waves iterate together (the size-class lists take one atomic per wave)
    int stride = gridDim.x * blockDim.x;
    for (int m0 = blockIdx.x * blockDim.x; m0 < n; m0 += stride) {
        int m = m0 + threadIdx.x;
        bool valid = m < n;
        int b1 = 0, b2 = 0;
        if (valid) {
            int slot = b.pair_base + w.live_local[b.pair_base + m];
            w.live_global[g0 + m] = make_int2(b.s, slot);
            int2 p = w.pairs[slot];
            b1 = b.b0 + p.x;
            b2 = b.b0 + p.y;
            write_live_records(w, g0 + m, b1, b2);
            w.need[b1] = 1;  // same value from every writer
            w.need[b2] = 1;
        }
        if (w.small_nv > 0) classify_pair(w, valid, g0 + m, b1, b2);
    }
}

void big_scatter(const DevWorld& w, const BigScene& b, hipStream_t st) {
    hipLaunchKernelGGL(k_big_scatter, dim3(grid_for(b.pair_cap, 256)), dim3(256), 0, st, w, b);
}

// ---------------------------------------------------------------------------
// Colouring of the frame's items
// ---------------------------------------------------------------------------
// Items t = 0..K-1: external constraints of the scene, then its live pairs
// (when collisions are on). item_level holds the colour, item_order the
// items grouped by colour, item_lstart the class starts.
__device__ __forceinline__ int big_items(const DevWorld& w, const BigScene& b, int collisions) {
    int next = w.scene_ext_off[b.s + 1] - w.scene_ext_off[b.s];
    return next + (collisions ? w.live_count[b.s] : 0);
}
// Scene-local non-fixed bodies of item t (-1 for a fixed one).
__device__ __forceinline__ int2 item_free_bodies(const DevWorld& w, const BigScene& b, int t) {
    int e0 = w.scene_ext_off[b.s];
    int next = w.scene_ext_off[b.s + 1] - e0;
    int i1, i2;
    if (t < next) {
        i1 = w.ext[e0 + t].b1 - b.b0;
        i2 = w.ext[e0 + t].b2 - b.b0;
    } else {
        int2 p = w.pairs[w.live_global[w.live_prefix[b.s] + t - next].y];
        i1 = p.x;
        i2 = p.y;
    }
    if (w.body[b.b0 + i1].fixed) i1 = -1;
    if (w.body[b.b0 + i2].fixed) i2 = -1;
    return make_int2(i1, i2);
}
// Encoded item as the solve kernels read it.
__device__ __forceinline__ int item_code(const DevWorld& w, const BigScene& b, int t) {
    int e0 = w.scene_ext_off[b.s];
    int next = w.scene_ext_off[b.s + 1] - e0;
    return (t < next) ? ~(e0 + t) : (w.live_prefix[b.s] + t - next);
}
__device__ __forceinline__ uint32_t item_priority(int t) {
    uint32_t x = (uint32_t)t * 0x9E3779B1u;
    x ^= x >> 16;
    x *= 0x85EBCA6Bu;
    x ^= x >> 13;
    return x;
}

__global__ void k_big_csr_count(DevWorld w, BigScene b, int collisions) {
    int K = big_items(w, b, collisions);
    if (K > b.item_cap) K = b.item_cap;
    int* colour = w.item_level + b.item_base;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < K; t += gridDim.x * blockDim.x) {
        int2 bb = item_free_bodies(w, b, t);
        if (bb.x >= 0) atomicAdd(&b.bi_cursor[bb.x], 1);
        if (bb.y >= 0 && bb.y != bb.x) atomicAdd(&b.bi_cursor[bb.y], 1);
        colour[t] = -1;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *b.remaining = K;
        if (big_items(w, b, collisions) > b.item_cap) atomicOr(&w.scene_err[b.s], (uint32_t)ERR_SOLVER_CAPACITY);
    }
}

__global__ void k_big_csr_fill(DevWorld w, BigScene b, int collisions) {
    int K = min(big_items(w, b, collisions), b.item_cap);
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < K; t += gridDim.x * blockDim.x) {
        int2 bb = item_free_bodies(w, b, t);
        if (bb.x >= 0) b.bi_items[atomicAdd(&b.bi_cursor[bb.x], 1)] = t;
        if (bb.y >= 0 && bb.y != bb.x) b.bi_items[atomicAdd(&b.bi_cursor[bb.y], 1)] = t;
    }
}

// One Jones-Plassmann round over the previous round's colours `src`
// (-1 = uncoloured), written with this round's to `dst`: an uncoloured item
// whose priority beats every uncoloured neighbour takes the smallest colour
// no coloured neighbour has. Reads see only the previous round, so the
// colouring does not depend on thread timing, and two neighbours never
// colour in the same round (one of them has the lower priority).
__global__ void k_big_colour_round(DevWorld w, BigScene b, int collisions, const int* __restrict__ src,
                                   int* __restrict__ dst) {
    int K = min(big_items(w, b, collisions), b.item_cap);
    // all coloured: copy. (The count falls during this launch too, but it
    // reaches 0 only once every item is coloured in src or by a block that
    // has run, so a block that sees 0 late copies what it would have kept.)
    bool done = __hip_atomic_load(b.remaining, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < K; t += gridDim.x * blockDim.x) {
        int ct = src[t];
        if (done || ct >= 0) {
            dst[t] = ct;
            continue;
        }
        uint32_t pt = item_priority(t);
        int2 bb = item_free_bodies(w, b, t);
        bool best = true;
        uint64_t used = 0;
        for (int side = 0; side < 2 && best; ++side) {
            int body = side ? bb.y : bb.x;
            if (body < 0 || (side && body == bb.x)) continue;
            for (int m = b.bi_start[body]; m < b.bi_start[body + 1]; ++m) {
                int u = b.bi_items[m];
                if (u == t) continue;
                int cu = src[u];
                if (cu < 0) {
                    uint32_t pu = item_priority(u);
                    if (pu > pt || (pu == pt && u > t)) { best = false; break; }
                } else if (cu < 64) {
                    used |= 1ull << cu;
                }
            }
        }
        if (!best) {
            dst[t] = -1;
            continue;
        }
        int c = __ffsll((unsigned long long)~used) - 1;
        if (c < 0 || c >= BIG_COLOURS) c = BIG_SERIAL;
        dst[t] = c;
        atomicSub(b.remaining, 1);
    }
}

// Items still uncoloured join the serial class; class sizes and starts.
__global__ void k_big_classes(DevWorld w, BigScene b, int collisions) {
    int K = min(big_items(w, b, collisions), b.item_cap);
    int* colour = w.item_level + b.item_base;
    int* ls = w.item_lstart + b.item_base + b.s;
    __shared__ int cnt[BIG_CLASSES];
    for (int c = threadIdx.x; c < BIG_CLASSES; c += blockDim.x) cnt[c] = 0;
    __syncthreads();
    for (int t = threadIdx.x; t < K; t += blockDim.x) {
        int c = colour[t];
        if (c < 0) colour[t] = c = BIG_SERIAL;
        atomicAdd(&cnt[c], 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0, used = 0;
        for (int c = 0; c < BIG_CLASSES; ++c) {
            ls[c] = run;
            run += cnt[c];
            if (cnt[c]) used = c + 1;
        }
        ls[BIG_CLASSES] = run;
        w.scene_nlevels[b.s] = used;
    }
}

// Serial-class flags over the item capacity (scanned into item-order ranks).
__global__ void k_big_serial_flags(DevWorld w, BigScene b, int collisions) {
    int K = min(big_items(w, b, collisions), b.item_cap);
    const int* colour = w.item_level + b.item_base;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < b.item_cap; t += gridDim.x * blockDim.x)
        b.tmp[t] = (t < K && colour[t] == BIG_SERIAL) ? 1 : 0;
}

// Groups items by class: the parallel classes in any order (their items are
// independent), the serial class in item order (ranks from the scan).
__global__ void k_big_order(DevWorld w, BigScene b, int collisions) {
    int K = min(big_items(w, b, collisions), b.item_cap);
    const int* colour = w.item_level + b.item_base;
    const int* ls = w.item_lstart + b.item_base + b.s;
    int* cur = w.item_cursor + b.item_base + b.s;
    int* order = w.item_order + b.item_base;
    int lane = lane_id();
    unsigned long long below = lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
    // whole waves iterate (the cursor atomics are aggregated per wave and colour)
    for (int t0 = blockIdx.x * blockDim.x; t0 < K; t0 += gridDim.x * blockDim.x) {
        int t = t0 + threadIdx.x;
        bool valid = t < K;
        int c = valid ? colour[t] : BIG_SERIAL;
        int pos = valid ? ls[BIG_SERIAL] + b.tmp[t] : 0;
        unsigned long long todo = __ballot(c != BIG_SERIAL);
        while (todo) {  // one round per colour present in the wave
            int leader = __ffsll((long long)todo) - 1;
            int cl = __shfl(c, leader);
            unsigned long long m = __ballot(c == cl);
            int base = 0;
            if (lane == leader) base = atomicAdd(&cur[cl], __popcll(m));
            base = __shfl(base, leader);
            if (c == cl) pos = base + __popcll(m & below);
            todo &= ~m;
        }
        if (valid) {
            int code = item_code(w, b, t);
            order[pos] = code;
            w.item_bodies[b.item_base + pos] = item_body_pair(w, code, b.b0);
        }
    }
}

__global__ void k_big_cursor(DevWorld w, BigScene b) {
    const int* ls = w.item_lstart + b.item_base + b.s;
    int* cur = w.item_cursor + b.item_base + b.s;
    if (threadIdx.x < BIG_CLASSES) cur[threadIdx.x] = ls[threadIdx.x];
}

void big_colour(const DevWorld& w, const BigScene& b, bool collisions, hipStream_t st) {
    int co = collisions ? 1 : 0;
    int gi = grid_for(b.item_cap, 256);
    hipLaunchKernelGGL(k_fill, dim3(grid_for(b.n, 256)), dim3(256), 0, st, b.bi_cursor, b.n, 0);
    hipLaunchKernelGGL(k_big_csr_count, dim3(gi), dim3(256), 0, st, w, b, co);
    scan(b.bi_cursor, b.bi_start, b.n, b.scan_blocks, st);
    hipLaunchKernelGGL(k_copy, dim3(grid_for(b.n, 256)), dim3(256), 0, st, b.bi_start, b.bi_cursor, b.n);
    hipLaunchKernelGGL(k_big_csr_fill, dim3(gi), dim3(256), 0, st, w, b, co);
    // rounds alternate between item_level and item_order (free until
    // k_big_order fills it); an even number of rounds ends in item_level
    static_assert(BIG_COLOUR_ROUNDS % 2 == 0, "the last colour round must write item_level");
    int* colour = w.item_level + b.item_base;
    int* spare = w.item_order + b.item_base;
    for (int r = 0; r < BIG_COLOUR_ROUNDS; ++r)
        hipLaunchKernelGGL(k_big_colour_round, dim3(gi), dim3(256), 0, st, w, b, co, (r & 1) ? spare : colour,
                           (r & 1) ? colour : spare);
    hipLaunchKernelGGL(k_big_classes, dim3(1), dim3(1024), 0, st, w, b, co);
    hipLaunchKernelGGL(k_big_cursor, dim3(1), dim3(64), 0, st, w, b);
    hipLaunchKernelGGL(k_big_serial_flags, dim3(gi), dim3(256), 0, st, w, b, co);
    scan(b.tmp, b.tmp, b.item_cap, b.scan_blocks, st);
    hipLaunchKernelGGL(k_big_order, dim3(gi), dim3(256), 0, st, w, b, co);
}

// ---------------------------------------------------------------------------
// Coloured solve
// ---------------------------------------------------------------------------
__global__ void k_big_ext_zero(DevWorld w, BigScene b) {
    int e0 = w.scene_ext_off[b.s], e1 = w.scene_ext_off[b.s + 1];
    for (int k = e0 + blockIdx.x * blockDim.x + threadIdx.x; k < e1; k += gridDim.x * blockDim.x) {
        w.ext_lambda[3 * k] = 0.0;
        w.ext_lambda[3 * k + 1] = 0.0;
        w.ext_lambda[3 * k + 2] = 0.0;
    }
}

template <bool kExt, bool kVel>
__global__ void __launch_bounds__(64) k_big_class(DevWorld w, BigScene b, double h, int c) {
    const int* ls = w.item_lstart + b.item_base + b.s;
    const int* order = w.item_order + b.item_base;
    const int2* bodies = w.item_bodies + b.item_base;
    int i0 = ls[c], i1 = ls[c + 1];
    uint32_t err = 0;
    if (c == BIG_SERIAL) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            for (int idx = i0; idx < i1; ++idx) {
                if (kVel) solve_item_vel(w, order[idx], bodies[idx], h);
                else solve_item_pos<kExt>(w, order[idx], bodies[idx], h, err);
            }
        }
    } else {
        for (int idx = i0 + blockIdx.x * blockDim.x + threadIdx.x; idx < i1; idx += gridDim.x * blockDim.x) {
            if (kVel) solve_item_vel(w, order[idx], bodies[idx], h);
            else solve_item_pos<kExt>(w, order[idx], bodies[idx], h, err);
        }
    }
    if (err) atomicOr(&w.scene_err[b.s], err);
}

// Colour classes c0.. and then the serial class, on one block: a block
// barrier between classes instead of a launch per class.
template <bool kExt, bool kVel>
__global__ void __launch_bounds__(1024) k_big_tail(DevWorld w, BigScene b, double h, int c0) {
    const int* ls = w.item_lstart + b.item_base + b.s;
    const int* order = w.item_order + b.item_base;
    const int2* bodies = w.item_bodies + b.item_base;
    uint32_t err = 0;
    for (int c = c0; c < BIG_COLOURS; ++c) {
        int i0 = ls[c], i1 = ls[c + 1];
        if (i0 == i1) continue;  // the same for every thread
        for (int idx = i0 + (int)threadIdx.x; idx < i1; idx += blockDim.x) {
            if (kVel) solve_item_vel(w, order[idx], bodies[idx], h);
            else solve_item_pos<kExt>(w, order[idx], bodies[idx], h, err);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int idx = ls[BIG_SERIAL]; idx < ls[BIG_SERIAL + 1]; ++idx) {
            if (kVel) solve_item_vel(w, order[idx], bodies[idx], h);
            else solve_item_pos<kExt>(w, order[idx], bodies[idx], h, err);
        }
    }
    if (err) atomicOr(&w.scene_err[b.s], err);
}

__global__ void k_big_velocity_update(DevWorld w, BigScene b, double h) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < b.n; i += gridDim.x * blockDim.x)
        velocity_update_body(w.body[b.b0 + i], h);
}

void big_solve(const DevWorld& w, const BigScene& b, const FrameParams& fp, hipStream_t st) {
    int gi = std::min(grid_for(b.item_cap, 64), 2048);
    bool ext = w.n_ext > 0;
    if (ext) hipLaunchKernelGGL(k_big_ext_zero, dim3(grid_for(w.n_ext, 256)), dim3(256), 0, st, w, b);
    // classes 0..L-1 one launch each (the large ones), the rest on one block
    int L = std::min(std::max(w.big_launch_classes, 0), BIG_CLASSES);
    for (unsigned it = 0; it < fp.iters; ++it) {
        for (int c = 0; c < L; ++c) {
            if (ext) hipLaunchKernelGGL((k_big_class<true, false>), dim3(gi), dim3(64), 0, st, w, b, fp.h, c);
            else hipLaunchKernelGGL((k_big_class<false, false>), dim3(gi), dim3(64), 0, st, w, b, fp.h, c);
        }
        if (L < BIG_CLASSES) {
            if (ext) hipLaunchKernelGGL((k_big_tail<true, false>), dim3(1), dim3(1024), 0, st, w, b, fp.h, L);
            else hipLaunchKernelGGL((k_big_tail<false, false>), dim3(1), dim3(1024), 0, st, w, b, fp.h, L);
        }
    }
    hipLaunchKernelGGL(k_big_velocity_update, dim3(grid_for(b.n, 256)), dim3(256), 0, st, w, b, fp.h);
    for (int c = 0; c < L; ++c)
        hipLaunchKernelGGL((k_big_class<false, true>), dim3(gi), dim3(64), 0, st, w, b, fp.h, c);
    if (L < BIG_CLASSES) hipLaunchKernelGGL((k_big_tail<false, true>), dim3(1), dim3(1024), 0, st, w, b, fp.h, L);
}

}  // namespace rp
