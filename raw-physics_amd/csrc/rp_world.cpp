// rp_world.cpp — host runtime of librawphys.so: hull construction, world
// upload to HBM, frame stepping, state I/O and the drop-in pbd_simulate.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/rawphys.h"
#include "rp_kernels.h"

namespace rp {

// Scene-range views of a world replayed concurrently, one stream each. With
// the runtime's default four hardware queues per process, three views ran
// the batched workload fastest on MI355X (3.47 ms/frame, against 3.70 for
// two, 4.45 for one; a fourth stream shared a queue and took 5.87).
constexpr int kMaxViews = 4;
constexpr int kAutoViews = 3;

static thread_local uint32_t g_last_err = 0;
static thread_local std::string g_last_msg;

static void set_msg(const std::string& m) { g_last_msg = m; }
// Error word and message of a failed call made outside this file (rp_obj.cpp).
void rp_set_last_error(uint32_t code, const std::string& msg) {
    g_last_err = code;
    g_last_msg = msg;
}

#define RP_CHECK(expr)                                                                    \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) {                                                           \
            set_msg(std::string(#expr) + ": " + hipGetErrorString(e_));                   \
            return false;                                                                 \
        }                                                                                 \
    } while (0)

// ---------------------------------------------------------------------------
// Hull construction (collider_convex_hull_create, collider.cpp:194-364)
// ---------------------------------------------------------------------------
struct HostHull {
    std::vector<V3> v, fn;
    std::vector<int32_t> fe_off, fe, vf_off, vf, vn_off, vn, ff_off, ff;
    bool ok = true;
};

struct Tri { int32_t x, y, z; };
struct Edge { int32_t x, y; };

static void flatten(const std::vector<std::vector<int32_t>>& lists, std::vector<int32_t>& off, std::vector<int32_t>& d) {
    off.assign(1, 0);
    d.clear();
    for (const auto& l : lists) {
        d.insert(d.end(), l.begin(), l.end());
        off.push_back((int32_t)d.size());
    }
}

// collect_faces_planar_to (collider.cpp:47-77): depth-first over triangle
// neighbours, gathering triangles whose normal matches `target`.
static void planar_flood(const std::vector<V3>& hv, const std::vector<Tri>& tris,
                         const std::vector<std::vector<int32_t>>& tri_nb, std::vector<char>& done, int32_t f, V3 target,
                         std::vector<Tri>& out) {
    const double EPSILON = 0.000001;
    Tri t = tris[f];
    V3 n = nrm(cross(sub(hv[t.y], hv[t.x]), sub(hv[t.z], hv[t.x])));
    if (done[f]) return;
    double projection = dot(n, target);
    if ((projection - 1.0) > -EPSILON && (projection - 1.0) < EPSILON) {
        out.push_back(t);
        done[f] = 1;
        for (int32_t nb : tri_nb[f]) planar_flood(hv, tris, tri_nb, done, nb, target, out);
    }
}

// create_convex_hull_face (collider.cpp:93-166): the boundary loop of a set
// of coplanar triangles, in the reference's edge order.
static bool face_elements(const std::vector<Tri>& tris, std::vector<int32_t>& el) {
    std::vector<Edge> edges;
    auto find = [&](Edge q) -> int {
        for (size_t i = 0; i < edges.size(); ++i)
            if ((edges[i].x == q.x && edges[i].y == q.y) || (edges[i].x == q.y && edges[i].y == q.x)) return (int)i;
        return -1;
    };
    for (const Tri& t : tris) {
        Edge es[3] = {{t.x, t.y}, {t.y, t.z}, {t.z, t.x}};
        for (Edge e : es) {
            int i = find(e);
            if (i >= 0) {  // array_remove: swap with last (light_array.h:146)
                edges[i] = edges.back();
                edges.pop_back();
            } else {
                edges.push_back(e);
            }
        }
    }
    for (size_t i = 0; i < edges.size(); ++i) {
        Edge cur = edges[i];
        for (size_t j = i + 1; j < edges.size(); ++j) {
            Edge cand = edges[j];
            if (cur.y != cand.x && cur.y != cand.y) continue;
            if (cur.y == cand.y) std::swap(cand.x, cand.y);
            Edge tmp = edges[i + 1];
            edges[i + 1] = cand;
            edges[j] = tmp;
        }
    }
    el.clear();
    for (const Edge& e : edges) el.push_back(e.x);
    return !edges.empty() && edges.front().x == edges.back().y;
}

static HostHull build_hull(const std::vector<V3>& corners) {
    HostHull H;
    // unique vertices in first-appearance order (collider.cpp:200-209);
    // keys compare with ==, so the two zeros are one key
    struct Key {
        uint64_t a, b, c;
        bool operator==(const Key& o) const { return a == o.a && b == o.b && c == o.c; }
    };
    struct KeyHash {
        size_t operator()(const Key& k) const { return (size_t)(k.a * 0x9E3779B97F4A7C15ull ^ (k.b + 0x632BE59BD9B4E019ull) * 31 ^ k.c); }
    };
    auto key = [](V3 p) {
        double x = p.x + 0.0, y = p.y + 0.0, z = p.z + 0.0;
        Key k;
        memcpy(&k.a, &x, 8); memcpy(&k.b, &y, 8); memcpy(&k.c, &z, 8);
        return k;
    };
    std::unordered_map<Key, int32_t, KeyHash> index;
    std::vector<int32_t> cidx(corners.size());
    for (size_t c = 0; c < corners.size(); ++c) {
        Key k = key(corners[c]);
        auto it = index.find(k);
        if (it == index.end()) {
            it = index.emplace(k, (int32_t)H.v.size()).first;
            H.v.push_back(corners[c]);
        }
        cidx[c] = it->second;
    }
    size_t nt = corners.size() / 3, nv = H.v.size();
    std::vector<Tri> tris(nt);
    for (size_t i = 0; i < nt; ++i) tris[i] = {cidx[3 * i], cidx[3 * i + 1], cidx[3 * i + 2]};
    // triangle neighbours and vertex neighbours (collider.cpp:248-282)
    std::vector<std::vector<int32_t>> tri_nb(nt), vn(nv), vf(nv);
    auto add_unique = [](std::vector<int32_t>& l, int32_t v) {
        if (std::find(l.begin(), l.end(), v) == l.end()) l.push_back(v);
    };
    for (size_t i = 0; i < nt; ++i) {
        Tri t = tris[i];
        for (size_t j = 0; j < nt; ++j) {
            if (i == j) continue;
            Tri u = tris[j];
            bool share = t.x == u.x || t.x == u.y || t.x == u.z || t.y == u.x || t.y == u.y || t.y == u.z ||
                         t.z == u.x || t.z == u.y || t.z == u.z;
            if (share) tri_nb[i].push_back((int32_t)j);
        }
        add_unique(vn[t.x], t.y); add_unique(vn[t.x], t.z);
        add_unique(vn[t.y], t.x); add_unique(vn[t.y], t.z);
        add_unique(vn[t.z], t.x); add_unique(vn[t.z], t.y);
    }
    // merge coplanar triangles into faces (collider.cpp:284-319)
    std::vector<char> done(nt, 0);
    std::vector<std::vector<int32_t>> felems;
    for (size_t i = 0; i < nt; ++i) {
        if (done[i]) continue;
        Tri t = tris[i];
        V3 n = nrm(cross(sub(H.v[t.y], H.v[t.x]), sub(H.v[t.z], H.v[t.x])));
        std::vector<Tri> planar;
        planar_flood(H.v, tris, tri_nb, done, (int32_t)i, n, planar);
        std::vector<int32_t> el;
        H.ok &= face_elements(planar, el);
        int32_t fi = (int32_t)felems.size();
        felems.push_back(el);
        H.fn.push_back(n);
        for (const Tri& p : planar) { vf[p.x].push_back(fi); vf[p.y].push_back(fi); vf[p.z].push_back(fi); }
    }
    // face neighbours: faces sharing any vertex (collider.cpp:321-341)
    size_t nf = felems.size();
    std::vector<std::vector<int32_t>> ff(nf);
    for (size_t i = 0; i < nf; ++i)
        for (size_t j = 0; j < nf; ++j) {
            if (i == j) continue;
            bool share = false;
            for (int32_t a : felems[i]) {
                for (int32_t b : felems[j])
                    if (a == b) { share = true; break; }
                if (share) break;
            }
            if (share) ff[i].push_back((int32_t)j);
        }
    flatten(felems, H.fe_off, H.fe);
    flatten(vf, H.vf_off, H.vf);
    flatten(vn, H.vn_off, H.vn);
    flatten(ff, H.ff_off, H.ff);
    H.ok &= nv > 0 && nf > 0;
    return H;
}

// One collider of an entity (collider.h:27-44): a hull instance or a sphere.
struct HostCollider {
    int32_t type;   // COL_HULL or COL_SPHERE
    int32_t hull;   // hull instance (COL_HULL)
    float radius;   // r32 radius (COL_SPHERE, collider.h:29)
};

// Point-mass inertia at the hull vertices (colliders_get_default_inertia_tensor,
// collider.cpp:448-494) and its inverse (gm_mat3_inverse, gm.h:347-368).
static M3 hull_inertia(const HostHull& h, double mass) {
    double mpv = mass / (uint32_t)h.v.size();
    M3 I;
    memset(&I, 0, sizeof(I));
    for (const V3& v : h.v) {
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
    return I;
}
static bool m3_inverse(const M3& M, M3& out) {
    const double(*a)[3] = M.m;
    double det = a[0][0] * (a[1][1] * a[2][2] - a[2][1] * a[1][2]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                 a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    if (det == 0.0) return false;
    double invdet = 1 / det;
    out.m[0][0] = (a[1][1] * a[2][2] - a[2][1] * a[1][2]) * invdet;
    out.m[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * invdet;
    out.m[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * invdet;
    out.m[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) * invdet;
    out.m[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * invdet;
    out.m[1][2] = (a[1][0] * a[0][2] - a[0][0] * a[1][2]) * invdet;
    out.m[2][0] = (a[1][0] * a[2][1] - a[2][0] * a[1][1]) * invdet;
    out.m[2][1] = (a[2][0] * a[0][1] - a[0][0] * a[2][1]) * invdet;
    out.m[2][2] = (a[0][0] * a[1][1] - a[1][0] * a[0][1]) * invdet;
    return true;
}
static double hull_radius(const HostHull& h) {  // collider.cpp:178-189, 510-521
    double maxd = 0.0;
    for (const V3& v : h.v) {
        double l = len(v);
        if (l > maxd) maxd = l;
    }
    return maxd;
}

// colliders_get_bounding_sphere_radius (collider.cpp:510-521)
static double colliders_radius(const std::vector<HostHull>& hulls, const HostCollider* c, int n) {
    double r = -DBL_MAX;
    for (int i = 0; i < n; ++i) {
        double ri = (c[i].type == COL_SPHERE) ? (double)c[i].radius : hull_radius(hulls[c[i].hull]);
        if (ri > r) r = ri;
    }
    return r;
}

// colliders_get_default_inertia_tensor (collider.cpp:448-494): a lone sphere
// is a solid ball; otherwise the mass is spread evenly over the vertices of
// every hull collider (a sphere among several colliders is an assert there).
static M3 colliders_inertia(const std::vector<HostHull>& hulls, const HostCollider* c, int n, double mass,
                            uint32_t& err) {
    M3 I;
    memset(&I, 0, sizeof(I));
    if (n == 1 && c[0].type == COL_SPHERE) {
        double Id = (2.0 / 5.0) * mass * (double)c[0].radius * (double)c[0].radius;
        I.m[0][0] = I.m[1][1] = I.m[2][2] = Id;
        return I;
    }
    if (n == 1) return hull_inertia(hulls[c[0].hull], mass);
    uint32_t total = 0;
    for (int i = 0; i < n; ++i) {
        if (c[i].type == COL_SPHERE) err |= ERR_HULL;
        else total += (uint32_t)hulls[c[i].hull].v.size();
    }
    double mpv = mass / total;
    for (int i = 0; i < n; ++i) {
        if (c[i].type == COL_SPHERE) continue;
        for (const V3& v : hulls[c[i].hull].v) {
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
    return I;
}

// ---------------------------------------------------------------------------
// Host-side world and its device image
// ---------------------------------------------------------------------------
// Scenes above this many bodies take the large-scene path (grid broad
// phase, parallel islands, graph-coloured solve) unless the desc says
// otherwise.
constexpr uint32_t kLargeSceneBodies = 4096;
// Pair capacity of a large scene per body (plus one per body for each large body).
constexpr int64_t kLargePairsPerBody = 32;

struct HostWorld {
    uint32_t large_scene_bodies = kLargeSceneBodies;
    std::vector<int32_t> scene_body_off{0}, scene_ext_off{0};
    std::vector<Body> body;
    std::vector<int32_t> body_hull;     // first collider's hull instance (-1: a sphere)
    std::vector<V3> body_scale;         // world scale (rendering only)
    std::vector<int32_t> body_col_off{0};
    std::vector<HostCollider> cols;     // every body's colliders, in order
    std::vector<HostHull> hulls;
    std::vector<ExtC> ext;
    std::vector<uint32_t> scene_err;  // errors found while building
};

struct DeviceImage {
    DevWorld d;
    std::vector<BigScene> big;  // large scenes (device buffers in allocs)
    std::vector<void*> allocs;
    std::vector<int32_t> body_tv_off;  // host copy for inspection
    std::vector<int32_t> pair_base;    // host copy of scene_pair_base
    std::vector<int32_t> body_xf_off;  // [nb+1] first transform-list entry of each body
    int64_t pair_capacity = 0;
    bool ok = false;

    template <class T>
    bool alloc(T*& p, size_t n) {
        void* q = nullptr;
        size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
        RP_CHECK(hipMalloc(&q, bytes));
        RP_CHECK(hipMemset(q, 0, bytes));
        allocs.push_back(q);
        p = (T*)q;
        return true;
    }
    template <class T>
    bool upload(const T*& p, const std::vector<T>& h) {
        T* q;
        if (!alloc(q, h.size())) return false;
        if (!h.empty()) RP_CHECK(hipMemcpy(q, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
        p = q;
        return true;
    }
    template <class T>
    bool upload_mut(T*& p, const std::vector<T>& h) {
        if (!alloc(p, h.size())) return false;
        if (!h.empty()) RP_CHECK(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
        return true;
    }
    void release() {
        for (void* p : allocs) (void)hipFree(p);
        allocs.clear();
        if (d.group_work) (void)hipFree(d.group_work);
        d.group_work = nullptr;
        d.group_work_slots = 0;
    }
    // Lane-group working memory for the current schedules (grown, never shrunk;
    // the caller has synchronized the stream).
    bool ensure_group_work(int64_t at_least = 0) {
        int64_t need = std::max(group_work_slots_needed(d), at_least);
        if (need <= d.group_work_slots) return true;
        if (d.group_work) (void)hipFree(d.group_work);
        d.group_work = nullptr;
        d.group_work_slots = 0;
        RP_CHECK(hipMalloc(&d.group_work, (size_t)need * GROUP_WORK_BYTES));
        d.group_work_slots = need;
        return true;
    }
};

static bool build_image(const HostWorld& hw, DeviceImage& im) {
    DevWorld& d = im.d;
    memset(&d, 0, sizeof(d));
    d.ns = (int)hw.scene_body_off.size() - 1;
    d.nb = (int)hw.body.size();
    d.n_ext = (int)hw.ext.size();
    // topology dedupe: bodies of the same shape share index lists
    std::map<std::vector<int32_t>, int32_t> topo_ids;
    std::vector<int32_t> topo_base, topo_pool;
    std::vector<int32_t> hull_nv, hull_nf, hull_v_off, hull_fn_off, hull_topo;
    std::vector<V3> hull_v, hull_fn;
    for (const HostHull& h : hw.hulls) {
        std::vector<int32_t> key;
        const std::vector<int32_t>* parts[8] = {&h.fe_off, &h.fe, &h.vf_off, &h.vf, &h.vn_off, &h.vn, &h.ff_off, &h.ff};
        for (auto* p : parts) {
            key.push_back((int32_t)p->size());
            key.insert(key.end(), p->begin(), p->end());
        }
        auto it = topo_ids.find(key);
        if (it == topo_ids.end()) {
            int32_t id = (int32_t)topo_ids.size();
            for (auto* p : parts) {
                topo_base.push_back((int32_t)topo_pool.size());
                topo_pool.insert(topo_pool.end(), p->begin(), p->end());
            }
            it = topo_ids.emplace(key, id).first;
        }
        hull_topo.push_back(it->second);
        hull_nv.push_back((int32_t)h.v.size());
        hull_nf.push_back((int32_t)h.fn.size());
        hull_v_off.push_back((int32_t)hull_v.size());
        hull_fn_off.push_back((int32_t)hull_fn.size());
        hull_v.insert(hull_v.end(), h.v.begin(), h.v.end());
        hull_fn.insert(hull_fn.end(), h.fn.begin(), h.fn.end());
    }
    // per-collider transformed geometry and the transform work list
    // (body, hull, vertex, tv slot); a body's first hull collider's vertices
    // start at body_tv_off
    std::vector<int4> xf;
    std::vector<int32_t> col_type, col_hull, col_tv_off, body_general(d.nb, 0);
    std::vector<float> col_radius;
    int64_t ntv = 0;
    im.body_tv_off.resize(d.nb);
    im.body_xf_off.assign(d.nb + 1, 0);
    for (int b = 0; b < d.nb; ++b) {
        int c0 = hw.body_col_off[b], c1 = hw.body_col_off[b + 1];
        im.body_tv_off[b] = (int32_t)ntv;
        im.body_xf_off[b] = (int32_t)xf.size();
        if (c1 - c0 != 1 || hw.cols[c0].type != COL_HULL) body_general[b] = 1;
        for (int c = c0; c < c1; ++c) {
            const HostCollider& hc = hw.cols[c];
            col_type.push_back(hc.type);
            col_hull.push_back(hc.hull);
            col_radius.push_back(hc.radius);
            col_tv_off.push_back((int32_t)ntv);
            if (hc.type != COL_HULL) continue;
            const HostHull& h = hw.hulls[hc.hull];
            for (size_t i = 0; i < h.v.size(); ++i) xf.push_back(make_int4(b, hc.hull, (int)i, (int)(ntv + i)));
            ntv += (int64_t)h.v.size();
        }
    }
    d.n_xf = (int)xf.size();
    im.body_xf_off[d.nb] = d.n_xf;
    d.has_general = 0;
    for (int32_t g : body_general) d.has_general |= g;
    // pair and constraint capacities per scene; large scenes get a bounded
    // pair capacity (their broad phase reports ERR_PAIR_CAPACITY beyond it)
    std::vector<int32_t> pair_base(d.ns + 1, 0), con_base(d.ns + 1, 0), scene_big(d.ns, 0);
    im.big.clear();
    int64_t pcap = 0, ccap = 0;
    for (int s = 0; s < d.ns; ++s) {
        int64_t n = hw.scene_body_off[s + 1] - hw.scene_body_off[s];
        int64_t next = hw.scene_ext_off[s + 1] - hw.scene_ext_off[s];
        pair_base[s] = (int32_t)pcap;
        con_base[s] = (int32_t)ccap;
        int64_t scene_pairs = n * (n - 1) / 2;
        // (at least 65 bodies: the class starts of the coloured solve need
        // more item slots than a tiny scene has)
        if (n > std::max<int64_t>(hw.large_scene_bodies, 64)) {
            BigScene bs;
            memset(&bs, 0, sizeof(bs));
            bs.s = s;
            bs.b0 = hw.scene_body_off[s];
            bs.n = (int)n;
            // bodies far above the typical size (a floor) are tested against
            // all others; the grid cell bounds every pair of the rest
            std::vector<double> r(n);
            for (int64_t i = 0; i < n; ++i) r[i] = hw.body[bs.b0 + i].radius;
            std::vector<double> sorted_r = r;
            std::nth_element(sorted_r.begin(), sorted_r.begin() + n / 2, sorted_r.end());
            double small_bound = 4.0 * sorted_r[n / 2];
            double rmax = 0.0;
            int64_t n_large = 0;
            for (double x : r) {
                if (x > small_bound) ++n_large;
                else rmax = std::max(rmax, x);
            }
            bs.n_large = (int)n_large;
            // 1% margin: cell coordinates of two bodies that pass the test
            // differ by at most one even after rounding
            bs.cell = (2.0 * rmax + 0.1) * 1.01 + 1e-9;
            int64_t T = 1024;
            while (T < 2 * (n - n_large)) T <<= 1;
            bs.T = (int)T;
            scene_pairs = std::min(scene_pairs, kLargePairsPerBody * n + n_large * n);
            bs.pair_base = (int32_t)pcap;
            bs.pair_cap = (int)scene_pairs;
            bs.item_base = (int32_t)ccap;
            bs.item_cap = (int)(next + scene_pairs);
            im.big.push_back(bs);
            scene_big[s] = 1;
        }
        pcap += scene_pairs;
        ccap += next + scene_pairs;
        if (pcap > INT32_MAX / MAXC || ccap > INT32_MAX - d.ns - 1) {
            set_msg("world too large for 32-bit pair/constraint indices");
            return false;
        }
    }
    pair_base[d.ns] = (int32_t)pcap;
    con_base[d.ns] = (int32_t)ccap;
    int max_nv = 1;
    for (const HostHull& h : hw.hulls) max_nv = std::max(max_nv, (int)h.v.size());
    d.stage_cap = 2 * max_nv;
    // LDS a narrow-phase block may stage hulls in: 64 KiB where a lane-group
    // schedule fits in it (16 lanes per pair takes the least), else the
    // device's limit per block (160 KiB on gfx950), for hulls that would
    // otherwise have no lane-group schedule; such worlds run lane groups in
    // the fused kernel only (large_lds_split)
    d.lds_limit = 64 * 1024;
    if ((size_t)(64 / 16) * d.stage_cap * sizeof(V3) > (size_t)d.lds_limit) {
        int dev = 0, lim = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
            lim > d.lds_limit) {
            d.lds_limit = lim;
            RP_CHECK(allow_narrow_lds(lim));
        }
    }
    // size classes: with hulls above SMALL_HULL_NV vertices, the pairs of two
    // smaller hulls get their own list and LDS slices sized for them (if the
    // large-hull pairs fit a 16-lane schedule)
    d.small_nv = 0;
    if (max_nv > SMALL_HULL_NV && (size_t)(64 / 16) * d.stage_cap * sizeof(V3) <= (size_t)d.lds_limit) {
        for (const HostHull& h : hw.hulls)
            if ((int)h.v.size() <= SMALL_HULL_NV) d.small_nv = std::max(d.small_nv, (int)h.v.size());
    }
    d.stage_small = 2 * d.small_nv;
    im.pair_capacity = pcap;
    im.pair_base = pair_base;
    // a few small scenes (the drop-in's one scene per call) leave the device
    // nearly empty, so a pair's latency sets the step: 16 lanes per pair
    // (1.09 ms per 16-body frame against 1.17 ms with 8, at a resting state);
    // many scenes fill the device and 8 lanes per pair fit more pairs per wave
    bool few = d.ns < 64 && im.big.empty();
    d.narrow_group = 1;
    for (int g : {few ? 16 : 8, few ? 8 : 16, 4})
        if (narrow_group_supported(d, g)) { d.narrow_group = g; break; }
    d.narrow_blocks = narrow_blocks_for(d, d.narrow_group, pcap);
    // manifolds built by the GJK + EPA kernel on its staged hulls where it
    // stages them, else by the one-lane-per-pair manifold kernel
    d.manifold_group = d.narrow_group > 1 ? 0 : 1;
    d.manifold_blocks = narrow_blocks_for(d, 1, pcap);
    if (!im.ensure_group_work()) return false;
    d.solve_lanes = few ? 32 : 64;
    d.big_launch_classes = BIG_LAUNCH_CLASSES;

    bool ok = true;
    ok &= im.upload(d.scene_body_off, hw.scene_body_off);
    ok &= im.upload(d.scene_pair_base, pair_base);
    ok &= im.upload(d.scene_ext_off, hw.scene_ext_off);
    ok &= im.upload(d.scene_item_base, con_base);
    ok &= im.upload_mut(d.body, hw.body);
    ok &= im.alloc(d.model, d.nb);
    ok &= im.alloc(d.pose, (size_t)d.nb * 7);
    std::vector<V3> scale = hw.body_scale;
    scale.resize(d.nb, v3(1.0, 1.0, 1.0));
    ok &= im.upload(d.body_scale, scale);
    ok &= im.upload(d.body_hull, hw.body_hull);
    const int32_t* p;
    ok &= im.upload(p, im.body_tv_off); d.body_tv_off = p;
    ok &= im.upload(d.hull_nv, hull_nv);
    ok &= im.upload(d.hull_nf, hull_nf);
    ok &= im.upload(d.hull_v_off, hull_v_off);
    ok &= im.upload(d.hull_fn_off, hull_fn_off);
    ok &= im.upload(d.hull_topo, hull_topo);
    ok &= im.upload(d.hull_v, hull_v);
    ok &= im.upload(d.hull_fn, hull_fn);
    ok &= im.upload(d.topo_base, topo_base);
    ok &= im.upload(d.topo_pool, topo_pool);
    ok &= im.alloc(d.tv, (size_t)ntv);
    ok &= im.alloc(d.need, d.nb);
    ok &= im.upload(d.xf, xf);
    ok &= im.upload(d.body_col_off, hw.body_col_off);
    ok &= im.upload(d.col_type, col_type);
    ok &= im.upload(d.col_hull, col_hull);
    ok &= im.upload(d.col_tv_off, col_tv_off);
    ok &= im.upload(d.col_radius, col_radius);
    ok &= im.upload(d.body_general, body_general);
    ok &= im.alloc(d.pairs, (size_t)pcap);
    ok &= im.alloc(d.pair_count, d.ns);
    ok &= im.alloc(d.live_local, (size_t)pcap);
    ok &= im.alloc(d.live_count, d.ns);
    ok &= im.alloc(d.live_prefix, d.ns + 1 + kMaxViews);  // a view's prefix ends one slot past its scenes
    ok &= im.alloc(d.live_global, (size_t)pcap);
    ok &= im.alloc(d.live_rec, (size_t)pcap);
    ok &= im.alloc(d.live_hull, (size_t)pcap);
    ok &= im.alloc(d.ncontact, (size_t)pcap);
    ok &= im.alloc(d.pair_normal, (size_t)pcap);
    ok &= im.alloc(d.mlist, (size_t)pcap);
    ok &= im.alloc(d.pair_lists, 2 * (size_t)pcap);
    d.list_stride = pcap;
    ok &= im.alloc(d.counters, CNT_COUNT * kMaxViews);
    ok &= im.alloc(d.contacts, (size_t)pcap * MAXC);
    ok &= im.upload(d.ext, hw.ext);
    ok &= im.alloc(d.ext_lambda, 3 * hw.ext.size());
    ok &= im.alloc(d.item_level, (size_t)ccap);
    ok &= im.alloc(d.item_order, (size_t)ccap);
    ok &= im.alloc(d.item_bodies, (size_t)ccap);
    ok &= im.alloc(d.item_lstart, (size_t)ccap + d.ns);
    ok &= im.alloc(d.item_cursor, (size_t)ccap + d.ns);
    ok &= im.alloc(d.scene_nlevels, d.ns);
    ok &= im.alloc(d.body_last, d.nb);
    ok &= im.alloc(d.uf_parent, d.nb);
    ok &= im.alloc(d.uf_flag, d.nb);
    uint32_t* errp;
    ok &= im.alloc(errp, d.ns);
    d.scene_err = errp;
    if (ok && !hw.scene_err.empty())
        ok &= hipMemcpy(errp, hw.scene_err.data(), d.ns * sizeof(uint32_t), hipMemcpyHostToDevice) == hipSuccess;
    ok &= im.alloc(d.stats, STAT_COUNT);
    ok &= im.upload(d.scene_big, scene_big);
    for (BigScene& bs : im.big) {
        std::vector<int32_t> large, is_large(bs.n, 0);
        std::vector<double> r(bs.n);
        for (int i = 0; i < bs.n; ++i) r[i] = hw.body[bs.b0 + i].radius;
        std::vector<double> sorted_r = r;
        std::nth_element(sorted_r.begin(), sorted_r.begin() + bs.n / 2, sorted_r.end());
        double small_bound = 4.0 * sorted_r[bs.n / 2];
        for (int i = 0; i < bs.n; ++i)
            if (r[i] > small_bound) { large.push_back(i); is_large[i] = 1; }
        const int32_t* cp;
        ok &= im.upload(cp, large); bs.large = const_cast<int32_t*>(cp);
        bs.large_chunks = (bs.n + LARGE_CHUNK - 1) / LARGE_CHUNK;
        int64_t nlc = (int64_t)bs.n_large * bs.large_chunks;
        if (nlc >= INT32_MAX) {
            set_msg("too many large bodies in one scene");
            return false;
        }
        ok &= im.alloc(bs.large_cnt, (size_t)nlc + 1);
        ok &= im.upload(cp, is_large); bs.is_large = const_cast<int32_t*>(cp);
        int64_t tmp_n = std::max<int64_t>(bs.pair_cap, bs.item_cap) + 1;
        ok &= im.alloc(bs.cell_start, (size_t)bs.T + 1);
        ok &= im.alloc(bs.cell_cursor, (size_t)bs.T);
        ok &= im.alloc(bs.sorted, (size_t)bs.n);
        ok &= im.alloc(bs.spos, (size_t)bs.n);
        ok &= im.alloc(bs.pair_off, (size_t)bs.n + 1);
        ok &= im.alloc(bs.tmp, (size_t)tmp_n);
        ok &= im.alloc(bs.bi_start, (size_t)bs.n + 1);
        ok &= im.alloc(bs.bi_cursor, (size_t)bs.n);
        ok &= im.alloc(bs.bi_items, 2 * (size_t)bs.item_cap);
        ok &= im.alloc(bs.remaining, 1);
        int64_t scan_n = std::max<int64_t>({tmp_n, (int64_t)bs.T, (int64_t)bs.n + 1, nlc + 1});
        ok &= im.alloc(bs.scan_blocks, (size_t)scan_block_count(scan_n));
    }
    im.ok = ok;
    return ok;
}

// Entity fields at creation (entity_create_ex, entity.cpp:24-65).
static Body make_body(const HostWorld& hw, int c0, int nc, V3 x, Q4 q, V3 v, V3 w, double mass, bool fixed,
                      const double mat[3], uint32_t& err) {
    Body e;
    memset(&e, 0, sizeof(e));
    e.x = x; e.q = q; e.v = v; e.w = w;
    e.radius = colliders_radius(hw.hulls, hw.cols.data() + c0, nc);
    e.fixed = fixed ? 1 : 0;
    if (!fixed) {
        e.inv_mass = 1.0 / mass;
        e.I = colliders_inertia(hw.hulls, hw.cols.data() + c0, nc, mass, err);
        if (!m3_inverse(e.I, e.Iinv)) err |= ERR_HULL;
    }
    e.active = 1;
    e.deact = 0.0;
    e.mu_s = mat[0];
    e.mu_d = mat[1];
    e.e = mat[2];
    return e;
}

static bool host_world_from_desc(const rp_world_desc* d, HostWorld& hw) {
    uint32_t ns = d->num_scenes;
    uint32_t nb = d->scene_body_offset[ns];
    hw.scene_err.assign(ns, 0);
    if (d->large_scene_bodies) hw.large_scene_bodies = d->large_scene_bodies;
    // hull instances: one per distinct (shape, scale)
    std::map<std::tuple<uint32_t, uint64_t, uint64_t, uint64_t>, int32_t> inst;
    for (uint32_t s = 0; s < ns; ++s) {
        for (uint32_t g = d->scene_body_offset[s]; g < d->scene_body_offset[s + 1]; ++g) {
            const double* sc = d->body_scale + 3 * g;
            uint64_t k0, k1, k2;
            memcpy(&k0, sc, 8); memcpy(&k1, sc + 1, 8); memcpy(&k2, sc + 2, 8);
            uint32_t cb = d->body_collider_offset ? d->body_collider_offset[g] : 0;
            uint32_t nc = d->body_collider_offset ? d->body_collider_offset[g + 1] - cb : 1;
            int c0 = (int)hw.cols.size();
            for (uint32_t ci = 0; ci < nc; ++ci) {
                uint32_t shp = d->body_collider_offset ? d->collider_shape[cb + ci] : d->body_shape[g];
                if (shp >= d->num_shapes) { set_msg("body shape out of range"); return false; }
                HostCollider col = {COL_HULL, -1, 0.0f};
                if (d->shape_sphere_radius && d->shape_sphere_radius[shp] > 0.0f) {
                    col.type = COL_SPHERE;  // collider_sphere_create (collider.cpp:12-18)
                    col.radius = d->shape_sphere_radius[shp];
                    hw.cols.push_back(col);
                    continue;
                }
                auto key = std::make_tuple(shp, k0, k1, k2);
                auto it = inst.find(key);
                if (it == inst.end()) {
                    std::vector<V3> corners;
                    for (uint32_t c = d->shape_corner_offset[shp]; c < d->shape_corner_offset[shp + 1]; ++c) {
                        // examples_util.cpp:7-16: float widened to double, then scaled
                        V3 p = v3((double)d->shape_corners[3 * c], (double)d->shape_corners[3 * c + 1],
                                  (double)d->shape_corners[3 * c + 2]);
                        p.x *= sc[0]; p.y *= sc[1]; p.z *= sc[2];
                        corners.push_back(p);
                    }
                    hw.hulls.push_back(build_hull(corners));
                    it = inst.emplace(key, (int32_t)hw.hulls.size() - 1).first;
                }
                if (!hw.hulls[it->second].ok) hw.scene_err[s] |= ERR_HULL;
                col.hull = it->second;
                hw.cols.push_back(col);
            }
            hw.body_col_off.push_back((int32_t)hw.cols.size());
            const double* P = d->body_position + 3 * g;
            const double* R = d->body_rotation + 4 * g;
            const double* LV = d->body_linear_velocity + 3 * g;
            const double* AV = d->body_angular_velocity + 3 * g;
            Q4 q = {R[0], R[1], R[2], R[3]};
            hw.body.push_back(make_body(hw, c0, (int)nc, v3(P[0], P[1], P[2]), q, v3(LV[0], LV[1], LV[2]),
                                        v3(AV[0], AV[1], AV[2]), d->body_mass[g], d->body_fixed[g] != 0,
                                        d->body_material + 3 * g, hw.scene_err[s]));
            hw.body_hull.push_back(hw.cols[c0].type == COL_HULL ? hw.cols[c0].hull : -1);
            hw.body_scale.push_back(v3(sc[0], sc[1], sc[2]));
        }
        hw.scene_body_off.push_back((int32_t)hw.body.size());
        if (d->scene_constraint_offset && d->constraints) {
            uint32_t b0 = d->scene_body_offset[s], n = d->scene_body_offset[s + 1] - b0;
            for (uint32_t k = d->scene_constraint_offset[s]; k < d->scene_constraint_offset[s + 1]; ++k) {
                const rp_constraint_desc& cd = d->constraints[k];
                if (cd.body1 >= n || cd.body2 >= n || cd.type == RP_COLLISION_CONSTRAINT || cd.type < 0 || cd.type > 4) {
                    set_msg("invalid external constraint");
                    return false;
                }
                ExtC c;
                memset(&c, 0, sizeof(c));
                c.type = cd.type;
                c.b1 = (int32_t)(b0 + cd.body1);
                c.b2 = (int32_t)(b0 + cd.body2);
                c.limited = cd.limited;
                for (int a = 0; a < 4; ++a) { c.axis[a] = cd.axis[a]; c.limits[a] = cd.limits[a]; }
                c.r1 = v3(cd.r1_lc[0], cd.r1_lc[1], cd.r1_lc[2]);
                c.r2 = v3(cd.r2_lc[0], cd.r2_lc[1], cd.r2_lc[2]);
                c.distance = v3(cd.distance[0], cd.distance[1], cd.distance[2]);
                c.compliance = cd.compliance;
                hw.ext.push_back(c);
            }
        }
        hw.scene_ext_off.push_back((int32_t)hw.ext.size());
    }
    (void)nb;
    return true;
}

}  // namespace rp

using namespace rp;

struct rp_world {
    HostWorld hw;
    DeviceImage im;
    hipStream_t stream = nullptr;
    std::vector<Body> host_bodies;  // scratch for state I/O
    KernelTimer timer;              // profiling mode
    double kc_ms[KC_COUNT] = {0};
    uint64_t kc_launches[KC_COUNT] = {0};
    // one frame's launches captured as a HIP graph per view and replayed per
    // frame; views are scene ranges of the world on streams of their own
    bool use_graphs = true;
    int want_views = 0;  // 0 = automatic
    int nviews = 0;
    DevWorld view[kMaxViews];
    std::vector<BigScene> view_big[kMaxViews];
    hipStream_t view_stream[kMaxViews] = {};
    hipGraphExec_t view_exec[kMaxViews] = {};
    hipEvent_t fork_ev = nullptr, join_ev[kMaxViews] = {};
    FrameParams graph_fp;
    uint64_t sched_gen = 1, graph_gen = 0;  // schedule changes invalidate the captured graphs
    // bounded queue: at most two chunks of frames in flight on each stream
    hipEvent_t chunk_done[2][kMaxViews] = {};
    // differentiable runs: per-frame body checkpoints of the last recorded
    // forward pass and the buffers of the backward sweep
    Body* ckpt = nullptr;
    uint32_t ck_frames = 0;
    FrameParams ck_fp;
    int32_t* scene_dyn = nullptr;
    Recorder rec = {nullptr, nullptr, nullptr, 0};
    double* abar = nullptr;
    double* gparam = nullptr;
    double* amax = nullptr;                 // [nb] largest state adjoint at a frame start
    std::vector<double> dl_norm;            // [nb] largest |dL/dstate| component at the run's end
    bool last_collisions = true;  // of the last frame run (its solve items include live pairs)
    int device = -1;              // HIP device the world was created on
};

static void adjoint_free(rp_world* w) {
    void* ps[] = {w->ckpt, w->scene_dyn, w->rec.body, w->rec.ncontact, w->rec.contacts, w->abar, w->gparam, w->amax};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    w->ckpt = nullptr;
    w->scene_dyn = nullptr;
    w->rec = {nullptr, nullptr, nullptr, 0};
    w->abar = w->gparam = w->amax = nullptr;
    w->ck_frames = 0;
}

// Frames per chunk of the bounded launch queue.
constexpr uint32_t kFramesPerChunk = 32;

static bool same_params(const FrameParams& a, const FrameParams& b) {
    return a.dt == b.dt && a.h == b.h && a.gravity == b.gravity && a.substeps == b.substeps && a.iters == b.iters &&
           a.collisions == b.collisions && a.forces_given == b.forces_given && a.keep_pose == b.keep_pose;
}

// Every world launches on one process-wide stream. The runtime keeps, per
// hardware queue, scratch memory sized for the largest kernel that queue has
// run; worlds on streams of their own spread over several queues and their
// retained scratch adds up until a later allocation fails
// (HSA_STATUS_ERROR_OUT_OF_RESOURCES). Calls on different worlds therefore
// run one after the other on the device. A world's views use streams 1..3
// of the same pool (the kernels they run keep no large private memory: see
// build_views).
static hipStream_t pool_stream(int k) {
    static std::mutex mu;
    static hipStream_t st[kMaxViews] = {};
    std::lock_guard<std::mutex> lock(mu);
    if (!st[k] && hipStreamCreate(&st[k]) != hipSuccess) st[k] = nullptr;
    return st[k];
}
static hipStream_t shared_stream() { return pool_stream(0); }

// Views for graph replay: with a grouped narrow schedule whose kernels keep
// no large private memory (no one-lane kernels, no sphere or compound pairs)
// and enough scenes, the world is cut into up to kMaxViews scene ranges of
// about equal body counts, each replayed on a stream of its own. A scene's
// work never touches another's, so the views only run concurrently: while
// one view's latency-bound solve runs, another's narrow phase fills the
// SIMDs. Otherwise one view: the whole world on the main stream. Worlds of
// 64 scenes or more get kAutoViews views unless set otherwise.
static bool build_views(rp_world* w) {
    const DevWorld& d = w->im.d;
    DeviceImage& im = w->im;
    int K = 1;
    if (d.narrow_group > 1 && d.manifold_group != 1 && !d.has_general) {
        int want = w->want_views > 0 ? w->want_views : (d.ns >= 64 ? kAutoViews : 1);
        K = std::max(1, std::min({want, kMaxViews, d.ns}));
    }
    for (int k = 0; k < K; ++k) {
        w->view_stream[k] = k == 0 ? w->stream : pool_stream(k);
        if (!w->view_stream[k]) { set_msg("hipStreamCreate failed"); return false; }
    }
    if (K == 1) {
        w->view[0] = d;
        w->view_big[0] = im.big;
        w->nviews = 1;
        return true;
    }
    // scene cuts at about equal body counts
    std::vector<int> cut(K + 1, 0);
    cut[K] = d.ns;
    for (int k = 1; k < K; ++k) {
        int64_t target = (int64_t)d.nb * k / K;
        int s = cut[k - 1];
        while (s < d.ns && w->hw.scene_body_off[s] < target) ++s;
        cut[k] = std::max(s, cut[k - 1]);
    }
    int64_t slots = 0;
    for (int k = 0; k < K; ++k) {
        int s0 = cut[k], s1 = cut[k + 1];
        DevWorld v = d;
        int b0 = w->hw.scene_body_off[s0], b1 = w->hw.scene_body_off[s1];
        int64_t P0 = im.pair_base[s0], P1 = im.pair_base[s1];
        v.ns = s1 - s0;
        v.body_base = b0;
        v.nb = b1 - b0;
        v.n_ext = w->hw.scene_ext_off[s1] - w->hw.scene_ext_off[s0];
        v.scene_body_off += s0; v.scene_pair_base += s0; v.scene_ext_off += s0; v.scene_item_base += s0;
        v.pair_count += s0; v.live_count += s0; v.live_prefix += s0 + k;
        v.scene_nlevels += s0; v.scene_big += s0; v.scene_err += s0;
        v.item_lstart += s0; v.item_cursor += s0;
        v.live_global += P0; v.live_rec += P0; v.live_hull += P0; v.ncontact += P0; v.pair_normal += P0; v.mlist += P0; v.pair_lists += P0;
        v.contacts += P0 * MAXC;
        v.counters += CNT_COUNT * k;
        v.xf += im.body_xf_off[b0];
        v.n_xf = im.body_xf_off[b1] - im.body_xf_off[b0];
        v.narrow_blocks = std::min(narrow_blocks_for(v, v.narrow_group, P1 - P0), std::max(1, d.narrow_blocks / K));
        v.manifold_blocks = std::min(narrow_blocks_for(v, v.manifold_group ? v.manifold_group : 1, P1 - P0),
                                     std::max(1, d.manifold_blocks / K));
        v.group_work = nullptr;
        w->view[k] = v;
        w->view_big[k].clear();
        for (const BigScene& bs : im.big)
            if (bs.s >= s0 && bs.s < s1) {
                BigScene c = bs;
                c.s -= s0;
                w->view_big[k].push_back(c);
            }
        slots += group_work_slots_needed(v);
    }
    if (!im.ensure_group_work(slots)) return false;
    int64_t at = 0;
    for (int k = 0; k < K; ++k) {
        w->view[k].group_work = im.d.group_work + (size_t)at * GROUP_WORK_BYTES;
        w->view[k].group_work_slots = group_work_slots_needed(w->view[k]);
        at += w->view[k].group_work_slots;
    }
    w->nviews = K;
    return true;
}

static void drop_graphs(rp_world* w) {
    for (hipGraphExec_t& e : w->view_exec)
        if (e) {
            (void)hipGraphExecDestroy(e);
            e = nullptr;
        }
}

// The graphs of one frame for these parameters and the current schedules,
// captured once per view and reused while they do not change.
static hipError_t frame_graph(rp_world* w, const FrameParams& fp) {
    if (w->nviews > 0 && w->view_exec[0] && same_params(w->graph_fp, fp) && w->graph_gen == w->sched_gen)
        return hipSuccess;
    (void)hipStreamSynchronize(w->stream);
    for (int k = 1; k < w->nviews; ++k) (void)hipStreamSynchronize(w->view_stream[k]);
    drop_graphs(w);
    if (!build_views(w)) return hipErrorOutOfMemory;
    for (int k = 0; k < w->nviews; ++k) {
        hipStream_t st = w->view_stream[k];
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return e;
        hipError_t le = launch_frame(w->view[k], fp, st, nullptr, w->view_big[k].data(), (int)w->view_big[k].size());
        e = hipStreamEndCapture(st, &g);
        if (le != hipSuccess) e = le;
        if (e == hipSuccess) e = hipGraphInstantiate(&w->view_exec[k], g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            drop_graphs(w);
            return e;
        }
    }
    w->graph_fp = fp;
    w->graph_gen = w->sched_gen;
    return hipSuccess;
}

static bool world_init(rp_world* w) {
    if (hipGetDevice(&w->device) != hipSuccess) {
        set_msg("hipGetDevice failed");
        return false;
    }
    w->stream = shared_stream();
    if (!w->stream) {
        set_msg("hipStreamCreate failed");
        return false;
    }
    bool ok = hipEventCreateWithFlags(&w->fork_ev, hipEventDisableTiming) == hipSuccess;
    for (int k = 0; k < kMaxViews; ++k) {
        ok &= hipEventCreateWithFlags(&w->join_ev[k], hipEventDisableTiming) == hipSuccess;
        for (int c = 0; c < 2; ++c)
            ok &= hipEventCreateWithFlags(&w->chunk_done[c][k], hipEventDisableTiming) == hipSuccess;
    }
    if (!ok) {
        set_msg("hipEventCreate failed");
        return false;
    }
    return build_image(w->hw, w->im);
}

static void world_free(rp_world* w) {
    if (!w) return;
    if (w->stream) (void)hipStreamSynchronize(w->stream);
    for (int k = 1; k < w->nviews; ++k) (void)hipStreamSynchronize(w->view_stream[k]);
    drop_graphs(w);
    adjoint_free(w);
    for (int k = 0; k < kMaxViews; ++k) {
        if (w->join_ev[k]) (void)hipEventDestroy(w->join_ev[k]);
        for (int c = 0; c < 2; ++c)
            if (w->chunk_done[c][k]) (void)hipEventDestroy(w->chunk_done[c][k]);
    }
    if (w->fork_ev) (void)hipEventDestroy(w->fork_ev);
    for (hipEvent_t e : w->timer.pool) (void)hipEventDestroy(e);
    w->im.release();
    delete w;
}

// Keeps the stream's queue bounded: after each chunk of frames an event is
// recorded, and the host waits for the chunk before the previous one, so at
// most two chunks are queued while the device always has one to run.
static hipError_t chunk_mark(rp_world* w, uint32_t f, int view, hipStream_t st) {
    if ((f + 1) % kFramesPerChunk != 0) return hipSuccess;
    uint32_t k = (f + 1) / kFramesPerChunk;
    hipError_t e = hipSuccess;
    if (k >= 2) e = hipEventSynchronize(w->chunk_done[k & 1][view]);
    if (e == hipSuccess) e = hipEventRecord(w->chunk_done[k & 1][view], st);
    return e;
}

static uint32_t world_run(rp_world* w, const FrameParams& fp, uint32_t frames, bool sync) {
    w->last_collisions = fp.collisions;
    if (w->use_graphs && !w->timer.on && frames > 0) {
        hipError_t e = frame_graph(w, fp);
        // the views start after the main stream's earlier work and the main
        // stream continues after theirs
        if (e == hipSuccess && w->nviews > 1) {
            e = hipEventRecord(w->fork_ev, w->stream);
            for (int k = 1; e == hipSuccess && k < w->nviews; ++k) e = hipStreamWaitEvent(w->view_stream[k], w->fork_ev, 0);
        }
        for (uint32_t f = 0; e == hipSuccess && f < frames; ++f) {
            for (int k = 0; e == hipSuccess && k < w->nviews; ++k) {
                e = hipGraphLaunch(w->view_exec[k], w->view_stream[k]);
                if (e == hipSuccess) e = chunk_mark(w, f, k, w->view_stream[k]);
            }
        }
        for (int k = 1; e == hipSuccess && k < w->nviews; ++k) {
            e = hipEventRecord(w->join_ev[k], w->view_stream[k]);
            if (e == hipSuccess) e = hipStreamWaitEvent(w->stream, w->join_ev[k], 0);
        }
        if (e != hipSuccess) {
            set_msg(std::string("frame graph: ") + hipGetErrorString(e));
            return RP_ERR_RUNTIME;
        }
        frames = 0;
    }
    for (uint32_t f = 0; f < frames; ++f) {
        w->timer.used = 0;
        hipError_t e = launch_frame(w->im.d, fp, w->stream, &w->timer, w->im.big.data(), (int)w->im.big.size());
        if (e == hipSuccess && !w->timer.on) e = chunk_mark(w, f, 0, w->stream);
        if (e != hipSuccess) {
            set_msg(std::string("kernel launch: ") + hipGetErrorString(e));
            return RP_ERR_RUNTIME;
        }
        if (w->timer.on) {  // profiling: wait for the frame, then read its kernel times
            if (hipStreamSynchronize(w->stream) != hipSuccess) return RP_ERR_RUNTIME;
            for (size_t k = 0; k < w->timer.used; ++k) {
                float ms = 0.0f;
                if (hipEventElapsedTime(&ms, w->timer.pool[2 * k], w->timer.pool[2 * k + 1]) == hipSuccess) {
                    w->kc_ms[w->timer.cls[k]] += ms;
                    ++w->kc_launches[w->timer.cls[k]];
                }
            }
        }
    }
    if (sync) {
        hipError_t e = hipStreamSynchronize(w->stream);
        if (e != hipSuccess) {
            set_msg(std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
            return RP_ERR_RUNTIME;
        }
    }
    return 0;
}

static bool download_bodies(rp_world* w) {
    w->host_bodies.resize(w->hw.body.size());
    if (w->host_bodies.empty()) return true;
    RP_CHECK(hipStreamSynchronize(w->stream));
    RP_CHECK(hipMemcpy(w->host_bodies.data(), w->im.d.body, w->host_bodies.size() * sizeof(Body), hipMemcpyDeviceToHost));
    return true;
}

static bool upload_bodies(rp_world* w) {
    if (w->host_bodies.empty()) return true;
    RP_CHECK(hipMemcpy(w->im.d.body, w->host_bodies.data(), w->host_bodies.size() * sizeof(Body), hipMemcpyHostToDevice));
    return true;
}

extern "C" {

rp_world* rp_world_create(const rp_world_desc* desc) {
    if (!desc) { set_msg("null descriptor"); return nullptr; }
    rp_world* w = new rp_world();
    if (!host_world_from_desc(desc, w->hw) || !world_init(w)) {
        world_free(w);
        return nullptr;
    }
    return w;
}

void rp_world_destroy(rp_world* w) { world_free(w); }
uint32_t rp_world_num_bodies(const rp_world* w) { return (uint32_t)w->hw.body.size(); }
uint32_t rp_world_num_scenes(const rp_world* w) { return (uint32_t)w->hw.scene_body_off.size() - 1; }
void* rp_world_stream(rp_world* w) { return (void*)w->stream; }

static uint32_t step_impl(rp_world* w, double dt, uint32_t frames, uint32_t substeps, uint32_t iters, int32_t coll,
                          double gravity, bool sync) {
    if (!(dt > 0.0)) return 0;  // pbd.cpp:471 (dt <= 0 is a no-op; NaN too)
    FrameParams fp;
    fp.dt = dt;
    fp.h = dt / substeps;
    fp.gravity = gravity;
    fp.substeps = substeps;
    fp.iters = iters;
    fp.collisions = coll != 0;
    return world_run(w, fp, frames, sync);
}

uint32_t rp_world_step(rp_world* w, double dt, uint32_t frames, uint32_t substeps, uint32_t iters, int32_t coll,
                       double gravity) {
    return step_impl(w, dt, frames, substeps, iters, coll, gravity, true);
}
uint32_t rp_world_step_async(rp_world* w, double dt, uint32_t frames, uint32_t substeps, uint32_t iters, int32_t coll,
                             double gravity) {
    return step_impl(w, dt, frames, substeps, iters, coll, gravity, false);
}
uint32_t rp_world_step_noalloc(rp_world* w, double dt, uint32_t frames, uint32_t substeps, uint32_t iters,
                               int32_t coll, double gravity) {
    g_last_err = 0;
    if (!(dt > 0.0)) return 0;  // pbd.cpp:471
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != w->device) {
        set_msg("rp_world_step_noalloc: the world was created on another device");
        return g_last_err = RP_ERR_RUNTIME;
    }
    unsigned char* ptr = noalloc_scratch();
    if (!ptr) {
        set_msg("rp_world_step_noalloc: hipGetSymbolAddress failed");
        return g_last_err = RP_ERR_RUNTIME;
    }
    // the world's kernels with the lane-group working memory in the static
    // array; the narrow-phase grids are capped to the groups it holds (lane
    // groups take the live pairs in grid-stride order: identical results)
    DevWorld d = w->im.d;
    int64_t slots = (int64_t)(NOALLOC_SCRATCH_BYTES / GROUP_WORK_BYTES);
    if (d.narrow_group > 1) d.narrow_blocks = (int)std::min<int64_t>(d.narrow_blocks, slots / (64 / d.narrow_group));
    if (d.manifold_group > 1)
        d.manifold_blocks = (int)std::min<int64_t>(d.manifold_blocks, slots / (64 / d.manifold_group));
    if (d.narrow_blocks < 1 || d.manifold_blocks < 1 || group_work_slots_needed(d) > slots) {
        set_msg("rp_world_step_noalloc: the schedule needs more lane-group working memory than the static array holds");
        return g_last_err = RP_ERR_RUNTIME;
    }
    d.group_work = (char*)ptr;
    d.group_work_slots = slots;
    FrameParams fp;
    fp.dt = dt;
    fp.h = dt / substeps;
    fp.gravity = gravity;
    fp.substeps = substeps;
    fp.iters = iters;
    fp.collisions = coll != 0;
    w->last_collisions = fp.collisions;
    for (uint32_t f = 0; f < frames; ++f) {
        hipError_t e = launch_frame(d, fp, w->stream, nullptr, w->im.big.data(), (int)w->im.big.size());
        if (e == hipSuccess) e = chunk_mark(w, f, 0, w->stream);
        if (e != hipSuccess) {
            set_msg(std::string("kernel launch: ") + hipGetErrorString(e));
            return g_last_err = RP_ERR_RUNTIME;
        }
    }
    hipError_t e = hipStreamSynchronize(w->stream);
    if (e != hipSuccess) {
        set_msg(std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
        return g_last_err = RP_ERR_RUNTIME;
    }
    return 0;
}

uint32_t rp_world_synchronize(rp_world* w) {
    hipError_t e = hipStreamSynchronize(w->stream);
    if (e != hipSuccess) { set_msg(hipGetErrorString(e)); return RP_ERR_RUNTIME; }
    return 0;
}

static void put3(double* p, size_t i, V3 v) { if (p) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; } }
static void put4(double* p, size_t i, Q4 q) { if (p) { p[4 * i] = q.x; p[4 * i + 1] = q.y; p[4 * i + 2] = q.z; p[4 * i + 3] = q.w; } }
static V3 get3(const double* p, size_t i) { return v3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
static Q4 get4(const double* p, size_t i) { Q4 q = {p[4 * i], p[4 * i + 1], p[4 * i + 2], p[4 * i + 3]}; return q; }

void rp_world_get_state(rp_world* w, const rp_state* o) {
    if (!download_bodies(w)) return;
    for (size_t i = 0; i < w->host_bodies.size(); ++i) {
        const Body& e = w->host_bodies[i];
        put3(o->position, i, e.x);
        put4(o->rotation, i, e.q);
        put3(o->linear_velocity, i, e.v);
        put3(o->angular_velocity, i, e.w);
        put3(o->previous_position, i, e.px);
        put4(o->previous_rotation, i, e.pq);
        put3(o->previous_linear_velocity, i, e.pv);
        put3(o->previous_angular_velocity, i, e.pw);
        if (o->active) o->active[i] = e.active;
        if (o->deactivation_time) o->deactivation_time[i] = e.deact;
    }
}

void rp_world_set_state(rp_world* w, const rp_state* in) {
    if (!download_bodies(w)) return;
    for (size_t i = 0; i < w->host_bodies.size(); ++i) {
        Body& e = w->host_bodies[i];
        if (in->position) e.x = get3(in->position, i);
        if (in->rotation) e.q = get4(in->rotation, i);
        if (in->linear_velocity) e.v = get3(in->linear_velocity, i);
        if (in->angular_velocity) e.w = get3(in->angular_velocity, i);
        if (in->previous_position) e.px = get3(in->previous_position, i);
        if (in->previous_rotation) e.pq = get4(in->previous_rotation, i);
        if (in->previous_linear_velocity) e.pv = get3(in->previous_linear_velocity, i);
        if (in->previous_angular_velocity) e.pw = get3(in->previous_angular_velocity, i);
        if (in->active) e.active = in->active[i];
        if (in->deactivation_time) e.deact = in->deactivation_time[i];
    }
    upload_bodies(w);
}

uint32_t rp_world_error(rp_world* w, uint32_t scene) {
    if (scene >= rp_world_num_scenes(w)) return 0;
    uint32_t e = 0;
    if (hipStreamSynchronize(w->stream) != hipSuccess ||
        hipMemcpy(&e, w->im.d.scene_err + scene, 4, hipMemcpyDeviceToHost) != hipSuccess)
        return RP_ERR_RUNTIME;
    return e;
}

uint32_t rp_world_errors(rp_world* w, uint32_t* out) {
    uint32_t ns = rp_world_num_scenes(w);
    if (!out || ns == 0) return 0;
    if (hipStreamSynchronize(w->stream) != hipSuccess ||
        hipMemcpy(out, w->im.d.scene_err, (size_t)ns * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return RP_ERR_RUNTIME;
    return 0;
}

static uint64_t read_stat(rp_world* w, int k) {
    unsigned long long v = 0;
    if (hipStreamSynchronize(w->stream) != hipSuccess) return 0;
    if (hipMemcpy(&v, w->im.d.stats + k, 8, hipMemcpyDeviceToHost) != hipSuccess) return 0;
    return v;
}
uint64_t rp_world_epa_failures(rp_world* w) { return read_stat(w, STAT_EPA_FAIL); }
uint64_t rp_world_gjk_cycle_exits(rp_world* w) { return read_stat(w, STAT_GJK_CYCLE); }

uint32_t rp_world_model_matrices(rp_world* w, void* out, int32_t as_float, int32_t out_on_device) {
    const DevWorld& d = w->im.d;
    if (!out || d.nb == 0) return 0;
    size_t bytes = (size_t)d.nb * 16 * (as_float ? sizeof(float) : sizeof(double));
    if (out_on_device) {
        if (launch_model_matrices(d, out, as_float != 0, w->stream) != hipSuccess) return RP_ERR_RUNTIME;
        return 0;
    }
    void* tmp = nullptr;
    uint32_t err = 0;
    if (hipMalloc(&tmp, bytes) != hipSuccess || launch_model_matrices(d, tmp, as_float != 0, w->stream) != hipSuccess ||
        hipMemcpyAsync(out, tmp, bytes, hipMemcpyDeviceToHost, w->stream) != hipSuccess ||
        hipStreamSynchronize(w->stream) != hipSuccess) {
        set_msg("model matrices: device error");
        err = RP_ERR_RUNTIME;
    }
    if (tmp) (void)hipFree(tmp);
    return err;
}

void rp_world_body_info(rp_world* w, double* inv_mass, double* inertia, double* inv_inertia, double* radius) {
    for (size_t i = 0; i < w->hw.body.size(); ++i) {
        const Body& e = w->hw.body[i];
        if (inv_mass) inv_mass[i] = e.inv_mass;
        if (inertia) memcpy(inertia + 9 * i, e.I.m, 72);
        if (inv_inertia) memcpy(inv_inertia + 9 * i, e.Iinv.m, 72);
        if (radius) radius[i] = e.radius;
    }
}

uint32_t rp_world_broad_pairs(rp_world* w, uint32_t scene, uint32_t* out, uint32_t cap) {
    const DevWorld& d = w->im.d;
    if ((int)scene >= d.ns) return 0;
    if (launch_broad(d, w->stream, w->im.big.data(), (int)w->im.big.size()) != hipSuccess ||
        hipStreamSynchronize(w->stream) != hipSuccess)
        return 0;
    int32_t n = 0, base = 0;
    (void)hipMemcpy(&n, d.pair_count + scene, 4, hipMemcpyDeviceToHost);
    (void)hipMemcpy(&base, d.scene_pair_base + scene, 4, hipMemcpyDeviceToHost);
    std::vector<int2> pr(std::max(n, 1));
    if (n > 0) (void)hipMemcpy(pr.data(), d.pairs + base, n * sizeof(int2), hipMemcpyDeviceToHost);
    for (int32_t k = 0; k < n && (uint32_t)k < cap; ++k) {
        out[2 * k] = (uint32_t)pr[k].x;
        out[2 * k + 1] = (uint32_t)pr[k].y;
    }
    return (uint32_t)n;
}

uint32_t rp_world_item_classes(rp_world* w, uint32_t scene, int32_t* out, uint32_t cap) {
    const DevWorld& d = w->im.d;
    const BigScene* bs = nullptr;
    for (const BigScene& b : w->im.big)
        if (b.s == (int)scene) bs = &b;
    if (!bs || (int)scene >= d.ns) return 0;
    if (hipStreamSynchronize(w->stream) != hipSuccess) return 0;
    int32_t live = 0;
    (void)hipMemcpy(&live, d.live_count + scene, 4, hipMemcpyDeviceToHost);
    int64_t next = w->hw.scene_ext_off[scene + 1] - w->hw.scene_ext_off[scene];
    int64_t K = std::min<int64_t>(next + (w->last_collisions ? live : 0), bs->item_cap);
    if (out && K > 0)
        (void)hipMemcpy(out, d.item_level + bs->item_base, std::min<int64_t>(K, cap) * 4, hipMemcpyDeviceToHost);
    return (uint32_t)K;
}

uint32_t rp_world_contacts(rp_world* w, uint32_t scene, uint32_t i, uint32_t j, double* out, uint32_t cap) {
    const DevWorld& d = w->im.d;
    if ((int)scene >= d.ns) return 0;
    int b0 = w->hw.scene_body_off[scene], n = w->hw.scene_body_off[scene + 1] - b0;
    if ((int)i >= n || (int)j >= n) return 0;
    double* dout = nullptr;
    int* dcount = nullptr;
    uint32_t* derr = nullptr;
    uint32_t nc = 0;
    if (hipMalloc(&dout, 9 * MAXC * sizeof(double)) == hipSuccess && hipMalloc(&dcount, 4) == hipSuccess &&
        hipMalloc(&derr, 4) == hipSuccess &&
        launch_pair_contacts(d, b0 + (int)i, b0 + (int)j, dout, dcount, derr, w->stream) == hipSuccess &&
        hipStreamSynchronize(w->stream) == hipSuccess) {
        int c = 0;
        uint32_t e = 0;
        double buf[9 * MAXC];
        (void)hipMemcpy(&c, dcount, 4, hipMemcpyDeviceToHost);
        (void)hipMemcpy(&e, derr, 4, hipMemcpyDeviceToHost);
        (void)hipMemcpy(buf, dout, sizeof(buf), hipMemcpyDeviceToHost);
        nc = (uint32_t)c;
        for (uint32_t k = 0; k < nc && k < cap; ++k) memcpy(out + 9 * k, buf + 9 * k, 72);
        g_last_err = e;
    }
    if (dout) (void)hipFree(dout);
    if (dcount) (void)hipFree(dcount);
    if (derr) (void)hipFree(derr);
    return nc;
}

void rp_world_hull(rp_world* w, uint32_t body, uint32_t counts[6], double* vertices, double* face_normals,
                   uint32_t* face_off, uint32_t* face_el, uint32_t* vf_off, uint32_t* vf, uint32_t* vn_off,
                   uint32_t* vn, uint32_t* ff_off, uint32_t* ff) {
    memset(counts, 0, 6 * sizeof(uint32_t));
    if (body >= w->hw.body.size() || w->hw.body_hull[body] < 0) return;
    const HostHull& h = w->hw.hulls[w->hw.body_hull[body]];  // the body's first collider
    counts[0] = (uint32_t)h.v.size(); counts[1] = (uint32_t)h.fn.size();
    counts[2] = (uint32_t)h.fe.size(); counts[3] = (uint32_t)h.vf.size();
    counts[4] = (uint32_t)h.vn.size(); counts[5] = (uint32_t)h.ff.size();
    auto cp = [](uint32_t* dst, const std::vector<int32_t>& src) { if (dst) for (size_t i = 0; i < src.size(); ++i) dst[i] = (uint32_t)src[i]; };
    if (vertices) memcpy(vertices, h.v.data(), h.v.size() * 24);
    if (face_normals) memcpy(face_normals, h.fn.data(), h.fn.size() * 24);
    cp(face_off, h.fe_off); cp(face_el, h.fe); cp(vf_off, h.vf_off); cp(vf, h.vf);
    cp(vn_off, h.vn_off); cp(vn, h.vn); cp(ff_off, h.ff_off); cp(ff, h.ff);
}

uint32_t rp_fp64_selftest(const double* a, const double* b, double* out, uint64_t n) {
    double *da = nullptr, *db = nullptr, *dout = nullptr;
    uint32_t rc = RP_ERR_RUNTIME;
    size_t bytes = std::max<uint64_t>(n, 1) * sizeof(double);
    if (hipMalloc(&da, bytes) == hipSuccess && hipMalloc(&db, bytes) == hipSuccess &&
        hipMalloc(&dout, 5 * bytes) == hipSuccess && hipMemcpy(da, a, n * 8, hipMemcpyHostToDevice) == hipSuccess &&
        hipMemcpy(db, b, n * 8, hipMemcpyHostToDevice) == hipSuccess &&
        launch_fp64_selftest(da, db, dout, (int64_t)n, nullptr) == hipSuccess && hipDeviceSynchronize() == hipSuccess &&
        hipMemcpy(out, dout, 5 * n * 8, hipMemcpyDeviceToHost) == hipSuccess)
        rc = 0;
    if (da) (void)hipFree(da);
    if (db) (void)hipFree(db);
    if (dout) (void)hipFree(dout);
    return rc;
}

uint64_t rp_world_adjoint_mismatches(rp_world* w) { return read_stat(w, STAT_ADJ_MISMATCH); }

uint32_t rp_world_set_body_params(rp_world* w, const double* mass, const double* mu_s, const double* mu_d) {
    if (!download_bodies(w)) return RP_ERR_RUNTIME;
    uint32_t err = 0;
    for (size_t b = 0; b < w->host_bodies.size(); ++b) {
        Body& e = w->host_bodies[b];
        Body& h = w->hw.body[b];
        if (mu_s) e.mu_s = h.mu_s = mu_s[b];
        if (mu_d) e.mu_d = h.mu_d = mu_d[b];
        if (mass && !e.fixed) {
            // as at creation (make_body): 1/m, the hull's inertia and its inverse
            int c0 = w->hw.body_col_off[b], nc = w->hw.body_col_off[b + 1] - c0;
            e.inv_mass = h.inv_mass = 1.0 / mass[b];
            e.I = h.I = colliders_inertia(w->hw.hulls, w->hw.cols.data() + c0, nc, mass[b], err);
            if (!m3_inverse(e.I, e.Iinv)) err |= ERR_HULL;
            h.Iinv = e.Iinv;
        }
    }
    if (!upload_bodies(w)) return RP_ERR_RUNTIME;
    return err;
}

uint32_t rp_world_adjoint_forward(rp_world* w, double dt, uint32_t frames, uint32_t substeps, uint32_t iters,
                                  double gravity) {
    adjoint_free(w);
    if (dt <= 0.0 || substeps == 0 || frames == 0) {
        set_msg("adjoint_forward: dt, substeps and frames must be positive");
        return RP_ERR_SELECT;
    }
    // scope: small scenes with one dynamic body and no external constraints
    const DevWorld& d = w->im.d;
    std::vector<int32_t> dyn(d.ns, -1);
    for (int s = 0; s < d.ns; ++s) {
        int nd = 0;
        for (int b = w->hw.scene_body_off[s]; b < w->hw.scene_body_off[s + 1]; ++b)
            if (!w->hw.body[b].fixed) { dyn[s] = b; ++nd; }
        bool big = false;
        for (const BigScene& bs : w->im.big) big |= bs.s == s;
        if (nd > 1 || big || w->hw.scene_ext_off[s + 1] != w->hw.scene_ext_off[s]) {
            set_msg("adjoint: a scene has more than one dynamic body, external constraints, or is a large scene");
            return RP_ERR_ADJOINT_SCOPE;
        }
    }
    size_t nb = (size_t)d.nb;
    if (hipMalloc(&w->ckpt, (frames + 1) * nb * sizeof(Body)) != hipSuccess ||
        hipMalloc(&w->scene_dyn, std::max<size_t>(1, dyn.size()) * 4) != hipSuccess ||
        hipMemcpy(w->scene_dyn, dyn.data(), dyn.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
        set_msg("adjoint_forward: device allocation failed");
        adjoint_free(w);
        return RP_ERR_RUNTIME;
    }
    FrameParams fp;
    fp.dt = dt;
    fp.h = dt / substeps;
    fp.gravity = gravity;
    fp.substeps = substeps;
    fp.iters = iters;
    fp.collisions = true;
    for (uint32_t f = 0; f < frames; ++f) {
        if (hipMemcpyAsync(w->ckpt + f * nb, d.body, nb * sizeof(Body), hipMemcpyDeviceToDevice, w->stream) !=
            hipSuccess)
            return RP_ERR_RUNTIME;
        uint32_t e = world_run(w, fp, 1, false);
        if (e) return e;
    }
    if (hipMemcpyAsync(w->ckpt + frames * nb, d.body, nb * sizeof(Body), hipMemcpyDeviceToDevice, w->stream) !=
            hipSuccess ||
        hipStreamSynchronize(w->stream) != hipSuccess)
        return RP_ERR_RUNTIME;
    w->ck_frames = frames;
    w->ck_fp = fp;
    return 0;
}

uint32_t rp_world_adjoint_backward(rp_world* w, const double* dl_dstate, double* grad_state, double* grad_param) {
    if (!w->ckpt || w->ck_frames == 0) {
        set_msg("adjoint_backward: no recorded forward pass");
        return RP_ERR_SELECT;
    }
    const DevWorld& d = w->im.d;
    size_t nb = (size_t)d.nb;
    const FrameParams& fp = w->ck_fp;
    if (!w->rec.body) {
        w->rec.pair_cap = std::max<int64_t>(1, w->im.pair_capacity);
        if (hipMalloc(&w->rec.body, fp.substeps * nb * sizeof(Body)) != hipSuccess ||
            hipMalloc(&w->rec.ncontact, fp.substeps * w->rec.pair_cap * sizeof(int32_t)) != hipSuccess ||
            hipMalloc(&w->rec.contacts, fp.substeps * w->rec.pair_cap * MAXC * sizeof(Contact)) != hipSuccess ||
            hipMalloc(&w->abar, nb * 13 * sizeof(double)) != hipSuccess ||
            hipMalloc(&w->gparam, nb * 3 * sizeof(double)) != hipSuccess ||
            hipMalloc(&w->amax, nb * sizeof(double)) != hipSuccess) {
            set_msg("adjoint_backward: device allocation failed");
            return RP_ERR_RUNTIME;
        }
    }
    w->dl_norm.assign(nb, 0.0);
    for (size_t i = 0; i < nb; ++i)
        for (int j = 0; j < 13; ++j) w->dl_norm[i] = std::max(w->dl_norm[i], std::fabs(dl_dstate[13 * i + j]));
    if (hipMemcpy(w->abar, dl_dstate, nb * 13 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(w->amax, w->dl_norm.data(), nb * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(w->gparam, 0, nb * 3 * sizeof(double)) != hipSuccess)
        return RP_ERR_RUNTIME;
    for (int f = (int)w->ck_frames - 1; f >= 0; --f) {
        hipError_t e = hipMemcpyAsync(d.body, w->ckpt + f * nb, nb * sizeof(Body), hipMemcpyDeviceToDevice, w->stream);
        if (e == hipSuccess)
            e = launch_frame(d, fp, w->stream, nullptr, w->im.big.data(), (int)w->im.big.size(), &w->rec);
        if (e == hipSuccess)
            e = launch_adjoint_frame(d, w->rec, w->scene_dyn, w->ckpt + (f + 1) * nb, w->abar, w->gparam, w->amax, fp,
                                     w->stream);
        if (e != hipSuccess) {
            set_msg(std::string("adjoint_backward: ") + hipGetErrorString(e));
            return RP_ERR_RUNTIME;
        }
    }
    // the world continues from the end of the recorded run
    if (hipMemcpyAsync(d.body, w->ckpt + w->ck_frames * nb, nb * sizeof(Body), hipMemcpyDeviceToDevice, w->stream) !=
            hipSuccess ||
        hipStreamSynchronize(w->stream) != hipSuccess)
        return RP_ERR_RUNTIME;
    if ((grad_state && hipMemcpy(grad_state, w->abar, nb * 13 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) ||
        (grad_param && hipMemcpy(grad_param, w->gparam, nb * 3 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess))
        return RP_ERR_RUNTIME;
    return 0;
}

uint32_t rp_world_adjoint_amplification(rp_world* w, double* out) {
    size_t nb = w->hw.body.size();
    if (!w->amax || w->dl_norm.size() != nb) {
        set_msg("adjoint_amplification: no backward pass");
        return RP_ERR_SELECT;
    }
    std::vector<double> m(nb);
    if (hipMemcpy(m.data(), w->amax, nb * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return RP_ERR_RUNTIME;
    for (size_t i = 0; i < nb; ++i) out[i] = w->dl_norm[i] > 0.0 ? m[i] / w->dl_norm[i] : 0.0;
    return 0;
}

// Worlds whose lane-group slices need more than 64 KiB of LDS per block (the
// larger limit applies only where no lane group fits in 64 KiB) run lane
// groups in the fused GJK + EPA + manifold kernel only: the split kernels
// were never run with such slices. The one-lane split schedule (1, 1) uses
// no LDS and stays available.
static bool large_lds_split(const DevWorld& d, int narrow_group, int manifold_group) {
    if (d.lds_limit <= 64 * 1024) return false;
    return (manifold_group != 0 && narrow_group > 1) || manifold_group > 1;
}

int32_t rp_world_set_narrow_group(rp_world* w, int32_t lanes_per_pair) {
    if (!narrow_group_supported(w->im.d, lanes_per_pair)) return -1;
    if (large_lds_split(w->im.d, lanes_per_pair, w->im.d.manifold_group)) return -1;
    (void)hipStreamSynchronize(w->stream);
    int old_g = w->im.d.narrow_group, old_b = w->im.d.narrow_blocks;
    w->im.d.narrow_group = lanes_per_pair;
    w->im.d.narrow_blocks = narrow_blocks_for(w->im.d, lanes_per_pair, w->im.pair_capacity);
    if (!w->im.ensure_group_work()) {
        w->im.d.narrow_group = old_g;
        w->im.d.narrow_blocks = old_b;
        return -1;
    }
    ++w->sched_gen;
    return 0;
}

int32_t rp_world_narrow_group(const rp_world* w) { return w->im.d.narrow_group; }

int32_t rp_world_set_manifold_group(rp_world* w, int32_t lanes_per_pair) {
    if (lanes_per_pair != 0 && !narrow_group_supported(w->im.d, lanes_per_pair)) return -1;
    if (large_lds_split(w->im.d, w->im.d.narrow_group, lanes_per_pair)) return -1;
    (void)hipStreamSynchronize(w->stream);
    int old_g = w->im.d.manifold_group, old_b = w->im.d.manifold_blocks;
    w->im.d.manifold_group = lanes_per_pair;
    w->im.d.manifold_blocks = narrow_blocks_for(w->im.d, lanes_per_pair ? lanes_per_pair : 1, w->im.pair_capacity);
    if (!w->im.ensure_group_work()) {
        w->im.d.manifold_group = old_g;
        w->im.d.manifold_blocks = old_b;
        return -1;
    }
    ++w->sched_gen;
    return 0;
}

int32_t rp_world_manifold_group(const rp_world* w) { return w->im.d.manifold_group; }

int32_t rp_world_lds_limit(const rp_world* w) { return w->im.d.lds_limit; }

int32_t rp_world_set_narrow_block_cap(rp_world* w, int32_t blocks) {
    if (blocks < 0 || blocks > 256 * 32) return -1;
    (void)hipStreamSynchronize(w->stream);
    DevWorld& d = w->im.d;
    d.narrow_block_cap = blocks;
    d.narrow_blocks = narrow_blocks_for(d, d.narrow_group, w->im.pair_capacity);
    d.manifold_blocks = narrow_blocks_for(d, d.manifold_group ? d.manifold_group : 1, w->im.pair_capacity);
    if (!w->im.ensure_group_work()) return -1;
    ++w->sched_gen;
    return 0;
}

int32_t rp_world_set_solve_lanes(rp_world* w, int32_t lanes_per_scene) {
    if (lanes_per_scene != 16 && lanes_per_scene != 32 && lanes_per_scene != 64) return -1;
    (void)hipStreamSynchronize(w->stream);
    w->im.d.solve_lanes = lanes_per_scene;
    ++w->sched_gen;
    return 0;
}

int32_t rp_world_solve_lanes(const rp_world* w) { return w->im.d.solve_lanes; }

int32_t rp_world_set_colour_launches(rp_world* w, int32_t classes) {
    if (classes < 0 || classes > BIG_CLASSES) return -1;
    (void)hipStreamSynchronize(w->stream);
    w->im.d.big_launch_classes = classes;
    ++w->sched_gen;
    return 0;
}

int32_t rp_world_set_views(rp_world* w, int32_t views) {
    if (views < 0 || views > kMaxViews) return -1;
    w->want_views = views;
    ++w->sched_gen;
    return 0;
}

int32_t rp_world_views(const rp_world* w) { return w->nviews; }

void rp_world_set_graphs(rp_world* w, int32_t on) {
    (void)hipStreamSynchronize(w->stream);
    w->use_graphs = on != 0;
}

void rp_world_set_profiling(rp_world* w, int32_t on) {
    (void)hipStreamSynchronize(w->stream);
    w->timer.on = on != 0;
    for (int k = 0; k < KC_COUNT; ++k) { w->kc_ms[k] = 0.0; w->kc_launches[k] = 0; }
}

void rp_world_kernel_times(rp_world* w, double* ms, uint64_t* launches) {
    for (int k = 0; k < KC_COUNT; ++k) {
        if (ms) ms[k] = w->kc_ms[k];
        if (launches) launches[k] = w->kc_launches[k];
    }
}

uint32_t rp_last_error(void) { return g_last_err; }
const char* rp_last_error_string(void) { return g_last_msg.c_str(); }

}  // extern "C"

// ---------------------------------------------------------------------------
// Drop-in step over the reference's Entity structs
// ---------------------------------------------------------------------------
static_assert(sizeof(rp_entity) == 504, "Entity layout (src/entity.h:17-47)");
static_assert(offsetof(rp_entity, world_position) == 24, "Entity layout");
static_assert(offsetof(rp_entity, world_rotation) == 48, "Entity layout");
static_assert(offsetof(rp_entity, colliders) == 136, "Entity layout");
static_assert(offsetof(rp_entity, inverse_mass) == 160, "Entity layout");
static_assert(offsetof(rp_entity, angular_velocity) == 312, "Entity layout");
static_assert(offsetof(rp_entity, fixed) == 360, "Entity layout");
static_assert(offsetof(rp_entity, deactivation_time) == 368, "Entity layout");
static_assert(offsetof(rp_entity, previous_world_position) == 400, "Entity layout");
static_assert(offsetof(rp_entity, previous_angular_velocity) == 480, "Entity layout");
static_assert(sizeof(rp_constraint) == 144, "Constraint layout (src/physics/pbd.h:79-91)");
static_assert(sizeof(rp_collider) == 64, "Collider layout (src/physics/collider.h:38-44)");
static_assert(sizeof(rp_physics_force) == 52, "Physics_Force layout (src/entity.h:11-15)");

static size_t la_len(const void* p) { return p ? ((const size_t*)p)[-1] : 0; }  // light_array.h:87-89
static V3 to_v3(const rp_vec3& v) { return v3(v.x, v.y, v.z); }
static rp_vec3 from_v3(V3 v) { rp_vec3 r = {v.x, v.y, v.z}; return r; }

// The drop-in keeps one device world for the last entity set it saw and only
// re-uploads body records while the set (entities, colliders, constraint
// endpoints) is unchanged.
struct DropIn {
    std::mutex mu;
    rp_world* w = nullptr;
    std::vector<const void*> key;
};
static DropIn g_dropin;

static Body entity_body(const rp_entity* e);

static bool dropin_build(rp_entity** ents, size_t n, rp_constraint* cons, HostWorld& hw, uint32_t& err) {
    hw.scene_err.assign(1, 0);
    std::unordered_map<uint64_t, int32_t> id2idx;
    for (size_t i = 0; i < n; ++i) id2idx[ents[i]->id] = (int32_t)i;
    for (size_t i = 0; i < n; ++i) {
        const rp_entity* e = ents[i];
        size_t ncol = la_len(e->colliders);
        if (ncol == 0) {
            set_msg("an entity has no collider");
            return false;
        }
        int32_t first_hull = -1;
        for (size_t ci = 0; ci < ncol; ++ci) {
            const rp_collider& col = e->colliders[ci];
            if (col.type == RP_COLLIDER_SPHERE) {
                hw.cols.push_back(HostCollider{COL_SPHERE, -1, col.u.sphere.radius});
                continue;
            }
            if (col.type != RP_COLLIDER_CONVEX_HULL) {
                set_msg("unknown collider type");
                return false;
            }
            // the reference's own hull topology (collider.cpp:194-364)
            const rp_convex_hull& ch = col.u.convex_hull;
            HostHull h;
            size_t nv = la_len(ch.vertices), nf = la_len(ch.faces);
            for (size_t k = 0; k < nv; ++k) h.v.push_back(to_v3(ch.vertices[k]));
            std::vector<std::vector<int32_t>> fe(nf), vf(nv), vn(nv), ff(nf);
            for (size_t f = 0; f < nf; ++f) {
                h.fn.push_back(to_v3(ch.faces[f].normal));
                for (size_t k = 0; k < la_len(ch.faces[f].elements); ++k) fe[f].push_back((int32_t)ch.faces[f].elements[k]);
                for (size_t k = 0; k < la_len(ch.face_to_neighbors[f]); ++k)
                    ff[f].push_back((int32_t)ch.face_to_neighbors[f][k]);
            }
            for (size_t v = 0; v < nv; ++v) {
                for (size_t k = 0; k < la_len(ch.vertex_to_faces[v]); ++k) vf[v].push_back((int32_t)ch.vertex_to_faces[v][k]);
                for (size_t k = 0; k < la_len(ch.vertex_to_neighbors[v]); ++k)
                    vn[v].push_back((int32_t)ch.vertex_to_neighbors[v][k]);
            }
            flatten(fe, h.fe_off, h.fe);
            flatten(vf, h.vf_off, h.vf);
            flatten(vn, h.vn_off, h.vn);
            flatten(ff, h.ff_off, h.ff);
            hw.hulls.push_back(h);
            hw.cols.push_back(HostCollider{COL_HULL, (int32_t)hw.hulls.size() - 1, 0.0f});
            if (ci == 0) first_hull = (int32_t)hw.hulls.size() - 1;
        }
        hw.body_col_off.push_back((int32_t)hw.cols.size());
        hw.body_hull.push_back(first_hull);
        hw.body_scale.push_back(to_v3(e->world_scale));
        // the entity's record, so that what the image derives from it at
        // build time (a large scene's large-body set and grid cell, from the
        // bounding radii) matches what the step then reads
        hw.body.push_back(entity_body(e));
    }
    hw.scene_body_off.push_back((int32_t)n);
    for (size_t k = 0; k < la_len(cons); ++k) {
        const rp_constraint& c = cons[k];
        auto i1 = id2idx.find(c.e1_id), i2 = id2idx.find(c.e2_id);
        if (i1 == id2idx.end() || i2 == id2idx.end() || c.type == RP_COLLISION_CONSTRAINT || c.type < 0 || c.type > 4) {
            set_msg("external constraint endpoint not in the entity array");
            err |= ERR_SELECT;
            return false;
        }
        ExtC x;
        memset(&x, 0, sizeof(x));
        x.type = c.type;
        x.b1 = i1->second;
        x.b2 = i2->second;
        switch (c.type) {
        case RP_POSITIONAL_CONSTRAINT:
            x.r1 = to_v3(c.u.positional_constraint.r1_lc); x.r2 = to_v3(c.u.positional_constraint.r2_lc);
            x.compliance = c.u.positional_constraint.compliance; x.distance = to_v3(c.u.positional_constraint.distance);
            break;
        case RP_MUTUAL_ORIENTATION_CONSTRAINT:
            x.compliance = c.u.mutual_orientation_constraint.compliance;
            break;
        case RP_HINGE_JOINT_CONSTRAINT: {
            const rp_hinge_joint_constraint& hj = c.u.hinge_joint_constraint;
            x.r1 = to_v3(hj.r1_lc); x.r2 = to_v3(hj.r2_lc); x.compliance = hj.compliance;
            x.axis[0] = hj.e1_aligned_axis; x.axis[1] = hj.e2_aligned_axis;
            x.axis[2] = hj.e1_limit_axis; x.axis[3] = hj.e2_limit_axis;
            x.limited = hj.limited; x.limits[0] = hj.lower_limit; x.limits[1] = hj.upper_limit;
        } break;
        case RP_SPHERICAL_JOINT_CONSTRAINT: {
            const rp_spherical_joint_constraint& sj = c.u.spherical_joint_constraint;
            x.r1 = to_v3(sj.r1_lc); x.r2 = to_v3(sj.r2_lc);
            x.axis[0] = sj.e1_swing_axis; x.axis[1] = sj.e2_swing_axis;
            x.axis[2] = sj.e1_twist_axis; x.axis[3] = sj.e2_twist_axis;
            x.limits[0] = sj.swing_lower_limit; x.limits[1] = sj.swing_upper_limit;
            x.limits[2] = sj.twist_lower_limit; x.limits[3] = sj.twist_upper_limit;
        } break;
        }
        hw.ext.push_back(x);
    }
    hw.scene_ext_off.push_back((int32_t)hw.ext.size());
    return true;
}

// Entity -> body record, with the frame's force / torque sums
// (calculate_external_force/torque, physics_util.cpp:5-23).
static Body entity_body(const rp_entity* e) {
    Body b;
    memset(&b, 0, sizeof(b));
    b.x = to_v3(e->world_position);
    b.q = {e->world_rotation.x, e->world_rotation.y, e->world_rotation.z, e->world_rotation.w};
    b.v = to_v3(e->linear_velocity);
    b.w = to_v3(e->angular_velocity);
    b.px = to_v3(e->previous_world_position);
    b.pq = {e->previous_world_rotation.x, e->previous_world_rotation.y, e->previous_world_rotation.z,
            e->previous_world_rotation.w};
    b.pv = to_v3(e->previous_linear_velocity);
    b.pw = to_v3(e->previous_angular_velocity);
    b.deact = e->deactivation_time;
    b.inv_mass = e->inverse_mass;
    memcpy(b.I.m, e->inertia_tensor.data, 72);
    memcpy(b.Iinv.m, e->inverse_inertia_tensor.data, 72);
    b.mu_s = e->static_friction_coefficient;
    b.mu_d = e->dynamic_friction_coefficient;
    b.e = e->restitution_coefficient;
    b.radius = e->bounding_sphere_radius;
    b.fixed = e->fixed;
    b.active = e->active;
    V3 F = v3(0.0, 0.0, 0.0), T = v3(0.0, 0.0, 0.0);
    for (size_t i = 0; e->forces && i < la_len(e->forces); ++i) F = add(F, to_v3(e->forces[i].force));
    for (size_t i = 0; e->forces && i < la_len(e->forces); ++i) {
        V3 dist = sub(to_v3(e->forces[i].position), v3(0.0, 0.0, 0.0));
        T = add(T, cross(dist, to_v3(e->forces[i].force)));
    }
    b.F = F;
    b.T = T;
    return b;
}

// The collision pass's side effect on the entities' colliders (pbd.cpp:598-599
// -> collider_update, collider.cpp:409-438): both entities of every pair the
// pass visits get their colliders transformed to the pose they have after the
// substep's integration. The visited set is fixed for the call (activity does
// not change within it), so after the call each visited entity's colliders
// hold its pose after the last substep's integration, which the device kept;
// the others are left as they were.
static bool dropin_write_colliders(rp_world* w, rp_entity** ents, size_t n) {
    std::vector<int32_t> visited(n);
    std::vector<double> pose(7 * n);
    RP_CHECK(hipMemcpy(visited.data(), w->im.d.need, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    RP_CHECK(hipMemcpy(pose.data(), w->im.d.pose, n * 7 * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) {
        if (!visited[i]) continue;
        const double* p = pose.data() + 7 * i;
        V3 x = v3(p[0], p[1], p[2]);
        Q4 q = {p[3], p[4], p[5], p[6]};
        M4 M = model_matrix(x, q);  // util_get_model_matrix_no_scale (util.cpp:60-74)
        rp_collider* cols = ents[i]->colliders;
        for (size_t ci = 0; ci < la_len(cols); ++ci) {
            if (cols[ci].type == RP_COLLIDER_SPHERE) {
                cols[ci].u.sphere.center = from_v3(x);
                continue;
            }
            rp_convex_hull& ch = cols[ci].u.convex_hull;
            for (size_t k = 0; k < la_len(ch.transformed_vertices); ++k)
                ch.transformed_vertices[k] = from_v3(transform_vertex(M, to_v3(ch.vertices[k])));
            for (size_t k = 0; k < la_len(ch.transformed_faces); ++k)
                ch.transformed_faces[k].normal = from_v3(transform_normal(M, to_v3(ch.faces[k].normal)));
        }
    }
    return true;
}

extern "C" uint32_t rp_pbd_simulate(double dt, rp_entity** ents, rp_constraint* cons, uint32_t substeps,
                                    uint32_t iters, int32_t collisions) {
    g_last_err = 0;
    if (!(dt > 0.0) || !ents) return 0;  // pbd.cpp:471
    size_t n = la_len(ents);
    std::lock_guard<std::mutex> lock(g_dropin.mu);
    std::vector<const void*> key;
    for (size_t i = 0; i < n; ++i) {
        key.push_back(ents[i]);
        key.push_back(ents[i]->colliders);
        uint64_t rbits;  // a large scene's grid is built from the radii
        memcpy(&rbits, &ents[i]->bounding_sphere_radius, 8);
        key.push_back((const void*)(uintptr_t)rbits);
        for (size_t ci = 0; ci < la_len(ents[i]->colliders); ++ci) {
            const rp_collider& c = ents[i]->colliders[ci];
            uint32_t rbits = 0;
            if (c.type != RP_COLLIDER_CONVEX_HULL) memcpy(&rbits, &c.u.sphere.radius, 4);
            key.push_back(c.type == RP_COLLIDER_CONVEX_HULL ? (const void*)c.u.convex_hull.vertices
                                                            : (const void*)(((uintptr_t)rbits << 8) | 0xA5u));
        }
    }
    key.push_back((const void*)(uintptr_t)la_len(cons));
    for (size_t k = 0; k < la_len(cons); ++k) {
        key.push_back((const void*)(uintptr_t)cons[k].e1_id);
        key.push_back((const void*)(uintptr_t)cons[k].e2_id);
        key.push_back((const void*)(uintptr_t)cons[k].type);
    }
    uint32_t err = 0;
    if (!g_dropin.w || key != g_dropin.key) {
        if (g_dropin.w) world_free(g_dropin.w);
        g_dropin.w = new rp_world();
        g_dropin.key.clear();
        if (!dropin_build(ents, n, cons, g_dropin.w->hw, err) || !world_init(g_dropin.w)) {
            world_free(g_dropin.w);
            g_dropin.w = nullptr;
            g_last_err = err | RP_ERR_RUNTIME;
            return g_last_err;
        }
        g_dropin.key = key;
    }
    rp_world* w = g_dropin.w;
    // constraint payloads may change between calls (compliance, limits)
    if (la_len(cons)) {
        HostWorld tmp;
        uint32_t e2 = 0;
        if (dropin_build(ents, n, cons, tmp, e2) && !tmp.ext.empty())
            (void)hipMemcpy((void*)w->im.d.ext, tmp.ext.data(), tmp.ext.size() * sizeof(ExtC), hipMemcpyHostToDevice);
    }
    w->host_bodies.resize(n);
    for (size_t i = 0; i < n; ++i) w->host_bodies[i] = entity_body(ents[i]);
    if (!upload_bodies(w)) return g_last_err = RP_ERR_RUNTIME;
    (void)hipMemset(w->im.d.scene_err, 0, 4);
    FrameParams fp;
    fp.dt = dt;
    fp.h = dt / substeps;
    fp.gravity = 0.0;
    fp.substeps = substeps;
    fp.iters = iters;
    fp.collisions = collisions != 0;
    fp.forces_given = true;
    fp.keep_pose = fp.collisions && substeps > 0;
    if (world_run(w, fp, 1, true)) return g_last_err = RP_ERR_RUNTIME;
    if (!download_bodies(w)) return g_last_err = RP_ERR_RUNTIME;
    if (fp.keep_pose && !dropin_write_colliders(w, ents, n)) return g_last_err = RP_ERR_RUNTIME;
    for (size_t i = 0; i < n; ++i) {
        rp_entity* e = ents[i];
        const Body& b = w->host_bodies[i];
        e->world_position = from_v3(b.x);
        e->world_rotation = {b.q.x, b.q.y, b.q.z, b.q.w};
        e->linear_velocity = from_v3(b.v);
        e->angular_velocity = from_v3(b.w);
        e->previous_world_position = from_v3(b.px);
        e->previous_world_rotation = {b.pq.x, b.pq.y, b.pq.z, b.pq.w};
        e->previous_linear_velocity = from_v3(b.pv);
        e->previous_angular_velocity = from_v3(b.pw);
        e->active = b.active;
        e->deactivation_time = b.deact;
    }
    g_last_err = rp_world_error(w, 0);
    return g_last_err;
}

extern "C" void physics_simulate(rp_entity** entities, double dt) { rp_pbd_simulate(dt, entities, nullptr, 20, 1, 1); }

// The reference's own C++ entry points (src/physics/pbd.h:93-94).
struct Entity;
struct Constraint;
__attribute__((visibility("default"))) void pbd_simulate(double dt, Entity** entities, unsigned int num_substeps,
                                                         unsigned int num_pos_iters, int enable_collisions) {
    rp_pbd_simulate(dt, (rp_entity**)entities, nullptr, num_substeps, num_pos_iters, enable_collisions);
}
__attribute__((visibility("default"))) void pbd_simulate_with_constraints(double dt, Entity** entities,
                                                                          Constraint* external_constraints,
                                                                          unsigned int num_substeps,
                                                                          unsigned int num_pos_iters,
                                                                          int enable_collisions) {
    rp_pbd_simulate(dt, (rp_entity**)entities, (rp_constraint*)external_constraints, num_substeps, num_pos_iters,
                    enable_collisions);
}
