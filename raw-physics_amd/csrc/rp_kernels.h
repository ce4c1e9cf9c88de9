// rp_kernels.h — device-resident world layout and kernel launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <vector>
#include "rp_step.h"

namespace rp {

enum { STAT_EPA_FAIL = 0, STAT_GJK_CYCLE = 1, STAT_ADJ_MISMATCH = 2, STAT_COUNT = 4 };
enum { CNT_MANIFOLD = 0, CNT_SMALL = 1, CNT_BIG = 2, CNT_COUNT = 4 };
// Hulls of up to this many vertices are "small": the live pairs of two small
// hulls are staged with a per-pair LDS slice sized for them.
constexpr int SMALL_HULL_NV = 64;

// All device pointers of one world. Passed by value to every kernel.
// A world or a view of it: views (scene ranges of one world, launched on
// streams of their own) shift the scene-, pair- and queue-indexed pointers to
// their range and keep body, hull, pair-slot and item indices absolute;
// body_base is the first body of the range for the kernels over bodies.
struct DevWorld {
    int ns, nb, n_xf, narrow_blocks, n_ext;
    int body_base;
    int narrow_group;               // lanes per pair in GJK + EPA: 1 (lane per pair) or 4, 8, 16
    int manifold_group;             // lanes per pair in the contact manifold: 1, 4, 8, 16
    int manifold_blocks;
    int narrow_block_cap;           // most blocks of a narrow-phase launch (0: the default, 8192)
    int lds_limit;                  // dynamic LDS bytes one block may use on the world's device
    int stage_cap;                  // LDS vertices per pair: 2 x the largest hull
    int small_nv;                   // > 0: pairs of two hulls of <= small_nv vertices go to their own list
    int stage_small;                // LDS vertices per pair of that list: 2 x small_nv
    int solve_lanes;                // lanes per scene in the solve: 64, 32 or 16
    int big_launch_classes;         // large scenes: colour classes solved one launch each, the rest in one block
    // scenes
    const int32_t* scene_body_off;  // [ns+1]
    const int32_t* scene_pair_base; // [ns+1]  candidate pair capacity prefix: n(n-1)/2 per scene
    const int32_t* scene_ext_off;   // [ns+1]
    const int32_t* scene_item_base; // [ns+1]  solve-item capacity prefix: external constraints + pair capacity
    // bodies
    Body* body;                     // [nb]
    M4* model;                      // [nb]  model matrix of the current substep
    double* pose;                   // [nb][7] (x, q) after the last substep's integration (FrameParams::keep_pose)
    const V3* body_scale;           // [nb]  world scale (rendering model matrices)
    const int32_t* body_hull;       // [nb]
    const int32_t* body_tv_off;     // [nb]
    // hull instances (vertices scaled, local frame)
    const int32_t *hull_nv, *hull_nf, *hull_v_off, *hull_fn_off, *hull_topo;
    const V3* hull_v;
    const V3* hull_fn;
    const int32_t* topo_base;       // 8 offsets into topo_pool per topology
    const int32_t* topo_pool;
    // transformed geometry
    V3* tv;
    const int4* xf;                 // [n_xf] (body, hull instance, vertex, tv slot)
    // colliders (bodies with one hull collider take the fast narrow path)
    int has_general;                // some body has a sphere or several colliders
    const int32_t* body_col_off;    // [nb+1]
    const int32_t* col_type;        // [colliders] COL_HULL / COL_SPHERE
    const int32_t* col_hull;        // [colliders] hull instance
    const int32_t* col_tv_off;      // [colliders] first transformed vertex
    const float* col_radius;        // [colliders] sphere radius (r32)
    const int32_t* body_general;    // [nb]
    int32_t* need;                  // [nb] body takes part in a live pair this frame
    // pairs
    int2* pairs;                    // [pair capacity] scene-local (i, j)
    int32_t* pair_count;            // [ns]
    int32_t* live_local;            // [pair capacity] pair index within the scene
    int32_t* live_count;            // [ns]
    int32_t* live_prefix;           // [ns+1]
    int2* live_global;              // [pair capacity] (scene, pair slot)
    int4* live_rec;                 // [pair capacity] (body 1, body 2, hull instance 1, hull instance 2)
    int4* live_hull;                // [pair capacity] (first local vertex, vertex count) of both hulls
    int32_t* ncontact;              // [pair capacity]
    V3* pair_normal;                // [pair capacity] EPA normal of the pairs queued for the manifold
    int32_t* mlist;                 // [pair capacity] manifold queue: live-pair slots
    int32_t* pair_lists;            // [2][pair capacity] live pairs of small hulls, of large hulls (by frame)
    int64_t list_stride;            // offset of the large-hull list
    int32_t* counters;              // [CNT_COUNT] queue lengths of the current substep
    char* group_work;               // [group_work_slots][GROUP_WORK_BYTES] lane-group working memory
    int64_t group_work_slots;       // >= blocks x groups per block of the grouped narrow / manifold kernels
    Contact* contacts;              // [pair capacity * MAXC]
    // external constraints
    const ExtC* ext;                // [n_ext]
    double* ext_lambda;             // [3 n_ext]
    // solver scratch
    int32_t* item_level;            // [items]       level of each solve item (frame)
    int32_t* item_order;            // [items]       items grouped by level
    int2* item_bodies;              // [items]       global body indices of each item_order entry
    int32_t* item_lstart;           // [items + ns]  level starts per scene
    int32_t* item_cursor;           // [items + ns]  fill cursors of the sort
    int32_t* scene_nlevels;         // [ns]
    int32_t* body_last;             // [nb]          last level per body (large scenes)
    int32_t *uf_parent, *uf_flag;   // [nb]
    const int32_t* scene_big;       // [ns] 1: large scene (grid broad phase, coloured solve)
    // diagnostics
    uint32_t* scene_err;            // [ns]
    unsigned long long* stats;      // [STAT_COUNT]
};

// A large scene: grid broad phase, parallel islands and a graph-coloured
// solve instead of the per-scene wave kernels. Device buffers are owned by
// the world.
// A settled pile of 10^5 bodies has about 4.9 items (body pairs) per body;
// greedy colouring of its item conflict graph needs up to about 40 colours,
// and 24 colours with 32 rounds left 60 000 of its 490 000 items to the
// serial class (one lane: seconds per frame).
constexpr int BIG_COLOURS = 63;             // parallel colours (one 64-bit mask with the serial class)
constexpr int BIG_SERIAL = BIG_COLOURS;     // class solved by one lane in item order
constexpr int BIG_CLASSES = BIG_COLOURS + 1;
constexpr int BIG_COLOUR_ROUNDS = 64;       // Jones-Plassmann rounds per frame (launches end at once when all are coloured)
// Default: every class launched one by one over the device (a settled
// 10^5-body pile leaves about 100 000 items beyond the first 16 colours,
// 4 ms per substep on the one-block tail; an empty class's launch costs a
// few microseconds).
constexpr int BIG_LAUNCH_CLASSES = BIG_CLASSES;
constexpr int LARGE_CHUNK = 256;            // bodies per wave in a large body's pair scan
struct BigScene {
    int s, b0, n;                // scene, first body, bodies
    int pair_base, pair_cap;     // slice of the pair arrays
    int item_base, item_cap;     // slice of the item arrays
    int T;                       // hash-table buckets (power of two)
    int n_large;                 // bodies tested against all others (radius above the cell bound)
    double cell;                 // grid cell edge
    int large_chunks;            // LARGE_CHUNK-body chunks per large body's scan
    int32_t* large;              // [n_large] scene-local indices, ascending
    int32_t* large_cnt;          // [n_large * large_chunks + 1] chunk pair counts, then offsets
    int32_t* is_large;           // [n]
    int32_t* cell_start;         // [T + 1]
    int32_t* cell_cursor;        // [T]
    int32_t* sorted;             // [n] small bodies grouped by bucket
    double4* spos;               // [n] their positions and radii, in the same order
    int32_t* pair_off;           // [n + 1] per-body pair counts, then offsets
    int32_t* tmp;                // [max(pair_cap, item_cap) + 1] scan buffer
    int32_t* bi_start;           // [n + 1] body -> items CSR
    int32_t* bi_cursor;          // [n]
    int32_t* bi_items;           // [2 item_cap]
    int32_t* remaining;          // [1] uncoloured items
    int32_t* scan_blocks;        // [scan block sums]
};

// Substep records of one frame for the adjoint's recomputation: body
// states at the start of each substep and the contacts the narrow phase made
// (the frozen contact set).
struct Recorder {
    Body* body;          // [substeps][nb]
    int32_t* ncontact;   // [substeps][pair capacity]
    Contact* contacts;   // [substeps][pair capacity * MAXC]
    int64_t pair_cap;
};

struct FrameParams {
    double dt = 0.0, h = 0.0, gravity = 0.0;
    unsigned substeps = 0, iters = 0;
    bool collisions = true;
    bool forces_given = false;  // body F/T already hold the frame's force sums
    bool keep_pose = false;     // record every body's pose after the last substep's integration
};

// Kernel classes timed by the profiling mode.
enum { KC_FRAME = 0, KC_INTEGRATE, KC_TRANSFORM, KC_NARROW, KC_SOLVE, KC_MANIFOLD, KC_COUNT = 8 };

// Records a pair of HIP events around each timed launch (profiling mode
// only); the host reads the elapsed times after the frame completes.
struct KernelTimer {
    bool on = false;
    std::vector<hipEvent_t> pool;
    std::vector<int> cls;  // kernel class of each recorded pair
    size_t used = 0;
    void begin(int c, hipStream_t st) {
        if (!on) return;
        if (pool.size() < 2 * (used + 1)) {
            hipEvent_t a, b;
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            pool.push_back(a);
            pool.push_back(b);
        }
        if (cls.size() <= used) cls.resize(used + 1);
        cls[used] = c;
        (void)hipEventRecord(pool[2 * used], st);
    }
    void end(hipStream_t st) {
        if (!on) return;
        (void)hipEventRecord(pool[2 * used + 1], st);
        ++used;
    }
};

hipError_t launch_frame(const DevWorld& w, const FrameParams& fp, hipStream_t st, KernelTimer* kt = nullptr,
                        const BigScene* bigs = nullptr, int nbig = 0, const Recorder* rec = nullptr);
// Adjoint of one recorded frame for scenes with one dynamic body
// (rp_adjoint.hip): abar [nb][13] is the adjoint of (x, q, v, w) at the end
// of the frame on entry and at its start on exit; gparam [nb][3] accumulates
// d/d(mass, static friction, dynamic friction); amax [nb] keeps the largest
// |abar| component of each dynamic body seen at a frame start.
hipError_t launch_adjoint_frame(const DevWorld& w, const Recorder& rec, const int32_t* scene_dyn, const Body* end_state,
                                double* abar, double* gparam, double* amax, const FrameParams& fp, hipStream_t st);
// Large-scene stages (rp_large.hip), in frame order.
void big_broad(const DevWorld& w, const BigScene& b, hipStream_t st);
void big_islands(const DevWorld& w, const BigScene& b, double dt, hipStream_t st);
void big_scatter(const DevWorld& w, const BigScene& b, hipStream_t st);
void big_colour(const DevWorld& w, const BigScene& b, bool collisions, hipStream_t st);
void big_solve(const DevWorld& w, const BigScene& b, const FrameParams& fp, hipStream_t st);
// Blocks of the device-wide scan for n elements (size of BigScene::scan_blocks).
int scan_block_count(int64_t n);
hipError_t launch_pair_contacts(const DevWorld& w, int b1, int b2, double* out, int* count, uint32_t* err,
                                hipStream_t st);
hipError_t launch_broad(const DevWorld& w, hipStream_t st, const BigScene* bigs = nullptr, int nbig = 0);
// entity_get_model_matrix of every body into out ([nb][4][4] row-major,
// double or float), device memory.
hipError_t launch_model_matrices(const DevWorld& w, void* out, bool f32, hipStream_t st);
// Working-memory slots the grouped narrow-phase kernels of this schedule use.
int64_t group_work_slots_needed(const DevWorld& w);
// rp_world_step_noalloc's lane-group working memory: a static device array
// of NOALLOC_SCRATCH_BYTES on each device (4519 slots of GROUP_WORK_BYTES).
constexpr size_t NOALLOC_SCRATCH_BYTES = 64u << 20;
// Its device address on the current device (looked up once per device), or
// null if the lookup failed.
unsigned char* noalloc_scratch();
// Lanes per pair the narrow phase can use for this world (LDS permitting).
bool narrow_group_supported(const DevWorld& w, int g);
// Lets the lane-group narrow-phase kernels take up to `bytes` of dynamic LDS
// per block (gfx950 gives a workgroup up to 160 KiB).
hipError_t allow_narrow_lds(int bytes);
int narrow_blocks_for(const DevWorld& w, int g, int64_t pair_capacity);
hipError_t launch_fp64_selftest(const double* a, const double* b, double* out, int64_t n, hipStream_t st);

}  // namespace rp
