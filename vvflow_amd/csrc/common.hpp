// common.hpp -- shared infrastructure of the gfx950 implementation: the
// device context, error handling, device buffers and small wave-level
// helpers.  Host and device code.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <map>
#include <string>
#include <vector>

#include "vvhd_gpu.h"

namespace vvhd {

// ----------------------------------------------------------------- errors --

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define VVHD_HIP(call)                                                                     \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            throw ::vvhd::Error(VVHD_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define VVHD_CHECK(cond, code, msg)                         \
    do {                                                    \
        if (!(cond)) throw ::vvhd::Error((code), (msg));    \
    } while (0)

// ---------------------------------------------------------------- buffers --

// Device buffer that only grows; contents are not preserved across grow().
template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t cap = 0;
    void grow(size_t n) {
        if (n <= cap) return;
        size_t nc = std::max<size_t>(n, cap + cap / 2);
        if (p) VVHD_HIP(hipFree(p));
        p = nullptr;
        cap = 0;
        if (hipMalloc(&p, nc * sizeof(T)) != hipSuccess) {
            p = nullptr;
            throw Error(VVHD_ENOMEM, "hipMalloc of " + std::to_string(nc * sizeof(T)) + " bytes failed");
        }
        cap = nc;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// ------------------------------------------------------------ device utils --

constexpr int kWave = 64;
constexpr int kLeafMaxList = 16;   // TSortedTree.cpp:10

// Doubles mapped to uint64 so that unsigned order equals numeric order
// (NaN excluded); used for atomic min/max of bounding boxes.
__host__ __device__ inline uint64_t ord_from_double(double d) {
    uint64_t u = __double_as_longlong(d);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__host__ __device__ inline double double_from_ord(uint64_t u) {
    uint64_t v = (u >> 63) ? (u & 0x7fffffffffffffffull) : ~u;
    return __longlong_as_double(v);
}

__device__ inline int lane_id() { return __lane_id(); }

template <typename T>
__device__ inline T wave_bcast(T v, int src) {
    return __shfl(v, src, kWave);
}

__device__ inline double wave_sum(double v) {
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
__device__ inline double wave_max(double v) {
    for (int o = kWave / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
    return v;
}
__device__ inline int wave_sum_i(int v) {
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
// Inclusive scan across the wave.
__device__ inline int wave_incl_scan(int v) {
    const int l = lane_id();
    for (int o = 1; o < kWave; o <<= 1) {
        int t = __shfl_up(v, o, kWave);
        if (l >= o) v += t;
    }
    return v;
}

// Maps a linear block index so that consecutive logical blocks land on the
// same XCD (blocks are dispatched round-robin over the 8 XCDs of MI355X):
// neighbouring leaves then share that XCD's L2.
__device__ inline int xcd_block(int b, int nblocks) {
    const int kXcd = 8;
    const int per = nblocks / kXcd;
    const int rem = nblocks % kXcd;
    const int xcd = b % kXcd;
    const int idx = b / kXcd;
    // XCD x owns logical blocks [start(x), start(x) + size(x))
    const int start = xcd * per + min(xcd, rem);
    return start + idx;
}

// ------------------------------------------------------------------ context --

struct TreeArrays {
    // node table (capacity ncap)
    DBuf<double> x, y, h, w;              // Stretch() geometry
    DBuf<uint64_t> blx, bly, trx, try_;   // ordered-uint64 bbox (atomics)
    DBuf<int32_t> v0, v1, ch1, parent, depth, nseg;
    // hRange / sRange (TSortedTree.hpp:39-40) in element numbering: the tree
    // is built over the vortices, then the heat particles (from hbase), then
    // the streak particles (from sbase)
    DBuf<int32_t> h0, h1, s0, s1;
    int32_t hbase = 0, sbase = 0, nelem = 0;
    DBuf<uint64_t> path;                  // path code (bit per level)
    DBuf<double> cmp_x, cmp_y, cmp_g, cmm_x, cmm_y, cmm_g;
    DBuf<float> cmp_ie, cmm_ie;           // _1_eps of monopoles: +inf or 0
    DBuf<int32_t> nleaves, arrive, leaf_rank;
    DBuf<int32_t> pre;                    // preorder (DFS) position of each node
    DBuf<int32_t> split;                  // per node: 0 undecided,1 leaf,2 split
    // levels
    DBuf<int32_t> counters;               // [0]=node count, [1]=leaf count, [2..] misc
    // leaves (DFS order)
    DBuf<int32_t> leaf_node;              // leaf -> node id
    DBuf<int64_t> near_off, far_off;      // CSR offsets per leaf
    DBuf<int32_t> near_list, far_list;    // node ids
    DBuf<int64_t> seg_off;                // per leaf
    DBuf<int32_t> seg_list;               // segment ids
    DBuf<int32_t> seg_leaf;               // per segment: leaf (DFS)
    DBuf<int32_t> seg_keys, seg_ids;      // sort scratch of the per-leaf segment lists
    DBuf<int32_t> vleaf;                  // per element (vortex, heat, streak): leaf (DFS)
    DBuf<int32_t> node_L, node_m, leaf_tmp, seg_node;   // build scratch
    DBuf<int32_t> cnt_near, cnt_far, err;               // list scratch
    DBuf<int64_t> slot_near, slot_far;                  // per (leaf, top entry) list offsets
    DBuf<uint8_t> entries;                              // top-of-tree entry tables of the list builder (per root)
    DBuf<uint8_t> trav;                                 // packed node records of the list builder
    DBuf<double4> pbox;                                 // packed node boxes (pack_boxes)
    DBuf<int32_t> groups;                               // list builder leaf groups: {first leaf, count, root}
    int ngroups = 0;
    DBuf<int32_t> leaf_sim;                             // per leaf: its simulation (root)
    DBuf<int32_t> root_first;                           // per root: first leaf (DFS), nroots + 1
    int32_t nnodes = 0, nleaves_h = 0;
    int64_t nnear = 0, nfar = 0;
    int max_depth = 0;
    std::vector<std::pair<int, int>> levels;   // node id range per depth
    bool built = false;
    int far_criteria = 8;
    double min_node = 0, max_node = DBL_MAX;
};

// A list of passive particles (Space::HeatList or StreakList, TSpace.hpp:
// 40-43): TObj records as structure of arrays, FP64, with double buffers
// for reordering and compaction.
struct PList {
    int64_t n = 0;
    DBuf<double> x, y, g, vx, vy, ie;
    DBuf<double> x2, y2, g2, vx2, vy2, ie2;
    void reserve(int64_t m, bool keep, hipStream_t s) {
        if ((size_t)m <= x.cap) return;
        const size_t cap = std::max<size_t>((size_t)m, x.cap + x.cap / 2);
        for (DBuf<double>* b : {&x, &y, &g, &vx, &vy, &ie}) {
            DBuf<double> nb;
            nb.grow(cap);
            if (keep && n > 0) VVHD_HIP(hipMemcpyAsync(nb.p, b->p, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
            VVHD_HIP(hipStreamSynchronize(s));
            b->release();
            *b = nb;
        }
        for (DBuf<double>* b : {&x2, &y2, &g2, &vx2, &vy2, &ie2}) {
            b->release();
            b->grow(cap);
        }
    }
    void swap_buffers() {
        std::swap(x.p, x2.p);
        std::swap(y.p, y2.p);
        std::swap(g.p, g2.p);
        std::swap(vx.p, vx2.p);
        std::swap(vy.p, vy2.p);
        std::swap(ie.p, ie2.p);
        std::swap(x.cap, x2.cap);
        std::swap(y.cap, y2.cap);
        std::swap(g.cap, g2.cap);
        std::swap(vx.cap, vx2.cap);
        std::swap(vy.cap, vy2.cap);
        std::swap(ie.cap, ie2.cap);
    }
    void release() {
        for (DBuf<double>* b : {&x, &y, &g, &vx, &vy, &ie, &x2, &y2, &g2, &vx2, &vy2, &ie2}) b->release();
        n = 0;
    }
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // A diffusive call right after a convective one runs on stream2 (default
    // priority), from the state the convective call started from, into
    // dvx/dvy; the main stream adds them after its convective kernels.
    hipStream_t stream2 = nullptr;
    hipStream_t stream3 = nullptr;   // the learned path's near field, beside its far field
    hipEvent_t ev_pre_conv = nullptr;
    hipEvent_t ev_side = nullptr;            // end of the learned path's side-stream work
    uint64_t api_calls = 0, conv_call = 0;   // entry-point calls so far; the last convective one
    bool overlap = true;                     // vvhd_set_overlap
    bool far_net_grad = true;                // vvhd_set_far_gradient
    DBuf<double> dvx, dvy;
    std::string err;
    double re = INFINITY, dt = 1, inf_g = 0, time = 0;

    // vortices, structure of arrays, FP64 master copy
    int64_t n = 0;
    DBuf<double> x, y, g, vx, vy, ie;
    DBuf<double> x2, y2, g2, vx2, vy2, ie2;   // double buffers
    DBuf<int32_t> id, id2;                   // index before the tree build
    DBuf<float2> src_ge;                     // {g, eps^2} FP32 pair-kernel copy
    DBuf<double2> near_v;                    // near-field sums per vortex
    DBuf<double4> taylor;                    // far-field Taylor terms per leaf
    // snapshot for benchmarks
    DBuf<double> snap;
    int64_t snap_n = -1;

    // independent simulations sharing the context (one tree root each):
    // simulation s owns vortices [sim_off_h[s], sim_off_h[s+1]) and bodies
    // [sim_body_h[s], sim_body_h[s+1]), Reynolds number sim_re_h[s]
    int nsim = 1;
    std::vector<int64_t> sim_off_h;
    std::vector<int32_t> sim_body_h;
    std::vector<double> sim_re_h;
    DBuf<int64_t> sim_off;                   // device copies, refreshed by the tree build
    DBuf<int32_t> sim_body;
    DBuf<double> sim_re;
    DBuf<int32_t> seg_root;                  // per segment: its simulation

    // heat and streak lists, the streak sources (host), Space::pr, and the
    // bodies' heat conditions: TBody::heat_condition per body (hc_t,
    // TBody.hpp:15), TAtt::heat_const and heat_layer_obj_no per segment
    PList heat, streak;
    std::vector<double> streak_src_h;        // StreakSourceList, 6 doubles per record
    double pr = 0;
    std::vector<int32_t> heat_cond_h;        // per body
    DBuf<int32_t> heat_cond;
    DBuf<float> heat_const;                  // per segment
    DBuf<int64_t> heat_layer;                // per segment
    DBuf<double> hres;                       // per heat particle: epsh result
    DBuf<uint8_t> hflag;
    DBuf<uint8_t> hdirty;                    // per leaf: a heat merge touched it

    // segments and bodies (host copy is authoritative between calls)
    std::vector<vvhd_seg> segs_h;
    std::vector<vvhd_body> bodies_h;
    DBuf<vvhd_seg> segs;
    DBuf<double> fric;                       // per segment friction accumulator
    DBuf<int32_t> body_range;                // per body: first segment, count
    DBuf<vvhd_body> bodies_d;                // device copy of the body records
    DBuf<double> body_geo;                   // per body: r2max, bbox (blx, bly, trx, try)

    // boundary-condition system (column-major, LAPACK convention)
    DBuf<double> slae_A, slae_Acached, slae_LU, slae_b, slae_x;
    DBuf<int32_t> slae_base, slae_ipiv, slae_info;
    int slae_n = -1;                         // order of the cached factorisation
    int slae_nsys = 1, slae_nsys_factored = 0;   // systems filled / factorised (one per simulation)
    int slae_filled = 0;                     // order of the last filled system
    void* rocblas = nullptr;                 // rocblas_handle

    TreeArrays t;

    // scratch
    DBuf<int32_t> iscratch0, iscratch1, iscratch2, iscratch3, iscratch4;
    DBuf<int64_t> lscratch0;
    DBuf<uint8_t> tmp;

    // epsilon results
    DBuf<int32_t> nn1, nn2;
    DBuf<double> d1, d2;
    DBuf<double> leaf_crit, leaf_restr;
    DBuf<uint8_t> eflag;                     // per target: bit0 processed, bit1 merge flag
    DBuf<int32_t> flag_list;                 // targets flagged by the parallel pass
    DBuf<int32_t> ecount;                    // [0]=flags, [1]=merged, ...
    DBuf<int64_t> rnear_off;                 // transposed near lists (repair)
    DBuf<int32_t> rnear_list, rnear_cnt, leaf_stamp;
    DBuf<uint64_t> leaf_maxd2;                       // per-leaf bound of the second-neighbour distance
    DBuf<uint8_t> leaf_mark;                         // leaves whose reverse near lists are built
    DBuf<int64_t> rep_state;                         // per simulation: resume target, commit count

    // learned far-field correction
    DBuf<double> fm;                         // per-node moments after epsilon: kMomArrays x node capacity
    size_t fm_stride = 0;
    DBuf<double2> far_v;                     // per-vortex far-field velocity of the learned path
    DBuf<uint8_t> far_nodes;                 // per-node feature records of the learned path
    DBuf<double> far_leaf;                   // per-leaf network correction sums at three points (6 doubles)
    DBuf<float> mlp_params;                  // FP32 master parameters (kMlpParams)
    DBuf<uint16_t> mlp_frags;                // bf16 MFMA fragments + FP32 tail packed from the parameters
    bool mlp_loaded = false;
    DBuf<float> tr_x, tr_y, tr_w;            // training set on the device
    int64_t tr_n = 0;
    DBuf<float> adam_m, adam_v, grad_part, grad_out;
    int64_t adam_t = 0;
    DBuf<int64_t> pair_off;
    DBuf<int32_t> pair_cnt, pair_t, pair_f;

    // field evaluators: per-vortex weights, the grid map, running gsum
    DBuf<double> fld_a, fld_b, fld_gtmp;
    DBuf<float> fld_map;

    // near-field work items and queue (persistent form), the autotuner's
    // choices per workload class, and the captured convective graphs
    DBuf<int2> near_items;
    DBuf<int32_t> near_q;
    int n_cu = 256;                          // compute units of the device
    bool autotune = true;                    // vvhd_set_tuning
    int near_form = 0;                       // the form used with the autotuner off
    int near_last = 0;                       // the form of the last near-field launch
    std::map<uint64_t, int> tune_near;
    hipEvent_t tune_ev[2] = {};
    bool use_graph = false;                  // vvhd_set_graph
    std::vector<std::pair<std::vector<uint8_t>, void*>> conv_graphs;
    int64_t graph_captures = 0, graph_launches = 0;

    // timing
    hipEvent_t ev[16] = {};
    // tree build: per-level counters copied asynchronously (pinned, two slots)
    static constexpr int kLevSlots = 2;   // tree levels in flight (pinned counter slots and events)
    hipEvent_t lev_ev[kLevSlots] = {};
    int32_t* lev_counters = nullptr;      // kLevSlots x 8 pinned counters
    int32_t* fm_counts = nullptr;         // pinned: flow-move survivors, removals in bodies, scan error
    bool phase_recorded[8] = {false, false, false, false, false, false, false, false};

    // work counts of the current tree
    int64_t work[7] = {0, 0, 0, 0, 0, 0, 0};
};

// Read-only device view of the built tree used by the physics kernels.
struct TreeDev {
    const double *nx, *ny, *nh, *nw;
    const uint64_t *blx, *bly, *trx, *try_;
    const int32_t *v0, *v1, *ch1, *leaf_rank, *leaf_node;
    const int64_t *near_off, *far_off, *seg_off;
    const int32_t *near_list, *far_list, *seg_list, *vleaf;
    const double *cmp_x, *cmp_y, *cmp_g, *cmm_x, *cmm_y, *cmm_g;
    const float *cmp_ie, *cmm_ie;
    int nleaf;
    const double4* pbox;   // per node {blx, bly, trx, try} as doubles (pack_boxes)
    const int32_t *h0, *h1, *s0, *s1;   // heat / streak ranges (element numbering)
    int hbase, sbase;                   // first heat and first streak element
};

inline TreeDev tree_dev(const TreeArrays& t) {
    return TreeDev{t.x.p,       t.y.p,      t.h.p,       t.w.p,       t.blx.p,     t.bly.p,     t.trx.p,
                   t.try_.p,    t.v0.p,     t.v1.p,      t.ch1.p,     t.leaf_rank.p, t.leaf_node.p, t.near_off.p,
                   t.far_off.p, t.seg_off.p, t.near_list.p, t.far_list.p, t.seg_list.p, t.vleaf.p,  t.cmp_x.p,
                   t.cmp_y.p,   t.cmp_g.p,  t.cmm_x.p,   t.cmm_y.p,   t.cmm_g.p,   t.cmp_ie.p,  t.cmm_ie.p,
                   t.nleaves_h, t.pbox.p, t.h0.p,   t.h1.p,      t.s0.p,      t.s1.p,      t.hbase,
                   t.sbase};
}

// box_dist2 from the packed copy of the boxes: one 32-byte load per node.
// Valid while the boxes are those pack_boxes last copied.
__device__ inline double box_dist2p(const TreeDev& T, int a, int b) {
    const double4 A = T.pbox[a], B = T.pbox[b];
    double dx = fmax(0.0, fmax(B.x - A.z, A.x - B.z));
    double dy = fmax(0.0, fmax(B.y - A.w, A.y - B.w));
    return dx * dx + dy * dy;
}

// Box-to-box squared distance of two nodes (0 when the boxes touch or
// overlap).  Rounded operations are monotone, so the result never exceeds the
// computed squared distance of any pair of points inside the two boxes.
__device__ inline double box_dist2(const TreeDev& T, int a, int b) {
    const double ablx = double_from_ord(T.blx[a]), atrx = double_from_ord(T.trx[a]);
    const double ably = double_from_ord(T.bly[a]), atry = double_from_ord(T.try_[a]);
    const double bblx = double_from_ord(T.blx[b]), btrx = double_from_ord(T.trx[b]);
    const double bbly = double_from_ord(T.bly[b]), btry = double_from_ord(T.try_[b]);
    double dx = fmax(0.0, fmax(bblx - atrx, ablx - btrx));
    double dy = fmax(0.0, fmax(bbly - atry, ably - btry));
    return dx * dx + dy * dy;
}

// host helpers implemented in ctx.hip
void ensure_vortex_capacity(Ctx& c, int64_t n);
// Refreshes the device copies of the simulation table (one simulation over
// the whole context unless a batch is declared).
void sim_sync(Ctx& c);
// Copies the node boxes (ordered-integer images) into TreeArrays::pbox.
void pack_boxes(Ctx& c);
void reserve_vortices(Ctx& c, int64_t n, bool preserve);
void sync(Ctx& c);

// phase entry points (one translation unit each)
void tree_build(Ctx& c, int far_criteria, double min_node, double max_node, int32_t* perm_out);
void tree_lists(Ctx& c);
void epsilon(Ctx& c, int merge, int32_t* nn1, int32_t* nn2, int32_t* merged);
void convective(Ctx& c, double inf_x, double inf_y, int kahan);
// beside: the near field runs beside the far-field kernels (tuned among the
// persistent forms only)
void convective_near(Ctx& c, int kahan, hipStream_t s, double2* near_out, bool beside);
// Allocation-free convective (vvhd_convective_noalloc): the scratch lives in a
// fixed device array of kNoallocScratchBytes; noalloc_init looks its address
// up once per device (the device must be current).
constexpr size_t kNoallocScratchBytes = size_t(64) << 20;
void noalloc_init(int device);
size_t convective_noalloc_bytes(int64_t n, int64_t nleaf);
void convective_noalloc(Ctx& c, double inf_x, double inf_y, int kahan);
void velocity_points(Ctx& c, const double* xy, int64_t n, double inf_x, double inf_y, double* v_out);
void diffusive(Ctx& c, bool overlap);
void bodies_prepare(Ctx& c);
void bodies_release(Ctx& c);
void farfield_release(Ctx& c);
void direct_sum(Ctx& c, const int64_t* targets, int64_t nt, double* v_out);
void pack_sources(Ctx& c);
// heat and streak lists (passive.hip)
int heat_epsilon(Ctx& c, int merge);
void passive_convective(Ctx& c, const double4* taylor, double inf_x, double inf_y);
void heat_diffusive(Ctx& c);
void passive_move_and_clean(Ctx& c);
void heat_shed(Ctx& c);
void heat_crop(Ctx& c, double scale);
void streak_shed(Ctx& c);
void heat_layer_reset(Ctx& c);
void list_append(Ctx& c, PList& L, const double* o6, int64_t m);
void list_download(Ctx& c, PList& L, double* o6);
void far_taylor(Ctx& c, double4* taylor);
void sync_bodies_to_host(Ctx& c);
__global__ void k_leaf_lists(TreeDev T, int32_t* out);
void eps2h_h2(Ctx& c, bool h2, const double* xy, const int32_t* leaf, int64_t n, double* out);
void grid_field(Ctx& c, int kind, const vvhd_grid& gr, double m, double dl, double inf_x, double inf_y, int with_ref,
                double ref_x, double ref_y, float* map);

}  // namespace vvhd
