// tree.hip -- the adaptive binary tree of TSortedTree.cpp rebuilt
// level-synchronously on the GPU.
//
// stree::build (TSortedTree.cpp:232-265) splits nodes recursively at the
// centre of their tight bounding box with an in-place Hoare partition
// (DistributeContent, :81-99).  That partition pairs the k-th element left
// of the split position that belongs right with the k-th element right of
// it, counted from the end, that belongs left; everything else stays.  Each
// level is therefore replayed exactly with one exclusive scan of the
// "goes left" flags: misplaced elements get their rank on their side and
// swap with the partner of complementary rank.  Bounding boxes are exact
// FP64 min/max (atomics on an order-preserving integer image), so node
// centres, leaf decisions and the resulting permutation are bit-identical to
// the reference.  Segments (body panels) follow a stable partition
// (:139-148), i.e. each segment descends like findNode and keeps its order.
// The heat and streak lists (hRange, sRange) are partitioned the same way
// as the vortex list, each in place within its own range: the elements of
// the build are the vortices, then the heat particles, then the streak
// particles, and every node holds one range per list.
//
// Node ids are allocated level by level, so every level is a contiguous id
// range.  After the levels: subtree leaf counts (bottom-up) and leaf DFS
// ranks (top-down) give the bottomNodes order; per-sign monopoles
// (CalculateCMass, :150-197) are combined bottom-up level by level; and the
// near/far lists of every leaf (FindNearNodes, :199-217) come from one
// traversal per wave of 64 DFS-consecutive leaves with a shared stack.

#include "common.hpp"

#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cmath>

#pragma clang fp contract(off)

namespace vvhd {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxLevels = 126;   // deepest tree accepted

inline int nblocks(int64_t n, int b = kBlock) { return (int)std::max<int64_t>(1, (n + b - 1) / b); }

// --------------------------------------------------------------- kernels --

struct NodeView {
    double *x, *y, *h, *w;
    uint64_t *blx, *bly, *trx, *try_;
    int32_t *v0, *v1, *ch1, *parent, *depth, *nseg, *split;
    uint64_t* path;
    int32_t *h0, *h1, *s0, *s1;
};

NodeView view(TreeArrays& t) {
    return NodeView{t.x.p,  t.y.p,   t.h.p,      t.w.p,     t.blx.p,   t.bly.p,  t.trx.p,   t.try_.p,
                    t.v0.p, t.v1.p,  t.ch1.p,    t.parent.p, t.depth.p, t.nseg.p, t.split.p, t.path.p,
                    t.h0.p, t.h1.p,  t.s0.p,     t.s1.p};
}

// List of element k: 0 vortex, 1 heat, 2 streak.
__device__ inline int elem_list(int k, int hbase, int sbase) { return k < hbase ? 0 : (k < sbase ? 1 : 2); }

// A node's range of list cls.
__device__ inline void list_range(const NodeView& nv, int node, int cls, int& r0, int& r1) {
    if (cls == 0) {
        r0 = nv.v0[node];
        r1 = nv.v1[node];
    } else if (cls == 1) {
        r0 = nv.h0[node];
        r1 = nv.h1[node];
    } else {
        r0 = nv.s0[node];
        r1 = nv.s1[node];
    }
}

__device__ inline void set_list_range(const NodeView& nv, int node, int cls, int r0, int r1) {
    int32_t* a = cls == 0 ? nv.v0 : (cls == 1 ? nv.h0 : nv.s0);
    int32_t* b = cls == 0 ? nv.v1 : (cls == 1 ? nv.h1 : nv.s1);
    a[node] = r0;
    b[node] = r1;
}

__device__ inline void init_node(NodeView nv, int id, int parent, int depth, uint64_t path, int v0, int v1, int h0,
                                 int h1, int s0, int s1) {
    nv.parent[id] = parent;
    nv.depth[id] = depth;
    nv.path[id] = path;
    nv.v0[id] = v0;
    nv.v1[id] = v1;
    nv.h0[id] = h0;
    nv.h1[id] = h1;
    nv.s0[id] = s0;
    nv.s1[id] = s1;
    nv.ch1[id] = -1;
    nv.nseg[id] = 0;
    nv.split[id] = 0;
    nv.blx[id] = ord_from_double(DBL_MAX);
    nv.bly[id] = ord_from_double(DBL_MAX);
    nv.trx[id] = ord_from_double(-DBL_MAX);
    nv.try_[id] = ord_from_double(-DBL_MAX);
}

// One root per simulation: root s holds the vortices [sim_off[s], sim_off[s+1]).
// The heat elements [hbase, sbase) and streak elements [sbase, nelem) belong
// to root 0 (one simulation only).
__global__ void k_init_roots(NodeView nv, int nroots, const int64_t* sim_off, int32_t* counters, int hbase,
                             int sbase, int nelem) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nroots) {
        if (s == 0)
            init_node(nv, s, -1, 0, 0, (int)sim_off[s], (int)sim_off[s + 1], hbase, sbase, sbase, nelem);
        else
            init_node(nv, s, -1, 0, 0, (int)sim_off[s], (int)sim_off[s + 1], nelem, nelem, nelem, nelem);
    }
    if (s == 0) {
        counters[0] = nroots;   // nodes
        counters[1] = 0;        // leaves
        counters[2] = 0;        // level begin
        counters[3] = nroots;   // level end
        counters[4] = 0;        // overflow flag
        counters[5] = 0;        // max depth
    }
}

// Element -> root (its simulation).
__global__ void k_fill_roots(int32_t* enode, int n, const int64_t* sim_off, int nroots) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int lo = 0, hi = nroots - 1;   // last root whose offset <= k
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sim_off[mid] <= k)
            lo = mid;
        else
            hi = mid - 1;
    }
    enode[k] = lo;
}

// Segments start at the root of their body's simulation.
__global__ void k_seg_roots(NodeView nv, const int32_t* seg_root, int nseg, int32_t* seg_node) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nseg) return;
    seg_node[i] = seg_root[i];
    atomicAdd(&nv.nseg[seg_root[i]], 1);
}

// Adds a point to a node's bounding box; NaN coordinates are ignored, as by
// std::max/std::min in Stretch.
__device__ inline void bbox_add(NodeView nv, int node, double px, double py) {
    if (!isnan(px)) {
        atomicMin((unsigned long long*)&nv.blx[node], (unsigned long long)ord_from_double(px));
        atomicMax((unsigned long long*)&nv.trx[node], (unsigned long long)ord_from_double(px));
    }
    if (!isnan(py)) {
        atomicMin((unsigned long long*)&nv.bly[node], (unsigned long long)ord_from_double(py));
        atomicMax((unsigned long long*)&nv.try_[node], (unsigned long long)ord_from_double(py));
    }
}

// Wave-aggregated bbox update: lanes targeting the same node reduce first and
// one lane performs the atomics, for up to four distinct nodes per wave; the
// rest fall back to per-lane atomics.  Called by all lanes of the wave.
__device__ inline void bbox_add_wave(NodeView nv, int node, double px, double py, bool active) {
    uint64_t left = __ballot(active);
    for (int round = 0; round < 4 && left; round++) {
        const int leader = (int)__builtin_ctzll(left);
        const int lead_node = __shfl(node, leader, kWave);
        const bool mine = active && ((left >> lane_id()) & 1ull) && node == lead_node;
        const uint64_t m_same = __ballot(mine);
        const bool vx = mine && !isnan(px), vy = mine && !isnan(py);
        double lx = vx ? px : DBL_MAX, hx = vx ? px : -DBL_MAX;
        double ly = vy ? py : DBL_MAX, hy = vy ? py : -DBL_MAX;
        for (int o = kWave / 2; o > 0; o >>= 1) {
            lx = fmin(lx, __shfl_xor(lx, o, kWave));
            hx = fmax(hx, __shfl_xor(hx, o, kWave));
            ly = fmin(ly, __shfl_xor(ly, o, kWave));
            hy = fmax(hy, __shfl_xor(hy, o, kWave));
        }
        if (lane_id() == leader) {
            if (lx <= hx) {
                atomicMin((unsigned long long*)&nv.blx[lead_node], (unsigned long long)ord_from_double(lx));
                atomicMax((unsigned long long*)&nv.trx[lead_node], (unsigned long long)ord_from_double(hx));
            }
            if (ly <= hy) {
                atomicMin((unsigned long long*)&nv.bly[lead_node], (unsigned long long)ord_from_double(ly));
                atomicMax((unsigned long long*)&nv.try_[lead_node], (unsigned long long)ord_from_double(hy));
            }
        }
        left &= ~m_same;
    }
    if (active && ((left >> lane_id()) & 1ull)) bbox_add(nv, node, px, py);
}

// Bounding boxes of the nodes of a level with at most kTopNodes nodes (the
// root, and the children of the first levels), from the elements in their
// final positions: every element k with enode[k] - base in [0, kTopNodes)
// adds to that node.  A node's elements form one contiguous range, so each
// thread keeps a running box for the node of its current element (order-
// preserving integer images, NaN coordinates skipped) and adds it to the
// block's LDS slot of that node only when the node changes; one lane per
// block and node then issues the global atomics.  Each block covers
// kTopChunk consecutive elements, kTopPer per thread, loaded before use.
constexpr int kTopNodes = 512;
constexpr int kTopPer = 4;
constexpr int kTopChunk = kBlock * kTopPer;

__global__ __launch_bounds__(kBlock) void k_bbox_top(NodeView nv, const int32_t* base_ptr, const int32_t* enode,
                                                     const double* x, const double* y, int n) {
    __shared__ unsigned long long s_lx[kTopNodes], s_ly[kTopNodes], s_hx[kTopNodes], s_hy[kTopNodes];
    const unsigned long long lo0 = ord_from_double(DBL_MAX), hi0 = ord_from_double(-DBL_MAX);
    for (int t = threadIdx.x; t < kTopNodes; t += blockDim.x) {
        s_lx[t] = lo0;
        s_ly[t] = lo0;
        s_hx[t] = hi0;
        s_hy[t] = hi0;
    }
    __syncthreads();
    const int base = base_ptr ? *base_ptr : 0;
    const int k0 = blockIdx.x * kTopChunk + threadIdx.x;
    int nodes[kTopPer];
    double px[kTopPer], py[kTopPer];
#pragma unroll
    for (int j = 0; j < kTopPer; j++) {
        const int k = k0 + j * kBlock;
        nodes[j] = -1;
        if (k < n) {
            nodes[j] = (enode ? enode[k] : 0) - base;
            px[j] = x[k];
            py[j] = y[k];
        }
    }
    int cur = -1;
    unsigned long long lx = lo0, hx = hi0, ly = lo0, hy = hi0;
    auto flush = [&]() {
        if (cur < 0) return;
        if (lx != lo0) atomicMin(&s_lx[cur], lx);
        if (hx != hi0) atomicMax(&s_hx[cur], hx);
        if (ly != lo0) atomicMin(&s_ly[cur], ly);
        if (hy != hi0) atomicMax(&s_hy[cur], hy);
    };
#pragma unroll
    for (int j = 0; j < kTopPer; j++) {
        const int node = nodes[j];
        if (node < 0 || node >= kTopNodes) continue;
        if (node != cur) {
            flush();
            cur = node;
            lx = ly = lo0;
            hx = hy = hi0;
        }
        if (!isnan(px[j])) {
            const unsigned long long o = ord_from_double(px[j]);
            lx = o < lx ? o : lx;
            hx = o > hx ? o : hx;
        }
        if (!isnan(py[j])) {
            const unsigned long long o = ord_from_double(py[j]);
            ly = o < ly ? o : ly;
            hy = o > hy ? o : hy;
        }
    }
    flush();
    __syncthreads();
    for (int t = threadIdx.x; t < kTopNodes; t += blockDim.x) {
        const int id = base + t;
        if (s_lx[t] != lo0) atomicMin((unsigned long long*)&nv.blx[id], s_lx[t]);
        if (s_hx[t] != hi0) atomicMax((unsigned long long*)&nv.trx[id], s_hx[t]);
        if (s_ly[t] != lo0) atomicMin((unsigned long long*)&nv.bly[id], s_ly[t]);
        if (s_hy[t] != hi0) atomicMax((unsigned long long*)&nv.try_[id], s_hy[t]);
    }
}

// Bounding boxes of a new level's nodes once k_swap has put every element in
// place: a node's elements are then the range [v0, v1), so LPN lanes per
// node (64, or 16 when nodes are small) reduce it (order-preserving integer
// images, NaN coordinates skipped, as the atomics do) without contending
// with the other nodes; segments add their atomics afterwards.
template <int LPN>
__global__ void k_child_boxes(NodeView nv, const int32_t* counters, const double* x, const double* y) {
    const int lb = counters[2], le = counters[3];
    constexpr int kPer = kWave / LPN;   // nodes per wave
    const int slots = gridDim.x * (blockDim.x / kWave) * kPer;
    const int sub = lane_id() % LPN;
    const uint64_t lo0 = ord_from_double(DBL_MAX), hi0 = ord_from_double(-DBL_MAX);
    const int first = lb + (blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave) * kPer + lane_id() / LPN;
    for (int base = lb; base < le; base += slots) {   // uniform trip count: shuffles stay converged
        const int id = first + (base - lb);
        const bool ok = id < le;
        uint64_t lx = lo0, hx = hi0, ly = lo0, hy = hi0;
        for (int cls = 0; cls < 3; cls++) {
            int r0 = 0, r1 = 0;
            if (ok) list_range(nv, id, cls, r0, r1);
            for (int k = r0 + sub; k < r1; k += LPN) {
                const double px = x[k], py = y[k];
                if (!isnan(px)) {
                    const uint64_t o = ord_from_double(px);
                    lx = o < lx ? o : lx;
                    hx = o > hx ? o : hx;
                }
                if (!isnan(py)) {
                    const uint64_t o = ord_from_double(py);
                    ly = o < ly ? o : ly;
                    hy = o > hy ? o : hy;
                }
            }
        }
        for (int o = LPN / 2; o > 0; o >>= 1) {
            const uint64_t a = __shfl_xor(lx, o, kWave), b = __shfl_xor(hx, o, kWave);
            const uint64_t c = __shfl_xor(ly, o, kWave), d = __shfl_xor(hy, o, kWave);
            lx = a < lx ? a : lx;
            hx = b > hx ? b : hx;
            ly = c < ly ? c : ly;
            hy = d > hy ? d : hy;
        }
        if (ok && sub == 0) {
            if (lx != lo0) atomicMin((unsigned long long*)&nv.blx[id], (unsigned long long)lx);
            if (hx != hi0) atomicMax((unsigned long long*)&nv.trx[id], (unsigned long long)hx);
            if (ly != lo0) atomicMin((unsigned long long*)&nv.bly[id], (unsigned long long)ly);
            if (hy != hi0) atomicMax((unsigned long long*)&nv.try_[id], (unsigned long long)hy);
        }
    }
}

// Segments of the root bounding boxes.
__global__ void k_root_segs(NodeView nv, const vvhd_seg* segs, const int32_t* seg_root, int nseg) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nseg) bbox_add(nv, seg_root[i], segs[i].x, segs[i].y);
}

// Root bounding boxes when there are more roots than k_bbox_top handles.
__global__ void k_bbox_elems(NodeView nv, const int32_t* enode, const double* x, const double* y, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const bool a = k < n;
    bbox_add_wave(nv, a ? enode[k] : 0, a ? x[k] : 0, a ? y[k] : 0, a);
}

// First leaf (DFS rank) of every root, from the subtree leaf counts.
// First leaf of every root and its preorder position: a tree of k leaves
// has 2k - 1 nodes, so root s starts at 2 first - s.
__global__ void k_root_first(const int32_t* nleaves, int nroots, int32_t* first, int32_t* root_first,
                             int32_t* pre) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int f = 0;
    for (int s = 0; s < nroots; s++) {
        first[s] = f;
        root_first[s] = f;
        pre[s] = 2 * f - s;
        f += nleaves[s];
    }
    root_first[nroots] = f;
}

// Leaf -> simulation.
__global__ void k_leaf_sim(const int32_t* root_first, int nroots, int nleaf, int32_t* leaf_sim) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nleaf) return;
    int lo = 0, hi = nroots - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (root_first[mid] <= l)
            lo = mid;
        else
            hi = mid - 1;
    }
    leaf_sim[l] = lo;
}

// snode::Stretch tail + DivideNode decision (TSortedTree.cpp:36-56, 122-125)
// for the nodes of the current level; split nodes get a pair of children.
__global__ void k_decide(NodeView nv, int32_t* counters, int32_t* leaf_tmp, int cap, double min_node,
                         double max_node) {
    const int lb = counters[2], le = counters[3];
    const int lane = lane_id();
    const uint64_t below = (1ull << lane) - 1;
    // whole waves step through the level, so the counters are claimed once
    // per wave (leaf slots, pairs of child ids) and lanes take their ranks
    const int stride = gridDim.x * blockDim.x;
    for (int b0 = lb + blockIdx.x * blockDim.x + threadIdx.x - lane; b0 < le; b0 += stride) {
        const int id = b0 + lane;
        const bool valid = id < le;
        bool leaf = false;
        int depth = 0;
        if (valid) {
            const double blx = double_from_ord(nv.blx[id]), bly = double_from_ord(nv.bly[id]);
            const double trx = double_from_ord(nv.trx[id]), try_ = double_from_ord(nv.try_[id]);
            const double x = (blx + trx) * 0.5;
            const double y = (bly + try_) * 0.5;
            const double h = try_ - bly;
            const double w = trx - blx;
            nv.x[id] = x;
            nv.y[id] = y;
            nv.h[id] = h;
            nv.w[id] = w;
            const double mx = (h < w) ? w : h;   // std::max(h, w)
            const double mn = (w < h) ? w : h;   // std::min(h, w)
            const int nvort = nv.v1[id] - nv.v0[id];
            const int nlist = max(max(nvort, nv.nseg[id]), max(nv.h1[id] - nv.h0[id], nv.s1[id] - nv.s0[id]));
            leaf = (mx < max_node && mn <= min_node) || (mx < max_node && nlist < kLeafMaxList);
            depth = nv.depth[id] + 1;
        }
        const uint64_t lm = __ballot(valid && leaf), sm = __ballot(valid && !leaf);
        int lbase = 0, sbase = 0, dmax = depth;
        for (int o = kWave / 2; o > 0; o >>= 1) dmax = max(dmax, __shfl_xor(dmax, o, kWave));
        if (lane == 0) {
            if (lm) lbase = atomicAdd(&counters[1], __popcll(lm));
            if (sm) {
                sbase = atomicAdd(&counters[0], 2 * __popcll(sm));
                atomicMax(&counters[5], dmax);
            }
        }
        lbase = __shfl(lbase, 0, kWave);
        sbase = __shfl(sbase, 0, kWave);
        if (!valid) continue;
        if (leaf) {
            nv.split[id] = 1;
            leaf_tmp[lbase + __popcll(lm & below)] = id;
            continue;
        }
        const int c = sbase + 2 * __popcll(sm & below);
        if (c + 2 > cap) {
            counters[4] = 1;
            nv.split[id] = 1;   // stop here; reported as an error by the host
            continue;
        }
        nv.split[id] = 2;
        nv.ch1[id] = c;
        init_node(nv, c, id, depth, nv.path[id] << 1, nv.v0[id], nv.v0[id], nv.h0[id], nv.h0[id], nv.s0[id], nv.s0[id]);
        init_node(nv, c + 1, id, depth, (nv.path[id] << 1) | 1ull, nv.v1[id], nv.v1[id], nv.h1[id], nv.h1[id],
                  nv.s1[id], nv.s1[id]);
    }
}

__device__ inline bool goes_left(const NodeView& nv, int node, double px, double py) {
    return (nv.h[node] < nv.w[node]) ? (px < nv.x[node]) : (py < nv.y[node]);
}

// Per element: "goes left" flag of elements in split nodes (else 0); 0 past
// the last element.
struct GoesLeftFlag {
    NodeView nv;
    const int32_t* enode;
    const double* x;
    const double* y;
    int n;
    __device__ int32_t operator()(int k) const {
        if (k >= n) return 0;
        const int node = enode[k];
        return (nv.split[node] == 2 && goes_left(nv, node, x[k], y[k])) ? 1 : 0;
    }
};

// Ranks of misplaced elements: left_pos[r0 + r] / right_pos[r0 + q], with
// [r0, r1) the range of the element's list in its node.  P is the exclusive
// prefix of the flags over all elements (n + 1 entries).  The first element
// of a split node's range also records that list's split position and number
// of misplaced pairs (node_L / node_m at list * cap + node) and sets the
// children's ranges.  Thread 0 advances the level range (the nodes of the
// next level were allocated by k_decide).
// The flags come from k_flag_count, before any swap of this level: k_swap
// reads them too and must not derive them from coordinates that the swaps
// may be rewriting.
__global__ void k_ranks(NodeView nv, const int32_t* enode, const int32_t* P, int n, int hbase, int sbase, int cap,
                        int32_t* node_L, int32_t* node_m, int32_t* left_pos, int32_t* right_pos, int32_t* counters,
                        const int32_t* flag) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) {
        counters[2] = counters[3];
        counters[3] = counters[0];
    }
    if (k >= n) return;
    const int node = enode[k];
    if (nv.split[node] != 2) return;
    const int cls = elem_list(k, hbase, sbase);
    int v0, v1;   // a split node's own range is unchanged
    list_range(nv, node, cls, v0, v1);
    const int P0 = P[v0];
    const int Lc = P[v1] - P0;
    const int Labs = v0 + Lc;
    if (k == v0) {
        node_L[cls * cap + node] = Labs;
        node_m[cls * cap + node] = Lc - (P[Labs] - P0);   // left-goers sitting right of Labs
        const int c = nv.ch1[node];
        set_list_range(nv, c, cls, v0, Labs);
        set_list_range(nv, c + 1, cls, Labs, v1);
    }
    const bool f = flag[k] != 0;
    if (k < Labs && !f) {
        const int r = (k - v0) - (P[k] - P0);
        left_pos[v0 + r] = k;
    } else if (k >= Labs && f) {
        const int q = P[k] - P[Labs];
        right_pos[v0 + q] = k;
    }
}

// Hoare swaps, child bounding boxes and element -> child node.
// With bbox false the child boxes are left to k_bbox_top.
__global__ void k_swap(NodeView nv, int32_t* enode, const int32_t* flag, const int32_t* P, int n, int hbase,
                       int sbase, int cap, const int32_t* node_L, const int32_t* node_m, const int32_t* left_pos,
                       const int32_t* right_pos, double* x, double* y, int32_t* id, bool bbox) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    bool active = false;
    int node = 0, Labs = 0, v0 = 0, f = 0, c = 0, slot = 0;
    if (k < n) {
        node = enode[k];
        active = nv.split[node] == 2;
    }
    if (active) {
        const int cls = elem_list(k, hbase, sbase);
        int v1;
        list_range(nv, node, cls, v0, v1);
        slot = cls * cap + node;
        Labs = node_L[slot];
        f = flag[k];
        c = nv.ch1[node];
    }
    const bool left_mis = active && k < Labs && !f;
    const bool right_mis = active && k >= Labs && f;
    // Bounding boxes by element identity: each element's own flag decides its
    // child.  A misplaced pair is accounted by its left thread (the right
    // thread's coordinates may already be swapped and are not used).
    const double px = active ? x[k] : 0, py = active ? y[k] : 0;
    double qx = 0, qy = 0;
    int partner = -1;
    if (left_mis) {
        const int r = (k - v0) - (P[k] - P[v0]);
        partner = right_pos[v0 + node_m[slot] - 1 - r];
        qx = x[partner];
        qy = y[partner];
    }
    if (bbox) {
        const bool own = active && !right_mis;
        bbox_add_wave(nv, c + (f ? 0 : 1), px, py, own);
        bbox_add_wave(nv, c, qx, qy, left_mis);
    }
    if (active) enode[k] = (k < Labs) ? c : c + 1;
    // Only the left thread of a misplaced pair writes either position.
    if (left_mis) {
        const int32_t ik = id[k], ip = id[partner];
        x[k] = qx;
        y[k] = qy;
        id[k] = ip;
        x[partner] = px;
        y[partner] = py;
        id[partner] = ik;
    }
}

// Segments descend one level (stable partition == per-segment descent).
__global__ void k_seg_step(NodeView nv, const vvhd_seg* segs, int nseg, int32_t* seg_node) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    const int node = seg_node[s];
    if (nv.split[node] != 2) return;
    const int c = nv.ch1[node] + (goes_left(nv, node, segs[s].x, segs[s].y) ? 0 : 1);
    seg_node[s] = c;
    atomicAdd(&nv.nseg[c], 1);
    bbox_add(nv, c, segs[s].x, segs[s].y);
}


// A heat or streak list in tree order: positions from the element arrays,
// the other fields gathered from the list's own arrays (id holds element
// numbers, base is the list's first element).
__global__ void k_gather_list(const int32_t* id, int n, int base, const double* ex, const double* ey, const double* g,
                              const double* vx, const double* vy, const double* ie, double* x2, double* y2, double* g2,
                              double* vx2, double* vy2, double* ie2) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int s = id[k] - base;
    x2[k] = ex[k];
    y2[k] = ey[k];
    g2[k] = g[s];
    vx2[k] = vx[s];
    vy2[k] = vy[s];
    ie2[k] = ie[s];
}

// Gathers the non-key vortex fields into tree order.
__global__ void k_gather(const int32_t* id, int n, const double* g, const double* vx, const double* vy,
                         const double* ie, double* g2, double* vx2, double* vy2, double* ie2) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int s = id[k];
    g2[k] = g[s];
    vx2[k] = vx[s];
    vy2[k] = vy[s];
    ie2[k] = ie[s];
}

struct CMView {
    double *px, *py, *pg, *mx, *my, *mg;
    float *pie, *mie;
};

// Bottom-up over one level: subtree leaf counts, and monopoles --
// CalculateCMassFromScratch (TSortedTree.cpp:176-197) for leaves (array
// order) and CalculateCMass (:150-174) from the children otherwise.
__global__ void k_up_level(NodeView nv, CMView cm, int lb, int le, const double* x, const double* y, const double* g,
                           int32_t* nleaves) {
    const int id = lb + blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= le) return;
    if (nv.split[id] != 2) {
        nleaves[id] = 1;
        double px = 0, py = 0, pg = 0, mx = 0, my = 0, mg = 0;
        for (int k = nv.v0[id]; k < nv.v1[id]; k++) {
            const double gk = g[k];
            if (gk > 0) {
                px += x[k] * gk;
                py += y[k] * gk;
                pg += gk;
            } else {
                mx += x[k] * gk;
                my += y[k] * gk;
                mg += gk;
            }
        }
        if (pg) {
            px *= (1. / pg);
            py *= (1. / pg);
        } else {
            px = nv.x[id];
            py = nv.y[id];
        }
        if (mg) {
            mx *= (1. / mg);
            my *= (1. / mg);
        } else {
            mx = nv.x[id];
            my = nv.y[id];
        }
        cm.px[id] = px;
        cm.py[id] = py;
        cm.pg[id] = pg;
        cm.mx[id] = mx;
        cm.my[id] = my;
        cm.mg[id] = mg;
        cm.pie[id] = INFINITY;
        cm.mie[id] = INFINITY;
        return;
    }
    const int a = nv.ch1[id], b = a + 1;
    nleaves[id] = nleaves[a] + nleaves[b];
    {
        const double ag = cm.pg[a], bg = cm.pg[b];
        const double s = ag + bg;
        if (s) {
            const double inv = 1. / s;
            cm.px[id] = (cm.px[a] * ag + cm.px[b] * bg) * inv;
            cm.py[id] = (cm.py[a] * ag + cm.py[b] * bg) * inv;
            cm.pg[id] = s;
            cm.pie[id] = INFINITY;
        } else {   // TObj(x, y, 0): _1_eps = 0
            cm.px[id] = nv.x[id];
            cm.py[id] = nv.y[id];
            cm.pg[id] = 0;
            cm.pie[id] = 0;
        }
    }
    {
        const double ag = cm.mg[a], bg = cm.mg[b];
        const double s = ag + bg;
        if (s) {
            const double inv = 1. / s;
            cm.mx[id] = (cm.mx[a] * ag + cm.mx[b] * bg) * inv;
            cm.my[id] = (cm.my[a] * ag + cm.my[b] * bg) * inv;
            cm.mg[id] = s;
            cm.mie[id] = INFINITY;
        } else {
            cm.mx[id] = nv.x[id];
            cm.my[id] = nv.y[id];
            cm.mg[id] = 0;
            cm.mie[id] = 0;
        }
    }
}

// Top-down over one level: first leaf rank of every subtree (DFS order).
__global__ void k_down_level(NodeView nv, int lb, int le, const int32_t* nleaves, int32_t* first, int32_t* leaf_rank,
                             int32_t* leaf_node, int32_t* vleaf, int32_t* pre) {
    const int id = lb + blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= le) return;
    const int f = first[id];
    if (nv.split[id] == 2) {
        const int a = nv.ch1[id];
        first[a] = f;
        first[a + 1] = f + nleaves[a];
        const int p = pre[id];   // the first child's subtree (2 nleaves - 1 nodes) follows the node
        pre[a] = p + 1;
        pre[a + 1] = p + 2 * nleaves[a];
    } else {
        leaf_rank[id] = f;
        leaf_node[f] = id;
        for (int k = nv.v0[id]; k < nv.v1[id]; k++) vleaf[k] = f;
        for (int k = nv.h0[id]; k < nv.h1[id]; k++) vleaf[k] = f;
        for (int k = nv.s0[id]; k < nv.s1[id]; k++) vleaf[k] = f;
    }
}

// FindNearNodes (TSortedTree.cpp:199-217) split over the top of the tree:
// the "entries" are the nodes at depth kSplitDepth plus the shallower
// leaves, in DFS order (at most 2^kSplitDepth).  One wave handles one entry
// for 64 DFS-consecutive leaves: each lane first tests the entry's
// ancestors for its leaf, then the wave walks the entry's subtree once (a
// node is expanded while any lane still needs it; a lane that classifies a
// node as far ignores the rest of that subtree).  The split gives
// 2^kSplitDepth times more independent waves than leaf groups, which hides
// the latency of the dependent node loads.  Emissions follow the
// reference's preorder: a far ancestor is emitted by the first entry beneath
// it only, and the per-(leaf, entry) segments are concatenated in entry
// order.
#ifndef VVHD_LIST_SPLIT_DEPTH
#define VVHD_LIST_SPLIT_DEPTH 6
#endif
constexpr int kSplitDepth = VVHD_LIST_SPLIT_DEPTH;
constexpr int kMaxEntries = 1 << kSplitDepth;

struct ListEntries {
    int n;                                 // number of entries
    int node[kMaxEntries];                 // entry node (preorder position)
    int depth[kMaxEntries];                // its depth (<= kSplitDepth)
    int anc[kMaxEntries][kSplitDepth];     // ancestors, root first (preorder positions)
    int lead[kMaxEntries];                 // bit d: first entry beneath anc[d]
};

// Entry table of every root, one wave per root: the frontier starts at the
// root and is expanded kSplitDepth times (an internal node above the split
// depth is replaced by its two children in place), which keeps it in DFS
// order; at most 2^kSplitDepth = 64 entries, one per lane.
__global__ void k_list_entries(NodeView nv, const int32_t* pre, int nroots, ListEntries* Etab) {
    __shared__ int fr[2][kMaxEntries];
    const int r = blockIdx.x;
    const int lane = threadIdx.x;
    ListEntries* E = Etab + r;
    if (lane == 0) fr[0][0] = r;
    int n = 1, cur = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int step = 0; step < kSplitDepth; step++) {
        const int T = lane < n ? fr[cur][lane] : -1;
        int c = -1;
        if (T >= 0 && nv.depth[T] < kSplitDepth) c = nv.ch1[T];
        const int out = T < 0 ? 0 : (c >= 0 ? 2 : 1);
        const int incl = wave_incl_scan(out);
        const int pos = incl - out;
        if (c >= 0) {
            fr[cur ^ 1][pos] = c;
            fr[cur ^ 1][pos + 1] = c + 1;
        } else if (T >= 0) {
            fr[cur ^ 1][pos] = T;
        }
        n = __shfl(incl, kWave - 1, kWave);
        cur ^= 1;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (lane < n) {
        const int T = fr[cur][lane];
        const int d = nv.depth[T];
        E->node[lane] = pre[T];
        E->depth[lane] = d;
        // ancestors, root first; T is the first entry beneath an ancestor when
        // every step from it down is to a first child
        int c = T, lead = 0;
        bool first = true;
        for (int k = d - 1; k >= 0; k--) {
            const int p = nv.parent[c];
            E->anc[lane][k] = pre[p];
            first = first && (nv.ch1[p] == c);
            if (first) lead |= 1 << k;
            c = p;
        }
        E->lead[lane] = lead;
    }
    if (lane == 0) E->n = n;
}

__global__ void k_pack_boxes(NodeView nv, int nnodes, double4* out) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nnodes) return;
    out[id] = make_double4(double_from_ord(nv.blx[id]), double_from_ord(nv.bly[id]), double_from_ord(nv.trx[id]),
                           double_from_ord(nv.try_[id]));
}

// What the list walk reads of a node, one 32-byte record per node stored at
// the node's preorder position: centre, h + w (added first, as the far test
// does), node id, and the position that follows its subtree (position + 1
// for a leaf).  A walk then only moves forward through memory: into the
// subtree at position + 1, or past it.
struct TravNode {
    double x, y, hw;
    int32_t id, skip;
};
static_assert(sizeof(TravNode) == 32, "TravNode layout");

__global__ void k_trav_nodes(NodeView nv, int nnodes, const int32_t* pre, const int32_t* nleaves, TravNode* out) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nnodes) return;
    TravNode r;
    r.x = nv.x[id];
    r.y = nv.y[id];
    r.hw = nv.h[id] + nv.w[id];
    r.id = id;
    const int p = pre[id];
    r.skip = p + 2 * nleaves[id] - 1;
    out[p] = r;
}

// Near and far lists of every leaf, two passes (counts, then entries at the
// scanned offsets).  A wave takes one entry of one root for a group of up to
// 64 DFS-consecutive leaves (one leaf per lane) and writes slot (leaf,
// entry).  The wave walks the entry's subtree in preorder without a stack:
// from a node it moves to the next position (the first child) when some lane
// still needs the node's subtree, otherwise past the subtree.  A lane that
// classifies a node as far (or reaches a near leaf) waits until the walk
// leaves that subtree.
__global__ __launch_bounds__(kBlock) void k_lists(NodeView nv, const TravNode* __restrict__ tn,
                                                  const ListEntries* __restrict__ Etab,
                                                  const int32_t* __restrict__ groups, int ngroups,
                                                  const int32_t* leaf_node, double fc, int nnodes,
                                                  const int64_t* near_off, const int64_t* far_off,
                                                  int32_t* near_list, int32_t* far_list, int32_t* near_cnt,
                                                  int32_t* far_cnt) {
    const int w = threadIdx.x / kWave;
    const int wave = blockIdx.x * (kBlock / kWave) + w;
    const int q = wave % kMaxEntries;
    const int grp = wave / kMaxEntries;
    if (grp >= ngroups) return;
    const int lane = lane_id();
    // a group: up to 64 DFS-consecutive leaves of one root
    const int l = groups[3 * grp] + lane;
    const ListEntries* E = Etab + groups[3 * grp + 2];
    if (q >= E->n) return;   // uniform over the wave
    const bool valid = lane < groups[3 * grp + 1];
    if (__ballot(valid) == 0) return;
    const int L = valid ? leaf_node[l] : 0;
    const double lx = nv.x[L], ly = nv.y[L], lh = nv.h[L], lw = nv.w[L];   // (h_T + w_T) + lh + lw below
    const size_t slot = (size_t)l * kMaxEntries + q;
    int64_t pn = (valid && near_off) ? near_off[slot] : 0, pf = (valid && far_off) ? far_off[slot] : 0;
    int cn = 0, cf = 0;
    // the position is uniform: records come through the scalar cache, all
    // 32 bytes in one load
    auto rec = [&](int p) {
        const uint4* q4 = reinterpret_cast<const uint4*>(tn + p);
        const uint4 u0 = q4[0], u1 = q4[1];
        TravNode r;
        __builtin_memcpy(&r, &u0, 16);
        __builtin_memcpy(reinterpret_cast<char*>(&r) + 16, &u1, 16);
        return r;
    };
    const int P0 = __builtin_amdgcn_readfirstlane(E->node[q]);
    const int dE = __builtin_amdgcn_readfirstlane(E->depth[q]);
    const int lead = __builtin_amdgcn_readfirstlane(E->lead[q]);
    // the entry's own record and its ancestors' are all loaded before any is
    // used, so their latencies overlap
    TravNode anc[kSplitDepth];
#pragma unroll
    for (int k = 0; k < kSplitDepth; k++)
        if (k < dE) anc[k] = rec(__builtin_amdgcn_readfirstlane(E->anc[q][k]));
    const int end = min(rec(P0).skip, nnodes);
    // a lane takes part at positions >= resume
    int resume = valid ? P0 : end;
    // ancestors of the entry: the first far one (if any) cuts the whole entry
#pragma unroll
    for (int k = 0; k < kSplitDepth; k++) {
        if (k >= dE || !valid || resume == end) continue;
        const TravNode& a = anc[k];
        const double dx = a.x - lx, dy = a.y - ly;
        const double hp = a.hw + lh + lw;
        if (dx * dx + dy * dy > fc * hp * hp) {
            if ((lead >> k) & 1) {
                if (far_list) far_list[pf++] = a.id;
                cf++;
            }
            resume = end;
        }
    }
    int P = __builtin_amdgcn_readfirstlane(__ballot(resume < end) != 0 ? P0 : end);
    while (P < end) {
        const TravNode r = rec(P);
        const int skip = __builtin_amdgcn_readfirstlane(r.skip);
        const int id = __builtin_amdgcn_readfirstlane(r.id);
        const bool act = P >= resume;
        const bool leaf = skip == P + 1;
        const double dx = r.x - lx, dy = r.y - ly;
        const double hp = r.hw + lh + lw;
        const bool far = act && (dx * dx + dy * dy > fc * hp * hp);
        if (far) {
            if (far_list) far_list[pf++] = id;
            cf++;
            resume = skip;
        } else if (act && leaf) {
            if (near_list) near_list[pn++] = id;
            cn++;
        }
        const bool need = act && !far && !leaf;
        P = __builtin_amdgcn_readfirstlane(__ballot(need) != 0 ? P + 1 : max(skip, P + 1));
    }
    if (valid && near_cnt) near_cnt[slot] = cn;
    if (valid && far_cnt) far_cnt[slot] = cf;
}

// Per-leaf CSR offsets from the per-(leaf, entry) offsets.
__global__ void k_leaf_offsets(const int64_t* slot_off, int nleaf, int64_t* leaf_off) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l <= nleaf) leaf_off[l] = slot_off[(size_t)l * kMaxEntries];
}

// Segments: leaf of each segment, then stable order by (leaf, index).
// Leaf (DFS rank) of every segment, and the segment numbers to sort by it.
__global__ void k_seg_leaf(const int32_t* seg_node, const int32_t* leaf_rank, int nseg, int32_t* seg_leaf,
                           int32_t* ids) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    seg_leaf[s] = leaf_rank[seg_node[s]];
    ids[s] = s;
}

__global__ void k_seg_off(const int32_t* seg_leaf, const int32_t* seg_list, int nseg, int nleaf, int64_t* seg_off) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l > nleaf) return;
    // first position whose leaf >= l
    int lo = 0, hi = nseg;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (seg_leaf[seg_list[mid]] < l)
            lo = mid + 1;
        else
            hi = mid;
    }
    seg_off[l] = lo;
}

__global__ void k_iota(int32_t* a, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] = k;
}

__global__ void k_fill_i32(int32_t* a, int n, int32_t v) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] = v;
}

template <typename T>
void swap_buf(DBuf<T>& a, DBuf<T>& b) {
    std::swap(a.p, b.p);
    std::swap(a.cap, b.cap);
}

// exclusive scan of int32 counts into int64 offsets (n+1 outputs)
void scan_i32_to_i64(Ctx& c, const int32_t* in, int64_t* out, int64_t n) {
    size_t bytes = 0;
    VVHD_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, (int64_t)0, (size_t)n + 1, rocprim::plus<int64_t>(),
                                     c.stream));
    c.tmp.grow(bytes + 16);
    VVHD_HIP(rocprim::exclusive_scan(c.tmp.p, bytes, in, out, (int64_t)0, (size_t)n + 1, rocprim::plus<int64_t>(),
                                     c.stream));
}

// Exclusive prefix of the goes-left flags in two passes over blocks of kScanSpan
// elements: k_flag_count stores every flag and each block's count,
// k_flag_prefix adds the counts of the blocks before it to a block scan.
// The flag and prefix arrays are padded to whole blocks (zero flags).
constexpr int kScanSpan = 2048;   // 256 threads x 8 elements

__device__ inline int block_sum_256(int v, int* s_w) {
    v = wave_sum_i(v);
    const int w = threadIdx.x / kWave;
    __syncthreads();
    if (lane_id() == 0) s_w[w] = v;
    __syncthreads();
    return s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

__global__ __launch_bounds__(256) void k_flag_count(GoesLeftFlag f, int64_t m, int32_t* flag, int32_t* bcount) {
    __shared__ int s_w[4];
    const int64_t base = (int64_t)blockIdx.x * kScanSpan;
    int cnt = 0;
    for (int i = 0; i < kScanSpan / 256; i++) {
        const int64_t k = base + i * 256 + threadIdx.x;
        const int v = k < m ? f((int)k) : 0;
        flag[k] = v;
        cnt += v;
    }
    cnt = block_sum_256(cnt, s_w);
    if (threadIdx.x == 0) bcount[blockIdx.x] = cnt;
}

__global__ __launch_bounds__(256) void k_flag_prefix(const int32_t* flag, const int32_t* bcount, int32_t* P) {
    __shared__ int s_w[4];
    const int b = blockIdx.x, t = threadIdx.x, w = t / kWave;
    int pre = 0;
    for (int j = t; j < b; j += 256) pre += bcount[j];
    pre = block_sum_256(pre, s_w);
    const int64_t k0 = (int64_t)b * kScanSpan + 8 * t;
    const int4 lo = *reinterpret_cast<const int4*>(flag + k0), hi = *reinterpret_cast<const int4*>(flag + k0 + 4);
    const int v[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
    int ex[8], tot = 0;
    for (int i = 0; i < 8; i++) {
        ex[i] = tot;
        tot += v[i];
    }
    const int incl = wave_incl_scan(tot);
    __syncthreads();   // s_w is reused
    if (lane_id() == kWave - 1) s_w[w] = incl;
    __syncthreads();
    int woff = 0;
    for (int q = 0; q < w; q++) woff += s_w[q];
    const int start = pre + woff + incl - tot;
    *reinterpret_cast<int4*>(P + k0) = make_int4(start + ex[0], start + ex[1], start + ex[2], start + ex[3]);
    *reinterpret_cast<int4*>(P + k0 + 4) = make_int4(start + ex[4], start + ex[5], start + ex[6], start + ex[7]);
}

// P[k] = number of flags before k, for k < m (flag is written too); the
// arrays hold scan_blocks(m) * kScanSpan entries, bcount scan_blocks(m).
int64_t scan_blocks(int64_t m) { return (m + kScanSpan - 1) / kScanSpan; }

void scan_flags(Ctx& c, const GoesLeftFlag& f, int32_t* flag, int32_t* P, int32_t* bcount, int64_t m) {
    const int64_t nb = scan_blocks(m);
    hipLaunchKernelGGL(k_flag_count, dim3((unsigned)nb), dim3(256), 0, c.stream, f, m, flag, bcount);
    hipLaunchKernelGGL(k_flag_prefix, dim3((unsigned)nb), dim3(256), 0, c.stream, (const int32_t*)flag,
                       (const int32_t*)bcount, P);
}

}  // namespace

void tree_build(Ctx& c, int far_criteria, double min_node, double max_node, int32_t* perm_out) {
    TreeArrays& t = c.t;
    const int nvort = (int)c.n;
    const int nh = (int)c.heat.n, nst = (int)c.streak.n;
    const int nseg = (int)c.segs_h.size();
    VVHD_CHECK(c.n + c.heat.n + c.streak.n < INT32_MAX / 2, VVHD_EINVAL, "tree_build: too many particles");
    VVHD_CHECK(far_criteria >= 0, VVHD_EINVAL, "tree_build: far_criteria must be non-negative");
    t.far_criteria = far_criteria;
    t.min_node = min_node;
    t.max_node = max_node;
    sim_sync(c);
    const int nroots = c.nsim;
    VVHD_CHECK(nroots == 1 || nh + nst == 0, VVHD_EUNSUPPORTED,
               "tree_build: heat and streak lists with several simulations in one context");
    // elements: the vortices, then the heat and the streak particles
    const int n = nvort + nh + nst;
    t.hbase = nvort;
    t.sbase = nvort + nh;
    t.nelem = n;
    if (nh + nst) {
        reserve_vortices(c, n + 1, true);
        if (nh) {
            VVHD_HIP(hipMemcpyAsync(c.x.p + nvort, c.heat.x.p, (size_t)nh * 8, hipMemcpyDeviceToDevice, c.stream));
            VVHD_HIP(hipMemcpyAsync(c.y.p + nvort, c.heat.y.p, (size_t)nh * 8, hipMemcpyDeviceToDevice, c.stream));
        }
        if (nst) {
            VVHD_HIP(hipMemcpyAsync(c.x.p + t.sbase, c.streak.x.p, (size_t)nst * 8, hipMemcpyDeviceToDevice, c.stream));
            VVHD_HIP(hipMemcpyAsync(c.y.p + t.sbase, c.streak.y.p, (size_t)nst * 8, hipMemcpyDeviceToDevice, c.stream));
        }
    }
    const int cap = 2 * (n + nseg) + 64 + nroots;
    for (auto* b : {&t.x, &t.y, &t.h, &t.w}) b->grow(cap);
    for (auto* b : {&t.blx, &t.bly, &t.trx, &t.try_, &t.path}) b->grow(cap);
    for (auto* b : {&t.v0, &t.v1, &t.ch1, &t.parent, &t.depth, &t.nseg, &t.split, &t.nleaves, &t.arrive, &t.leaf_rank,
                    &t.pre, &t.h0, &t.h1, &t.s0, &t.s1})
        b->grow(cap);
    for (auto* b : {&t.cmp_x, &t.cmp_y, &t.cmp_g, &t.cmm_x, &t.cmm_y, &t.cmm_g}) b->grow(cap);
    t.cmp_ie.grow(cap);
    t.cmm_ie.grow(cap);
    t.counters.grow(16);
    t.leaf_node.grow(cap);
    t.vleaf.grow(n + 1);
    c.iscratch0.grow(n + 1);   // element node
    const int64_t scan_pad = scan_blocks((int64_t)n + 1) * kScanSpan;
    c.iscratch1.grow(std::max<int64_t>(n + 2, scan_pad));   // flags
    c.iscratch2.grow(std::max<int64_t>(n + 2, scan_pad));   // prefix
    c.tmp.grow(scan_blocks((int64_t)n + 1) * sizeof(int32_t) + 16);   // per-block flag counts
    c.iscratch3.grow(2 * (size_t)n + 2);   // left/right positions
    DBuf<int32_t>& enode = c.iscratch0;
    int32_t* flag = c.iscratch1.p;
    int32_t* P = c.iscratch2.p;
    int32_t* left_pos = c.iscratch3.p;
    int32_t* right_pos = c.iscratch3.p + n + 1;
    DBuf<int32_t>& node_L = t.node_L;
    DBuf<int32_t>& node_m = t.node_m;
    DBuf<int32_t>& leaf_tmp = t.leaf_tmp;
    DBuf<int32_t>& seg_node = t.seg_node;
    node_L.grow(3 * (size_t)cap);
    node_m.grow(3 * (size_t)cap);
    leaf_tmp.grow(cap);
    seg_node.grow(nseg + 1);
    NodeView nv = view(t);
    hipStream_t s = c.stream;

    hipLaunchKernelGGL(k_init_roots, dim3(nblocks(nroots)), dim3(kBlock), 0, s, nv, nroots, (const int64_t*)c.sim_off.p,
                       t.counters.p, t.hbase, t.sbase, n);
    hipLaunchKernelGGL(k_iota, dim3(nblocks(n)), dim3(kBlock), 0, s, c.id.p, n);
    hipLaunchKernelGGL(k_fill_roots, dim3(nblocks(n)), dim3(kBlock), 0, s, enode.p, n, (const int64_t*)c.sim_off.p,
                       nroots);
    if (nseg)
        hipLaunchKernelGGL(k_seg_roots, dim3(nblocks(nseg)), dim3(kBlock), 0, s, nv, (const int32_t*)c.seg_root.p,
                           nseg, seg_node.p);
    if (n && nroots <= kTopNodes)
        hipLaunchKernelGGL(k_bbox_top, dim3(nblocks(n, kTopChunk)), dim3(kBlock), 0, s, nv, (const int32_t*)nullptr,
                           (const int32_t*)enode.p, c.x.p, c.y.p, n);
    else if (n)
        hipLaunchKernelGGL(k_bbox_elems, dim3(nblocks(n)), dim3(kBlock), 0, s, nv, (const int32_t*)enode.p, c.x.p,
                           c.y.p, n);
    if (nseg)
        hipLaunchKernelGGL(k_root_segs, dim3(nblocks(nseg)), dim3(kBlock), 0, s, nv, c.segs.p,
                           (const int32_t*)c.seg_root.p, nseg);

    std::vector<std::pair<int, int>>& levels = t.levels;   // node id range of every level
    levels.clear();
    levels.push_back({0, nroots});
    // One level per iteration.  The kernels take their node range from the
    // device counters; the host reads a level's counters (copied to a pinned
    // slot) only to record the range and to stop, and it keeps kLevSlots - 1
    // further levels enqueued meanwhile, so the device does not wait for the
    // host between levels.  Levels after the last one find an empty range and
    // do nothing.
    auto launch_level = [&](int d) {
        const int node_blocks = 1024;
        hipLaunchKernelGGL(k_decide, dim3(node_blocks), dim3(kBlock), 0, s, nv, t.counters.p, leaf_tmp.p, cap,
                           min_node, max_node);
        scan_flags(c, GoesLeftFlag{nv, enode.p, c.x.p, c.y.p, n}, flag, P, reinterpret_cast<int32_t*>(c.tmp.p), n + 1);
        hipLaunchKernelGGL(k_ranks, dim3(nblocks(n + 1)), dim3(kBlock), 0, s, nv, enode.p, P, n, t.hbase, t.sbase, cap,
                           node_L.p, node_m.p, left_pos, right_pos, t.counters.p, (const int32_t*)flag);
        // the children of level d number at most nroots 2^(d+1)
        const bool top = ((int64_t)nroots << std::min<int>(d + 1, 30)) <= kTopNodes;
        hipLaunchKernelGGL(k_swap, dim3(nblocks(n)), dim3(kBlock), 0, s, nv, enode.p, flag, P, n, t.hbase, t.sbase,
                           cap, node_L.p, node_m.p, left_pos, right_pos, c.x.p, c.y.p, c.id.p, false);
        if (top && n)
            hipLaunchKernelGGL(k_bbox_top, dim3(nblocks(n, kTopChunk)), dim3(kBlock), 0, s, nv,
                               (const int32_t*)(t.counters.p + 2), (const int32_t*)enode.p, c.x.p, c.y.p, n);
        else if (n) {
            // at most nroots 2^(d+1) children; 16 lanes each once they average under 256 elements
            const int64_t kids = std::min<int64_t>((int64_t)nroots << std::min<int>(d + 1, 30), cap);
            const bool small = (int64_t)n < 256 * kids;
            const int64_t waves = small ? (kids + 3) / 4 : kids;
            const int grid = (int)std::min<int64_t>(2048, std::max<int64_t>(1, (waves + 3) / 4));
            if (small)
                hipLaunchKernelGGL(k_child_boxes<16>, dim3(grid), dim3(kBlock), 0, s, nv,
                                   (const int32_t*)t.counters.p, c.x.p, c.y.p);
            else
                hipLaunchKernelGGL(k_child_boxes<64>, dim3(grid), dim3(kBlock), 0, s, nv,
                                   (const int32_t*)t.counters.p, c.x.p, c.y.p);
        }
        if (nseg)
            hipLaunchKernelGGL(k_seg_step, dim3(nblocks(nseg)), dim3(kBlock), 0, s, nv, c.segs.p, nseg, seg_node.p);
        int32_t* slot = c.lev_counters + 8 * (d % Ctx::kLevSlots);
        VVHD_HIP(hipMemcpyAsync(slot, t.counters.p, 8 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        VVHD_HIP(hipEventRecord(c.lev_ev[d % Ctx::kLevSlots], s));
    };
    int32_t hc[8];
    constexpr int ahead = Ctx::kLevSlots - 1;
    for (int d = 0; d < ahead; d++) launch_level(d);
    for (int d = 0;; d++) {
        launch_level(d + ahead);
        VVHD_HIP(hipEventSynchronize(c.lev_ev[d % Ctx::kLevSlots]));
        std::memcpy(hc, c.lev_counters + 8 * (d % Ctx::kLevSlots), sizeof(hc));
        VVHD_CHECK(hc[4] == 0, VVHD_ENOMEM, "tree_build: node capacity exceeded");
        if (hc[3] == hc[2]) break;   // level d made no new nodes; the levels after it did nothing
        levels.push_back({hc[2], hc[3]});
        VVHD_CHECK((int)levels.size() < kMaxLevels, VVHD_EUNSUPPORTED, "tree_build: tree deeper than supported");
    }
    t.nnodes = hc[0];
    t.nleaves_h = hc[1];
    t.max_depth = hc[5];
    const int nl = t.nleaves_h;

    // tree order of the remaining fields
    hipLaunchKernelGGL(k_gather, dim3(nblocks(nvort)), dim3(kBlock), 0, s, c.id.p, nvort, c.g.p, c.vx.p, c.vy.p,
                       c.ie.p, c.g2.p, c.vx2.p, c.vy2.p, c.ie2.p);
    swap_buf(c.g, c.g2);
    swap_buf(c.vx, c.vx2);
    swap_buf(c.vy, c.vy2);
    swap_buf(c.ie, c.ie2);
    if (perm_out && nvort) VVHD_HIP(hipMemcpyAsync(perm_out, c.id.p, (size_t)nvort * 4, hipMemcpyDeviceToHost, s));
    // the heat and streak lists in their new order
    for (int q = 0; q < 2; q++) {
        PList& L = q == 0 ? c.heat : c.streak;
        const int base = q == 0 ? t.hbase : t.sbase, m = q == 0 ? nh : nst;
        if (!m) continue;
        hipLaunchKernelGGL(k_gather_list, dim3(nblocks(m)), dim3(kBlock), 0, s, c.id.p + base, m, base, c.x.p + base,
                           c.y.p + base, L.g.p, L.vx.p, L.vy.p, L.ie.p, L.x2.p, L.y2.p, L.g2.p, L.vx2.p, L.vy2.p,
                           L.ie2.p);
        L.swap_buffers();
    }

    // leaf counts and monopoles bottom-up, DFS ranks top-down
    CMView cm{t.cmp_x.p, t.cmp_y.p, t.cmp_g.p, t.cmm_x.p, t.cmm_y.p, t.cmm_g.p, t.cmp_ie.p, t.cmm_ie.p};
    for (int d = (int)levels.size() - 1; d >= 0; d--) {
        const int lb = levels[d].first, le = levels[d].second;
        hipLaunchKernelGGL(k_up_level, dim3(nblocks(le - lb)), dim3(kBlock), 0, s, nv, cm, lb, le, c.x.p, c.y.p,
                           c.g.p, t.nleaves.p);
    }
    int32_t* first = t.arrive.p;
    t.root_first.grow(nroots + 1);
    hipLaunchKernelGGL(k_root_first, dim3(1), dim3(1), 0, s, t.nleaves.p, nroots, first, t.root_first.p, t.pre.p);
    for (int d = 0; d < (int)levels.size(); d++) {
        const int lb = levels[d].first, le = levels[d].second;
        hipLaunchKernelGGL(k_down_level, dim3(nblocks(le - lb)), dim3(kBlock), 0, s, nv, lb, le, t.nleaves.p, first,
                           t.leaf_rank.p, t.leaf_node.p, t.vleaf.p, t.pre.p);
    }
    // leaf -> simulation and the list builder's leaf groups (64 leaves of one root)
    t.leaf_sim.grow(nl + 1);
    std::vector<int32_t> grp;
    if (nroots == 1) {
        if (nl) hipLaunchKernelGGL(k_fill_i32, dim3(nblocks(nl)), dim3(kBlock), 0, s, t.leaf_sim.p, nl, 0);
        for (int f = 0; f < nl; f += kWave) grp.insert(grp.end(), {f, std::min(kWave, nl - f), 0});
    } else {
        std::vector<int32_t> rf(nroots + 1);
        VVHD_HIP(hipMemcpyAsync(rf.data(), t.root_first.p, (nroots + 1) * 4, hipMemcpyDeviceToHost, s));
        VVHD_HIP(hipStreamSynchronize(s));
        if (nl)
            hipLaunchKernelGGL(k_leaf_sim, dim3(nblocks(nl)), dim3(kBlock), 0, s, (const int32_t*)t.root_first.p,
                               nroots, nl, t.leaf_sim.p);
        for (int r = 0; r < nroots; r++)
            for (int f = rf[r]; f < rf[r + 1]; f += kWave) grp.insert(grp.end(), {f, std::min(kWave, rf[r + 1] - f), r});
    }
    t.ngroups = (int)grp.size() / 3;
    t.groups.grow(grp.size() + 3);
    if (!grp.empty())
        VVHD_HIP(hipMemcpyAsync(t.groups.p, grp.data(), grp.size() * 4, hipMemcpyHostToDevice, s));
    // segments per leaf
    t.seg_off.grow(nl + 1);
    t.seg_leaf.grow(nseg + 1);
    t.seg_list.grow(nseg + 1);
    if (nseg) {
        // segments of every leaf in their own order (the stable partition of
        // the segment list, TSortedTree.cpp:139-148): a stable radix sort of
        // the segment numbers by leaf
        t.seg_keys.grow(nseg + 1);
        t.seg_ids.grow(nseg + 1);
        hipLaunchKernelGGL(k_seg_leaf, dim3(nblocks(nseg)), dim3(kBlock), 0, s, seg_node.p, t.leaf_rank.p, nseg,
                           t.seg_leaf.p, t.seg_ids.p);
        int bits = 1;
        while (bits < 31 && (1 << bits) <= nl) bits++;
        size_t bytes = 0;
        VVHD_HIP(rocprim::radix_sort_pairs(nullptr, bytes, t.seg_leaf.p, t.seg_keys.p, t.seg_ids.p, t.seg_list.p,
                                           (unsigned)nseg, 0, bits, s));
        c.tmp.grow(bytes + 16);
        VVHD_HIP(rocprim::radix_sort_pairs(c.tmp.p, bytes, t.seg_leaf.p, t.seg_keys.p, t.seg_ids.p, t.seg_list.p,
                                           (unsigned)nseg, 0, bits, s));
        hipLaunchKernelGGL(k_seg_off, dim3(nblocks(nl + 1)), dim3(kBlock), 0, s, t.seg_leaf.p, t.seg_list.p, nseg, nl,
                           t.seg_off.p);
    } else {
        VVHD_HIP(hipMemsetAsync(t.seg_off.p, 0, (size_t)(nl + 1) * 8, s));
    }
    t.built = true;
    tree_lists(c);
    pack_boxes(c);
}

void pack_boxes(Ctx& c) {
    TreeArrays& t = c.t;
    t.pbox.grow(std::max(t.nnodes, 1));
    if (t.nnodes)
        hipLaunchKernelGGL(k_pack_boxes, dim3(nblocks(t.nnodes)), dim3(kBlock), 0, c.stream, view(t), t.nnodes,
                           t.pbox.p);
}

void tree_lists(Ctx& c) {
    TreeArrays& t = c.t;
    const int nl = t.nleaves_h;
    NodeView nv = view(t);
    hipStream_t s = c.stream;
    const size_t nslot = (size_t)nl * kMaxEntries;
    DBuf<int32_t>& cnt_n = t.cnt_near;
    DBuf<int32_t>& cnt_f = t.cnt_far;
    cnt_n.grow(nslot + 1);
    cnt_f.grow(nslot + 1);
    t.slot_near.grow(nslot + 1);
    t.slot_far.grow(nslot + 1);
    t.near_off.grow(nl + 1);
    t.far_off.grow(nl + 1);
    const int nroots = c.nsim;
    t.trav.grow((size_t)t.nnodes * sizeof(TravNode));
    TravNode* tn = reinterpret_cast<TravNode*>(t.trav.p);
    hipLaunchKernelGGL(k_trav_nodes, dim3(nblocks(t.nnodes)), dim3(kBlock), 0, s, nv, t.nnodes, t.pre.p, t.nleaves.p,
                       tn);
    t.entries.grow(nroots * sizeof(ListEntries));
    ListEntries* E = reinterpret_cast<ListEntries*>(t.entries.p);
    hipLaunchKernelGGL(k_list_entries, dim3(nroots), dim3(kWave), 0, s, nv, t.pre.p, nroots, E);
    VVHD_HIP(hipMemsetAsync(cnt_n.p, 0, (nslot + 1) * 4, s));
    VVHD_HIP(hipMemsetAsync(cnt_f.p, 0, (nslot + 1) * 4, s));
    const double fc = (double)t.far_criteria;
    const int64_t waves = (int64_t)t.ngroups * kMaxEntries;
    const int grid = (int)std::max<int64_t>(1, (waves + kBlock / kWave - 1) / (kBlock / kWave));
    hipLaunchKernelGGL(k_lists, dim3(grid), dim3(kBlock), 0, s, nv, (const TravNode*)tn, (const ListEntries*)E,
                       (const int32_t*)t.groups.p, t.ngroups, t.leaf_node.p, fc, t.nnodes, (const int64_t*)nullptr,
                       (const int64_t*)nullptr, (int32_t*)nullptr, (int32_t*)nullptr, cnt_n.p, cnt_f.p);
    scan_i32_to_i64(c, cnt_n.p, t.slot_near.p, (int)nslot);
    scan_i32_to_i64(c, cnt_f.p, t.slot_far.p, (int)nslot);
    hipLaunchKernelGGL(k_leaf_offsets, dim3(nblocks(nl + 1)), dim3(kBlock), 0, s, t.slot_near.p, nl, t.near_off.p);
    hipLaunchKernelGGL(k_leaf_offsets, dim3(nblocks(nl + 1)), dim3(kBlock), 0, s, t.slot_far.p, nl, t.far_off.p);
    int64_t tot[2];
    VVHD_HIP(hipMemcpyAsync(&tot[0], t.slot_near.p + nslot, 8, hipMemcpyDeviceToHost, s));
    VVHD_HIP(hipMemcpyAsync(&tot[1], t.slot_far.p + nslot, 8, hipMemcpyDeviceToHost, s));
    VVHD_HIP(hipStreamSynchronize(s));
    t.nnear = tot[0];
    t.nfar = tot[1];
    t.near_list.grow(t.nnear + 1);
    t.far_list.grow(t.nfar + 1);
    hipLaunchKernelGGL(k_lists, dim3(grid), dim3(kBlock), 0, s, nv, (const TravNode*)tn, (const ListEntries*)E,
                       (const int32_t*)t.groups.p, t.ngroups, t.leaf_node.p, fc, t.nnodes, (const int64_t*)t.slot_near.p,
                       (const int64_t*)t.slot_far.p, t.near_list.p, t.far_list.p, (int32_t*)nullptr, (int32_t*)nullptr);
}

}  // namespace vvhd
