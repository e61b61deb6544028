// epsilon.hip -- MEpsilonFast::CalcEpsilonFast (MEpsilonFast.cpp:11-63) on
// the GPU, with the reference's sequential merge semantics.
//
// The reference visits targets leaf by leaf in DFS order, i.e. in array
// order, and for each searches the two nearest vortices (g != 0, not itself)
// of its near leaves; ties go to the first found.  Because leaves in DFS
// order hold increasing array ranges, "first found" is "smaller index", so
// the result is the top-2 under the order (d^2, index): order independent,
// hence computable in parallel.  Distances are FP64 with the reference's
// operation order (no contraction), so indices are bit-exact.
//
// A merge (:111-126) mutates the target and zeroes its partner; every later
// target sees the mutated state.  Pass 1 computes every target on the
// unmodified state and flags the targets that would merge.  When flags exist,
// a single-workgroup pass commits merges in target order: after each commit
// it updates the targets that come later and have the merged leaves in their
// near lists -- in O(1) when the moved vortex only enters their top-2, by a
// full rescan when one of their two neighbours changed -- and re-evaluates
// their merge flags, exactly reproducing the sequential outcome.
//
// Pass 1 prunes near leaves by box distance: leaves touching the target leaf
// are scanned first, then only leaves whose box distance does not exceed the
// largest second-neighbour distance found.  Box distances are computed with
// the same monotone rounded operations, so pruning never drops a candidate.

#include "common.hpp"

#include <cstring>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cmath>

#pragma clang fp contract(off)

namespace vvhd {
namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
inline int nblocks(int64_t n, int b = kBlock) { return (int)std::max<int64_t>(1, (n + b - 1) / b); }

__device__ inline double sgn(double x) { return (x > 0) ? 1 : ((x < 0) ? -1 : 0); }

// Top-2 state under the order (d, index); index -1 marks an empty slot whose
// distance is DBL_MAX (the reference's initial res1/res2).
struct Top2 {
    double d1, d2;
    int i1, i2;
};

__device__ inline bool less_di(double d, int j, double dc, int ic) {
    return d < dc || (d == dc && ic >= 0 && j < ic);
}

__device__ inline void top2_insert(Top2& t, double d, int j) {
    if (less_di(d, j, t.d1, t.i1)) {
        t.d2 = t.d1;
        t.i2 = t.i1;
        t.d1 = d;
        t.i1 = j;
    } else if (less_di(d, j, t.d2, t.i2)) {
        t.d2 = d;
        t.i2 = j;
    }
}

__device__ inline void top2_merge(Top2& t, const Top2& o) {
    if (o.i1 >= 0) top2_insert(t, o.d1, o.i1);
    if (o.i2 >= 0) top2_insert(t, o.d2, o.i2);
}

__device__ inline Top2 top2_shfl_xor(const Top2& t, int mask) {
    Top2 o;
    o.d1 = __shfl_xor(t.d1, mask, kWave);
    o.d2 = __shfl_xor(t.d2, mask, kWave);
    o.i1 = __shfl_xor(t.i1, mask, kWave);
    o.i2 = __shfl_xor(t.i2, mask, kWave);
    return o;
}

__device__ inline Top2 top2_empty() { return Top2{DBL_MAX, DBL_MAX, -1, -1}; }

// (lv->r - lobj->r).abs2() of MEpsilonFast.cpp:141
__device__ inline double dist2(double tx, double ty, double sx, double sy) {
    const double dx = tx - sx, dy = ty - sy;
    return dx * dx + dy * dy;
}

// nearestBodySegment (MEpsilonFast.cpp:214-253) for every leaf centre, then
// the merge criterion and the epsilon floor (:29-47).  The bodies searched
// are those of the leaf's simulation, [sim_body[s], sim_body[s+1]).
__global__ void k_leaf_criteria(TreeDev T, const vvhd_seg* segs, const int32_t* body_range, const int32_t* sim_body,
                                const int32_t* leaf_sim, int merge, double* crit_out, double* restr_out) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= T.nleaf) return;
    const int sim = leaf_sim[l];
    const int b0 = sim_body[sim], b1 = sim_body[sim + 1];
    if (b0 == b1) {   // no segment in the simulation: criterion 0 (or NaN), floor 0
        crit_out[l] = merge ? 0.0 : __builtin_nan("");
        restr_out[l] = 0;
        return;
    }
    const int L = T.leaf_node[l];
    const double px = T.nx[L], py = T.ny[L];
    int att = -1;
    double res = DBL_MAX;
    for (int64_t e = T.near_off[l]; e < T.near_off[l + 1]; e++) {
        const int m = T.leaf_rank[T.near_list[e]];
        for (int64_t q = T.seg_off[m]; q < T.seg_off[m + 1]; q++) {
            const int s = T.seg_list[q];
            const double d2 = dist2(px, py, segs[s].x, segs[s].y);
            if (d2 < res) {
                res = d2;
                att = s;
            }
        }
    }
    if (att < 0) {
        for (int b = b0; b < b1; b++) {
            const int f = body_range[2 * b], cnt = body_range[2 * b + 1];
            for (int k = 0; k < cnt; k++) {
                const double d2 = dist2(px, py, segs[f + k].x, segs[f + k].y);
                if (d2 < res) {
                    res = d2;
                    att = f + k;
                } else {
                    k += 9;   // the reference's "speed up" skip
                }
            }
        }
    }
    double crit;
    if (merge) {
        if (att >= 0) {
            const double dlx = segs[att].dlx, dly = segs[att].dly;
            const double ddx = px - segs[att].x, ddy = py - segs[att].y;
            crit = 0.16 * (dlx * dlx + dly * dly) * (1. + sqrt(ddx * ddx + ddy * ddy));
        } else {
            crit = 0;
        }
    } else {
        crit = __builtin_nan("");
    }
    double restr = 0;
    if (att >= 0) {
        const double dlx = segs[att].dlx, dly = segs[att].dly;
        restr = sqrt(dlx * dlx + dly * dly) * (1.0 / 3.0);
    }
    crit_out[l] = crit;
    restr_out[l] = restr;
}

constexpr int kStageE = 384;   // staged sources per wave (pass 1)

// Pass 1: top-2 of every target on the unmodified state, one wave per leaf
// and chunk of up to 16 targets.  The sources (g != 0) of the selected near
// leaves are staged in LDS, FP64, and lane (target t = l % ntc, slot
// s = l / ntc) runs over every nslot-th of them; the slots' top-2s combine
// afterwards (the top-2 under (d^2, index) does not depend on the order).
// First the leaves touching the target leaf, then those whose box lies
// within the largest second-neighbour distance found.
__global__ __launch_bounds__(kBlock) void k_eps_top2(TreeDev T, const double* x, const double* y, const double* g,
                                                     int merge, const double* leaf_crit, int32_t* nn1, int32_t* nn2,
                                                     double* d1o, double* d2o, uint8_t* eflag, int32_t* ecount,
                                                     int32_t* flag_list) {
    __shared__ double s_x[kWavesPerBlock][kStageE], s_y[kWavesPerBlock][kStageE];
    __shared__ int s_j[kWavesPerBlock][kStageE];
    __shared__ int s_off[kWavesPerBlock][kWave], s_j0[kWavesPerBlock][kWave];
    const int w = threadIdx.x / kWave;
    const int l = xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock + w;
    if (l >= T.nleaf) return;
    const int lane = lane_id();
    const int L = T.leaf_node[l];
    const int v0 = T.v0[L], v1 = T.v1[L];
    const double crit = leaf_crit[l];
    const int64_t e0 = T.near_off[l], e1 = T.near_off[l + 1];
    double* sx = s_x[w];
    double* sy = s_y[w];
    int* sj = s_j[w];
    int* soff = s_off[w];
    int* sj0 = s_j0[w];
    auto wave_sync = [] {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    for (int base = v0; base < v1; base += 16) {
        const int ntc = min(16, v1 - base);
        const int nslot = kWave / ntc;
        const int t = lane % ntc, slot = lane / ntc;
        const int i = base + t;
        const double tx = x[i], ty = y[i], tg = g[i];
        const bool active = tg != 0 && slot < nslot;
        Top2 st = top2_empty();
        auto compute = [&](int nsrc) {
            wave_sync();
            if (active)
                for (int k = slot; k < nsrc; k += nslot) {
                    const int j = sj[k];
                    if (j >= 0 && j != i) top2_insert(st, dist2(tx, ty, sx[k], sy[k]), j);
                }
            wave_sync();
        };
        // sources of the near leaves with lo < box distance <= hi
        auto stage_scan = [&](double lo, double hi) {
            int fill = 0;
            for (int64_t eb = e0; eb < e1; eb += kWave) {
                const int64_t e = eb + lane;
                int cnt = 0, j0 = 0;
                if (e < e1) {
                    const int M = T.near_list[e];
                    const double bd = box_dist2p(T, L, M);
                    if (bd > lo && bd <= hi) {
                        j0 = T.v0[M];
                        cnt = T.v1[M] - j0;
                    }
                }
                const int incl = wave_incl_scan(cnt);
                const int tot = __shfl(incl, kWave - 1, kWave);
                if (tot == 0) continue;
                soff[lane] = incl - cnt;
                sj0[lane] = j0;
                wave_sync();
                for (int done = 0; done < tot;) {
                    const int take = min(tot - done, kStageE - fill);
                    for (int q = lane; q < take; q += kWave) {
                        const int gidx = done + q;
                        int a = 0, b = kWave - 1;   // last lane whose offset <= gidx
                        while (a < b) {
                            const int mid = (a + b + 1) >> 1;
                            if (soff[mid] <= gidx)
                                a = mid;
                            else
                                b = mid - 1;
                        }
                        const int j = sj0[a] + (gidx - soff[a]);
                        sx[fill + q] = x[j];
                        sy[fill + q] = y[j];
                        sj[fill + q] = g[j] != 0 ? j : -1;
                    }
                    fill += take;
                    done += take;
                    if (fill == kStageE) {
                        compute(fill);
                        fill = 0;
                    }
                }
                wave_sync();
            }
            if (fill) compute(fill);
        };
        auto combine = [&](Top2 s) {
            Top2 c = top2_empty();
            for (int q = 0; q < nslot; q++) {
                Top2 o;
                const int src = t + ntc * q;
                o.d1 = __shfl(s.d1, src, kWave);
                o.d2 = __shfl(s.d2, src, kWave);
                o.i1 = __shfl(s.i1, src, kWave);
                o.i2 = __shfl(s.i2, src, kWave);
                top2_merge(c, o);
            }
            return c;
        };
        stage_scan(-1.0, 0.0);
        const Top2 c1 = combine(st);
        double bound = (tg != 0) ? (c1.i2 >= 0 ? c1.d2 : INFINITY) : -INFINITY;
        bound = wave_max(bound);
        if (bound > 0) stage_scan(0.0, bound);
        const Top2 fin = combine(st);
        if (lane < ntc) {
            uint8_t f = 0;
            int n1 = -1, n2 = -1;
            if (tg != 0) {
                f = 1;
                n1 = fin.i1;
                n2 = fin.i2;
                if (n1 >= 0 && n2 >= 0 && !isnan(crit)) {
                    const double s1 = sgn(g[n1]), s2 = sgn(g[n2]), si = sgn(tg);
                    if ((fin.d1 < crit) || ((s1 == s2) && (s1 != si))) {
                        f |= 2;
                        flag_list[atomicAdd(&ecount[0], 1)] = i;
                    }
                }
            }
            nn1[i] = n1;
            nn2[i] = n2;
            d1o[i] = fin.d1;
            d2o[i] = fin.d2;
            eflag[i] = f;
        }
    }
}

// Squared distance from a point to a node's box (0 inside); never exceeds
// the computed squared distance to any point of the box.
__device__ inline double point_box_dist2(const TreeDev& T, int M, double px, double py) {
    const double dx = fmax(0.0, fmax(double_from_ord(T.blx[M]) - px, px - double_from_ord(T.trx[M])));
    const double dy = fmax(0.0, fmax(double_from_ord(T.bly[M]) - py, py - double_from_ord(T.try_[M])));
    return dx * dx + dy * dy;
}

// Sources of the near leaves of leaf l selected by sel(M) for one target i,
// lanes over each selected leaf's vortices.
// Chunks of kWave entries starting at chunk `first`, every `stride`-th.
template <typename Sel>
__device__ inline void scan_one(const TreeDev& T, int l, const double* x, const double* y, const double* g, int i,
                                double tx, double ty, Top2& st, Sel sel, int first = 0, int stride = 1) {
    const int64_t e0 = T.near_off[l], e1 = T.near_off[l + 1];
    for (int64_t eb = e0 + (int64_t)first * kWave; eb < e1; eb += (int64_t)stride * kWave) {
        const int64_t e = eb + lane_id();
        int M = -1;
        bool q = false;
        if (e < e1) {
            M = T.near_list[e];
            q = sel(M);
        }
        uint64_t mask = __ballot(q);
        while (mask) {
            const int b = (int)__builtin_ctzll(mask);
            mask &= mask - 1;
            const int Mb = __shfl(M, b, kWave);
            const int s0 = T.v0[Mb], s1 = T.v1[Mb];
            for (int j = s0 + lane_id(); j < s1; j += kWave) {
                const double gj = g[j];
                if (gj != 0 && j != i) top2_insert(st, dist2(tx, ty, x[j], y[j]), j);
            }
        }
    }
}

__device__ inline Top2 wave_top2(Top2 st) {
    for (int o = kWave / 2; o > 0; o >>= 1) top2_merge(st, top2_shfl_xor(st, o));
    return st;
}

// Full top-2 searches of the targets tg[0..nt) over their near leaves by
// the NW waves of a workgroup: leaves touching the target's leaf first, then
// those whose box lies within the second-neighbour distance found.  Groups
// of G waves (G the largest
// power of two with G * nt <= NW, at least 1) take one target each; within
// a group, wave k scans every G-th chunk of the near list, and the partial
// top-2s combine through LDS (the top-2 under (d, index) does not depend on
// the order of the candidates).  out(i, st) runs on one lane per target.
// Every thread of the workgroup must call it; it starts and ends with a
// barrier.
template <int NW, typename Out>
__device__ void group_rescans(const TreeDev& T, const double* x, const double* y, const double* g, const int* tg,
                              int nt, Top2* s_part, Out out) {
    int G = 1;
    while (2 * G * nt <= NW && 2 * G <= NW) G *= 2;
    const int ngroups = NW / G;
    const int wave = threadIdx.x / kWave, grp = wave / G, k = wave % G;
    for (int r0 = 0; r0 < nt; r0 += ngroups) {
        const int h = r0 + grp;
        const int i = h < nt ? tg[h] : -1;
        int l = 0, L = 0;
        double tx = 0, ty = 0;
        if (i >= 0) {
            l = T.vleaf[i];
            L = T.leaf_node[l];
            tx = x[i];
            ty = y[i];
        }
        Top2 st = top2_empty();
        if (i >= 0) scan_one(T, l, x, y, g, i, tx, ty, st, [&](int M) { return box_dist2(T, L, M) <= 0.0; }, k, G);
        st = wave_top2(st);
        __syncthreads();
        if (lane_id() == 0) s_part[wave] = st;
        __syncthreads();
        Top2 c = top2_empty();
        for (int w = grp * G; w < grp * G + G; w++) top2_merge(c, s_part[w]);
        const double bound = c.i2 >= 0 ? c.d2 : INFINITY;
        st = top2_empty();
        if (i >= 0)
            scan_one(T, l, x, y, g, i, tx, ty, st,
                     [&](int M) { return box_dist2(T, L, M) > 0.0 && point_box_dist2(T, M, tx, ty) <= bound; }, k, G);
        st = wave_top2(st);
        __syncthreads();
        if (lane_id() == 0) s_part[wave] = st;
        __syncthreads();
        if (i >= 0 && k == 0 && lane_id() == 0) {
            for (int w = grp * G; w < grp * G + G; w++) top2_merge(c, s_part[w]);
            out(i, c);
        }
    }
    __syncthreads();
}

__device__ inline bool merge_flag(const double* g, int i, const Top2& st, double crit) {
    if (st.i1 < 0 || st.i2 < 0 || isnan(crit)) return false;
    const double s1 = sgn(g[st.i1]), s2 = sgn(g[st.i2]), si = sgn(g[i]);
    return (st.d1 < crit) || ((s1 == s2) && (s1 != si));
}

// Transposed near lists: for every leaf M the leaves L with M in near(L).
// With mark set, only the lists of marked leaves are built (the others stay
// empty).
__global__ void k_rnear_count(TreeDev T, int32_t* cnt, const uint8_t* mark) {
    const int l = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    if (l >= T.nleaf) return;
    for (int64_t e = T.near_off[l] + lane_id(); e < T.near_off[l + 1]; e += kWave) {
        const int m = T.leaf_rank[T.near_list[e]];
        if (!mark || mark[m]) atomicAdd(&cnt[m], 1);
    }
}

// The same lists for a few marked leaves, flat over all near-list entries
// (one thread per entry, coalesced; the owner leaf of a kept entry is found
// by binary search over the list offsets).
__global__ void k_rnear_count_flat(TreeDev T, int64_t nnear, int32_t* cnt, const uint8_t* mark) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnear; e += (int64_t)gridDim.x * blockDim.x) {
        const int m = T.leaf_rank[T.near_list[e]];
        if (mark[m]) atomicAdd(&cnt[m], 1);
    }
}
__global__ void k_rnear_fill_flat(TreeDev T, int64_t nnear, const int64_t* roff, int32_t* fillc, int32_t* rlist,
                                  const uint8_t* mark) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnear; e += (int64_t)gridDim.x * blockDim.x) {
        const int m = T.leaf_rank[T.near_list[e]];
        if (!mark[m]) continue;
        int lo = 0, hi = T.nleaf;   // the leaf l with near_off[l] <= e < near_off[l + 1]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (T.near_off[mid] <= e)
                lo = mid;
            else
                hi = mid;
        }
        rlist[roff[m] + atomicAdd(&fillc[m], 1)] = lo;
    }
}

// Leaves of the flagged targets and of their first neighbours: the commits
// of pass 1's flags need their reverse lists.
__global__ void k_mark_flag_leaves(TreeDev T, const int32_t* flag_list, const int32_t* ecount, const int32_t* nn1,
                                   uint8_t* mark) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ecount[0]) return;
    const int q = flag_list[k];
    mark[T.vleaf[q]] = 1;
    mark[T.vleaf[nn1[q]]] = 1;
}

__global__ void k_rnear_fill(TreeDev T, const int64_t* roff, int32_t* fillc, int32_t* rlist, const uint8_t* mark) {
    const int l = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    if (l >= T.nleaf) return;
    for (int64_t e = T.near_off[l] + lane_id(); e < T.near_off[l + 1]; e += kWave) {
        const int m = T.leaf_rank[T.near_list[e]];
        if (!mark || mark[m]) rlist[roff[m] + atomicAdd(&fillc[m], 1)] = l;
    }
}

constexpr int kRepairThreads = 1024;
constexpr int kMaxHard = 2048;
constexpr int kMaxPend = 8192;
constexpr int kMaxLive = 2048;
constexpr int kMarkedFlags = 256;   // up to this many flags, reverse lists of their leaves only

struct RepairArgs {
    TreeDev T;
    double *x, *y, *g;
    const double* leaf_crit;
    const int64_t* roff;
    const int32_t* rlist;
    int32_t* stamp;
    int32_t *nn1, *nn2;
    double *d1a, *d2a;
    uint8_t* eflag;
    int32_t* ecount;
    uint64_t *box_lo_x, *box_lo_y, *box_hi_x, *box_hi_y;
    uint64_t* maxd2;   // per leaf: upper bound of d2 over its searched targets (ordered image)
};

struct CommitShared {
    int nhard;
    int overflow;
    int nlive;
    int hard[kMaxHard];
    int live[kMaxLive];
    double qx0, qy0, px0, py0;
    Top2 part[kRepairThreads / kWave];
};

// Upper bound of the second-neighbour distance over the searched targets of
// every leaf, as an ordered integer (-1 for a leaf without one; NaN counts
// as DBL_MAX).  Commits prune leaves whose box is farther from the merged
// vortices than this bound: none of their targets can change.
__device__ inline uint64_t d2_key(double d2) { return ord_from_double(isnan(d2) ? DBL_MAX : d2); }

__global__ void k_leaf_maxd2(TreeDev T, const double* d2a, const uint8_t* eflag, uint64_t* maxd2) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= T.nleaf) return;
    const int L = T.leaf_node[l];
    uint64_t m = ord_from_double(-1.0);
    for (int i = T.v0[L]; i < T.v1[L]; i++)
        if (eflag[i] & 1) m = max(m, d2_key(d2a[i]));
    maxd2[l] = m;
}

// One merge, MergeVortexes(q, nn1[q]) (MEpsilonFast.cpp:111-126), by a
// workgroup of NT threads, and its consequences for the targets that come
// later: q itself is searched again without merging; every later target
// whose near list holds leaf(q) takes q into its top-2 in O(1), and those
// that had q or its partner p among their two neighbours (or see only
// leaf(p)) are searched again.  Of the leaves whose near lists hold leaf(q)
// or leaf(p), only those whose box lies within their targets' largest
// second-neighbour distance of the old or new position of q or of p are
// visited: no target elsewhere can take q or have had q or p as neighbour.
// Targets whose merge flag becomes set are passed to on_flag(i).  `mark`
// tags the leaves of leaf(q)'s reverse near list in stamp[] for this
// commit.  Starts and ends with a barrier.
template <int NT, typename OnFlag>
__device__ void commit_one(const RepairArgs& A, int q, int mark, CommitShared& sh, OnFlag on_flag) {
    const TreeDev& T = A.T;
    const int tid = threadIdx.x;
    constexpr int kWaves = NT / kWave;
    double* x = A.x;
    double* y = A.y;
    double* g = A.g;
    const int p = A.nn1[q];
    __syncthreads();
    if (tid == 0) {
        sh.qx0 = x[q];
        sh.qy0 = y[q];
        sh.px0 = x[p];
        sh.py0 = y[p];
        const double gq = g[q], gp = g[p];
        if (sgn(gq) == sgn(gp)) {
            const double inv = 1. / (gq + gp);
            const double nx = (x[q] * gq + x[p] * gp) * inv;
            const double ny = (y[q] * gq + y[p] * gp) * inv;
            x[q] = nx;
            y[q] = ny;
        } else if (fabs(gq) < fabs(gp)) {
            x[q] = x[p];
            y[q] = y[p];
        }
        g[q] = gq + gp;
        g[p] = 0;
        atomicAdd(&A.ecount[1], 1);
        // the merged vortex may leave its leaf's box: grow the box so box
        // distances stay lower bounds (pruned rescans, diffusive cutoff)
        const int Lq = T.leaf_node[T.vleaf[q]];
        if (!isnan(x[q])) {
            atomicMin((unsigned long long*)&A.box_lo_x[Lq], (unsigned long long)ord_from_double(x[q]));
            atomicMax((unsigned long long*)&A.box_hi_x[Lq], (unsigned long long)ord_from_double(x[q]));
        }
        if (!isnan(y[q])) {
            atomicMin((unsigned long long*)&A.box_lo_y[Lq], (unsigned long long)ord_from_double(y[q]));
            atomicMax((unsigned long long*)&A.box_hi_y[Lq], (unsigned long long)ord_from_double(y[q]));
        }
        if (p > q) {   // skipped later by the reference (g == 0)
            A.eflag[p] = 0;
            A.nn1[p] = -1;
            A.nn2[p] = -1;
        }
        sh.nhard = 0;
    }
    __threadfence_block();
    __syncthreads();
    // the target itself: search again without merging
    group_rescans<kWaves>(T, x, y, g, &q, 1, sh.part, [&](int i, const Top2& st) {
        A.nn1[i] = st.i1;
        A.nn2[i] = st.i2;
        A.d1a[i] = st.d1;
        A.d2a[i] = st.d2;
        A.eflag[i] = 1;
        atomicMax((unsigned long long*)&A.maxd2[T.vleaf[i]], (unsigned long long)d2_key(st.d2));
    });
    // mark the leaves whose near list contains leaf(q)
    const int lq = T.vleaf[q], lp = T.vleaf[p];
    for (int64_t e = A.roff[lq] + tid; e < A.roff[lq + 1]; e += NT) A.stamp[A.rlist[e]] = mark;
    __threadfence_block();
    __syncthreads();
    const double qx = x[q], qy = y[q];
    const double qx0 = sh.qx0, qy0 = sh.qy0, px0 = sh.px0, py0 = sh.py0;
    // later targets that see leaf(q) (then those that see only leaf(p))
    for (int pass = 0; pass < 2; pass++) {
        const int lm = pass == 0 ? lq : lp;
        if (pass == 1 && lp == lq) break;
        const int64_t r0 = A.roff[lm], r1 = A.roff[lm + 1];
        // leaves that can hold a changed target, one thread per leaf
        if (tid == 0) sh.nlive = 0;
        __syncthreads();
        for (int64_t e = r0 + tid; e < r1; e += NT) {
            const int L = A.rlist[e];
            if (pass == 1 && A.stamp[L] == mark) continue;
            const int Ln = T.leaf_node[L];
            const double m = double_from_ord(A.maxd2[L]);
            bool need = !(point_box_dist2(T, Ln, qx0, qy0) > m) || !(point_box_dist2(T, Ln, px0, py0) > m);
            if (pass == 0) need = need || !(point_box_dist2(T, Ln, qx, qy) > m);
            if (need) {
                const int k = atomicAdd(&sh.nlive, 1);
                if (k < kMaxLive) sh.live[k] = L;
            }
        }
        __syncthreads();
        const int nlive = sh.nlive;
        const bool all = nlive > kMaxLive;   // list overflow: visit every leaf
        const int64_t nvis = all ? r1 - r0 : nlive;
        // NT / 16 leaves at a time, 16 threads per leaf
        for (int64_t e0 = 0; e0 < nvis; e0 += NT / 16) {
            const int64_t e = e0 + tid / 16;
            if (e >= nvis) continue;
            const int L = all ? A.rlist[r0 + e] : sh.live[e];
            if (all && pass == 1 && A.stamp[L] == mark) continue;
            const int Ln = T.leaf_node[L];
            const int a0 = T.v0[Ln], a1 = T.v1[Ln];
            for (int i = a0 + (tid & 15); i < a1; i += 16) {
                if (i <= q || !(A.eflag[i] & 1)) continue;
                const bool hard = A.nn1[i] == q || A.nn1[i] == p || A.nn2[i] == q || A.nn2[i] == p;
                if (hard) {
                    const int h = atomicAdd(&sh.nhard, 1);
                    if (h < kMaxHard)
                        sh.hard[h] = i;
                    else
                        sh.overflow = 1;
                    continue;
                }
                if (pass == 1) continue;   // q is not in the near set of i
                Top2 st{A.d1a[i], A.d2a[i], A.nn1[i], A.nn2[i]};
                top2_insert(st, dist2(x[i], y[i], qx, qy), q);
                A.nn1[i] = st.i1;
                A.nn2[i] = st.i2;
                A.d1a[i] = st.d1;
                A.d2a[i] = st.d2;   // only decreases: the leaf bound holds
                const bool f = merge_flag(g, i, st, A.leaf_crit[L]);
                A.eflag[i] = 1 | (f ? 2 : 0);
                if (f) on_flag(i);
            }
        }
        __syncthreads();
    }
    __threadfence_block();
    __syncthreads();
    const int nh = min(sh.nhard, kMaxHard);
    group_rescans<kWaves>(T, x, y, g, sh.hard, nh, sh.part, [&](int i, const Top2& st) {
        A.nn1[i] = st.i1;
        A.nn2[i] = st.i2;
        A.d1a[i] = st.d1;
        A.d2a[i] = st.d2;
        const bool f = merge_flag(g, i, st, A.leaf_crit[T.vleaf[i]]);
        A.eflag[i] = 1 | (f ? 2 : 0);
        atomicMax((unsigned long long*)&A.maxd2[T.vleaf[i]], (unsigned long long)d2_key(st.d2));
        if (f) on_flag(i);
    });
    __threadfence_block();
    __syncthreads();
}

// Pass 2: commits merges in target order, one workgroup per simulation
// (targets [sim_off[s], sim_off[s+1]); simulations share no leaf).
// Candidates are kept in an LDS list (the flags of pass 1 plus targets
// flagged by later updates); the next commit is the smallest listed index at
// or after the cursor whose flag is still set.  If the list overflows, the
// remaining commits find their target by scanning the flags.
// With mark set, only the reverse near lists of marked leaves exist: a commit
// that needs another one stops the simulation's pass (state[2s] = its
// target, state[2s+1] = the commit count, ecount[3] set); a second launch
// with resume set continues from there once all lists are built, taking its
// candidates from the flags at or after the cursor.  state[2s] = -1 marks a
// finished simulation.
__global__ __launch_bounds__(kRepairThreads) void k_eps_repair(RepairArgs A, const int64_t* sim_off,
                                                               const int32_t* flag_list, const uint8_t* mark,
                                                               int64_t* state, int resume) {
    __shared__ int s_q;
    __shared__ int s_np;
    __shared__ int s_pend[kMaxPend];
    __shared__ CommitShared sh;
    int32_t* ecount = A.ecount;
    const int tid = threadIdx.x;
    const int64_t lo = sim_off[blockIdx.x], n = sim_off[blockIdx.x + 1];
    int64_t cursor = lo;
    int commit = 0;
    if (resume) {
        cursor = state[2 * blockIdx.x];
        commit = (int)state[2 * blockIdx.x + 1];
        if (cursor < 0) return;   // finished in the first launch
    }
    if (tid == 0) {
        sh.overflow = 0;
        s_np = 0;
    }
    __syncthreads();
    if (!resume) {
        for (int k = tid; k < ecount[0]; k += kRepairThreads) {
            const int i = flag_list[k];
            if (i >= lo && i < n) {
                const int idx = atomicAdd(&s_np, 1);
                if (idx < kMaxPend) s_pend[idx] = i;
            }
        }
    } else {
        for (int64_t i = cursor + tid; i < n; i += kRepairThreads)
            if (A.eflag[i] & 2) {
                const int idx = atomicAdd(&s_np, 1);
                if (idx < kMaxPend) s_pend[idx] = (int)i;
            }
    }
    __syncthreads();
    const bool scan_mode0 = s_np > kMaxPend;
    while (true) {
        if (tid == 0) s_q = INT32_MAX;
        __syncthreads();
        const bool scan_mode = scan_mode0 || s_np > kMaxPend;
        if (!scan_mode) {
            for (int k = tid; k < s_np; k += kRepairThreads) {
                const int i = s_pend[k];
                if (i >= cursor && (A.eflag[i] & 2)) atomicMin(&s_q, i);
            }
            __syncthreads();
        } else {
            for (int64_t b = cursor; b < n; b += kRepairThreads) {
                const int64_t i = b + tid;
                if (i < n && (A.eflag[i] & 2)) atomicMin(&s_q, (int)i);
                __syncthreads();
                if (s_q != INT32_MAX) break;
                __syncthreads();
            }
            __syncthreads();
        }
        const int q = s_q;
        if (q == INT32_MAX) break;
        if (mark && !(mark[A.T.vleaf[q]] && mark[A.T.vleaf[A.nn1[q]]])) {
            if (tid == 0) {
                state[2 * blockIdx.x] = q;
                state[2 * blockIdx.x + 1] = commit;
                atomicExch(&ecount[3], 1);
            }
            return;   // uniform: q comes from LDS
        }
        commit++;
        commit_one<kRepairThreads>(A, q, commit, sh, [&](int i) {
            const int k = atomicAdd(&s_np, 1);
            if (k < kMaxPend) s_pend[k] = i;
        });
        cursor = q + 1;
    }
    if (tid == 0) {
        state[2 * blockIdx.x] = -1;
        if (sh.overflow) ecount[2] = 1;
    }
}

// Applies _1_eps = 1/max(eps, floor) to every processed target (:49-54) and
// packs the FP32 {g, eps^2} copy the pair kernels read.  Does nothing while
// the merge pass still has to resume (ecount[3]): it must see the final
// state only once.
__global__ void k_eps_apply(const TreeDev T, const double* g, double* ie, int64_t n, const int32_t* nn1,
                            const int32_t* nn2, const double* d1a, const double* d2a, const uint8_t* eflag,
                            const double* leaf_restr, float2* src_ge, const int32_t* ecount) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || ecount[3]) return;
    double iv = ie[i];
    if (eflag[i] & 1) {
        double eps;
        if (nn1[i] < 0)
            eps = DBL_MIN;
        else if (nn2[i] < 0)
            eps = sqrt(d1a[i]);
        else
            eps = sqrt(d2a[i]);
        const double r = leaf_restr[T.vleaf[i]];
        iv = 1.0 / ((eps < r) ? r : eps);
        ie[i] = iv;
    }
    const double gi = g[i];
    if (gi != 0) {
        const double e = 1. / iv;
        src_ge[i] = make_float2((float)gi, (float)(e * e));
    } else {
        src_ge[i] = make_float2(0.f, 1.f);
    }
}

__global__ void k_pack(const double* g, const double* ie, int64_t n, float2* src_ge) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double gi = g[i];
    if (gi != 0) {
        const double e = 1. / ie[i];
        src_ge[i] = make_float2((float)gi, (float)(e * e));
    } else {
        src_ge[i] = make_float2(0.f, 1.f);
    }
}

}  // namespace

void pack_sources(Ctx& c) {
    if (c.n)
        hipLaunchKernelGGL(k_pack, dim3(nblocks(c.n)), dim3(kBlock), 0, c.stream, c.g.p, c.ie.p, c.n, c.src_ge.p);
}

void epsilon(Ctx& c, int merge, int32_t* nn1_out, int32_t* nn2_out, int32_t* merged_out) {
    const int64_t n = c.n;
    const TreeDev T = tree_dev(c.t);
    const int nl = T.nleaf;
    hipStream_t s = c.stream;
    c.nn1.grow(n + 1);
    c.nn2.grow(n + 1);
    c.d1.grow(n + 1);
    c.d2.grow(n + 1);
    c.eflag.grow(n + 1);
    c.flag_list.grow(n + 1);
    c.ecount.grow(8);
    c.leaf_crit.grow(nl + 1);
    c.leaf_restr.grow(nl + 1);
    VVHD_HIP(hipMemsetAsync(c.ecount.p, 0, 8 * 4, s));
    hipLaunchKernelGGL(k_leaf_criteria, dim3(nblocks(nl)), dim3(kBlock), 0, s, T, c.segs.p, c.body_range.p,
                       (const int32_t*)c.sim_body.p, (const int32_t*)c.t.leaf_sim.p,
                       merge, c.leaf_crit.p, c.leaf_restr.p);
    if (nl > 0)
        hipLaunchKernelGGL(k_eps_top2, dim3((nl + kWavesPerBlock - 1) / kWavesPerBlock), dim3(kBlock), 0, s, T, c.x.p,
                           c.y.p, c.g.p, merge, c.leaf_crit.p, c.nn1.p, c.nn2.p, c.d1.p, c.d2.p, c.eflag.p,
                           c.ecount.p, c.flag_list.p);
    int32_t hc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    VVHD_HIP(hipMemcpyAsync(hc, c.ecount.p, 16, hipMemcpyDeviceToHost, s));
    VVHD_HIP(hipStreamSynchronize(s));
    // reverse near lists: all of them, or with few flags only those of the
    // flagged targets' leaves (a commit that needs another one resumes after
    // the full build)
    auto build_rnear = [&](const uint8_t* mark) {
        VVHD_HIP(hipMemsetAsync(c.rnear_cnt.p, 0, (size_t)(nl + 1) * 4, s));
        const int flat_grid = (int)std::min<int64_t>(8192, std::max<int64_t>(1, (c.t.nnear + kBlock - 1) / kBlock));
        if (mark)
            hipLaunchKernelGGL(k_rnear_count_flat, dim3(flat_grid), dim3(kBlock), 0, s, T, (int64_t)c.t.nnear,
                               c.rnear_cnt.p, mark);
        else
            hipLaunchKernelGGL(k_rnear_count, dim3(nblocks(nl, kWavesPerBlock)), dim3(kBlock), 0, s, T,
                               c.rnear_cnt.p, mark);
        size_t bytes = 0;
        VVHD_HIP(rocprim::exclusive_scan(nullptr, bytes, c.rnear_cnt.p, c.rnear_off.p, (int64_t)0, (size_t)nl + 1,
                                         rocprim::plus<int64_t>(), s));
        c.tmp.grow(bytes + 16);
        VVHD_HIP(rocprim::exclusive_scan(c.tmp.p, bytes, c.rnear_cnt.p, c.rnear_off.p, (int64_t)0, (size_t)nl + 1,
                                         rocprim::plus<int64_t>(), s));
        VVHD_HIP(hipMemsetAsync(c.rnear_cnt.p, 0, (size_t)(nl + 1) * 4, s));
        if (mark)
            hipLaunchKernelGGL(k_rnear_fill_flat, dim3(flat_grid), dim3(kBlock), 0, s, T, (int64_t)c.t.nnear,
                               c.rnear_off.p, c.rnear_cnt.p, c.rnear_list.p, mark);
        else
            hipLaunchKernelGGL(k_rnear_fill, dim3(nblocks(nl, kWavesPerBlock)), dim3(kBlock), 0, s, T,
                               c.rnear_off.p, c.rnear_cnt.p, c.rnear_list.p, mark);
    };
    // (buffers grown first: the arguments hold their pointers)
    auto repair_args = [&] {
        return RepairArgs{T,         c.x.p,         c.y.p,       c.g.p,     c.leaf_crit.p, c.rnear_off.p,
                          c.rnear_list.p, c.leaf_stamp.p, c.nn1.p, c.nn2.p, c.d1.p,        c.d2.p,
                          c.eflag.p, c.ecount.p,    c.t.blx.p,   c.t.bly.p, c.t.trx.p,     c.t.try_.p,
                          c.leaf_maxd2.p};
    };
    const bool marked = hc[0] > 0 && hc[0] <= kMarkedFlags;
    if (hc[0] > 0) {
        c.rnear_off.grow(nl + 1);
        c.rnear_cnt.grow(nl + 1);
        c.leaf_stamp.grow(nl + 1);
        c.rnear_list.grow(c.t.nnear + 1);
        c.rep_state.grow(2 * (size_t)c.nsim);
        c.leaf_mark.grow(nl + 1);
        c.leaf_maxd2.grow(nl + 1);
        VVHD_HIP(hipMemsetAsync(c.leaf_stamp.p, 0, (size_t)(nl + 1) * 4, s));
        if (marked) {
            VVHD_HIP(hipMemsetAsync(c.leaf_mark.p, 0, (size_t)(nl + 1), s));
            hipLaunchKernelGGL(k_mark_flag_leaves, dim3(nblocks(hc[0])), dim3(kBlock), 0, s, T,
                               (const int32_t*)c.flag_list.p, (const int32_t*)c.ecount.p, (const int32_t*)c.nn1.p,
                               c.leaf_mark.p);
        }
        build_rnear(marked ? c.leaf_mark.p : nullptr);
        hipLaunchKernelGGL(k_leaf_maxd2, dim3(nblocks(nl)), dim3(kBlock), 0, s, T, c.d2.p, c.eflag.p, c.leaf_maxd2.p);
        hipLaunchKernelGGL(k_eps_repair, dim3(c.nsim), dim3(kRepairThreads), 0, s, repair_args(),
                           (const int64_t*)c.sim_off.p,
                           c.flag_list.p, marked ? (const uint8_t*)c.leaf_mark.p : (const uint8_t*)nullptr,
                           c.rep_state.p, 0);
    }
    auto finish = [&] {
        hipLaunchKernelGGL(k_eps_apply, dim3(nblocks(n)), dim3(kBlock), 0, s, T, c.g.p, c.ie.p, n, c.nn1.p, c.nn2.p,
                           c.d1.p, c.d2.p, c.eflag.p, c.leaf_restr.p, c.src_ge.p, (const int32_t*)c.ecount.p);
        VVHD_HIP(hipMemcpyAsync(hc, c.ecount.p, 32, hipMemcpyDeviceToHost, s));
        VVHD_HIP(hipStreamSynchronize(s));
    };
    finish();
    if (marked && hc[3]) {   // a merge reached a leaf without its reverse list
        VVHD_HIP(hipMemsetAsync(c.ecount.p + 3, 0, 4, s));
        build_rnear(nullptr);
        hipLaunchKernelGGL(k_eps_repair, dim3(c.nsim), dim3(kRepairThreads), 0, s, repair_args(),
                           (const int64_t*)c.sim_off.p,
                           c.flag_list.p, (const uint8_t*)nullptr, c.rep_state.p, 1);
        finish();
    }
    if (nn1_out && n) VVHD_HIP(hipMemcpyAsync(nn1_out, c.nn1.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
    if (nn2_out && n) VVHD_HIP(hipMemcpyAsync(nn2_out, c.nn2.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
    VVHD_HIP(hipStreamSynchronize(s));
    VVHD_CHECK(hc[2] == 0, VVHD_EUNSUPPORTED, "epsilon: too many targets to rescan after one merge");
    if (hc[1] > 0) pack_boxes(c);   // merges grew leaf boxes
    const int heat_merged = heat_epsilon(c, merge);   // epsh on the heat list (MEpsilonFast.cpp:53-60)
    if (merged_out) *merged_out = hc[1] + heat_merged;
}

}  // namespace vvhd
