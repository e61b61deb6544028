// passive.hip -- the heat and streak lists (Space::HeatList, StreakList,
// StreakSourceList, TSpace.hpp:40-43) on the GPU.
//
// The tree build partitions both lists with the vortices (one range per list
// and node, tree.hip).  On the built tree:
//  * epsh (MEpsilonFast.cpp:175-212): the two nearest heat particles of every
//    heat particle among the heat particles of its leaf's near leaves, with
//    the reference's sequential merges.  A parallel pass evaluates every
//    target on the unmodified list and flags the targets that would merge;
//    when flags exist, one wave walks the leaves in DFS order and re-evaluates,
//    target by target, exactly the leaves that hold a flagged target or have
//    a near leaf touched by an earlier merge; every other leaf's targets
//    keep their parallel result, which the merges before them did not change.
//  * process_all_lists on the hRange and sRange targets (MConvectiveFast.cpp:
//    87-110): FP64 near sum over the near leaves' vortices plus the leaf's
//    Taylor far terms, one wave per target.
//  * process_heat_list (MDiffusiveFast.cpp:50-89): FP64, one wave per target.
//  * move_and_clean's heat and streak parts (MFlowmove.cpp:107-198):
//    classification in parallel, the order-dependent parts (the first heat
//    particle of a segment's heat layer, the segments' hsum in list order)
//    on one lane, compaction by an exclusive scan.
//  * heat_shed, heat_crop, streak_shed (MFlowmove.cpp:237-291).
// The lists are small next to the vortex list (a few hundred particles shed
// per step), so these kernels favour exactness over throughput.

#include "api.hpp"
#include "body_geom.hpp"
#include "common.hpp"

#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>

#pragma clang fp contract(off)

namespace vvhd {
namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr double kPi = 3.14159265358979323846;
constexpr double k1_2Pi = 1. / (2. * kPi);
constexpr double kExpArgLimit = -8.;   // MDiffusiveFast.cpp:6
constexpr double kRemoveEps = 1e-10;   // MFlowmove.hpp:8

inline int nblocks(int64_t n, int b = kBlock) { return (int)std::max<int64_t>(1, (n + b - 1) / b); }

__device__ inline double sgn(double x) { return (x > 0) ? 1 : ((x < 0) ? -1 : 0); }

// ------------------------------------------------------------------ epsh --

// Two nearest under the order (d^2, index): the reference keeps the first
// found on ties, and the near leaves' heat ranges come in increasing index
// order.
struct Near2 {
    double d1, d2;
    int i1, i2;
};

__device__ inline bool before(double d, int j, double dc, int ic) { return d < dc || (d == dc && ic >= 0 && j < ic); }

__device__ inline void near2_add(Near2& t, double d, int j) {
    if (before(d, j, t.d1, t.i1)) {
        t.d2 = t.d1;
        t.i2 = t.i1;
        t.d1 = d;
        t.i1 = j;
    } else if (before(d, j, t.d2, t.i2)) {
        t.d2 = d;
        t.i2 = j;
    }
}

// The wave's two nearest heat particles (g != 0, not j itself) to (px, py)
// among the heat ranges of leaf l's near leaves; every lane gets the result.
__device__ Near2 wave_near2(const TreeDev& T, const double* hx, const double* hy, const double* hg, int l, int j,
                            double px, double py) {
    Near2 t{DBL_MAX, DBL_MAX, -1, -1};
    for (int64_t e = T.near_off[l]; e < T.near_off[l + 1]; e++) {
        const int M = T.near_list[e];
        const int k1 = T.h1[M] - T.hbase;
        for (int k = T.h0[M] - T.hbase + lane_id(); k < k1; k += kWave) {
            if (k == j || hg[k] == 0) continue;
            const double dx = px - hx[k], dy = py - hy[k];
            near2_add(t, dx * dx + dy * dy, k);
        }
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        Near2 u;
        u.d1 = __shfl_xor(t.d1, o, kWave);
        u.d2 = __shfl_xor(t.d2, o, kWave);
        u.i1 = __shfl_xor(t.i1, o, kWave);
        u.i2 = __shfl_xor(t.i2, o, kWave);
        if (u.i1 >= 0) near2_add(t, u.d1, u.i1);
        if (u.i2 >= 0) near2_add(t, u.d2, u.i2);
    }
    return t;
}

// epsh's value for a target whose search found t and that does not merge:
// DBL_MIN without candidates, sqrt(d1) with one, sqrt(d2) otherwise.
__device__ inline double epsh_value(const Near2& t) {
    if (t.i1 < 0) return DBL_MIN;
    if (t.i2 < 0) return sqrt(t.d1);
    return sqrt(t.d2);
}

// Parallel pass, one wave per heat particle.  res = NaN marks "keep _1_eps"
// (g = 0) or "to be decided by the sequential pass" (flagged).
__global__ void k_heat_eps(TreeDev T, const double* hx, const double* hy, const double* hg, int64_t nh,
                           const double* crit, double* res, uint8_t* flag, int32_t* counters) {
    const int64_t j = (int64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    if (j >= nh) return;
    if (hg[j] == 0) {
        if (lane_id() == 0) {
            res[j] = __builtin_nan("");
            flag[j] = 0;
        }
        return;
    }
    const int l = T.vleaf[T.hbase + j];
    const Near2 t = wave_near2(T, hx, hy, hg, l, (int)j, hx[j], hy[j]);
    if (lane_id() != 0) return;
    const bool merge = t.i1 >= 0 && t.i2 >= 0 && t.d1 < crit[l];
    flag[j] = merge ? 1 : 0;
    res[j] = merge ? __builtin_nan("") : epsh_value(t);
    if (merge) atomicAdd(&counters[0], 1);
}

// Sequential pass (one wave): the reference's order, leaf by leaf (DFS),
// target by target, for the leaves that can differ from the parallel pass.
__global__ __launch_bounds__(kWave) void k_heat_eps_seq(TreeDev T, double* hx, double* hy, double* hg,
                                                        const double* crit, double* res, const uint8_t* flag,
                                                        uint8_t* dirty, int32_t* counters) {
    const int lane = lane_id();
    int merged = 0;
    for (int l = 0; l < T.nleaf; l++) {
        const int L = T.leaf_node[l];
        const int jb = T.h0[L] - T.hbase, je = T.h1[L] - T.hbase;
        if (jb == je) continue;
        bool need = false;
        for (int j = jb + lane; j < je; j += kWave) need = need || flag[j];
        for (int64_t e = T.near_off[l] + lane; e < T.near_off[l + 1]; e += kWave)
            need = need || dirty[T.leaf_rank[T.near_list[e]]];
        if (!__ballot(need)) continue;
        for (int j = jb; j < je; j++) {
            if (hg[j] == 0) {
                if (lane == 0) res[j] = __builtin_nan("");
                continue;
            }
            Near2 t = wave_near2(T, hx, hy, hg, l, j, hx[j], hy[j]);
            if (t.i1 >= 0 && t.i2 >= 0 && t.d1 < crit[l]) {
                // MergeVortexes (MEpsilonFast.cpp:111-126), then epsh without merging
                if (lane == 0) {
                    const int k = t.i1;
                    const double ga = hg[j], gb = hg[k];
                    if (sgn(ga) == sgn(gb)) {
                        const double inv = 1. / (ga + gb);
                        const double nx = (hx[j] * ga + hx[k] * gb) * inv;
                        const double ny = (hy[j] * ga + hy[k] * gb) * inv;
                        hx[j] = nx;
                        hy[j] = ny;
                    } else if (fabs(ga) < fabs(gb)) {
                        hx[j] = hx[k];
                        hy[j] = hy[k];
                    }
                    hg[j] = ga + gb;
                    hg[k] = 0;
                    dirty[l] = 1;
                    dirty[T.vleaf[T.hbase + k]] = 1;
                }
                merged++;
                __threadfence();   // lane 0's stores, seen by every lane's loads
                __builtin_amdgcn_wave_barrier();
                t = wave_near2(T, hx, hy, hg, l, j, hx[j], hy[j]);
            }
            if (lane == 0) res[j] = epsh_value(t);
        }
    }
    if (lane == 0) counters[1] = merged;
}

// _1_eps = 1 / max(eps, floor) (MEpsilonFast.cpp:55-60); NaN keeps _1_eps.
__global__ void k_heat_eps_apply(TreeDev T, const double* res, const double* restr, double* hie, int64_t nh) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nh) return;
    const double e = res[j];
    if (isnan(e)) return;
    const double r = restr[T.vleaf[T.hbase + j]];
    hie[j] = 1.0 / ((e < r) ? r : e);   // std::max(e, r)
}

// ------------------------------------------------------------ convective --

// process_all_lists on one heat or streak particle per wave: FP64 near sum,
// then the terms in the reference's order (:89-99, 102-110).
__global__ void k_passive_conv(TreeDev T, const double* x, const double* y, const double* g, const double* ie,
                               const double4* taylor, const double* px_, const double* py_, const double* pg,
                               double* pvx, double* pvy, int64_t np, int base, int skip_zero, double inf_x,
                               double inf_y) {
    const int64_t j = (int64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    if (j >= np) return;
    if (skip_zero && pg[j] == 0) return;
    const int l = T.vleaf[base + j];
    const int L = T.leaf_node[l];
    const double px = px_[j], py = py_[j];
    double sx = 0, sy = 0;
    for (int64_t e = T.near_off[l]; e < T.near_off[l + 1]; e++) {
        const int M = T.near_list[e];
        for (int k = T.v0[M] + lane_id(); k < T.v1[M]; k += kWave) {
            const double gk = g[k];
            if (!gk) continue;
            const double dx = px - x[k], dy = py - y[k];
            const double e1 = 1. / ie[k];
            const double f = gk / (dx * dx + dy * dy + e1 * e1);
            sx += -dy * f;
            sy += dx * f;
        }
    }
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    if (lane_id() != 0) return;
    const double4 Tl = taylor[l];
    double ax = pvx[j], ay = pvy[j];
    ax += inf_x;
    ay += inf_y;
    ax += sx * k1_2Pi;
    ay += sy * k1_2Pi;
    ax += 0.0;   // sinks: none
    ay += 0.0;
    ax += 0.0;   // bodies: rigid and static
    ay += 0.0;
    ax += Tl.x;
    ay += Tl.y;
    const double dxl = px - T.nx[L], dyl = py - T.ny[L];
    ax += Tl.z * dxl + Tl.w * dyl;
    ay += Tl.w * dxl + (-Tl.z) * dyl;
    pvx[j] = ax;
    pvy[j] = ay;
}

// ------------------------------------------------------------- diffusive --

// process_heat_list for one heat particle per wave (vortex_influence and
// segment_influence without friction, MDiffusiveFast.cpp:93-123), FP64.
__global__ void k_heat_diffusive(TreeDev T, const double* hx, const double* hy, const double* hg, const double* hie,
                                 double* hvx, double* hvy, int64_t nh, const vvhd_seg* segs, double re, double pr) {
    const int64_t j = (int64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    if (j >= nh) return;
    const double gi = hg[j];
    if (gi == 0) return;
    const int l = T.vleaf[T.hbase + j];
    const double px = hx[j], py = hy[j], ei = hie[j], si = sgn(gi);
    double s2x = 0, s2y = 0, s1 = 0, s3x = 0, s3y = 0, s0 = 0;
    for (int64_t e = T.near_off[l]; e < T.near_off[l + 1]; e++) {
        const int M = T.near_list[e];
        const int k1 = T.h1[M] - T.hbase;
        for (int k = T.h0[M] - T.hbase + lane_id(); k < k1; k += kWave) {
            const double gk = hg[k];
            if (!gk || sgn(gk) != si) continue;
            const double dx = px - hx[k], dy = py - hy[k];
            if (fabs(dx) + fabs(dy) < 1E-10) continue;
            const double drabs = sqrt(dx * dx + dy * dy);
            const double exparg = -drabs * ei;
            if (exparg < kExpArgLimit) continue;
            const double i1 = gk * exp(exparg);
            const double f = i1 / drabs;
            s2x += dx * f;
            s2y += dy * f;
            s1 += i1;
        }
        const int m = T.leaf_rank[M];
        for (int64_t q = T.seg_off[m] + lane_id(); q < T.seg_off[m + 1]; q += kWave) {
            const vvhd_seg sg = segs[T.seg_list[q]];
            const double dx = px - sg.x, dy = py - sg.y;
            const double d2 = dx * dx + dy * dy;
            const double drabs = sqrt(d2);
            const double exparg = -drabs * ei;
            if (exparg < kExpArgLimit) continue;
            const double ex = exp(exparg);
            const double dSx = -sg.dly, dSy = sg.dlx;
            s3x += dSx * ex;
            s3y += dSy * ex;
            s0 += (drabs * ei + 1) / d2 * (dx * dSx + dy * dSy) * ex;
        }
    }
    s2x = wave_sum(s2x);
    s2y = wave_sum(s2y);
    s1 = wave_sum(s1);
    s3x = wave_sum(s3x);
    s3y = wave_sum(s3y);
    s0 = wave_sum(s0);
    if (lane_id() != 0) return;
    if ((sgn(s1) != si) || (fabs(s1) < fabs(0.1 * gi))) s1 = 0.1 * gi;
    const double nx = s2x * ei + s3x, ny = s2y * ei + s3y;
    const double a = 1. / (s1 + s0 / (ei * ei));
    const double b = 1. / (pr * re);
    hvx[j] += (nx * a) * b;
    hvy[j] += (ny * a) * b;
}

// ------------------------------------------------------------- flow-move --

// TBody::isPointInHeatLayer (TBody.cpp:222-233): isPointInContour over the
// points r + rotl(dl) of the segments, with the body's own box and disc as
// the pre-test; the nearest segment, or -1.
__device__ inline int point_in_heat_layer(const vvhd_body& B, const vvhd_seg* segs, const double* geo, double px,
                                          double py) {
    bool in = B.area <= 0;
    if (!in && (px < geo[1] || py < geo[2] || px > geo[3] || py > geo[4] ||
                (px - B.cofm[0]) * (px - B.cofm[0]) + (py - B.cofm[1]) * (py - B.cofm[1]) > geo[0]))
        return -1;
    const int f = B.first_seg, n = B.nseg;
    for (int i = 0, j = n - 1; i < n; j = i++) {
        const vvhd_seg& a = segs[f + i];
        const vvhd_seg& b = segs[f + j];
        const double vix = a.x + (-a.dly), viy = a.y + a.dlx;
        const double vjx = b.x + (-b.dly), vjy = b.y + b.dlx;
        if (((viy < vjy) && (viy < py) && (py <= vjy) && ((vjy - viy) * (px - vix) > (vjx - vix) * (py - viy))) ||
            ((viy > vjy) && (viy > py) && (py >= vjy) && ((vjy - viy) * (px - vix) < (vjx - vix) * (py - viy))))
            in = !in;
    }
    if (!in) return -1;
    int best = -1;
    double bd = DBL_MAX;
    for (int k = 0; k < n; k++) {
        const double dx = segs[f + k].x - px, dy = segs[f + k].y - py;
        const double d2 = dx * dx + dy * dy;
        if (d2 < bd) {
            best = f + k;
            bd = d2;
        }
    }
    return best;
}

struct PMove {
    double *x, *y, *g, *vx, *vy;
    int64_t n;
    double dt;
    const vvhd_body* bodies;
    int nbody;
    const vvhd_seg* segs;
    const double* geo;
    const int32_t* cond;
};

// Heat particle j after the move (MFlowmove.cpp:107-186): cls 0 removed
// (|g| < remove_eps), 1 in a body (seg: nearest segment), 2 in the heat layer
// of a const_t body (seg: nearest segment), 3 kept.  Positions move in place.
// Streak particles (heat = false): 1 in a body, 3 kept.
__global__ void k_passive_classify(PMove A, bool heat, int32_t* cls, int32_t* seg, int32_t* first_in_layer) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= A.n) return;
    const double px = A.x[j] + A.vx[j] * A.dt, py = A.y[j] + A.vy[j] * A.dt;
    A.x[j] = px;
    A.y[j] = py;
    int c = 3, s = -1;
    if (heat && fabs(A.g[j]) < kRemoveEps) c = 0;
    if (c == 3)
        for (int b = 0; b < A.nbody; b++) {
            s = point_invalid(A.bodies[b], A.segs, A.geo + 5 * b, px, py);
            if (s >= 0) {
                c = 1;
                break;
            }
        }
    if (heat && c == 3) {
        for (int b = 0; b < A.nbody; b++) {
            const int q = point_in_heat_layer(A.bodies[b], A.segs, A.geo + 5 * b, px, py);
            if (q >= 0) {
                if (A.cond[b] == 2) {
                    c = 2;
                    s = q;
                    atomicMin(&first_in_layer[q], (int32_t)j);
                }
                break;
            }
        }
    }
    cls[j] = c;
    seg[j] = s;
}

// Keep flags: class 3, or the first class-2 particle of a segment whose
// layer particle is not set yet.
__global__ void k_passive_keep(int64_t n, const int32_t* cls, const int32_t* seg, const int32_t* first_in_layer,
                               const int64_t* layer, int32_t* keep) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j > n) return;
    if (j == n) {
        keep[j] = 0;
        return;
    }
    const int c = cls[j];
    keep[j] = (c == 3 || (c == 2 && layer[seg[j]] < 0 && first_in_layer[seg[j]] == j)) ? 1 : 0;
}

__global__ void k_passive_scatter(PMove A, const int32_t* keep, const int32_t* pos, const double* ie, double* x2,
                                  double* y2, double* g2, double* vx2, double* vy2, double* ie2, const int32_t* cls,
                                  const int32_t* seg, int64_t* layer) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= A.n || !keep[j]) return;
    const int p = pos[j];
    x2[p] = A.x[j];
    y2[p] = A.y[j];
    g2[p] = A.g[j];
    vx2[p] = 0.;
    vy2[p] = 0.;
    ie2[p] = ie[j];
    if (cls[j] == 2) layer[seg[j]] = p;   // heat_layer_obj_no: its index in the new list
}

// Heat particles absorbed by a segment (removed in a body or in a layer
// that has its particle already).
__global__ void k_absorbed(int64_t n, const int32_t* cls, const int32_t* keep, int32_t* flag) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j > n) return;
    flag[j] = (j < n && (cls[j] == 1 || cls[j] == 2) && !keep[j]) ? 1 : 0;
}

__global__ void k_absorbed_list(int64_t n, const int32_t* flag, const int32_t* pos, int32_t* list) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n && flag[j]) list[pos[j]] = (int32_t)j;
}

// hsum of the absorbing segments, in list order (MFlowmove.cpp:159, 177).
__global__ void k_heat_absorb(const int32_t* list, const int32_t* count, const int32_t* seg, const double* g,
                              vvhd_seg* segs) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int m = *count;
    for (int q = 0; q < m; q++) {
        const int j = list[q];
        segs[seg[j]].hsum -= g[j];
    }
}

// Stable compaction of a list by keep flags (all six fields).
__global__ void k_compact(int64_t n, const int32_t* keep, const int32_t* pos, const double* x, const double* y,
                          const double* g, const double* vx, const double* vy, const double* ie, double* x2,
                          double* y2, double* g2, double* vx2, double* vy2, double* ie2) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n || !keep[j]) return;
    const int p = pos[j];
    x2[p] = x[j];
    y2[p] = y[j];
    g2[p] = g[j];
    vx2[p] = vx[j];
    vy2[p] = vy[j];
    ie2[p] = ie[j];
}

// heat_crop's keep flags: r.x > scale goes (MFlowmove.cpp:275-282).
__global__ void k_crop_keep(int64_t n, const double* x, double scale, int32_t* keep) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j > n) return;
    keep[j] = (j < n && !(x[j] > scale)) ? 1 : 0;
}

// TObj records (6 doubles) appended to a list at offset n0.
__global__ void k_append_aos(const double* o6, int64_t m, int64_t n0, double* x, double* y, double* g, double* vx,
                             double* vy, double* ie) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const double* r = o6 + 6 * k;
    x[n0 + k] = r[0];
    y[n0 + k] = r[1];
    g[n0 + k] = r[2];
    vx[n0 + k] = r[3];
    vy[n0 + k] = r[4];
    ie[n0 + k] = r[5];
}

__global__ void k_to_aos(int64_t n, const double* x, const double* y, const double* g, const double* vx,
                         const double* vy, const double* ie, double* o6) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double* r = o6 + 6 * k;
    r[0] = x[k];
    r[1] = y[k];
    r[2] = g[k];
    r[3] = vx[k];
    r[4] = vy[k];
    r[5] = ie[k];
}

__global__ void k_fill_i64v(int64_t* a, int64_t n, int64_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_fill_i32v(int32_t* a, int64_t n, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// ----------------------------------------------------------- heat_shed --

// MFlowmove::heat_shed (MFlowmove.cpp:237-273), bodies and segments in
// order, appending at nh (capacity checked by the host); the new count goes
// to out_n.
__global__ void k_heat_shed(const vvhd_body* bodies, int nbody, vvhd_seg* segs, const int32_t* cond,
                            const float* heat_const, const int64_t* layer, double* hx, double* hy, double* hg,
                            double* hvx, double* hvy, double* hie, int64_t nh, int64_t* out_n, int32_t* err) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int64_t n = nh;
    auto push = [&](double px, double py, double g) {
        hx[n] = px;
        hy[n] = py;
        hg[n] = g;
        hvx[n] = 0;
        hvy[n] = 0;
        hie[n] = 0;
        n++;
    };
    for (int b = 0; b < nbody; b++) {
        const int c = cond[b];
        for (int k = 0; k < bodies[b].nseg; k++) {
            const int i = bodies[b].first_seg + k;
            vvhd_seg& s = segs[i];
            const double px = s.x + (-s.dly) * 0.5, py = s.y + s.dlx * 0.5;
            if (c == 1) {
                if (!s.hsum) continue;
                push(px, py, -s.hsum);
                s.hsum = 0;
            } else if (c == 2) {
                const double tg = (s.dlx * s.dlx + s.dly * s.dly) * (double)heat_const[i];
                const int64_t no = layer[i];
                if (no >= 0) {
                    if (no >= n) {
                        *err = 1;
                        continue;
                    }
                    s.hsum += tg - hg[no];
                    hg[no] = tg;
                } else {
                    s.hsum += tg;
                    push(px, py, tg);
                }
            } else if (c == 3) {
                const double tg = (s.dlx * s.dlx + s.dly * s.dly) * (double)heat_const[i];
                s.hsum += tg;
                push(px, py, tg);
            }
        }
    }
    *out_n = n;
}

void scan_i32(Ctx& c, const int32_t* in, int32_t* out, size_t n) {
    size_t bytes = 0;
    VVHD_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, (int32_t)0, n, rocprim::plus<int32_t>(), c.stream));
    c.tmp.grow(bytes + 16);
    VVHD_HIP(rocprim::exclusive_scan(c.tmp.p, bytes, in, out, (int32_t)0, n, rocprim::plus<int32_t>(), c.stream));
}

}  // namespace

// Per leaf (DFS order): hRange and sRange as list indices.
__global__ void k_leaf_lists(TreeDev T, int32_t* out) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= T.nleaf) return;
    const int L = T.leaf_node[l];
    out[4 * l] = T.h0[L] - T.hbase;
    out[4 * l + 1] = T.h1[L] - T.hbase;
    out[4 * l + 2] = T.s0[L] - T.sbase;
    out[4 * l + 3] = T.s1[L] - T.sbase;
}

int heat_epsilon(Ctx& c, int merge) {
    const int64_t nh = c.heat.n;
    if (nh == 0) return 0;
    const TreeDev T = tree_dev(c.t);
    hipStream_t s = c.stream;
    c.hres.grow(nh + 1);
    c.hflag.grow(nh + 1);
    c.iscratch4.grow(4);
    int32_t* counters = c.iscratch4.p;
    hipLaunchKernelGGL(k_fill_i32v, dim3(1), dim3(64), 0, s, counters, (int64_t)4, 0);
    hipLaunchKernelGGL(k_heat_eps, dim3(nblocks(nh, kWavesPerBlock)), dim3(kBlock), 0, s, T, c.heat.x.p, c.heat.y.p,
                       c.heat.g.p, nh, (const double*)c.leaf_crit.p, c.hres.p, c.hflag.p, counters);
    int32_t hc[2] = {0, 0};
    VVHD_HIP(hipMemcpyAsync(hc, counters, 8, hipMemcpyDeviceToHost, s));
    VVHD_HIP(hipStreamSynchronize(s));
    if (hc[0] > 0 && merge) {
        c.hdirty.grow(T.nleaf + 1);
        VVHD_HIP(hipMemsetAsync(c.hdirty.p, 0, (size_t)T.nleaf + 1, s));
        hipLaunchKernelGGL(k_heat_eps_seq, dim3(1), dim3(kWave), 0, s, T, c.heat.x.p, c.heat.y.p, c.heat.g.p,
                           (const double*)c.leaf_crit.p, c.hres.p, (const uint8_t*)c.hflag.p, c.hdirty.p, counters);
        VVHD_HIP(hipMemcpyAsync(hc, counters, 8, hipMemcpyDeviceToHost, s));
    }
    hipLaunchKernelGGL(k_heat_eps_apply, dim3(nblocks(nh)), dim3(kBlock), 0, s, T, (const double*)c.hres.p,
                       (const double*)c.leaf_restr.p, c.heat.ie.p, nh);
    VVHD_HIP(hipStreamSynchronize(s));
    VVHD_HIP(hipGetLastError());
    return hc[0] > 0 && merge ? hc[1] : 0;
}

void passive_convective(Ctx& c, const double4* taylor, double inf_x, double inf_y) {
    const TreeDev T = tree_dev(c.t);
    for (int q = 0; q < 2; q++) {
        PList& L = q == 0 ? c.heat : c.streak;
        if (!L.n) continue;
        hipLaunchKernelGGL(k_passive_conv, dim3(nblocks(L.n, kWavesPerBlock)), dim3(kBlock), 0, c.stream, T, c.x.p,
                           c.y.p, c.g.p, c.ie.p, taylor, L.x.p, L.y.p, L.g.p, L.vx.p, L.vy.p, L.n,
                           q == 0 ? T.hbase : T.sbase, q == 0 ? 1 : 0, inf_x, inf_y);
    }
    VVHD_HIP(hipGetLastError());
}

void heat_diffusive(Ctx& c) {
    if (!c.heat.n) return;
    const TreeDev T = tree_dev(c.t);
    hipLaunchKernelGGL(k_heat_diffusive, dim3(nblocks(c.heat.n, kWavesPerBlock)), dim3(kBlock), 0, c.stream, T,
                       c.heat.x.p, c.heat.y.p, c.heat.g.p, c.heat.ie.p, c.heat.vx.p, c.heat.vy.p, c.heat.n, c.segs.p,
                       c.re, c.pr);
    VVHD_HIP(hipGetLastError());
}

// The heat and streak parts of move_and_clean, after the bodies moved.
void passive_move_and_clean(Ctx& c) {
    const int nb = (int)c.bodies_h.size();
    const int nseg = (int)c.segs_h.size();
    hipStream_t s = c.stream;
    c.heat_cond.grow(std::max(nb, 1));
    c.heat_layer.grow(std::max(nseg, 1));
    for (int q = 0; q < 2; q++) {
        PList& L = q == 0 ? c.heat : c.streak;
        const int64_t n = L.n;
        if (!n) continue;
        VVHD_CHECK(n < (int64_t(1) << 31), VVHD_EUNSUPPORTED, "move_and_clean: more than 2^31 - 1 particles");
        c.iscratch0.grow(n + 1);   // class
        c.iscratch1.grow(n + 1);   // segment
        c.iscratch2.grow(n + 2);   // keep
        c.iscratch3.grow(n + 2);   // position
        c.iscratch4.grow(std::max(nseg, 1) + 1);   // first particle in each segment's layer
        if (nseg) hipLaunchKernelGGL(k_fill_i32v, dim3(nblocks(nseg)), dim3(kBlock), 0, s, c.iscratch4.p,
                                     (int64_t)nseg, INT32_MAX);
        const PMove A{L.x.p, L.y.p, L.g.p, L.vx.p, L.vy.p, n, c.dt, c.bodies_d.p, nb, c.segs.p, c.body_geo.p,
                      c.heat_cond.p};
        hipLaunchKernelGGL(k_passive_classify, dim3(nblocks(n)), dim3(kBlock), 0, s, A, q == 0, c.iscratch0.p,
                           c.iscratch1.p, c.iscratch4.p);
        hipLaunchKernelGGL(k_passive_keep, dim3(nblocks(n + 1)), dim3(kBlock), 0, s, n, (const int32_t*)c.iscratch0.p,
                           (const int32_t*)c.iscratch1.p, (const int32_t*)c.iscratch4.p,
                           (const int64_t*)c.heat_layer.p, c.iscratch2.p);
        scan_i32(c, c.iscratch2.p, c.iscratch3.p, (size_t)n + 1);
        hipLaunchKernelGGL(k_passive_scatter, dim3(nblocks(n)), dim3(kBlock), 0, s, A, (const int32_t*)c.iscratch2.p,
                           (const int32_t*)c.iscratch3.p, (const double*)L.ie.p, L.x2.p, L.y2.p, L.g2.p, L.vx2.p,
                           L.vy2.p, L.ie2.p, (const int32_t*)c.iscratch0.p, (const int32_t*)c.iscratch1.p,
                           c.heat_layer.p);
        if (q == 0 && nseg) {
            // absorbed particles in list order (their flags reuse the layer scratch)
            c.lscratch0.grow(2 * ((size_t)n + 2));
            int32_t* aflag = reinterpret_cast<int32_t*>(c.lscratch0.p);
            int32_t* apos = aflag + n + 2;
            c.iscratch4.grow(std::max<int64_t>(n + 1, nseg + 1));
            hipLaunchKernelGGL(k_absorbed, dim3(nblocks(n + 1)), dim3(kBlock), 0, s, n, (const int32_t*)c.iscratch0.p,
                               (const int32_t*)c.iscratch2.p, aflag);
            scan_i32(c, aflag, apos, (size_t)n + 1);
            hipLaunchKernelGGL(k_absorbed_list, dim3(nblocks(n)), dim3(kBlock), 0, s, n, (const int32_t*)aflag,
                               (const int32_t*)apos, c.iscratch4.p);
            hipLaunchKernelGGL(k_heat_absorb, dim3(1), dim3(1), 0, s, (const int32_t*)c.iscratch4.p,
                               (const int32_t*)(apos + n), (const int32_t*)c.iscratch1.p, (const double*)L.g.p,
                               c.segs.p);
        }
        int32_t kept = 0;
        VVHD_HIP(hipMemcpyAsync(&kept, c.iscratch3.p + n, 4, hipMemcpyDeviceToHost, s));
        VVHD_HIP(hipStreamSynchronize(s));
        L.swap_buffers();
        L.n = kept;
    }
    VVHD_HIP(hipGetLastError());
}

void heat_shed(Ctx& c) {
    const int nb = (int)c.bodies_h.size();
    if (!nb) return;
    const int nseg = (int)c.segs_h.size();
    bool any = false;
    for (int b = 0; b < nb && b < (int)c.heat_cond_h.size(); b++) any = any || c.heat_cond_h[b] != 0;
    if (!any) return;
    hipStream_t s = c.stream;
    c.heat.reserve(c.heat.n + nseg + 1, true, s);
    c.lscratch0.grow(2);
    c.iscratch4.grow(2);
    hipLaunchKernelGGL(k_fill_i32v, dim3(1), dim3(64), 0, s, c.iscratch4.p, (int64_t)2, 0);
    hipLaunchKernelGGL(k_heat_shed, dim3(1), dim3(1), 0, s, c.bodies_d.p, nb, c.segs.p, (const int32_t*)c.heat_cond.p,
                       (const float*)c.heat_const.p, (const int64_t*)c.heat_layer.p, c.heat.x.p, c.heat.y.p,
                       c.heat.g.p, c.heat.vx.p, c.heat.vy.p, c.heat.ie.p, c.heat.n, c.lscratch0.p, c.iscratch4.p);
    int64_t n = 0;
    int32_t err = 0;
    VVHD_HIP(hipMemcpyAsync(&n, c.lscratch0.p, 8, hipMemcpyDeviceToHost, s));
    VVHD_HIP(hipMemcpyAsync(&err, c.iscratch4.p, 4, hipMemcpyDeviceToHost, s));
    VVHD_HIP(hipStreamSynchronize(s));
    VVHD_CHECK(err == 0, VVHD_EINVAL, "heat_shed: heat layer particle index beyond the heat list");
    c.heat.n = n;
    c.t.built = false;
}

void heat_crop(Ctx& c, double scale) {
    const int64_t n = c.heat.n;
    if (!n) return;
    hipStream_t s = c.stream;
    c.iscratch2.grow(n + 2);
    c.iscratch3.grow(n + 2);
    PList& L = c.heat;
    hipLaunchKernelGGL(k_crop_keep, dim3(nblocks(n + 1)), dim3(kBlock), 0, s, n, (const double*)L.x.p, scale,
                       c.iscratch2.p);
    scan_i32(c, c.iscratch2.p, c.iscratch3.p, (size_t)n + 1);
    hipLaunchKernelGGL(k_compact, dim3(nblocks(n)), dim3(kBlock), 0, s, n, (const int32_t*)c.iscratch2.p,
                       (const int32_t*)c.iscratch3.p, (const double*)L.x.p, (const double*)L.y.p, (const double*)L.g.p,
                       (const double*)L.vx.p, (const double*)L.vy.p, (const double*)L.ie.p, L.x2.p, L.y2.p, L.g2.p,
                       L.vx2.p, L.vy2.p, L.ie2.p);
    int32_t kept = 0;
    VVHD_HIP(hipMemcpyAsync(&kept, c.iscratch3.p + n, 4, hipMemcpyDeviceToHost, s));
    VVHD_HIP(hipStreamSynchronize(s));
    L.swap_buffers();
    L.n = kept;
    c.t.built = false;
}

// Appends m TObj records (host) to a list.
void list_append(Ctx& c, PList& L, const double* o6, int64_t m) {
    if (!m) return;
    hipStream_t s = c.stream;
    L.reserve(L.n + m + 1, true, s);
    c.tmp.grow((size_t)m * 48);
    VVHD_HIP(hipMemcpyAsync(c.tmp.p, o6, (size_t)m * 48, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_append_aos, dim3(nblocks(m)), dim3(kBlock), 0, s, (const double*)c.tmp.p, m, L.n, L.x.p,
                       L.y.p, L.g.p, L.vx.p, L.vy.p, L.ie.p);
    VVHD_HIP(hipStreamSynchronize(s));
    L.n += m;
    c.t.built = false;
}

void list_download(Ctx& c, PList& L, double* o6) {
    if (!L.n) return;
    hipStream_t s = c.stream;
    c.tmp.grow((size_t)L.n * 48);
    hipLaunchKernelGGL(k_to_aos, dim3(nblocks(L.n)), dim3(kBlock), 0, s, L.n, (const double*)L.x.p,
                       (const double*)L.y.p, (const double*)L.g.p, (const double*)L.vx.p, (const double*)L.vy.p,
                       (const double*)L.ie.p, (double*)c.tmp.p);
    VVHD_HIP(hipMemcpyAsync(o6, c.tmp.p, (size_t)L.n * 48, hipMemcpyDeviceToHost, s));
    VVHD_HIP(hipStreamSynchronize(s));
}

void streak_shed(Ctx& c) {
    const int64_t m = (int64_t)c.streak_src_h.size() / 6;
    list_append(c, c.streak, c.streak_src_h.data(), m);
}

// TAtt::heat_layer_obj_no = -1 for every segment (TAtt constructor,
// Space::zero_forces, TSpace.cpp:131).
void heat_layer_reset(Ctx& c) {
    const int nseg = (int)c.segs_h.size();
    c.heat_layer.grow(std::max(nseg, 1));
    if (nseg)
        hipLaunchKernelGGL(k_fill_i64v, dim3(nblocks(nseg)), dim3(kBlock), 0, c.stream, c.heat_layer.p, (int64_t)nseg,
                           (int64_t)-1);
}

}  // namespace vvhd

// ------------------------------------------------------------ C entry points

namespace {
vvhd::PList* plist(vvhd::Ctx& c, int list) {
    if (list == VVHD_LIST_HEAT) return &c.heat;
    if (list == VVHD_LIST_STREAK) return &c.streak;
    return nullptr;
}
}  // namespace

extern "C" {

int vvhd_upload_list(vvhd_ctx* ctx, int list, const double* objs6, int64_t n) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    vvhd::Ctx& c = ctx->c;
    VVHD_CHECK(n >= 0 && (n == 0 || objs6), VVHD_EINVAL, "upload_list: bad arguments");
    VVHD_HIP(hipSetDevice(c.device));
    if (list == VVHD_LIST_STREAK_SOURCE) {
        c.streak_src_h.assign(objs6, objs6 + 6 * n);
    } else {
        vvhd::PList* L = plist(c, list);
        VVHD_CHECK(L, VVHD_EINVAL, "upload_list: unknown list");
        L->n = 0;
        vvhd::list_append(c, *L, objs6, n);
    }
    c.t.built = false;
    API_END(ctx)
}

int vvhd_download_list(vvhd_ctx* ctx, int list, double* objs6, int64_t n) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    vvhd::Ctx& c = ctx->c;
    if (list == VVHD_LIST_STREAK_SOURCE) {
        VVHD_CHECK(n == (int64_t)c.streak_src_h.size() / 6 && (n == 0 || objs6), VVHD_EINVAL,
                   "download_list: n does not match the list");
        if (n) std::memcpy(objs6, c.streak_src_h.data(), (size_t)n * 48);
    } else {
        vvhd::PList* L = plist(c, list);
        VVHD_CHECK(L, VVHD_EINVAL, "download_list: unknown list");
        VVHD_CHECK(n == L->n && (n == 0 || objs6), VVHD_EINVAL, "download_list: n does not match the list");
        vvhd::list_download(c, *L, objs6);
    }
    API_END(ctx)
}

int64_t vvhd_list_size(vvhd_ctx* ctx, int list) {
    if (!ctx) return -1;
    if (list == VVHD_LIST_VORTEX) return ctx->c.n;
    if (list == VVHD_LIST_STREAK_SOURCE) return (int64_t)ctx->c.streak_src_h.size() / 6;
    vvhd::PList* L = plist(ctx->c, list);
    return L ? L->n : -1;
}

int vvhd_set_pr(vvhd_ctx* ctx, double pr) {
    REQUIRE_CTX(ctx);
    ctx->c.pr = pr;
    return VVHD_OK;
}

int vvhd_set_heat(vvhd_ctx* ctx, const int32_t* heat_condition, const float* heat_const) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    vvhd::Ctx& c = ctx->c;
    const int nb = (int)c.bodies_h.size(), nseg = (int)c.segs_h.size();
    VVHD_CHECK((nb == 0 || heat_condition) && (nseg == 0 || heat_const), VVHD_EINVAL, "set_heat: null arguments");
    for (int b = 0; b < nb; b++)
        VVHD_CHECK(heat_condition[b] >= VVHD_HC_NEGLECT && heat_condition[b] <= VVHD_HC_CONST_W, VVHD_EINVAL,
                   "set_heat: unknown heat condition");
    c.heat_cond_h.assign(heat_condition, heat_condition + nb);
    c.heat_cond.grow(std::max(nb, 1));
    c.heat_const.grow(std::max(nseg, 1));
    if (nb)
        VVHD_HIP(hipMemcpyAsync(c.heat_cond.p, heat_condition, (size_t)nb * 4, hipMemcpyHostToDevice, c.stream));
    if (nseg) VVHD_HIP(hipMemcpyAsync(c.heat_const.p, heat_const, (size_t)nseg * 4, hipMemcpyHostToDevice, c.stream));
    VVHD_HIP(hipStreamSynchronize(c.stream));
    API_END(ctx)
}

int vvhd_heat_layer(vvhd_ctx* ctx, int64_t* obj_no, int32_t nseg) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    vvhd::Ctx& c = ctx->c;
    VVHD_CHECK(nseg == (int32_t)c.segs_h.size() && (nseg == 0 || obj_no), VVHD_EINVAL,
               "heat_layer: nseg does not match the segment count");
    if (nseg) {
        VVHD_HIP(hipMemcpyAsync(obj_no, c.heat_layer.p, (size_t)nseg * 8, hipMemcpyDeviceToHost, c.stream));
        VVHD_HIP(hipStreamSynchronize(c.stream));
    }
    API_END(ctx)
}

int vvhd_set_heat_layer(vvhd_ctx* ctx, const int64_t* obj_no, int32_t nseg) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    vvhd::Ctx& c = ctx->c;
    VVHD_CHECK(nseg == (int32_t)c.segs_h.size() && (nseg == 0 || obj_no), VVHD_EINVAL,
               "set_heat_layer: nseg does not match the segment count");
    if (nseg) {
        VVHD_HIP(hipMemcpyAsync(c.heat_layer.p, obj_no, (size_t)nseg * 8, hipMemcpyHostToDevice, c.stream));
        VVHD_HIP(hipStreamSynchronize(c.stream));
    }
    API_END(ctx)
}

int vvhd_diffusive_heat(vvhd_ctx* ctx) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    REQUIRE_ONE_SIM(ctx->c, "diffusive_heat");
    vvhd::Ctx& c = ctx->c;
    VVHD_CHECK(c.t.built, VVHD_ESTATE, "MDiffusiveFast::process_heat_list(): tree is not built");
    vvhd::heat_diffusive(c);
    VVHD_HIP(hipStreamSynchronize(c.stream));
    API_END(ctx)
}

int vvhd_heat_shed(vvhd_ctx* ctx) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    REQUIRE_ONE_SIM(ctx->c, "heat_shed");
    vvhd::heat_shed(ctx->c);
    vvhd::sync_bodies_to_host(ctx->c);
    API_END(ctx)
}

int vvhd_heat_crop(vvhd_ctx* ctx, double scale) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    REQUIRE_ONE_SIM(ctx->c, "heat_crop");
    vvhd::heat_crop(ctx->c, scale);
    API_END(ctx)
}

int vvhd_streak_shed(vvhd_ctx* ctx) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    REQUIRE_ONE_SIM(ctx->c, "streak_shed");
    vvhd::streak_shed(ctx->c);
    API_END(ctx)
}

int vvhd_tree_leaf_lists(vvhd_ctx* ctx, int32_t* ranges4) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    vvhd::Ctx& c = ctx->c;
    VVHD_CHECK(c.t.built, VVHD_ESTATE, "tree_leaf_lists: tree is not built");
    VVHD_CHECK(ranges4, VVHD_EINVAL, "tree_leaf_lists: null output");
    const int nl = c.t.nleaves_h;
    if (nl) {
        c.iscratch0.grow(4 * (size_t)nl);
        hipLaunchKernelGGL(vvhd::k_leaf_lists, dim3((nl + 255) / 256), dim3(256), 0, c.stream, vvhd::tree_dev(c.t),
                           c.iscratch0.p);
        VVHD_HIP(hipMemcpyAsync(ranges4, c.iscratch0.p, (size_t)nl * 16, hipMemcpyDeviceToHost, c.stream));
        VVHD_HIP(hipStreamSynchronize(c.stream));
    }
    API_END(ctx)
}

}  // extern "C"
