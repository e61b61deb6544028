// bodies.hip -- body geometry, flow-move, vortex shedding, force integrals
// and the FP64 boundary-condition system, on the device.
//
//  * TBody::doUpdateSegments / doFillProperties (TBody.cpp:200-215, 283-343)
//    and TBody::move (:179-198): one thread per body, sequential in segment
//    order, so every FP64 quantity rounds as in the reference.
//  * MFlowmove::move_and_clean (MFlowmove.cpp:11-215) for rigid bodies
//    without collisions: two passes (k_fm_count, k_fm_scatter) advect, flag
//    |g| < 1e-10 and in-body vortices (TBody::isPointInvalid, TBody.cpp:
//    217-281) and compact the survivors stably; then the bodies' dead force,
//    moment and circulation accumulated in vortex order and the zero/slip
//    segment loop.
//  * MFlowmove::vortex_shed (MFlowmove.cpp:217-235): appended behind the
//    list in segment order.
//  * Space::calc_forces / zero_forces (TSpace.cpp:94-143).
//  * MConvectiveFast::calc_circulation (MConvectiveFast.cpp:235-273, fill
//    :459-1017): FP64 fill of the column-major matrix and right-hand side
//    (influence cells one thread each; the vortex influence on a slip row one
//    wave per segment, summed in the reference's order; body rows one
//    thread), LU factorisation and solve with rocSOLVER.  The factors are
//    reused while the matrix is bit-identical to the factorised one, the
//    condition under which the reference reuses its inverse
//    (TMatrix.cpp:225-259).

#include "api.hpp"
#include "body_geom.hpp"

#include <cstring>
#include <rocprim/device/device_scan.hpp>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>

#pragma clang fp contract(off)

namespace vvhd {
namespace {

constexpr int kBlock = 256;
constexpr double kPi = 3.14159265358979323846;
constexpr double k2Pi = 2. * kPi;
constexpr double k1_2Pi = 1. / (2. * kPi);
constexpr double kRemoveEps = 1e-10;   // MFlowmove.hpp:8

inline int nblocks(int64_t n, int b = kBlock) { return (int)std::max<int64_t>(1, (n + b - 1) / b); }

__device__ inline double dot2(double ax, double ay, double bx, double by) { return ax * bx + ay * by; }

// ------------------------------------------------------------ body geometry

// doUpdateSegments + doFillProperties of one body; geo receives the
// bounding disc radius squared and the bounding box used by isPointInvalid.
__device__ void body_properties(vvhd_body& B, vvhd_seg* segs, double* geo) {
    const int f = B.first_seg, n = B.nseg;
    for (int k = 0; k < n; k++) {
        vvhd_seg& s = segs[f + k];
        const vvhd_seg& nx = segs[f + (k + 1) % n];
        s.dlx = nx.cx - s.cx;
        s.dly = nx.cy - s.cy;
        s.inv_eps = 3.0 / sqrt(s.dlx * s.dlx + s.dly * s.dly);
        s.x = (nx.cx + s.cx) * 0.5;
        s.y = (nx.cy + s.cy) * 0.5;
    }
    double slen = 0, area = 0;
    for (int k = 0; k < n; k++) {
        const vvhd_seg& s = segs[f + k];
        slen += sqrt(s.dlx * s.dlx + s.dly * s.dly);
        area += s.y * s.dlx;
    }
    double cx, cy, moi;
    if (area == 0) {
        double Lx = 0., Ly = 0.;
        for (int k = 0; k < n; k++) {
            const vvhd_seg& s = segs[f + k];
            const double a = sqrt(s.dlx * s.dlx + s.dly * s.dly);
            Lx += s.x * a;
            Ly += s.y * a;
        }
        const double inv = 1. / slen;
        cx = Lx * inv;
        cy = Ly * inv;
        moi = 0.;
    } else {
        double Sx = 0., Sy = 0., m12 = 0.;
        for (int k = 0; k < n; k++) {
            const vvhd_seg& s = segs[f + k];
            const vvhd_seg& nx = segs[f + (k + 1) % n];
            const double cr = dot2(-s.cy, s.cx, nx.cx, nx.cy);
            Sx -= s.x * cr;
            Sy -= s.y * cr;
            m12 -= (dot2(s.cx, s.cy, s.cx, s.cy) + dot2(s.cx, s.cy, nx.cx, nx.cy) + dot2(nx.cx, nx.cy, nx.cx, nx.cy)) *
                   cr;
        }
        const double inv = 1. / (3 * area);
        cx = Sx * inv;
        cy = Sy * inv;
        moi = m12 / 12. - area * (cx * cx + cy * cy);
    }
    B.slen = slen;
    B.area = area;
    B.cofm[0] = cx;
    B.cofm[1] = cy;
    B.moi_cofm = moi;
    double r2max = 0, blx = INFINITY, bly = INFINITY, trx = -INFINITY, try_ = -INFINITY;
    for (int k = 0; k < n; k++) {
        const vvhd_seg& s = segs[f + k];
        const double r2 = (s.cx - cx) * (s.cx - cx) + (s.cy - cy) * (s.cy - cy);
        if (r2 > r2max) r2max = r2;
        if (s.cx > trx) trx = s.cx;
        if (s.cy > try_) try_ = s.cy;
        if (s.cx < blx) blx = s.cx;
        if (s.cy < bly) bly = s.cy;
    }
    geo[0] = r2max;
    geo[1] = blx;
    geo[2] = bly;
    geo[3] = trx;
    geo[4] = try_;
}

__global__ void k_body_properties(vvhd_body* bodies, int nbody, vvhd_seg* segs, double* geo) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nbody) body_properties(bodies[b], segs, geo + 5 * b);
}

// TBody::move for a body with no prescribed motion, moved by the speed of
// the last solution (TBody.cpp:179-198, MFlowmove.cpp:200-246).
__global__ void k_body_move(vvhd_body* bodies, int nbody, vvhd_seg* segs, double* geo, double dt) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbody) return;
    vvhd_body& B = bodies[b];
    const double dH[3] = {0. * dt, 0. * dt, 0. * dt};
    const double dB[3] = {B.speed_slae[0] * dt, B.speed_slae[1] * dt, B.speed_slae[2] * dt};
    const double ax = B.holder[0] + B.dpos[0], ay = B.holder[1] + B.dpos[1];
    const double cs = cos(dB[2]), sn = sin(dB[2]);
    for (int k = 0; k < B.nseg; k++) {
        vvhd_seg& s = segs[B.first_seg + k];
        const double drx = s.cx - ax, dry = s.cy - ay;
        const double nx = ax + dB[0] + drx * cs + (-dry) * sn;
        const double ny = ay + dB[1] + dry * cs + drx * sn;
        s.cx = nx;
        s.cy = ny;
    }
    for (int i = 0; i < 3; i++) {
        B.holder[i] += dH[i];
        B.dpos[i] += dB[i] - dH[i];
        B.speed_slae_prev[i] = B.speed_slae[i];
    }
    body_properties(B, segs, geo + 5 * b);
}

__device__ inline int body_of_seg(const vvhd_body* bodies, int nbody, int s) {
    int b = 0;
    while (b + 1 < nbody && bodies[b + 1].first_seg <= s) b++;
    return b;
}

// ---------------------------------------------------------------- flow-move

// Flow-move over the vortex list (MFlowmove.cpp:104-145) in two passes with
// no waits between blocks: k_fm_count advects each tile of kFmTile vortices
// and counts its survivors and in-body removals (wave ballots); k_fm_scatter
// recomputes the same positions and flags, adds the counts of the tiles
// before its own (a block reduction over the tile counts) and writes the
// survivors, velocity zeroed, stably into the double buffers, and the in-body
// removals in vortex order as dead records {x, y, g, segment} for the dead
// force.  The last tile writes both totals to counts[0..1].
constexpr int kFmBlock = 256;
constexpr int kFmItems = 8;
constexpr int kFmTile = kFmBlock * kFmItems;
constexpr int kFmWaves = kFmBlock / 64;

struct FmArgs {
    const double *x, *y, *vx, *vy, *g, *ie;
    int64_t n;
    double dt;
    int remove;
    const vvhd_body* bodies;
    int nbody;
    const vvhd_seg* segs;
    const double* geo;
    // simulations sharing the context: vortex i of simulation s meets only
    // the bodies [sim_body[s], sim_body[s + 1]) (nsim = 1: all bodies)
    const int64_t* sim_off;
    const int32_t* sim_body;
    int nsim;
};

// Simulation of vortex i: the last s with sim_off[s] <= i.
__device__ inline int sim_of(const int64_t* sim_off, int nsim, int64_t i) {
    int lo = 0, hi = nsim - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sim_off[mid] <= i)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Moved position and removal class of vortex i: 1 kept, 2 removed in a body
// (segment in *bs), 0 removed for |g| < remove_eps or past the list.
__device__ inline int fm_classify(const FmArgs& A, int64_t i, double& px, double& py, int& bs) {
    bs = -1;
    if (i >= A.n) return 0;
    px = A.x[i] + A.vx[i] * A.dt;
    py = A.y[i] + A.vy[i] * A.dt;
    if (fabs(A.g[i]) < kRemoveEps) return 0;
    int b0 = 0, b1 = A.nbody;
    if (A.nsim > 1) {
        const int sm = sim_of(A.sim_off, A.nsim, i);
        b0 = A.sim_body[sm];
        b1 = A.sim_body[sm + 1];
    }
    if (A.remove)
        for (int b = b0; b < b1; b++) {
            const int sgi = point_invalid(A.bodies[b], A.segs, A.geo + 5 * b, px, py);
            if (sgi >= 0) {
                bs = sgi;
                return 2;
            }
        }
    return 1;
}

__global__ __launch_bounds__(kFmBlock) void k_fm_count(FmArgs A, int32_t* tile_cnt) {
    __shared__ uint32_t s_cnt[2][kFmWaves];
    const int t = threadIdx.x, lane = lane_id(), w = t >> 6;
    const int64_t base = (int64_t)blockIdx.x * kFmTile;
    uint32_t ck = 0, cd = 0;
#pragma unroll
    for (int k = 0; k < kFmItems; k++) {
        double px, py;
        int bs;
        const int cls = fm_classify(A, base + k * kFmBlock + t, px, py, bs);
        ck += (uint32_t)__popcll(__ballot(cls == 1));
        cd += (uint32_t)__popcll(__ballot(cls == 2));
    }
    if (lane == 0) {
        s_cnt[0][w] = ck;
        s_cnt[1][w] = cd;
    }
    __syncthreads();
    if (t == 0) {
        uint32_t a = 0, b = 0;
        for (int v = 0; v < kFmWaves; v++) {
            a += s_cnt[0][v];
            b += s_cnt[1][v];
        }
        tile_cnt[2 * blockIdx.x] = (int32_t)a;
        tile_cnt[2 * blockIdx.x + 1] = (int32_t)b;
    }
}

__global__ __launch_bounds__(kFmBlock) void k_fm_scatter(FmArgs A, const int32_t* tile_cnt, int ntiles,
                                                         double* __restrict__ x2, double* __restrict__ y2,
                                                         double* __restrict__ g2, double* __restrict__ vx2,
                                                         double* __restrict__ vy2, double* __restrict__ ie2,
                                                         double4* __restrict__ dead, int32_t* counts,
                                                         int32_t* __restrict__ kept_before,
                                                         int32_t* __restrict__ dead_before) {
    __shared__ uint32_t s_cnt[2][kFmItems][kFmWaves];
    __shared__ uint32_t s_red[2][kFmWaves];
    const int t = threadIdx.x, lane = lane_id(), w = t >> 6;
    const int tile = blockIdx.x;
    const int64_t base = (int64_t)tile * kFmTile;
    // counts of the tiles before this one
    uint32_t pk = 0, pd = 0;
    for (int q = t; q < tile; q += kFmBlock) {
        pk += (uint32_t)tile_cnt[2 * q];
        pd += (uint32_t)tile_cnt[2 * q + 1];
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        pk += __shfl_xor(pk, o, kWave);
        pd += __shfl_xor(pd, o, kWave);
    }
    double px[kFmItems], py[kFmItems];
    int bs[kFmItems];
    uint64_t mk[kFmItems], md[kFmItems];
#pragma unroll
    for (int k = 0; k < kFmItems; k++) {
        const int cls = fm_classify(A, base + k * kFmBlock + t, px[k], py[k], bs[k]);
        mk[k] = __ballot(cls == 1);
        md[k] = __ballot(cls == 2);
        if (lane == 0) {
            s_cnt[0][k][w] = (uint32_t)__popcll(mk[k]);
            s_cnt[1][k][w] = (uint32_t)__popcll(md[k]);
        }
    }
    if (lane == 0) {
        s_red[0][w] = pk;
        s_red[1][w] = pd;
    }
    __syncthreads();
    uint32_t ok = 0, od = 0;
    for (int v = 0; v < kFmWaves; v++) {
        ok += s_red[0][v];
        od += s_red[1][v];
    }
    if (t == 0 && tile == ntiles - 1) {
        uint32_t ak = ok, ad = od;
        for (int k = 0; k < kFmItems; k++)
            for (int v = 0; v < kFmWaves; v++) {
                ak += s_cnt[0][k][v];
                ad += s_cnt[1][k][v];
            }
        counts[0] = (int32_t)ak;
        counts[1] = (int32_t)ad;
    }
    const uint64_t lt = (1ull << lane) - 1;
#pragma unroll
    for (int k = 0; k < kFmItems; k++) {
        uint32_t rk = ok, rd = od;
        for (int v = 0; v < w; v++) {
            rk += s_cnt[0][k][v];
            rd += s_cnt[1][k][v];
        }
        const int64_t i = base + k * kFmBlock + t;
        // survivors before vortex i (the new offset of a simulation starting at i)
        if (kept_before && i <= A.n) {
            kept_before[i] = (int32_t)(rk + (uint32_t)__popcll(mk[k] & lt));
            dead_before[i] = (int32_t)(rd + (uint32_t)__popcll(md[k] & lt));
        }
        if (mk[k] >> lane & 1) {
            const uint32_t q = rk + (uint32_t)__popcll(mk[k] & lt);
            x2[q] = px[k];
            y2[q] = py[k];
            g2[q] = A.g[i];
            vx2[q] = 0.;
            vy2[q] = 0.;
            ie2[q] = A.ie[i];
        } else if (md[k] >> lane & 1) {
            const uint32_t q = rd + (uint32_t)__popcll(md[k] & lt);
            dead[q] = make_double4(px[k], py[k], A.g[i], (double)bs[k]);
        }
        for (int v = 0; v < kFmWaves; v++) {
            ok += s_cnt[0][k][v];
            od += s_cnt[1][k][v];
        }
    }
}

__global__ void k_fill_u64(uint64_t* p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Dead force, moment and circulation of the removed vortices in vortex order
// (MFlowmove.cpp:124-144), then the zero/slip segment loop (:205-214).
// One thread per body: the dead records of its simulation's vortices (a
// contiguous run of the records, in vortex order) whose segment is its own,
// then its zero and slip segments.  With one simulation the run is every
// record.  dead_off: per simulation, the records before its first vortex
// (nullptr: one simulation).
__global__ void k_dead(vvhd_body* bodies, int nbody, vvhd_seg* segs, const double4* dead, const int32_t* counts,
                       const int32_t* body_sim, const int32_t* dead_off) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbody) return;
    vvhd_body& B = bodies[b];
    const int f0 = B.first_seg, f1 = B.first_seg + B.nseg;
    int q0 = 0, q1 = counts[1];
    if (dead_off) {
        q0 = dead_off[body_sim[b]];
        q1 = dead_off[body_sim[b] + 1];
    }
    for (int q = q0; q < q1; q++) {
        const double4 r = dead[q];
        const int s = (int)r.w;
        if (s < f0 || s >= f1) continue;
        const double gi = r.z, xi = r.x, yi = r.y;
        B.fdt_dead[0] += (-yi) * gi;
        B.fdt_dead[1] += xi * gi;
        const double ax = B.holder[0] + B.dpos[0], ay = B.holder[1] + B.dpos[1];
        B.fdt_dead[2] += ((xi - ax) * (xi - ax) + (yi - ay) * (yi - ay)) * gi;
        segs[s].gsum -= gi;
        B.g_dead += gi;
    }
    const double ax = B.holder[0] + B.dpos[0], ay = B.holder[1] + B.dpos[1];
    for (int k = 0; k < B.nseg; k++) {
        vvhd_seg& s = segs[B.first_seg + k];
        if (fabs(s.g) < kRemoveEps || s.slip) {
            B.fdt_dead[0] += (-s.cy) * s.g;
            B.fdt_dead[1] += s.cx * s.g;
            B.fdt_dead[2] += ((s.cx - ax) * (s.cx - ax) + (s.cy - ay) * (s.cy - ay)) * s.g;
            s.gsum -= s.g;
            B.g_dead += s.g;
        }
    }
}

// Per simulation s: the dead records before its first vortex, from the
// per-vortex counts the scatter wrote (dead_before), and the total last.
__global__ void k_dead_off(const int64_t* sim_off, int nsim, const int32_t* dead_before, const int32_t* counts,
                           int64_t n, int32_t* dead_off) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s > nsim) return;
    const int64_t i = sim_off[s];
    dead_off[s] = (s == nsim || i >= n) ? counts[1] : dead_before[i];
}

// Where each segment's shed vortex lands: a flag per segment, positions by an
// exclusive scan, so the new vortices follow the segment order.
__global__ void k_shed_flags(vvhd_seg* segs, int nseg, int32_t* flag) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s > nseg) return;
    if (s == nseg) {
        flag[s] = 0;
        return;
    }
    vvhd_seg& r = segs[s];
    r.gsum += r.g;
    flag[s] = (fabs(r.g) >= kRemoveEps && !r.slip) ? 1 : 0;
}

// vortex_shed with several simulations: every vortex of simulation s moves
// up by the shed count of the simulations before it (pos at the first
// segment of s), and segment k of simulation s writes its vortex after the
// (moved) vortices of s: at sim_off[s + 1] + pos[k].
__global__ void k_shed_batch_move(const int64_t* sim_off, int nsim, const int32_t* sim_seg0, const int32_t* pos,
                                  int64_t n, const double* x, const double* y, const double* g, const double* vx,
                                  const double* vy, const double* ie, double* x2, double* y2, double* g2,
                                  double* vx2, double* vy2, double* ie2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int sm = sim_of(sim_off, nsim, i);
    const int64_t w = i + pos[sim_seg0[sm]];
    x2[w] = x[i];
    y2[w] = y[i];
    g2[w] = g[i];
    vx2[w] = vx[i];
    vy2[w] = vy[i];
    ie2[w] = ie[i];
}

__global__ void k_shed_batch_write(const vvhd_seg* segs, int nseg, const int32_t* seg_sim, const int64_t* sim_off,
                                   const int32_t* flag, const int32_t* pos, double* x, double* y, double* g,
                                   double* vx, double* vy, double* ie) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nseg || !flag[k]) return;
    const vvhd_seg& r = segs[k];
    const int64_t w = sim_off[seg_sim[k] + 1] + pos[k];
    x[w] = r.cx + (-r.dly) * 1e-4;
    y[w] = r.cy + r.dlx * 1e-4;
    g[w] = r.g;
    vx[w] = 0;
    vy[w] = 0;
    ie[w] = 0;
}

__global__ void k_shed_write(const vvhd_seg* segs, int nseg, const int32_t* flag, const int32_t* pos, int64_t n,
                             double* x, double* y, double* g, double* vx, double* vy, double* ie) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg || !flag[s]) return;
    const vvhd_seg& r = segs[s];
    const int64_t w = n + pos[s];
    x[w] = r.cx + (-r.dly) * 1e-4;
    y[w] = r.cy + r.dlx * 1e-4;
    g[w] = r.g;
    vx[w] = 0;
    vy[w] = 0;
    ie[w] = 0;
}

// Space::calc_forces (TSpace.cpp:94-120); fric holds the diffusive friction.
__global__ void k_calc_forces(vvhd_body* bodies, int nbody, vvhd_seg* segs, const double* fric, double re_one,
                              double dt, double pr, const double* sim_re, const int32_t* body_sim) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbody) return;
    const double re = sim_re ? sim_re[body_sim[b]] : re_one;   // the body's simulation's Reynolds number
    const double C_NyuDt_Pi = dt / (M_PI * re);
    const double C_Nyu_Pi = 1. / (M_PI * re);
    vvhd_body& B = bodies[b];
    double tmp = 0;
    B.friction[0] = B.friction[1] = B.friction[2] = 0;
    for (int k = 0; k < B.nseg; k++) {
        vvhd_seg& s = segs[B.first_seg + k];
        s.fric = fric[B.first_seg + k];
        const double a = sqrt(s.dlx * s.dlx + s.dly * s.dly);
        tmp += s.gsum;
        s.Cp += tmp;
        s.Fr += s.fric * C_NyuDt_Pi;
        s.Nu += s.hsum * (re * pr / a);
        const double c = s.fric * C_Nyu_Pi / a;
        B.friction[0] -= s.dlx * c;
        B.friction[1] -= s.dly * c;
        B.friction[2] -= dot2(-s.y, s.x, s.dlx, s.dly) * c;
        B.nusselt += s.hsum * (re * pr);
    }
    B.nusselt /= dt;
}

// Space::zero_forces (TSpace.cpp:122-143).
__global__ void k_zero_forces(vvhd_body* bodies, int nbody, vvhd_seg* segs, double* fric) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbody) return;
    vvhd_body& B = bodies[b];
    for (int k = 0; k < B.nseg; k++) {
        vvhd_seg& s = segs[B.first_seg + k];
        s.gsum = s.fric = s.hsum = 0;
        fric[B.first_seg + k] = 0;
    }
    for (int i = 0; i < 3; i++) {
        B.force_hydro[i] = 0;
        B.force_holder[i] = 0;
        B.friction_prev[i] = B.friction[i];
        B.friction[i] = 0;
        B.fdt_dead[i] = 0;
    }
    B.g_dead = 0;
    B.nusselt = 0.;
}

// Stepdata::write's reset of the per-segment pressure, friction and Nusselt
// accumulators after writing the profiles (MStepdata.cpp:166-176).
__global__ void k_zero_profiles(vvhd_seg* segs, int nseg) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nseg) return;
    segs[i].Cp = 0;
    segs[i].Fr = 0;
    segs[i].Nu = 0;
}

// ------------------------------------------------------------------- SLAE

// MConvectiveFast::_2PI_Xi_g and its near / dist parts (MConvectiveFast.cpp:
// 275-310), without the 2 pi factor.
__device__ inline double xi_g_near(double px, double py, double cx, double cy, double dlx, double dly, double rd) {
    return dot2(cx - px, cy - py, dlx, dly) / (rd * rd);
}
__device__ inline double xi_g_dist(double px, double py, double x1, double y1, double x2, double y2) {
    const double a = (px - x2) * (px - x2) + (py - y2) * (py - y2);
    const double b = (px - x1) * (px - x1) + (py - y1) * (py - y1);
    return 0.5 * log(a / b);
}
__device__ double xi_g(double px, double py, const vvhd_seg& s, double rd) {
    const double rd_sqr = rd * rd;
    const double dr1 = (px - s.cx) * (px - s.cx) + (py - s.cy) * (py - s.cy);
    const double ex = px - s.cx - s.dlx, ey = py - s.cy - s.dly;
    const double dr2 = ex * ex + ey * ey;
    if (dr1 >= rd_sqr && dr2 >= rd_sqr) return 0.5 * log(dr2 / dr1);
    if (dr1 <= rd_sqr && dr2 <= rd_sqr) return xi_g_near(px, py, s.x, s.y, s.dlx, s.dly, rd);
    const double a0 = s.dlx * s.dlx + s.dly * s.dly;
    const double b0 = dot2(px - s.x, py - s.y, s.dlx, s.dly);
    const double d = sqrt(b0 * b0 - a0 * (((px - s.x) * (px - s.x) + (py - s.y) * (py - s.y)) - rd * rd));
    double k = (b0 + d) / a0;
    if ((k <= -0.5) || (k >= 0.5)) k = (b0 - d) / a0;
    const double p3x = s.x + k * s.dlx, p3y = s.y + k * s.dly;
    const double c2x = s.cx + s.dlx, c2y = s.cy + s.dly;
    if (dr1 < rd_sqr)
        return xi_g_near(px, py, 0.5 * (p3x + s.cx), 0.5 * (p3y + s.cy), p3x - s.cx, p3y - s.cy, rd) +
               xi_g_dist(px, py, p3x, p3y, c2x, c2y);
    return xi_g_dist(px, py, s.cx, s.cy, p3x, p3y) +
           xi_g_near(px, py, 0.5 * (c2x + p3x), 0.5 * (c2y + p3y), c2x - p3x, c2y - p3y, rd);
}

// One system per simulation (one in all without a batch), each of order n,
// stored one after another (column-major, stride n * n; right-hand sides
// and solutions stride n).  base holds each body's first equation within
// its own system.
struct SlaeMap {
    const vvhd_body* bodies;
    const int32_t* base;      // first equation (and column) of each body in its system
    int nbody;
    int n;                    // order of every system
    const int32_t* sys_body;  // per system: first body (nsys + 1 entries)
    const int32_t* seg_sys;   // per segment: its system
    int nsys;
    int max_sys_segs;         // most segments in one system
};

__device__ inline int body_of_seg_in(const vvhd_body* bodies, int b0, int b1, int s) {
    int b = b0;
    while (b + 1 < b1 && bodies[b + 1].first_seg <= s) b++;
    return b;
}

// Influence cells of the slip rows: A(row of segment i, column of segment m)
// = Xi_g(corner_m; segment i, |dl_m| / 4) / 2 pi  (:459-482).
__global__ void k_slip_cells(SlaeMap M, const vvhd_seg* segs, int nseg, double* A) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nseg * M.max_sys_segs) return;
    const int i = (int)(t / M.max_sys_segs), mm = (int)(t % M.max_sys_segs);
    const int sys = M.seg_sys[i];
    const int b0 = M.sys_body[sys], b1 = M.sys_body[sys + 1];
    const int f0 = M.bodies[b0].first_seg, f1 = M.bodies[b1 - 1].first_seg + M.bodies[b1 - 1].nseg;
    const int m = f0 + mm;
    if (m >= f1) return;
    const int bi = body_of_seg_in(M.bodies, b0, b1, i);
    const vvhd_body& Bi = M.bodies[bi];
    if (i - Bi.first_seg == Bi.special_segment_no) return;
    const int row = M.base[bi] + (i - Bi.first_seg);
    const int bm = body_of_seg_in(M.bodies, b0, b1, m);
    const int col = M.base[bm] + (m - M.bodies[bm].first_seg);
    const vvhd_seg& sm = segs[m];
    const double rd = sqrt(sm.dlx * sm.dlx + sm.dly * sm.dly) * 0.25;
    A[(size_t)sys * M.n * M.n + (size_t)col * M.n + row] = xi_g(sm.cx, sm.cy, segs[i], rd) * k1_2Pi;
}

// Right-hand side of the slip rows, one wave (= one block) per segment:
// rotl(U) . dl minus NodeInfluence (:398-418) of the leaf holding the
// segment centre (stree::findNode, TSortedTree.cpp:284-303).  Lanes
// evaluate 64 terms at a time; lane 0 adds them in the reference's order
// (near nodes and their vortices in list order, then each far node's
// positive and negative monopole).  Also the speed cells (_2PI_A123, :375-396)
// of a rigid body with no prescribed rotation.
__global__ __launch_bounds__(64) void k_slip_rhs(SlaeMap M, TreeDev T, const vvhd_seg* segs, int nseg,
                                                 const double* x, const double* y, const double* g, const double* ie,
                                                 double infx, double infy, double* A, double* rhs) {
    __shared__ double buf[2 * kWave];
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    if (i >= nseg) return;
    const int sys = M.seg_sys[i];
    const int b0 = M.sys_body[sys], b1 = M.sys_body[sys + 1];
    const int bi = body_of_seg_in(M.bodies, b0, b1, i);
    const vvhd_body& Bi = M.bodies[bi];
    if (i - Bi.first_seg == Bi.special_segment_no) return;
    const int row = M.base[bi] + (i - Bi.first_seg);
    A += (size_t)sys * M.n * M.n;
    rhs += (size_t)sys * M.n;
    const vvhd_seg s = segs[i];
    int node = sys;   // the root of the segment's simulation
    while (T.ch1[node] >= 0) {
        const int c = T.ch1[node];
        if (T.nh[node] < T.nw[node])
            node = (s.x < T.nx[node]) ? c : c + 1;
        else
            node = (s.y < T.ny[node]) ? c : c + 1;
    }
    const int l = T.leaf_rank[node];
    // res starts at +0 and a sum of two reals rounds to -0 only when both are
    // -0, so res is never -0 and adding +0 for a skipped vortex is exact.
    double res = 0;
    for (int64_t e = T.near_off[l]; e < T.near_off[l + 1]; e++) {
        const int Mn = T.near_list[e];
        const int j1 = T.v1[Mn];
        for (int j0 = T.v0[Mn]; j0 < j1; j0 += kWave) {
            const int j = j0 + lane;
            double t = 0;
            if (j < j1) {
                const double gj = g[j];
                if (gj != 0) t = xi_g(x[j], y[j], s, 1. / ie[j]) * gj;
            }
            buf[lane] = t;
            __syncthreads();
            if (lane == 0) {
                const int cnt = min(kWave, j1 - j0);
                for (int q = 0; q < cnt; q++) res += buf[q];
            }
            __syncthreads();
        }
    }
    const double c2x = s.cx + s.dlx, c2y = s.cy + s.dly;
    const int64_t f1 = T.far_off[l + 1];
    for (int64_t f0 = T.far_off[l]; f0 < f1; f0 += kWave) {
        const int64_t e = f0 + lane;
        double tp = 0, tm = 0;
        if (e < f1) {
            const int F = T.far_list[e];
            tp = xi_g_dist(T.cmp_x[F], T.cmp_y[F], s.cx, s.cy, c2x, c2y) * T.cmp_g[F];
            tm = xi_g_dist(T.cmm_x[F], T.cmm_y[F], s.cx, s.cy, c2x, c2y) * T.cmm_g[F];
        }
        buf[2 * lane] = tp;
        buf[2 * lane + 1] = tm;
        __syncthreads();
        if (lane == 0) {
            const int cnt = (int)min((int64_t)kWave, f1 - f0);
            for (int q = 0; q < 2 * cnt; q++) res += buf[q];
        }
        __syncthreads();
    }
    if (lane == 0) {
        double r = dot2(-infy, infx, s.dlx, s.dly);
        r += dot2(-0.0, 0.0, s.dlx, s.dly);   // rotl(body speed 0 / 2 pi) . dl
        r -= res * k1_2Pi;
        rhs[row] = r;
        for (int jb = b0; jb < b1; jb++) {
            const double A1 = (jb == bi) ? k2Pi * s.dly : 0;
            const double A2 = (jb == bi) ? -k2Pi * s.dlx : 0;
            const double A3 = 0;
            const int sc = M.base[jb] + M.bodies[jb].nseg;
            A[(size_t)(sc + 0) * M.n + row] = A1 * k1_2Pi;
            A[(size_t)(sc + 1) * M.n + row] = A2 * k1_2Pi;
            A[(size_t)(sc + 2) * M.n + row] = A3 * k1_2Pi;
        }
    }
}

// Special-segment rows and the nine body rows (:484-862, :896-975), one
// thread per system: the steady-circulation rows depend on a counter over
// the system's bodies.  gsum_all holds each system's vortex circulation.
__global__ void k_body_rows(SlaeMap M, const vvhd_seg* segs, const double* gsum_all, double inf_g, double dt,
                            double* A, double* rhs) {
    const int sys = blockIdx.x * blockDim.x + threadIdx.x;
    if (sys >= M.nsys) return;
    const int n = M.n;
    A += (size_t)sys * n * n;
    rhs += (size_t)sys * n;
    gsum_all += sys;
    const int b0 = M.sys_body[sys], b1 = M.sys_body[sys + 1];
    auto cell = [&](int row, int col) -> double& { return A[(size_t)col * n + row]; };
    int once = 0;
    for (int b = b0; b < b1; b++) {
        const vvhd_body& B = M.bodies[b];
        const int base = M.base[b], ns = B.nseg;
        const int cs = base + ns, ch = base + ns + 3, ck = base + ns + 6;   // speed, hydro, holder columns
        {
            const int row = base + B.special_segment_no;
            if (B.kutta) {   // fillZeroEquationForSegment (:484-492)
                rhs[row] = 0;
                cell(row, base + B.special_segment_no) = 1;
            } else if (!once++) {   // fillInfSteadyEquationForSegment (:508-522)
                rhs[row] = -(*gsum_all) - inf_g;
                for (int jb = b0; jb < b1; jb++) {
                    for (int m = 0; m < M.bodies[jb].nseg; m++) cell(row, M.base[jb] + m) = 1;
                    cell(row, M.base[jb] + M.bodies[jb].nseg + 2) = 2 * M.bodies[jb].area;
                }
            } else {   // fillSteadyEquationForSegment (:494-506)
                rhs[row] = B.g_dead + 2 * B.area * B.speed_slae_prev[2];
                for (int m = 0; m < ns; m++) cell(row, base + m) = 1;
                cell(row, cs + 2) = 2 * B.area;
            }
        }
        int eq = base + ns;
        for (int c = 0; c < 3; c++, eq++) {   // fillSpeed?Equation (:803-862), rigid
            rhs[eq] = 0.0;
            cell(eq, cs + c) = 1;
        }
        const double _1_dt = 1 / dt, _1_2dt = 0.5 / dt;
        const double axx = B.holder[0] + B.dpos[0], axy = B.holder[1] + B.dpos[1];
        const double rx = B.cofm[0] - axx, ry = B.cofm[1] - axy;
        const double area = B.area, rho = 1.0;
        const double* sp = B.speed_slae_prev;
        const double moi_axis =
            B.moi_cofm + ((axx - B.cofm[0]) * (axx - B.cofm[0]) + (axy - B.cofm[1]) * (axy - B.cofm[1])) * area;
        // fillHydroXEquation (:533-562)
        rhs[eq] = +area * _1_dt * sp[0] + area * _1_dt * (-(2 * B.cofm[1] + ry)) * sp[2] + (sp[2] * sp[2]) * area * rx -
                  B.friction_prev[0] + B.fdt_dead[0] * _1_dt;
        for (int m = 0; m < ns; m++) cell(eq, base + m) = _1_dt * (-segs[B.first_seg + m].cy);
        cell(eq, cs + 0) = area * _1_dt;
        cell(eq, cs + 1) = 0;
        cell(eq, cs + 2) = area * _1_dt * (-2 * B.cofm[1] - ry);
        cell(eq, ch + 0) = -1;
        eq++;
        // fillHydroYEquation (:564-593)
        rhs[eq] = +area * _1_dt * sp[1] + area * _1_dt * (2 * B.cofm[0] + rx) * sp[2] + (sp[2] * sp[2]) * area * ry -
                  B.friction_prev[1] + B.fdt_dead[1] * _1_dt;
        for (int m = 0; m < ns; m++) cell(eq, base + m) = _1_dt * (segs[B.first_seg + m].cx);
        cell(eq, cs + 0) = 0;
        cell(eq, cs + 1) = area * _1_dt;
        cell(eq, cs + 2) = area * _1_dt * (2 * B.cofm[0] + rx);
        cell(eq, ch + 1) = -1;
        eq++;
        // fillHydroOEquation (:595-623)
        {
            const double c = area * _1_dt;
            rhs[eq] = +((-ry) * c * sp[0] + rx * c * sp[1]) + 2 * moi_axis * _1_dt * sp[2] - B.friction_prev[2] +
                      B.fdt_dead[2] * _1_2dt;
        }
        for (int m = 0; m < ns; m++) {
            const vvhd_seg& sm = segs[B.first_seg + m];
            cell(eq, base + m) = _1_2dt * ((sm.cx - axx) * (sm.cx - axx) + (sm.cy - axy) * (sm.cy - axy));
        }
        cell(eq, cs + 0) = area * _1_dt * (-ry);
        cell(eq, cs + 1) = area * _1_dt * (rx);
        cell(eq, cs + 2) = 2 * moi_axis * _1_dt;
        cell(eq, ch + 2) = -1;
        eq++;
        // fillNewtonXEquation (:634-667); no gravity, no child bodies
        rhs[eq] = -area * _1_dt * rho * sp[0] + area * _1_dt * rho * ry * sp[2] - area * rho * rx * (sp[2] * sp[2]) -
                  (rho - 1.0) * 0. * area;
        cell(eq, cs + 0) = -area * _1_dt * rho;
        cell(eq, cs + 1) = 0;
        cell(eq, cs + 2) = area * _1_dt * rho * ry;
        cell(eq, ch + 0) = 1;
        cell(eq, ck + 0) = -1;
        eq++;
        // fillNewtonYEquation (:669-702)
        rhs[eq] = -area * _1_dt * rho * sp[1] - area * _1_dt * rho * rx * sp[2] - area * rho * ry * (sp[2] * sp[2]) -
                  (rho - 1.0) * 0. * area;
        cell(eq, cs + 0) = 0;
        cell(eq, cs + 1) = -area * _1_dt * rho;
        cell(eq, cs + 2) = -area * _1_dt * rho * rx;
        cell(eq, ch + 1) = 1;
        cell(eq, ck + 1) = -1;
        eq++;
        // fillNewtonOEquation (:704-739)
        rhs[eq] = -area * _1_dt * rho * ((-ry) * sp[0] + rx * sp[1]) - moi_axis * _1_dt * rho * sp[2] -
                  (rho - 1.0) * ((-ry) * 0.0 + rx * 0.0) * area;
        cell(eq, cs + 0) = +area * _1_dt * rho * ry;
        cell(eq, cs + 1) = -area * _1_dt * rho * rx;
        cell(eq, cs + 2) = -moi_axis * _1_dt * rho;
        cell(eq, ch + 2) = 1;
        cell(eq, ck + 2) = -1;
    }
}

// Total vortex circulation (Space::gsum, TSpace.hpp:93-99) of every
// simulation, one block each over its range [sim_off[s], sim_off[s + 1]):
// per-thread Neumaier sums over contiguous chunks combined in a fixed tree,
// so the result is deterministic and at least as accurate as the serial sum.
__global__ __launch_bounds__(1024) void k_gsum(const double* g, const int64_t* sim_off, double* out) {
    __shared__ double ss[1024], sc[1024];
    const int t = threadIdx.x;
    g += sim_off[blockIdx.x];
    const int64_t n = sim_off[blockIdx.x + 1] - sim_off[blockIdx.x];
    out += blockIdx.x;
    const int64_t chunk = (n + 1023) / 1024;
    const int64_t a = min(n, (int64_t)t * chunk), b = min(n, a + chunk);
    double s = 0, c = 0;
    for (int64_t i = a; i < b; i++) {
        const double v = g[i];
        const double u = s + v;
        c += (fabs(s) >= fabs(v)) ? ((s - u) + v) : ((v - u) + s);
        s = u;
    }
    ss[t] = s;
    sc[t] = c;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (t < o) {
            const double s1 = ss[t], s2 = ss[t + o];
            const double u = s1 + s2;
            const double e = (fabs(s1) >= fabs(s2)) ? ((s1 - u) + s2) : ((s2 - u) + s1);
            ss[t] = u;
            sc[t] = sc[t] + sc[t + o] + e;
        }
        __syncthreads();
    }
    if (t == 0) *out = ss[0] + sc[0];
}

__global__ void k_differ(const double* a, const double* b, size_t n, int32_t* differ) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && __double_as_longlong(a[i]) != __double_as_longlong(b[i])) atomicOr(differ, 1);
}

__global__ void k_writeback(SlaeMap M, vvhd_body* bodies, vvhd_seg* segs, const double* xsol,
                            const int32_t* body_sys) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= M.nbody) return;
    vvhd_body& B = bodies[b];
    const int base = M.base[b];
    xsol += (size_t)body_sys[b] * M.n;
    for (int k = 0; k < B.nseg; k++) segs[B.first_seg + k].g = xsol[base + k];
    for (int i = 0; i < 3; i++) {
        B.speed_slae[i] = xsol[base + B.nseg + i];
        B.force_hydro[i] = xsol[base + B.nseg + 3 + i];
        B.force_holder[i] = xsol[base + B.nseg + 6 + i];
    }
}

}  // namespace

void sync_bodies_to_host(Ctx& c) {
    const int ns = (int)c.segs_h.size(), nb = (int)c.bodies_h.size();
    if (ns)
        VVHD_HIP(hipMemcpyAsync(c.segs_h.data(), c.segs.p, (size_t)ns * sizeof(vvhd_seg), hipMemcpyDeviceToHost,
                                c.stream));
    if (nb)
        VVHD_HIP(hipMemcpyAsync(c.bodies_h.data(), c.bodies_d.p, (size_t)nb * sizeof(vvhd_body),
                                hipMemcpyDeviceToHost, c.stream));
    VVHD_HIP(hipStreamSynchronize(c.stream));
}

static void exclusive_scan_i32(Ctx& c, const int32_t* in, int32_t* out, size_t n) {
    size_t bytes = 0;
    VVHD_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, (int32_t)0, n, rocprim::plus<int32_t>(), c.stream));
    c.tmp.grow(bytes + 16);
    VVHD_HIP(rocprim::exclusive_scan(c.tmp.p, bytes, in, out, (int32_t)0, n, rocprim::plus<int32_t>(), c.stream));
}

// Geometry of every uploaded body from its corners (the reference computes
// it when a body is loaded: TBody::doUpdateSegments + doFillProperties).
void bodies_prepare(Ctx& c) {
    const int nb = (int)c.bodies_h.size();
    c.bodies_d.grow(std::max(nb, 1));
    c.body_geo.grow(5 * (size_t)std::max(nb, 1));
    if (!nb) return;
    VVHD_HIP(hipMemcpyAsync(c.bodies_d.p, c.bodies_h.data(), (size_t)nb * sizeof(vvhd_body), hipMemcpyHostToDevice,
                            c.stream));
    hipLaunchKernelGGL(k_body_properties, dim3(nblocks(nb, 64)), dim3(64), 0, c.stream, c.bodies_d.p, nb, c.segs.p,
                       c.body_geo.p);
    sync_bodies_to_host(c);
}

void bodies_release(Ctx& c) {
    c.bodies_d.release();
    c.body_geo.release();
    for (auto* b : {&c.slae_A, &c.slae_Acached, &c.slae_LU, &c.slae_b, &c.slae_x}) b->release();
    for (auto* b : {&c.slae_base, &c.slae_ipiv, &c.slae_info, &c.iscratch4}) b->release();
    if (c.rocblas) (void)rocblas_destroy_handle((rocblas_handle)c.rocblas);
    c.rocblas = nullptr;
    c.slae_n = -1;
}

void move_and_clean(Ctx& c, int remove, int64_t* cleaned) {
    const int nb = (int)c.bodies_h.size();
    hipStream_t s = c.stream;
    if (c.nsim > 1) sim_sync(c);   // the device copy of the simulations' ranges
    if (nb)
        hipLaunchKernelGGL(k_body_move, dim3(nblocks(nb, 64)), dim3(64), 0, s, c.bodies_d.p, nb, c.segs.p,
                           c.body_geo.p, c.dt);
    const int64_t n = c.n;
    VVHD_CHECK(n < (int64_t(1) << 31), VVHD_EUNSUPPORTED, "move_and_clean: more than 2^31 - 1 vortices");
    const int ntiles = (int)std::max<int64_t>(1, (n + kFmTile - 1) / kFmTile);
    c.iscratch0.grow(4);
    int32_t* counts = c.iscratch0.p;
    c.iscratch1.grow(2 * (size_t)ntiles + 2);
    int32_t* tile_cnt = c.iscratch1.p;
    c.tmp.grow((size_t)std::max<int64_t>(n, 1) * sizeof(double4));
    double4* dead = (double4*)c.tmp.p;
    if (n == 0) {
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(kBlock), 0, s, (uint64_t*)counts, (int64_t)1);
    } else {
        const bool batch = c.nsim > 1;
        if (batch) {
            c.iscratch2.grow((size_t)ntiles * kFmTile + 1);
            c.iscratch3.grow((size_t)ntiles * kFmTile + 1);
        }
        const FmArgs A{c.x.p, c.y.p, c.vx.p, c.vy.p, c.g.p, c.ie.p, n, c.dt, remove, c.bodies_d.p, nb, c.segs.p,
                       c.body_geo.p, (const int64_t*)c.sim_off.p, (const int32_t*)c.sim_body.p, c.nsim};
        hipLaunchKernelGGL(k_fm_count, dim3(ntiles), dim3(kFmBlock), 0, s, A, tile_cnt);
        hipLaunchKernelGGL(k_fm_scatter, dim3(ntiles), dim3(kFmBlock), 0, s, A, (const int32_t*)tile_cnt, ntiles,
                           c.x2.p, c.y2.p, c.g2.p, c.vx2.p, c.vy2.p, c.ie2.p, dead, counts,
                           batch ? c.iscratch2.p : (int32_t*)nullptr, batch ? c.iscratch3.p : (int32_t*)nullptr);
    }
    if (nb) {
        const int32_t* body_sim = nullptr;
        int32_t* dead_off = nullptr;
        if (c.nsim > 1 && n) {
            std::vector<int32_t> bs(nb);
            for (int q = 0; q < c.nsim; q++)
                for (int b = c.sim_body_h[q]; b < c.sim_body_h[q + 1]; b++) bs[b] = q;
            c.iscratch4.grow((size_t)nb + c.nsim + 2);
            VVHD_HIP(hipMemcpyAsync(c.iscratch4.p, bs.data(), (size_t)nb * 4, hipMemcpyHostToDevice, s));
            body_sim = c.iscratch4.p;
            dead_off = c.iscratch4.p + nb;
            hipLaunchKernelGGL(k_dead_off, dim3(nblocks(c.nsim + 1)), dim3(kBlock), 0, s, (const int64_t*)c.sim_off.p,
                               c.nsim, (const int32_t*)c.iscratch3.p, (const int32_t*)counts, n, dead_off);
        }
        hipLaunchKernelGGL(k_dead, dim3(nblocks(nb, 64)), dim3(64), 0, s, c.bodies_d.p, nb, c.segs.p, dead, counts,
                           body_sim, (const int32_t*)dead_off);
    }
    VVHD_HIP(hipGetLastError());
    VVHD_HIP(hipEventRecord(c.ev[9], s));   // end of the device work of the phase
    VVHD_HIP(hipMemcpyAsync(c.fm_counts, counts, 8, hipMemcpyDeviceToHost, s));
    VVHD_HIP(hipStreamSynchronize(s));
    if (n) {
        std::swap(c.x.p, c.x2.p);
        std::swap(c.y.p, c.y2.p);
        std::swap(c.g.p, c.g2.p);
        std::swap(c.vx.p, c.vx2.p);
        std::swap(c.vy.p, c.vy2.p);
        std::swap(c.ie.p, c.ie2.p);
    }
    c.n = c.fm_counts[0];
    if (cleaned) *cleaned = c.fm_counts[1];
    c.t.built = false;
    if (c.nsim > 1) {   // each simulation's new range: the survivors before its first vortex
        std::vector<int32_t> kb(c.nsim);
        for (int q = 1; q < c.nsim; q++)
            if (c.sim_off_h[q] < n)
                VVHD_HIP(hipMemcpyAsync(&kb[q], c.iscratch2.p + c.sim_off_h[q], 4, hipMemcpyDeviceToHost, s));
        VVHD_HIP(hipStreamSynchronize(s));
        for (int q = 1; q < c.nsim; q++) c.sim_off_h[q] = c.sim_off_h[q] < n ? kb[q] : c.n;
        c.sim_off_h[c.nsim] = c.n;
    }
    if (c.heat.n || c.streak.n) passive_move_and_clean(c);
    if (nb) sync_bodies_to_host(c);
}

void vortex_shed(Ctx& c) {
    const int nb = (int)c.bodies_h.size();
    if (!nb) return;
    const int nseg = (int)c.segs_h.size();
    reserve_vortices(c, c.n + nseg + 1, true);
    c.iscratch0.grow(nseg + 1);
    c.iscratch1.grow(nseg + 1);
    hipLaunchKernelGGL(k_shed_flags, dim3(nblocks(nseg + 1)), dim3(kBlock), 0, c.stream, c.segs.p, nseg,
                       c.iscratch0.p);
    exclusive_scan_i32(c, c.iscratch0.p, c.iscratch1.p, (size_t)nseg + 1);
    if (c.nsim > 1) {
        // segments come in simulation order: simulation s sheds pos[seg0[s + 1]] - pos[seg0[s]]
        sim_sync(c);
        const int ns = c.nsim;
        std::vector<int32_t> seg0(ns + 1, nseg);
        for (int q = ns - 1; q >= 0; q--)
            seg0[q] = c.sim_body_h[q] < c.sim_body_h[q + 1] ? c.bodies_h[c.sim_body_h[q]].first_seg : seg0[q + 1];
        c.iscratch2.grow(ns + 1);
        VVHD_HIP(hipMemcpyAsync(c.iscratch2.p, seg0.data(), (size_t)(ns + 1) * 4, hipMemcpyHostToDevice, c.stream));
        if (c.n)
            hipLaunchKernelGGL(k_shed_batch_move, dim3(nblocks(c.n)), dim3(kBlock), 0, c.stream,
                               (const int64_t*)c.sim_off.p, ns, (const int32_t*)c.iscratch2.p,
                               (const int32_t*)c.iscratch1.p, c.n, c.x.p, c.y.p, c.g.p, c.vx.p, c.vy.p, c.ie.p, c.x2.p,
                               c.y2.p, c.g2.p, c.vx2.p, c.vy2.p, c.ie2.p);
        hipLaunchKernelGGL(k_shed_batch_write, dim3(nblocks(nseg)), dim3(kBlock), 0, c.stream, c.segs.p, nseg,
                           (const int32_t*)c.seg_root.p, (const int64_t*)c.sim_off.p, (const int32_t*)c.iscratch0.p,
                           (const int32_t*)c.iscratch1.p, c.x2.p, c.y2.p, c.g2.p, c.vx2.p, c.vy2.p, c.ie2.p);
        std::vector<int32_t> pos(nseg + 1);
        VVHD_HIP(hipMemcpyAsync(pos.data(), c.iscratch1.p, (size_t)(nseg + 1) * 4, hipMemcpyDeviceToHost, c.stream));
        VVHD_HIP(hipStreamSynchronize(c.stream));
        std::swap(c.x.p, c.x2.p);
        std::swap(c.y.p, c.y2.p);
        std::swap(c.g.p, c.g2.p);
        std::swap(c.vx.p, c.vx2.p);
        std::swap(c.vy.p, c.vy2.p);
        std::swap(c.ie.p, c.ie2.p);
        for (int q = 1; q <= ns; q++) c.sim_off_h[q] += pos[seg0[q]];
        c.n += pos[nseg];
        c.t.built = false;
        sync_bodies_to_host(c);
        return;
    }
    hipLaunchKernelGGL(k_shed_write, dim3(nblocks(nseg)), dim3(kBlock), 0, c.stream, c.segs.p, nseg, c.iscratch0.p,
                       c.iscratch1.p, c.n, c.x.p, c.y.p, c.g.p, c.vx.p, c.vy.p, c.ie.p);
    int32_t added = 0;
    VVHD_HIP(hipMemcpyAsync(&added, c.iscratch1.p + nseg, 4, hipMemcpyDeviceToHost, c.stream));
    VVHD_HIP(hipStreamSynchronize(c.stream));
    c.n += added;
    c.t.built = false;
    sync_bodies_to_host(c);
}

void calc_forces(Ctx& c) {
    const int nb = (int)c.bodies_h.size();
    if (!nb) return;
    const bool batch = c.nsim > 1;
    int32_t* body_sim = nullptr;
    if (batch) {
        sim_sync(c);
        std::vector<int32_t> bs(nb);
        for (int q = 0; q < c.nsim; q++)
            for (int b = c.sim_body_h[q]; b < c.sim_body_h[q + 1]; b++) bs[b] = q;
        c.iscratch3.grow(nb);
        body_sim = c.iscratch3.p;
        VVHD_HIP(hipMemcpyAsync(body_sim, bs.data(), (size_t)nb * 4, hipMemcpyHostToDevice, c.stream));
    }
    hipLaunchKernelGGL(k_calc_forces, dim3(nblocks(nb, 64)), dim3(64), 0, c.stream, c.bodies_d.p, nb, c.segs.p,
                       c.fric.p, c.re, c.dt, c.pr, batch ? (const double*)c.sim_re.p : (const double*)nullptr,
                       (const int32_t*)body_sim);
    sync_bodies_to_host(c);
}

void zero_forces(Ctx& c) {
    const int nb = (int)c.bodies_h.size();
    if (!nb) return;
    hipLaunchKernelGGL(k_zero_forces, dim3(nblocks(nb, 64)), dim3(64), 0, c.stream, c.bodies_d.p, nb, c.segs.p,
                       c.fric.p);
    heat_layer_reset(c);
    sync_bodies_to_host(c);
}

void calc_circulation(Ctx& c, double infx, double infy) {
    const int nb = (int)c.bodies_h.size();
    const int nseg = (int)c.segs_h.size();
    VVHD_CHECK(nb > 0, VVHD_ESTATE, "calc_circulation: no bodies");
    hipStream_t s = c.stream;
    // one system per simulation, all of the same order
    sim_sync(c);
    const int nsys = c.nsim;
    std::vector<int32_t> base(nb), body_sys(nb);
    int n = -1, max_segs = 0;
    for (int q = 0; q < nsys; q++) {
        const int b0 = c.sim_body_h[q], b1 = c.sim_body_h[q + 1];
        VVHD_CHECK(b1 > b0, VVHD_EUNSUPPORTED, "calc_circulation: a simulation without bodies in the batch");
        int acc = 0, segs_q = 0;
        for (int b = b0; b < b1; b++) {
            base[b] = acc;
            body_sys[b] = q;
            acc += c.bodies_h[b].nseg + 9;
            segs_q += c.bodies_h[b].nseg;
        }
        VVHD_CHECK(n < 0 || acc == n, VVHD_EUNSUPPORTED,
                   "calc_circulation: the simulations of a batch must have systems of one order");
        n = acc;
        max_segs = std::max(max_segs, segs_q);
    }
    c.slae_base.grow(2 * (size_t)nb + nsys + 1);
    int32_t* d_base = c.slae_base.p;
    int32_t* d_body_sys = c.slae_base.p + nb;
    int32_t* d_sys_body = c.slae_base.p + 2 * nb;
    VVHD_HIP(hipMemcpyAsync(d_base, base.data(), (size_t)nb * 4, hipMemcpyHostToDevice, s));
    VVHD_HIP(hipMemcpyAsync(d_body_sys, body_sys.data(), (size_t)nb * 4, hipMemcpyHostToDevice, s));
    VVHD_HIP(hipMemcpyAsync(d_sys_body, c.sim_body_h.data(), (size_t)(nsys + 1) * 4, hipMemcpyHostToDevice, s));
    const size_t nn = (size_t)n * n, nall = nn * nsys;
    c.slae_A.grow(nall);
    c.slae_b.grow((size_t)n * nsys + 1);
    c.slae_x.grow((size_t)n * nsys + nsys + 1);
    VVHD_HIP(hipMemsetAsync(c.slae_A.p, 0, nall * 8, s));
    VVHD_HIP(hipMemsetAsync(c.slae_b.p, 0, ((size_t)n * nsys + 1) * 8, s));
    SlaeMap M{c.bodies_d.p, d_base, nb, n, d_sys_body, c.seg_root.p, nsys, max_segs};
    double* gsum = c.slae_x.p + (size_t)n * nsys;   // per system, after the solutions
    hipLaunchKernelGGL(k_slip_cells, dim3(nblocks((int64_t)nseg * max_segs)), dim3(kBlock), 0, s, M, c.segs.p, nseg,
                       c.slae_A.p);
    hipLaunchKernelGGL(k_slip_rhs, dim3(nseg), dim3(kWave), 0, s, M, tree_dev(c.t), c.segs.p, nseg, c.x.p, c.y.p,
                       c.g.p, c.ie.p, infx, infy, c.slae_A.p, c.slae_b.p);
    hipLaunchKernelGGL(k_gsum, dim3(nsys), dim3(1024), 0, s, c.g.p, (const int64_t*)c.sim_off.p, gsum);
    hipLaunchKernelGGL(k_body_rows, dim3(nblocks(nsys, 64)), dim3(64), 0, s, M, c.segs.p, (const double*)gsum,
                       c.inf_g, c.dt, c.slae_A.p, c.slae_b.p);
    c.slae_filled = n;
    c.slae_nsys = nsys;
    bool refactor = !(c.time > 0) || c.slae_n != n || c.slae_nsys_factored != nsys;
    c.slae_info.grow((size_t)nsys + 2);
    if (!refactor) {
        VVHD_HIP(hipMemsetAsync(c.slae_info.p + nsys, 0, 4, s));
        hipLaunchKernelGGL(k_differ, dim3(nblocks((int64_t)nall)), dim3(kBlock), 0, s, c.slae_A.p, c.slae_Acached.p,
                           nall, c.slae_info.p + nsys);
        int32_t differ = 0;
        VVHD_HIP(hipMemcpyAsync(&differ, c.slae_info.p + nsys, 4, hipMemcpyDeviceToHost, s));
        VVHD_HIP(hipStreamSynchronize(s));
        refactor = differ != 0;
    }
    if (!c.rocblas) {
        rocblas_handle h = nullptr;
        VVHD_CHECK(rocblas_create_handle(&h) == rocblas_status_success, VVHD_EHIP, "rocblas_create_handle failed");
        c.rocblas = h;
    }
    rocblas_handle h = (rocblas_handle)c.rocblas;
    VVHD_CHECK(rocblas_set_stream(h, s) == rocblas_status_success, VVHD_EHIP, "rocblas_set_stream failed");
    if (refactor) {
        c.slae_n = -1;
        c.slae_Acached.grow(nall);
        c.slae_LU.grow(nall);
        c.slae_ipiv.grow((size_t)n * nsys + 1);
        VVHD_HIP(hipMemcpyAsync(c.slae_Acached.p, c.slae_A.p, nall * 8, hipMemcpyDeviceToDevice, s));
        VVHD_HIP(hipMemcpyAsync(c.slae_LU.p, c.slae_A.p, nall * 8, hipMemcpyDeviceToDevice, s));
        VVHD_CHECK(rocsolver_dgetrf_strided_batched(h, n, n, c.slae_LU.p, n, (rocblas_stride)nn, c.slae_ipiv.p,
                                                    (rocblas_stride)n, c.slae_info.p, nsys) == rocblas_status_success,
                   VVHD_EHIP, "rocsolver_dgetrf_strided_batched failed");
        std::vector<int32_t> info(nsys);
        VVHD_HIP(hipMemcpyAsync(info.data(), c.slae_info.p, (size_t)nsys * 4, hipMemcpyDeviceToHost, s));
        VVHD_HIP(hipStreamSynchronize(s));
        for (int q = 0; q < nsys; q++)
            if (info[q] != 0)
                throw Error(VVHD_ESINGULAR, "calc_circulation: singular matrix of simulation " + std::to_string(q) +
                                                ", column " + std::to_string(info[q]));
        c.slae_n = n;
        c.slae_nsys_factored = nsys;
    }
    VVHD_HIP(hipMemcpyAsync(c.slae_x.p, c.slae_b.p, (size_t)n * nsys * 8, hipMemcpyDeviceToDevice, s));
    VVHD_CHECK(rocsolver_dgetrs_strided_batched(h, rocblas_operation_none, n, 1, c.slae_LU.p, n, (rocblas_stride)nn,
                                                c.slae_ipiv.p, (rocblas_stride)n, c.slae_x.p, n, (rocblas_stride)n,
                                                nsys) == rocblas_status_success,
               VVHD_EHIP, "rocsolver_dgetrs_strided_batched failed");
    hipLaunchKernelGGL(k_writeback, dim3(nblocks(nb, 64)), dim3(64), 0, s, M, c.bodies_d.p, c.segs.p, c.slae_x.p,
                       (const int32_t*)d_body_sys);
    sync_bodies_to_host(c);
}

}  // namespace vvhd

using namespace vvhd;

extern "C" {

int vvhd_move_and_clean(vvhd_ctx* ctx, int remove, int64_t* cleaned) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    Ctx& c = ctx->c;
    VVHD_HIP(hipSetDevice(c.device));
    VVHD_HIP(hipEventRecord(c.ev[8], c.stream));
    move_and_clean(c, remove, cleaned);   // records ev[9] when its device work is queued
    c.phase_recorded[4] = true;
    API_END(ctx)
}

int vvhd_vortex_shed(vvhd_ctx* ctx) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    VVHD_HIP(hipSetDevice(ctx->c.device));
    vortex_shed(ctx->c);
    API_END(ctx)
}

int vvhd_calc_circulation(vvhd_ctx* ctx, double inf_x, double inf_y) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    Ctx& c = ctx->c;
    VVHD_HIP(hipSetDevice(c.device));
    VVHD_CHECK(c.t.built, VVHD_ESTATE, "calc_circulation: tree is not built");
    VVHD_HIP(hipEventRecord(c.ev[10], c.stream));
    calc_circulation(c, inf_x, inf_y);
    VVHD_HIP(hipEventRecord(c.ev[11], c.stream));
    c.phase_recorded[5] = true;
    API_END(ctx)
}

int vvhd_slae_system(vvhd_ctx* ctx, int32_t* n, double* A, double* rhs) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    Ctx& c = ctx->c;
    VVHD_CHECK(n, VVHD_EINVAL, "slae_system: null order");
    VVHD_CHECK(c.slae_filled > 0, VVHD_ESTATE, "slae_system: no system has been filled");
    const int m = c.slae_filled;
    *n = m;
    VVHD_HIP(hipSetDevice(c.device));
    if (A) {
        std::vector<double> cm((size_t)m * m);
        VVHD_HIP(hipMemcpyAsync(cm.data(), c.slae_A.p, cm.size() * 8, hipMemcpyDeviceToHost, c.stream));
        VVHD_HIP(hipStreamSynchronize(c.stream));
        for (int r = 0; r < m; r++)
            for (int q = 0; q < m; q++) A[(size_t)r * m + q] = cm[(size_t)q * m + r];
    }
    if (rhs) {
        VVHD_HIP(hipMemcpyAsync(rhs, c.slae_b.p, (size_t)m * 8, hipMemcpyDeviceToHost, c.stream));
        VVHD_HIP(hipStreamSynchronize(c.stream));
    }
    API_END(ctx)
}

int vvhd_calc_forces(vvhd_ctx* ctx) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    VVHD_HIP(hipSetDevice(ctx->c.device));
    calc_forces(ctx->c);
    API_END(ctx)
}

int vvhd_zero_forces(vvhd_ctx* ctx) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    VVHD_HIP(hipSetDevice(ctx->c.device));
    zero_forces(ctx->c);
    API_END(ctx)
}

int vvhd_zero_profiles(vvhd_ctx* ctx) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    Ctx& c = ctx->c;
    VVHD_HIP(hipSetDevice(c.device));
    const int nseg = (int)c.segs_h.size();
    if (nseg) {
        hipLaunchKernelGGL(k_zero_profiles, dim3(nblocks(nseg)), dim3(kBlock), 0, c.stream, c.segs.p, nseg);
        sync_bodies_to_host(c);
    }
    API_END(ctx)
}

}  // extern "C"
