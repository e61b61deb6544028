// fields.hip -- point and grid evaluations over a built tree:
//  * MEpsilonFast::eps2h and MEpsilonFast::h2 (MEpsilonFast.cpp:66-107) for
//    points or for every vortex;
//  * the grid fields of XVelocity (XVelocity.cpp:24-73), XVorticity
//    (XVorticity.cpp:26-94), XStreamfunction (XStreamfunction.cpp:47-118,
//    160-205) and XPressure (XPressure.cpp:32-146).
// One wave per point or grid node: lanes run over the sources of the near
// leaves (and over the body segments or all vortices where the reference sums
// them), then a wave reduction.  FP64 throughout.  The grid maps are float,
// as the reference's XField::map.  The reference evaluators work on a copy of
// the Space with their own tree; here the caller (the evaluator classes in
// vvhd_modules.hpp) builds that tree in a context of its own first.

#include "api.hpp"
#include "body_geom.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

#pragma clang fp contract(off)

namespace vvhd {
namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr double kPi = 3.14159265358979323846;
constexpr double k1_2Pi = 1. / (2. * kPi);
constexpr double k1_4Pi = 1. / (4. * kPi);
constexpr double k1_Pi = 1. / kPi;

inline int nblocks(int64_t n, int b = kBlock) { return (int)std::max<int64_t>(1, (n + b - 1) / b); }
inline int wave_blocks(int64_t n) { return (int)std::max<int64_t>(1, (n + kWavesPerBlock - 1) / kWavesPerBlock); }

// stree::findNode (TSortedTree.cpp:284-303): the leaf node containing p.
__device__ inline int find_leaf_node(const TreeDev& T, double px, double py) {
    int node = 0;
    while (T.ch1[node] >= 0) {
        const int c = T.ch1[node];
        if (T.nh[node] < T.nw[node])
            node = (px < T.nx[node]) ? c : c + 1;
        else
            node = (py < T.ny[node]) ? c : c + 1;
    }
    return node;
}

// The two smallest values of a multiset (insertion as MEpsilonFast::eps2h).
__device__ inline void top2_insert(double d, double& r1, double& r2) {
    if (d < r1) {
        r2 = r1;
        r1 = d;
    } else if (d < r2) {
        r2 = d;
    }
}

__device__ inline void wave_top2(double& r1, double& r2) {
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const double a = __shfl_xor(r1, o, kWave), b = __shfl_xor(r2, o, kWave);
        top2_insert(a, r1, r2);
        top2_insert(b, r1, r2);
    }
}

__device__ inline double wave_min(double v) {
    for (int o = kWave / 2; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, kWave));
    return v;
}

// MEpsilonFast::eps2h(node, p): the second smallest non-zero squared distance
// from p to the vortices of the near leaves of node (the smallest when there
// is one, the lowest double when there is none), on the calling wave.
__device__ double wave_eps2h(const TreeDev& T, const double* x, const double* y, int l, double px, double py) {
    double r1 = INFINITY, r2 = INFINITY;
    for (int64_t e = T.near_off[l]; e < T.near_off[l + 1]; e++) {
        const int M = T.near_list[e];
        for (int j = T.v0[M] + lane_id(); j < T.v1[M]; j += kWave) {
            const double dx = px - x[j], dy = py - y[j];
            const double d2 = dx * dx + dy * dy;
            if (d2 != 0) top2_insert(d2, r1, r2);
        }
    }
    wave_top2(r1, r2);
    if (isfinite(r2)) return r2;
    if (isfinite(r1)) return r1;
    return -DBL_MAX;
}

// MEpsilonFast::h2(node, p): the smallest squared distance from p to the body
// segments (their centres) of the near leaves of node; +inf when none.
__device__ double wave_h2(const TreeDev& T, const vvhd_seg* segs, int l, double px, double py) {
    double r = INFINITY;
    if (T.seg_off) {
        for (int64_t e = T.near_off[l]; e < T.near_off[l + 1]; e++) {
            const int lm = T.leaf_rank[T.near_list[e]];
            for (int64_t k = T.seg_off[lm] + lane_id(); k < T.seg_off[lm + 1]; k += kWave) {
                const vvhd_seg& s = segs[T.seg_list[k]];
                const double dx = px - s.x, dy = py - s.y;
                r = fmin(r, dx * dx + dy * dy);
            }
        }
    }
    return wave_min(r);
}

__device__ inline bool in_any_body(const vvhd_body* bodies, int nbody, const vvhd_seg* segs, const double* geo,
                                   double px, double py) {
    for (int b = 0; b < nbody; b++)
        if (point_invalid(bodies[b], segs, geo + 5 * b, px, py) >= 0) return true;
    return false;
}

// MConvectiveFast::velocity(p) (MConvectiveFast.cpp:20-34, 139-151) on the
// calling wave: near Lamb sum, far per-sign monopoles, U.
__device__ void wave_velocity(const TreeDev& T, const double* x, const double* y, const double* g, const double* ie,
                              double px, double py, double inf_x, double inf_y, double& ox, double& oy) {
    const int l = T.leaf_rank[find_leaf_node(T, px, py)];
    double nx_ = 0, ny_ = 0, fx = 0, fy = 0;
    for (int64_t e = T.near_off[l]; e < T.near_off[l + 1]; e++) {
        const int M = T.near_list[e];
        for (int j = T.v0[M] + lane_id(); j < T.v1[M]; j += kWave) {
            const double gj = g[j];
            if (!gj) continue;
            const double dx = px - x[j], dy = py - y[j];
            const double e1 = 1. / ie[j];
            const double f = gj / (dx * dx + dy * dy + e1 * e1);
            nx_ += -dy * f;
            ny_ += dx * f;
        }
    }
    for (int64_t e = T.far_off[l] + lane_id(); e < T.far_off[l + 1]; e += kWave) {
        const int F = T.far_list[e];
        {
            const double dx = px - T.cmp_x[F], dy = py - T.cmp_y[F];
            const double e1 = 1. / (double)T.cmp_ie[F];
            const double f = T.cmp_g[F] / (dx * dx + dy * dy + e1 * e1);
            fx += -dy * f;
            fy += dx * f;
        }
        {
            const double dx = px - T.cmm_x[F], dy = py - T.cmm_y[F];
            const double e1 = 1. / (double)T.cmm_ie[F];
            const double f = T.cmm_g[F] / (dx * dx + dy * dy + e1 * e1);
            fx += -dy * f;
            fy += dx * f;
        }
    }
    nx_ = wave_sum(nx_);
    ny_ = wave_sum(ny_);
    fx = wave_sum(fx);
    fy = wave_sum(fy);
    ox = nx_ * k1_2Pi + fx * k1_2Pi + inf_x;
    oy = ny_ * k1_2Pi + fy * k1_2Pi + inf_y;
}

struct Grid {
    double xmin, ymin, dxdy;
    int xres, yres;
    __device__ void point(int64_t k, double& px, double& py) const {
        const int xi = (int)(k % xres), yj = (int)(k / xres);
        px = xmin + dxdy * (double)xi;
        py = ymin + dxdy * (double)yj;
    }
};

// Per point (pts) or per vortex (pts == nullptr): eps2h or h2.
template <bool kH2>
__global__ __launch_bounds__(kBlock) void k_eps2h_h2(TreeDev T, const double* x, const double* y,
                                                     const vvhd_seg* segs, const double* pts, const int32_t* leaf,
                                                     int64_t n, double* out) {
    const int64_t p = (int64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    if (p >= n) return;
    double px, py;
    int l;
    if (pts) {
        px = pts[2 * p];
        py = pts[2 * p + 1];
        l = leaf ? leaf[p] : T.leaf_rank[find_leaf_node(T, px, py)];
    } else {
        px = x[p];
        py = y[p];
        l = T.vleaf[p];
    }
    const double r = kH2 ? wave_h2(T, segs, l, px, py) : wave_eps2h(T, x, y, l, px, py);
    if (lane_id() == 0) out[p] = r;
}

// Per-vortex weights of the vorticity and streamfunction fields from eps2h
// over the vortex's own leaf: vorticity a = 1/(m^2 max(eps2h, (0.6 dl)^2)),
// b = a g (XVorticity.cpp:52-55); streamfunction a = m^2 max(eps2h / 4,
// (0.2 dl)^2) (XStreamfunction.cpp:84-87).
template <bool kStream>
__global__ __launch_bounds__(kBlock) void k_field_weights(TreeDev T, const double* x, const double* y,
                                                          const double* g, int64_t n, double m, double dl, double* a,
                                                          double* b) {
    const int64_t i = (int64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    if (i >= n) return;
    const double e = wave_eps2h(T, x, y, T.vleaf[i], x[i], y[i]);
    if (lane_id() != 0) return;
    if (kStream) {
        a[i] = (m * m) * fmax(e * 0.25, (0.2 * dl) * (0.2 * dl));
    } else {
        const double ai = 1. / ((m * m) * fmax(e, (0.6 * dl) * (0.6 * dl)));
        a[i] = ai;
        b[i] = ai * g[i];
    }
}

struct BodyArgs {
    const vvhd_seg* segs;
    int nseg;
    const vvhd_body* bodies;
    int nbody;
    const double* geo;
    const double* gtmp;   // running sum of gsum along each body (pressure)
};

// XVorticity::vorticity (XVorticity.cpp:74-94): Gaussian blobs of the near
// leaves' vortices (exponent cut at -6) over pi, plus the wall term
// 0.5 (1 - erf(h2 / (dl m)^2)) below 3; 0 inside a body.
__global__ __launch_bounds__(kBlock) void k_vorticity_grid(TreeDev T, const double* x, const double* y,
                                                           const double* a, const double* b, BodyArgs B, Grid G,
                                                           double m, double dl, float* map) {
    const int64_t k = (int64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    if (k >= (int64_t)G.xres * G.yres) return;
    double px, py;
    G.point(k, px, py);
    if (in_any_body(B.bodies, B.nbody, B.segs, B.geo, px, py)) {
        if (lane_id() == 0) map[k] = 0.f;
        return;
    }
    const int l = T.leaf_rank[find_leaf_node(T, px, py)];
    double res = 0;
    for (int64_t e = T.near_off[l]; e < T.near_off[l + 1]; e++) {
        const int M = T.near_list[e];
        for (int j = T.v0[M] + lane_id(); j < T.v1[M]; j += kWave) {
            const double dx = px - x[j], dy = py - y[j];
            const double exparg = -(dx * dx + dy * dy) * a[j];
            res += (exparg > -6) ? b[j] * exp(exparg) : 0;
        }
    }
    res = wave_sum(res);
    const double h2 = wave_h2(T, B.segs, l, px, py);
    if (lane_id() == 0) {
        res *= k1_Pi;
        const double erfarg = h2 / ((dl * m) * (dl * m));
        res += (erfarg < 3) ? 0.5 * (1 - erf(erfarg)) : 0;
        map[k] = (float)res;
    }
}

// XStreamfunction::streamfunction(node, p) (XStreamfunction.cpp:160-205):
// 4 pi psi from the segments' bound circulations (and, for a moving body,
// their surface speed as sources and vortices), the far nodes' monopoles and
// the near vortices (log kernel smoothed by rd2 or by the vortex's weight),
// 2 pi psi from the sources, plus p . rotl(U - U_ref).
__global__ __launch_bounds__(kBlock) void k_streamfunction_grid(TreeDev T, const double* x, const double* y,
                                                                const double* g, const double* a, BodyArgs B, Grid G,
                                                                double rd2, double wx, double wy, float* map) {
    const int64_t k = (int64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    if (k >= (int64_t)G.xres * G.yres) return;
    double px, py;
    G.point(k, px, py);
    double psi_g = 0, psi_q = 0;
    for (int b = 0; b < B.nbody; b++) {
        const vvhd_body& Bd = B.bodies[b];
        for (int s = Bd.first_seg + lane_id(); s < Bd.first_seg + Bd.nseg; s += kWave) {
            const vvhd_seg& sg = B.segs[s];
            const double dx = px - sg.x, dy = py - sg.y;
            psi_g += -sg.g * log(dx * dx + dy * dy + rd2);
        }
        const bool moving = fabs(Bd.speed_slae[0]) + fabs(Bd.speed_slae[1]) + fabs(Bd.speed_slae[2]) >= 1e-10;
        if (!moving) continue;
        const double ax = Bd.holder[0] + Bd.dpos[0], ay = Bd.holder[1] + Bd.dpos[1];
        for (int s = Bd.first_seg + lane_id(); s < Bd.first_seg + Bd.nseg; s += kWave) {
            const vvhd_seg& sg = B.segs[s];
            const double vsx = Bd.speed_slae[0] + Bd.speed_slae[2] * (-(sg.y - ay));
            const double vsy = Bd.speed_slae[1] + Bd.speed_slae[2] * (sg.x - ax);
            const double gs = -(vsx * sg.dlx + vsy * sg.dly);
            const double qs = -((-vsy) * sg.dlx + vsx * sg.dly);
            const double dx = px - sg.x, dy = py - sg.y;
            psi_g += -gs * log(dx * dx + dy * dy + rd2);
            const double v1x = Bd.cofm[0] - px, v1y = Bd.cofm[1] - py;
            const double v2x = sg.x - px, v2y = sg.y - py;
            psi_q += -qs * atan2((-v2y) * v1x + v2x * v1y, v2x * v1x + v2y * v1y);
        }
    }
    const int l = T.leaf_rank[find_leaf_node(T, px, py)];
    for (int64_t e = T.far_off[l] + lane_id(); e < T.far_off[l + 1]; e += kWave) {
        const int F = T.far_list[e];
        double dx = px - T.cmp_x[F], dy = py - T.cmp_y[F];
        psi_g += -T.cmp_g[F] * log(dx * dx + dy * dy + rd2);
        dx = px - T.cmm_x[F];
        dy = py - T.cmm_y[F];
        psi_g += -T.cmm_g[F] * log(dx * dx + dy * dy + rd2);
    }
    for (int64_t e = T.near_off[l]; e < T.near_off[l + 1]; e++) {
        const int M = T.near_list[e];
        for (int j = T.v0[M] + lane_id(); j < T.v1[M]; j += kWave) {
            const double dx = px - x[j], dy = py - y[j];
            psi_g += -g[j] * log(dx * dx + dy * dy + a[j]);
        }
    }
    psi_g = wave_sum(psi_g);
    psi_q = wave_sum(psi_q);
    if (lane_id() == 0) map[k] = (float)(psi_g * k1_4Pi + psi_q * k1_2Pi + (px * (-wy) + py * wx));
}

// XPressure::pressure (XPressure.cpp:109-146): body surface terms, the
// vortices' convective term over every vortex, and the Bernoulli terms of the
// local velocity; 0 inside a body.
__global__ __launch_bounds__(kBlock) void k_pressure_grid(TreeDev T, const double* x, const double* y,
                                                          const double* g, const double* ie, const double* vx,
                                                          const double* vy, int64_t n, BodyArgs B, Grid G, double dt,
                                                          double inf_x, double inf_y, int with_ref, double ref_x,
                                                          double ref_y, float* map) {
    const int64_t k = (int64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    if (k >= (int64_t)G.xres * G.yres) return;
    double px, py;
    G.point(k, px, py);
    if (in_any_body(B.bodies, B.nbody, B.segs, B.geo, px, py)) {
        if (lane_id() == 0) map[k] = 0.f;
        return;
    }
    double cp = 0;
    for (int b = 0; b < B.nbody; b++) {
        const vvhd_body& Bd = B.bodies[b];
        const double ax = Bd.holder[0] + Bd.dpos[0], ay = Bd.holder[1] + Bd.dpos[1];
        for (int s = Bd.first_seg + lane_id(); s < Bd.first_seg + Bd.nseg; s += kWave) {
            const vvhd_seg& sg = B.segs[s];
            const double vsx = Bd.speed_slae[0] + Bd.speed_slae[2] * (-(sg.y - ay));
            const double vsy = Bd.speed_slae[1] + Bd.speed_slae[2] * (sg.x - ax);
            const double gs = -(vsx * sg.dlx + vsy * sg.dly);
            const double qs = -((-vsy) * sg.dlx + vsx * sg.dly);
            double kx = px - sg.x, ky = py - sg.y;
            const double r2 = kx * kx + ky * ky;
            kx = kx / r2;
            ky = ky / r2;
            // (rotl(K) g + K q) . Vs
            cp += ((-ky) * gs + kx * qs) * vsx + (kx * gs + ky * qs) * vsy;
            // - (dl / dt . rotl(K)) gtmp
            cp -= ((sg.dlx / dt) * (-ky) + (sg.dly / dt) * kx) * B.gtmp[s];
        }
    }
    for (int64_t j = lane_id(); j < n; j += kWave) {
        double kx = px - x[j], ky = py - y[j];
        const double r2 = kx * kx + ky * ky;
        kx = kx / r2;
        ky = ky / r2;
        cp += (vx[j] * (-ky) + vy[j] * kx) * g[j];
    }
    cp = wave_sum(cp);
    double ux, uy;
    wave_velocity(T, x, y, g, ie, px, py, inf_x, inf_y, ux, uy);
    if (lane_id() == 0) {
        double c = k1_2Pi * cp + 0.5 * ((inf_x * inf_x + inf_y * inf_y) - (ux * ux + uy * uy));
        if (with_ref) c += 0.5 * ((ux - ref_x) * (ux - ref_x) + (uy - ref_y) * (uy - ref_y));
        map[k] = (float)c;
    }
}

__global__ __launch_bounds__(kBlock) void k_velocity_grid(TreeDev T, const double* x, const double* y,
                                                          const double* g, const double* ie, Grid G, double inf_x,
                                                          double inf_y, double ref_x, double ref_y, int comp,
                                                          float* map) {
    const int64_t k = (int64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    if (k >= (int64_t)G.xres * G.yres) return;
    double px, py, ux, uy;
    G.point(k, px, py);
    wave_velocity(T, x, y, g, ie, px, py, inf_x, inf_y, ux, uy);
    if (lane_id() == 0) map[k] = (float)(comp == 0 ? ux - ref_x : uy - ref_y);
}

// Running sum of gsum along each body's segments (XPressure.cpp:128-133).
__global__ void k_gsum_running(const vvhd_body* bodies, int nbody, const vvhd_seg* segs, double* gtmp) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbody) return;
    double t = 0;
    for (int s = bodies[b].first_seg; s < bodies[b].first_seg + bodies[b].nseg; s++) {
        t += segs[s].gsum;
        gtmp[s] = t;
    }
}

}  // namespace

void eps2h_h2(Ctx& c, bool h2, const double* xy, const int32_t* leaf, int64_t n, double* out) {
    const TreeDev T = tree_dev(c.t);
    const int64_t m = xy ? n : c.n;
    if (m == 0) return;
    c.tmp.grow((size_t)m * (xy ? 28 : 8));
    double* dout = (double*)c.tmp.p;
    double* dpts = xy ? dout + m : nullptr;
    int32_t* dleaf = (xy && leaf) ? (int32_t*)(dout + 3 * m) : nullptr;
    if (xy) VVHD_HIP(hipMemcpyAsync(dpts, xy, (size_t)m * 16, hipMemcpyHostToDevice, c.stream));
    if (dleaf) {
        for (int64_t k = 0; k < m; k++)
            VVHD_CHECK(leaf[k] >= 0 && leaf[k] < c.t.nleaves_h, VVHD_EINVAL, "eps2h/h2: leaf index out of range");
        VVHD_HIP(hipMemcpyAsync(dleaf, leaf, (size_t)m * 4, hipMemcpyHostToDevice, c.stream));
    }
    if (h2)
        hipLaunchKernelGGL(k_eps2h_h2<true>, dim3(wave_blocks(m)), dim3(kBlock), 0, c.stream, T, c.x.p, c.y.p,
                           c.segs.p, dpts, (const int32_t*)dleaf, m, dout);
    else
        hipLaunchKernelGGL(k_eps2h_h2<false>, dim3(wave_blocks(m)), dim3(kBlock), 0, c.stream, T, c.x.p, c.y.p,
                           c.segs.p, dpts, (const int32_t*)dleaf, m, dout);
    VVHD_HIP(hipMemcpyAsync(out, dout, (size_t)m * 8, hipMemcpyDeviceToHost, c.stream));
    VVHD_HIP(hipStreamSynchronize(c.stream));
}

void grid_field(Ctx& c, int kind, const vvhd_grid& gr, double m, double dl, double inf_x, double inf_y, int with_ref,
                double ref_x, double ref_y, float* map) {
    const TreeDev T = tree_dev(c.t);
    const int64_t npts = (int64_t)gr.xres * gr.yres;
    const Grid G{gr.xmin, gr.ymin, gr.dxdy, gr.xres, gr.yres};
    const int nb = (int)c.bodies_h.size();
    c.fld_a.grow((size_t)std::max<int64_t>(c.n, 1));
    c.fld_b.grow((size_t)std::max<int64_t>(c.n, 1));
    c.fld_map.grow((size_t)std::max<int64_t>(npts, 1));
    c.fld_gtmp.grow((size_t)std::max<int>((int)c.segs_h.size(), 1));
    const BodyArgs B{c.segs.p, (int)c.segs_h.size(), c.bodies_d.p, nb, c.body_geo.p, c.fld_gtmp.p};
    hipStream_t s = c.stream;
    const int grid = wave_blocks(npts);
    switch (kind) {
        case VVHD_FIELD_VX:
        case VVHD_FIELD_VY:
            hipLaunchKernelGGL(k_velocity_grid, dim3(grid), dim3(kBlock), 0, s, T, c.x.p, c.y.p, c.g.p, c.ie.p, G,
                               inf_x, inf_y, ref_x, ref_y, kind == VVHD_FIELD_VX ? 0 : 1, c.fld_map.p);
            break;
        case VVHD_FIELD_VORTICITY:
            if (c.n)
                hipLaunchKernelGGL(k_field_weights<false>, dim3(wave_blocks(c.n)), dim3(kBlock), 0, s, T, c.x.p,
                                   c.y.p, c.g.p, c.n, m, dl, c.fld_a.p, c.fld_b.p);
            hipLaunchKernelGGL(k_vorticity_grid, dim3(grid), dim3(kBlock), 0, s, T, c.x.p, c.y.p, c.fld_a.p,
                               c.fld_b.p, B, G, m, dl, c.fld_map.p);
            break;
        case VVHD_FIELD_STREAMFUNCTION:
            if (c.n)
                hipLaunchKernelGGL(k_field_weights<true>, dim3(wave_blocks(c.n)), dim3(kBlock), 0, s, T, c.x.p,
                                   c.y.p, c.g.p, c.n, m, dl, c.fld_a.p, c.fld_b.p);
            hipLaunchKernelGGL(k_streamfunction_grid, dim3(grid), dim3(kBlock), 0, s, T, c.x.p, c.y.p, c.g.p,
                               c.fld_a.p, B, G, (0.2 * dl) * (0.2 * dl), inf_x - ref_x, inf_y - ref_y, c.fld_map.p);
            break;
        case VVHD_FIELD_PRESSURE:
            if (nb)
                hipLaunchKernelGGL(k_gsum_running, dim3(nblocks(nb, 64)), dim3(64), 0, s, c.bodies_d.p, nb, c.segs.p,
                                   c.fld_gtmp.p);
            hipLaunchKernelGGL(k_pressure_grid, dim3(grid), dim3(kBlock), 0, s, T, c.x.p, c.y.p, c.g.p, c.ie.p,
                               c.vx.p, c.vy.p, c.n, B, G, c.dt, inf_x, inf_y, with_ref, ref_x, ref_y, c.fld_map.p);
            break;
        default:
            VVHD_CHECK(false, VVHD_EINVAL, "field: unknown kind");
    }
    VVHD_HIP(hipGetLastError());
    if (npts) VVHD_HIP(hipMemcpyAsync(map, c.fld_map.p, (size_t)npts * 4, hipMemcpyDeviceToHost, s));
    VVHD_HIP(hipStreamSynchronize(s));
}

}  // namespace vvhd

using namespace vvhd;

extern "C" {

int vvhd_eps2h(vvhd_ctx* ctx, const double* xy, const int32_t* leaf, int64_t n, double* out) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    REQUIRE_ONE_SIM(ctx->c, "eps2h");
    Ctx& c = ctx->c;
    VVHD_CHECK(c.t.built, VVHD_ESTATE, "eps2h: tree is not built");
    VVHD_CHECK(n >= 0 && out && (xy || n == c.n), VVHD_EINVAL, "eps2h: bad arguments");
    VVHD_HIP(hipSetDevice(c.device));
    eps2h_h2(c, false, xy, leaf, n, out);
    API_END(ctx)
}

int vvhd_h2(vvhd_ctx* ctx, const double* xy, const int32_t* leaf, int64_t n, double* out) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    REQUIRE_ONE_SIM(ctx->c, "h2");
    Ctx& c = ctx->c;
    VVHD_CHECK(c.t.built, VVHD_ESTATE, "h2: tree is not built");
    VVHD_CHECK(n >= 0 && out && (xy || n == c.n), VVHD_EINVAL, "h2: bad arguments");
    VVHD_HIP(hipSetDevice(c.device));
    eps2h_h2(c, true, xy, leaf, n, out);
    API_END(ctx)
}

int vvhd_field(vvhd_ctx* ctx, int kind, const vvhd_grid* grid, double eps_mult, double dl, double inf_x,
               double inf_y, int with_ref, double ref_x, double ref_y, float* map) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    REQUIRE_ONE_SIM(ctx->c, "field");
    Ctx& c = ctx->c;
    VVHD_CHECK(grid && map, VVHD_EINVAL, "field: null grid or map");
    VVHD_CHECK(grid->xres > 1 && grid->yres > 1, VVHD_EINVAL, "XField(): xres and yres must be > 1");
    VVHD_CHECK(std::isfinite(grid->dxdy) && grid->dxdy > 0, VVHD_EINVAL, "XField(): dxdy must be finite and positive");
    VVHD_CHECK(kind == VVHD_FIELD_VX || kind == VVHD_FIELD_VY || eps_mult > 0 || kind == VVHD_FIELD_PRESSURE,
               VVHD_EINVAL, "field: eps_mult must be positive");
    VVHD_CHECK(c.t.built, VVHD_ESTATE, "field: tree is not built");
    VVHD_HIP(hipSetDevice(c.device));
    grid_field(c, kind, *grid, eps_mult, dl, inf_x, inf_y, with_ref, ref_x, ref_y, map);
    API_END(ctx)
}

}  // extern "C"
