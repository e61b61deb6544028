// diffusive.hip -- MDiffusiveFast::process_vort_list (MDiffusiveFast.cpp:
// 8-48, vortex_influence :93-105, segment_influence :107-123).
//
// One wave per leaf, lane = (target k of 16, slot s of 4).  For target i the
// vortex terms run over same-sign vortices j of the near leaves with
// |r_i - r_j| * (1/eps_i) <= 8 (the reference's exp-argument cutoff), so near
// leaves whose box lies farther than 8 eps from the whole target leaf are
// skipped without loss.  Vortex sums (S1, S2) are FP32 with Kahan
// compensation in leaf-local coordinates; body-segment terms (S0, S3 and the
// segment friction) are FP64.  The clamp and the velocity update are FP64.

#include "common.hpp"

#include <algorithm>
#include <cmath>

#pragma clang fp contract(off)

namespace vvhd {
namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr double kPi = 3.14159265358979323846;
constexpr double k2Pi = 2. * kPi;
constexpr double kExpArgLimit = -8.;   // MDiffusiveFast.cpp:6

__device__ inline double sgn(double x) { return (x > 0) ? 1 : ((x < 0) ? -1 : 0); }

struct Kahan {
    float s = 0.f, c = 0.f;
    __device__ inline void add(float v) {
        const float yv = v - c;
        const float t = s + yv;
        c = (t - s) - yv;
        s = t;
    }
    __device__ inline double value() const { return (double)s - (double)c; }
};

__device__ inline double slot_sum(double v) {
    v += __shfl_xor(v, 16, kWave);
    v += __shfl_xor(v, 32, kWave);
    return v;
}

__global__ __launch_bounds__(kBlock) void k_diffusive(TreeDev T, const double* __restrict__ x,
                                                      const double* __restrict__ y, const double* __restrict__ g,
                                                      const double* __restrict__ ie, const vvhd_seg* segs, double* vx,
                                                      double* vy, double* fric, double re) {
    const int w = threadIdx.x / kWave;
    const int l = xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock + w;
    if (l >= T.nleaf) return;
    const int lane = lane_id();
    const int k = lane & 15, slot = lane >> 4;
    const int L = T.leaf_node[l];
    const int v0 = T.v0[L], v1 = T.v1[L];
    if (v0 == v1) return;
    const double cx = T.nx[L], cy = T.ny[L];
    const int64_t e0 = T.near_off[l], e1 = T.near_off[l + 1];

    for (int base = v0; base < v1; base += 16) {
        const int i = base + k;
        const bool valid = i < v1;
        const double gi = valid ? g[i] : 0;
        const bool active = valid && gi != 0;
        const double iei = active ? ie[i] : 0;
        const double xi = valid ? x[i] : 0, yi = valid ? y[i] : 0;
        const float tx = (float)(xi - cx), ty = (float)(yi - cy);
        const float ief = (float)iei;
        const double si = sgn(gi);
        // cutoff radius of the chunk: 8 eps of its targets
        double R = active ? (-kExpArgLimit) / iei : 0.0;
        R = wave_max(R);
        const double R2 = R * R * (1.0 + 1e-6);
        Kahan s2x, s2y, s1;
        double s3x = 0, s3y = 0, s0 = 0;
        for (int64_t eb = e0; eb < e1; eb += kWave) {
            const int64_t e = eb + lane;
            int M = -1;
            bool q = false;
            if (e < e1) {
                M = T.near_list[e];
                q = box_dist2(T, L, M) <= R2;
            }
            uint64_t mask = __ballot(q);
            while (mask) {
                const int b = (int)__builtin_ctzll(mask);
                mask &= mask - 1;
                const int Mb = __shfl(M, b, kWave);
                const int j0 = T.v0[Mb], j1 = T.v1[Mb];
                for (int j = j0 + slot; j < j1; j += 4) {
                    const double gj = g[j];
                    if (!active || gj == 0 || sgn(gj) != si) continue;
                    const float dx = tx - (float)(x[j] - cx);
                    const float dy = ty - (float)(y[j] - cy);
                    if (fabsf(dx) + fabsf(dy) < 1e-10f) continue;
                    const float drabs = sqrtf(__builtin_fmaf(dx, dx, dy * dy));
                    const float exparg = -drabs * ief;
                    if (exparg < (float)kExpArgLimit) continue;
                    const float i1 = (float)gj * __expf(exparg);
                    const float f = i1 / drabs;
                    s2x.add(dx * f);
                    s2y.add(dy * f);
                    s1.add(i1);
                }
                // body segments of that leaf (FP64)
                const int m = T.leaf_rank[Mb];
                for (int64_t qq = T.seg_off[m] + slot; qq < T.seg_off[m + 1]; qq += 4) {
                    if (!active) continue;
                    const int sid = T.seg_list[qq];
                    const vvhd_seg sg = segs[sid];
                    const double drx = xi - sg.x, dry = yi - sg.y;
                    const double drabs2 = drx * drx + dry * dry;
                    const double drabs = sqrt(drabs2);
                    const double exparg = -drabs * iei;
                    if (exparg < kExpArgLimit) continue;
                    const double expres = exp(exparg);
                    const double dSx = -sg.dly, dSy = sg.dlx;
                    s3x += dSx * expres;
                    s3y += dSy * expres;
                    s0 += (drabs * iei + 1) / drabs2 * (drx * dSx + dry * dSy) * expres;
                    atomicAdd(&fric[sid], iei * iei * gi * expres * sqrt(dSx * dSx + dSy * dSy));
                }
            }
        }
        double S2x = slot_sum(s2x.value()), S2y = slot_sum(s2y.value()), S1 = slot_sum(s1.value());
        double S3x = slot_sum(s3x), S3y = slot_sum(s3y), S0 = slot_sum(s0);
        if (slot == 0 && active) {
            if ((sgn(S1) != si) || (fabs(S1) < fabs(0.1 * gi))) S1 = 0.1 * gi;
            double ax = vx[i], ay = vy[i];
            const double c1 = iei / (re * S1);
            ax += S2x * c1;
            ay += S2y * c1;
            if (S0 > kPi) S0 = kPi;
            const double c2 = (iei * iei) / (re * (k2Pi - S0));
            ax += S3x * c2;
            ay += S3y * c2;
            vx[i] = ax;
            vy[i] = ay;
        }
    }
}

}  // namespace

void diffusive(Ctx& c) {
    const TreeDev T = tree_dev(c.t);
    if (T.nleaf == 0) return;
    const int grid = (T.nleaf + kWavesPerBlock - 1) / kWavesPerBlock;
    hipLaunchKernelGGL(k_diffusive, dim3(grid), dim3(kBlock), 0, c.stream, T, c.x.p, c.y.p, c.g.p, c.ie.p, c.segs.p,
                       c.vx.p, c.vy.p, c.fric.p, c.re);
    VVHD_HIP(hipGetLastError());
}

}  // namespace vvhd
