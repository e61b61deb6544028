// diffusive.hip -- MDiffusiveFast::process_vort_list (MDiffusiveFast.cpp:
// 8-48, vortex_influence :93-105, segment_influence :107-123).
//
// One wave per leaf and chunk of up to 16 targets.  For target i the vortex
// terms run over same-sign vortices j of the near leaves with
// |r_i - r_j| * (1/eps_i) <= 8 (the reference's exp-argument cutoff), so near
// leaves whose box lies farther than 8 eps from the whole chunk are skipped
// without loss; the sources of the others are staged in LDS and shared by
// all lanes.  Vortex sums (S1, S2) are FP32, block-compensated (BlockSum), in
// leaf-local coordinates; body-segment terms (S0, S3) are FP64.  The clamp
// and the velocity update are FP64.
//
// Segment friction (segment_influence's fric term, :120-122) runs as its own
// pass so that every segment adds its contributions in the reference's
// serial order: target leaves in DFS order, their targets in list order.
// The leaves that have a segment's leaf in their near list are collected
// per segment leaf (a transposed near list restricted to leaves holding
// segments) and sorted; one wave per segment evaluates 64 terms at a time
// and lane 0 adds them in order.

#include "common.hpp"

#include <cstring>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_segmented_radix_sort.hpp>

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

// Block-compensated sum: terms go by FMA (or an add) into the low-order
// part q, which a compensated (Fast2Sum) step folds into the running sum s
// every few terms; q keeps that step's rounding error, so every rounding is
// relative to a few terms, never to the running sum.
struct BlockSum {
    float s = 0.f, q = 0.f;
    __device__ inline void fold() {
        const float t = s + q;
        q = q - (t - s);
        s = t;
    }
    __device__ inline double value() const { return (double)s + (double)q; }
};

// Sum over the slots of target t (lanes t, t + ntc, ..., in slot order).
__device__ inline double slot_sum(double v, int t, int ntc, int nslot) {
    double s = 0;
    for (int q = 0; q < nslot; q++) s += __shfl(v, t + ntc * q, kWave);
    return s;
}

constexpr int kStage = 512;   // staged sources per wave

// One wave per leaf and chunk of up to 16 targets.  The sources of the near
// leaves within the cutoff are staged in LDS (FP32, relative to the leaf
// centre; circulation as FP32), then lane (target t = l % ntc, slot
// s = l / ntc) runs over every nslot-th staged source: all 64 lanes work
// whatever the leaf's occupancy.  Per pair: same sign, |dr|_1 >= 1e-10,
// |dr| / eps_i <= 8; |dr| and 1/|dr| from one reciprocal square root.
// kSeg: the FP64 segment terms are compiled in (their library exp and sqrt
// take registers that would cost the vortex-only launches occupancy).
// dvx/dvy non-null (no segments only): the increment is stored there for
// k_add_diffusive instead of being added to vx/vy.
template <bool kSeg>
__global__ __launch_bounds__(kBlock) void k_diffusive(TreeDev T, const double* __restrict__ x,
                                                      const double* __restrict__ y, const double* __restrict__ g,
                                                      const double* __restrict__ ie, const vvhd_seg* segs, int nseg,
                                                      double* vx, double* vy, double re_one, const double* sim_re,
                                                      const int32_t* leaf_sim, double* dvx, double* dvy) {
    __shared__ float s_x[kWavesPerBlock][kStage], s_y[kWavesPerBlock][kStage], s_g[kWavesPerBlock][kStage];
    __shared__ int s_idx[kWavesPerBlock][kStage];   // vortex index of each slot of the chunk being staged
    const int w = threadIdx.x / kWave;
    const int l = xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock + w;
    if (l >= T.nleaf) return;
    const int lane = lane_id();
    const int L = T.leaf_node[l];
    const int v0 = T.v0[L], v1 = T.v1[L];
    if (v0 == v1) return;
    const double cx = T.nx[L], cy = T.ny[L];
    const int64_t e0 = T.near_off[l], e1 = T.near_off[l + 1];
    const double re = sim_re ? sim_re[leaf_sim[l]] : re_one;   // Reynolds number of the leaf's simulation
    float* sx = s_x[w];
    float* sy = s_y[w];
    float* sg = s_g[w];
    int* sidx = s_idx[w];

    for (int base = v0; base < v1; base += 16) {
        const int ntc = min(16, v1 - base);
        const int nslot = kWave / ntc;
        const int t = lane % ntc, slot = lane / ntc;
        const int i = base + t;
        const double gi = g[i];
        const bool active = gi != 0 && slot < nslot;
        const double iei = gi != 0 ? ie[i] : 0;
        const double xi = x[i], yi = y[i];
        const float tx = (float)(xi - cx), ty = (float)(yi - cy);
        const float ief = (float)iei;
        const bool pos = gi > 0;
        // cutoff radius of the chunk: 8 eps of its targets
        double R = gi != 0 ? (-kExpArgLimit) / iei : 0.0;
        R = wave_max(R);
        const double R2 = R * R * (1.0 + 1e-6);
        BlockSum s2x, s2y, s1;
        // staged sources: positive ones from the front of the buffer, negative
        // ones from the back, each in staging order; a target runs over its
        // own sign's part only
        int npos = 0, nneg = 0;
        auto compute = [&]() {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (active) {
                const int k0 = pos ? 0 : kStage - nneg, k1 = pos ? npos : kStage;
                int nk = 0;
                for (int k = k0 + slot; k < k1; k += nslot) {
                    // the partial sums fold every eight sources of this lane
                    if ((++nk & 7) == 0) {
                        s2x.fold();
                        s2y.fold();
                        s1.fold();
                    }
                    const float gj = sg[k];
                    const float dx = tx - sx[k];
                    const float dy = ty - sy[k];
                    if (fabsf(dx) + fabsf(dy) < 1e-10f) continue;
                    const float d2 = __builtin_fmaf(dx, dx, dy * dy);
                    const float rinv = __builtin_amdgcn_rsqf(d2);
                    const float exparg = -(d2 * rinv) * ief;
                    if (exparg < (float)kExpArgLimit) continue;
                    const float i1 = gj * __expf(exparg);
                    const float f = i1 * rinv;
                    s2x.q = __builtin_fmaf(dx, f, s2x.q);
                    s2y.q = __builtin_fmaf(dy, f, s2y.q);
                    s1.q += i1;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        };
        int fill = 0;
        for (int64_t eb = e0; eb < e1; eb += kWave) {
            const int64_t e = eb + lane;
            int cnt = 0, j0 = 0;
            if (e < e1) {
                const int M = T.near_list[e];
                if (box_dist2p(T, L, M) <= R2) {
                    j0 = T.v0[M];
                    cnt = T.v1[M] - j0;
                }
            }
            const int incl = wave_incl_scan(cnt);
            const int off = incl - cnt;
            const int tot = __shfl(incl, kWave - 1, kWave);
            if (tot == 0) continue;
            for (int done = 0; done < tot;) {
                const int take = min(tot - done, kStage - fill);
                // each lane lists the vortices of its leaf that fall in this chunk
                for (int k = max(off, done); k < min(off + cnt, done + take); k++) sidx[k - done] = j0 + (k - off);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                for (int q0 = 0; q0 < take; q0 += kWave) {
                    const int q = q0 + lane;
                    float gv = 0.f, xv = 0.f, yv = 0.f;
                    if (q < take) {
                        const int j = sidx[q];
                        gv = (float)g[j];
                        xv = (float)(x[j] - cx);
                        yv = (float)(y[j] - cy);
                    }
                    const uint64_t pm = __ballot(gv > 0.f), nm = __ballot(gv < 0.f);
                    const uint64_t below = (1ull << lane) - 1;
                    int k = -1;
                    if (gv > 0.f) k = npos + __popcll(pm & below);
                    if (gv < 0.f) k = kStage - 1 - (nneg + __popcll(nm & below));
                    if (k >= 0) {
                        sx[k] = xv;
                        sy[k] = yv;
                        sg[k] = gv;
                    }
                    npos += __popcll(pm);
                    nneg += __popcll(nm);
                }
                fill = npos + nneg;
                done += take;
                if (fill == kStage) {
                    compute();
                    npos = nneg = fill = 0;
                }
                // the next chunk rewrites the index list
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        }
        if (fill) compute();
        // body segments of the near leaves within the cutoff (FP64)
        double s3x = 0, s3y = 0, s0 = 0;
        if (kSeg && nseg) {
            for (int64_t eb = e0; eb < e1; eb += kWave) {
                const int64_t e = eb + lane;
                int M = -1;
                bool q = false;
                if (e < e1) {
                    M = T.near_list[e];
                    const int m = T.leaf_rank[M];
                    q = T.seg_off[m + 1] > T.seg_off[m] && box_dist2p(T, L, M) <= R2;
                }
                uint64_t mask = __ballot(q);
                while (mask) {
                    const int b = (int)__builtin_ctzll(mask);
                    mask &= mask - 1;
                    const int m = T.leaf_rank[__shfl(M, b, kWave)];
                    for (int64_t qq = T.seg_off[m] + slot; qq < T.seg_off[m + 1]; qq += nslot) {
                        if (!active) continue;
                        const int sid = T.seg_list[qq];
                        const vvhd_seg sgm = segs[sid];
                        const double drx = xi - sgm.x, dry = yi - sgm.y;
                        const double drabs2 = drx * drx + dry * dry;
                        const double drabs = sqrt(drabs2);
                        const double exparg = -drabs * iei;
                        if (exparg < kExpArgLimit) continue;
                        const double expres = exp(exparg);
                        const double dSx = -sgm.dly, dSy = sgm.dlx;
                        s3x += dSx * expres;
                        s3y += dSy * expres;
                        s0 += (drabs * iei + 1) / drabs2 * (drx * dSx + dry * dSy) * expres;
                    }
                }
            }
        }
        double S2x = slot_sum(s2x.value(), t, ntc, nslot), S2y = slot_sum(s2y.value(), t, ntc, nslot);
        double S1 = slot_sum(s1.value(), t, ntc, nslot);
        double S3x = slot_sum(s3x, t, ntc, nslot), S3y = slot_sum(s3y, t, ntc, nslot), S0 = slot_sum(s0, t, ntc, nslot);
        if (lane < ntc && gi != 0) {
            const double si = sgn(gi);
            if ((sgn(S1) != si) || (fabs(S1) < fabs(0.1 * gi))) S1 = 0.1 * gi;
            const double c1 = iei / (re * S1);
            if (!kSeg && dvx) {
                dvx[i] = S2x * c1;
                dvy[i] = S2y * c1;
                continue;
            }
            double ax = vx[i], ay = vy[i];
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

// Transposed near lists restricted to leaves that hold segments: for every
// near entry (leaf l, near leaf M) with segments in M, record l under M.
__global__ void k_fric_count(TreeDev T, int32_t* cnt, const int64_t* off, int32_t* fillc, int32_t* list) {
    const int l = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    if (l >= T.nleaf) return;
    for (int64_t e = T.near_off[l] + lane_id(); e < T.near_off[l + 1]; e += kWave) {
        const int m = T.leaf_rank[T.near_list[e]];
        if (T.seg_off[m + 1] == T.seg_off[m]) continue;
        if (off)
            list[off[m] + atomicAdd(&fillc[m], 1)] = l;
        else
            atomicAdd(&cnt[m], 1);
    }
}

// Friction of one segment per wave, summed in (leaf, target) order.
__global__ __launch_bounds__(kBlock) void k_friction(TreeDev T, const double* __restrict__ x,
                                                     const double* __restrict__ y, const double* __restrict__ g,
                                                     const double* __restrict__ ie, const vvhd_seg* segs,
                                                     const int32_t* seg_leaf, const int64_t* off,
                                                     const int32_t* list, int nseg, double* fric) {
    __shared__ double buf[kWavesPerBlock][kWave];
    const int w = threadIdx.x / kWave;
    const int sid = blockIdx.x * kWavesPerBlock + w;
    const int lane = lane_id();
    if (sid >= nseg) return;
    const vvhd_seg sg = segs[sid];
    const int m = seg_leaf[sid];
    const double dSx = -sg.dly, dSy = sg.dlx;
    const double dlen = sqrt(dSx * dSx + dSy * dSy);
    double acc = fric[sid];
    for (int64_t e = off[m]; e < off[m + 1]; e++) {
        const int L = T.leaf_node[list[e]];
        const int v0 = T.v0[L], v1 = T.v1[L];
        for (int b = v0; b < v1; b += kWave) {
            const int i = b + lane;
            double term = 0;
            bool has = false;
            if (i < v1) {
                const double gi = g[i];
                if (gi != 0) {
                    const double iei = ie[i];
                    const double drx = x[i] - sg.x, dry = y[i] - sg.y;
                    const double drabs = sqrt(drx * drx + dry * dry);
                    const double exparg = -drabs * iei;
                    if (!(exparg < kExpArgLimit)) {
                        term = iei * iei * gi * exp(exparg) * dlen;
                        has = true;
                    }
                }
            }
            buf[w][lane] = term;
            const uint64_t any = __ballot(has);
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            if (lane == 0) {
                const int cnt = min(kWave, v1 - b);
                for (int q = 0; q < cnt; q++)
                    if ((any >> q) & 1) acc += buf[w][q];
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
    }
    if (lane == 0) fric[sid] = acc;
}

}  // namespace

// Adds the increments of a diffusive pass that ran beside the convective
// kernels, in the order the in-place pass adds them (its segment term is +0).
__global__ void k_add_diffusive(const double* g, const double* dvx, const double* dvy, double* vx, double* vy,
                                int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n || g[i] == 0) return;
    double ax = vx[i] + dvx[i], ay = vy[i] + dvy[i];
    ax += 0.0;
    ay += 0.0;
    vx[i] = ax;
    vy[i] = ay;
}

void diffusive(Ctx& c, bool overlap) {
    const TreeDev T = tree_dev(c.t);
    if (T.nleaf == 0) {
        VVHD_HIP(hipEventRecord(c.ev[7], c.stream));
        return;
    }
    const int nl = T.nleaf;
    const int grid = (nl + kWavesPerBlock - 1) / kWavesPerBlock;
    const int nseg = (int)c.segs_h.size();
    const bool batch = c.nsim > 1;
    const double* sim_re = batch ? (const double*)c.sim_re.p : (const double*)nullptr;
    if (overlap && nseg == 0) {
        c.dvx.grow(c.n + 1);
        c.dvy.grow(c.n + 1);
        hipLaunchKernelGGL(k_diffusive<false>, dim3(grid), dim3(kBlock), 0, c.stream2, T, c.x.p, c.y.p, c.g.p, c.ie.p,
                           c.segs.p, nseg, c.vx.p, c.vy.p, c.re, sim_re, (const int32_t*)c.t.leaf_sim.p, c.dvx.p,
                           c.dvy.p);
        VVHD_HIP(hipEventRecord(c.ev[7], c.stream2));
        VVHD_HIP(hipStreamWaitEvent(c.stream, c.ev[7], 0));
        hipLaunchKernelGGL(k_add_diffusive, dim3((unsigned)std::max<int64_t>(1, (c.n + kBlock - 1) / kBlock)),
                           dim3(kBlock), 0, c.stream, c.g.p, c.dvx.p, c.dvy.p,
                           c.vx.p, c.vy.p, c.n);
        return;
    }
    if (nseg)
        hipLaunchKernelGGL(k_diffusive<true>, dim3(grid), dim3(kBlock), 0, c.stream, T, c.x.p, c.y.p, c.g.p, c.ie.p,
                           c.segs.p, nseg, c.vx.p, c.vy.p, c.re, sim_re, (const int32_t*)c.t.leaf_sim.p,
                           (double*)nullptr, (double*)nullptr);
    else
        hipLaunchKernelGGL(k_diffusive<false>, dim3(grid), dim3(kBlock), 0, c.stream, T, c.x.p, c.y.p, c.g.p, c.ie.p,
                           c.segs.p, nseg, c.vx.p, c.vy.p, c.re, sim_re, (const int32_t*)c.t.leaf_sim.p,
                           (double*)nullptr, (double*)nullptr);
    if (nseg) {
        hipStream_t s = c.stream;
        c.iscratch0.grow(nl + 1);
        c.iscratch1.grow(nl + 1);
        c.lscratch0.grow(nl + 1);
        int32_t* cnt = c.iscratch0.p;
        int32_t* fillc = c.iscratch1.p;
        int64_t* off = c.lscratch0.p;
        VVHD_HIP(hipMemsetAsync(cnt, 0, (size_t)(nl + 1) * 4, s));
        VVHD_HIP(hipMemsetAsync(fillc, 0, (size_t)(nl + 1) * 4, s));
        hipLaunchKernelGGL(k_fric_count, dim3(grid), dim3(kBlock), 0, s, T, cnt, (const int64_t*)nullptr, fillc,
                           (int32_t*)nullptr);
        size_t bytes = 0;
        VVHD_HIP(rocprim::exclusive_scan(nullptr, bytes, cnt, off, (int64_t)0, (size_t)nl + 1,
                                         rocprim::plus<int64_t>(), s));
        c.tmp.grow(bytes + 16);
        VVHD_HIP(rocprim::exclusive_scan(c.tmp.p, bytes, cnt, off, (int64_t)0, (size_t)nl + 1,
                                         rocprim::plus<int64_t>(), s));
        int64_t total = 0;
        VVHD_HIP(hipMemcpyAsync(&total, off + nl, 8, hipMemcpyDeviceToHost, s));
        VVHD_HIP(hipStreamSynchronize(s));
        c.iscratch2.grow(total + 1);
        c.iscratch3.grow(total + 1);
        hipLaunchKernelGGL(k_fric_count, dim3(grid), dim3(kBlock), 0, s, T, cnt, (const int64_t*)off, fillc,
                           c.iscratch2.p);
        if (total) {
            bytes = 0;
            VVHD_HIP(rocprim::segmented_radix_sort_keys(nullptr, bytes, c.iscratch2.p, c.iscratch3.p, (unsigned)total,
                                                        (unsigned)nl, off, off + 1, 0, 32, s));
            c.tmp.grow(bytes + 16);
            VVHD_HIP(rocprim::segmented_radix_sort_keys(c.tmp.p, bytes, c.iscratch2.p, c.iscratch3.p,
                                                        (unsigned)total, (unsigned)nl, off, off + 1, 0, 32, s));
        }
        hipLaunchKernelGGL(k_friction, dim3((nseg + kWavesPerBlock - 1) / kWavesPerBlock), dim3(kBlock), 0, s, T,
                           c.x.p, c.y.p, c.g.p, c.ie.p, c.segs.p, c.t.seg_leaf.p, off, c.iscratch3.p, nseg, c.fric.p);
    }
    VVHD_HIP(hipEventRecord(c.ev[7], c.stream));
    VVHD_HIP(hipGetLastError());
}

}  // namespace vvhd
