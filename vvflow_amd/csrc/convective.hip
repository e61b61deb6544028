// convective.hip -- MConvectiveFast::process_all_lists (MConvectiveFast.cpp:
// 36-114), MConvectiveFast::velocity (:20-34) and the FP64 direct sum used as
// the accuracy yardstick.
//
// process_all_lists runs as three kernels:
//  1. k_far_taylor, one wave per leaf: the reference's first-order Taylor
//     expansion about the leaf centre of the per-sign monopoles of the far
//     nodes (:48-69), FP64, lanes over the far list.
//  2. k_near, one wave per leaf: the Lamb-kernel sum
//     g_j rotl(r_i - r_j) / (|r_i - r_j|^2 + eps_j^2) over every vortex of the
//     near leaves (:116-137).  The near leaves' vortex ranges are concatenated
//     into one virtual source list that the 64 lanes walk 64 sources at a
//     time (a lane finds its leaf with a 6-step binary search over prefix
//     sums held across the wave).  Each lane converts its source once to FP32
//     coordinates relative to the target leaf centre (the FP64 difference is
//     taken first, so FP32 keeps ~1e-7 relative accuracy of the offset even
//     60 units from the origin) and accumulates the pair terms of up to 16
//     targets whose leaf-relative coordinates sit in scalar registers.
//     Accumulation is FP32, Kahan-compensated or plain; the 64 lane partials
//     are combined in FP64.
//  3. k_conv_finish, one thread per vortex: v += U_inf + near + far terms in
//     the reference's order, FP64.
//
// The near kernel has a persistent form: a fixed grid whose waves take work
// items -- a leaf and the first of up to 16 of its targets -- from a device
// queue, so that a leaf of many targets spreads over many waves and no block
// waits for one long leaf.  Each item runs exactly the arithmetic of the
// one-wave-per-leaf form for its targets, so both give the same bits.  A
// small autotuner times the one-wave-per-leaf form and the persistent form
// at several grid sizes on the first call of every workload class (number of
// simulations, leaves and targets per leaf, in powers of two) and keeps the
// fastest.  With the persistent form every size the phase needs is read from
// device memory, so the whole phase (queue reset, work items, Taylor terms,
// near field, finish) is captured once as a HIP graph and replayed while its
// buffers stay where they are.

#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#pragma clang fp contract(off)

namespace vvhd {
namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
// Sources of one batch of 64 near-list entries whose global indices fit the
// per-wave LDS table of the near kernel (a larger batch searches the entry
// prefix sums instead).
constexpr int kNearTab = 1024;
constexpr int kTargets = 16;
constexpr double kPi = 3.14159265358979323846;
constexpr double k1_2Pi = 1. / (2. * kPi);
constexpr double k1_Pi = 1. / kPi;

inline int nblocks(int64_t n, int b = kBlock) { return (int)std::max<int64_t>(1, (n + b - 1) / b); }

__device__ inline float readlane_f(float v, int k) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

// Far-field Taylor coefficients of every leaf (MConvectiveFast.cpp:48-69).
__global__ __launch_bounds__(kBlock) void k_far_taylor(TreeDev T, double4* taylor) {
    const int w = threadIdx.x / kWave;
    const int l = blockIdx.x * kWavesPerBlock + w;
    if (l >= T.nleaf) return;
    const int L = T.leaf_node[l];
    const double cx = T.nx[L], cy = T.ny[L];
    double a1 = 0, a2 = 0, a3 = 0, a4 = 0;
    for (int64_t e = T.far_off[l] + lane_id(); e < T.far_off[l + 1]; e += kWave) {
        const int F = T.far_list[e];
        const double dPx = cx - T.cmp_x[F], dPy = cy - T.cmp_y[F];
        const double dMx = cx - T.cmm_x[F], dMy = cy - T.cmm_y[F];
        const double aP = dPx * dPx + dPy * dPy, aM = dMx * dMx + dMy * dMy;
        const double fP1 = T.cmp_g[F] / aP, fM1 = T.cmm_g[F] / aM;
        const double fP2 = T.cmp_g[F] / (aP * aP), fM2 = T.cmm_g[F] / (aM * aM);
        a1 += (fP1 * dPy + fM1 * dMy);
        a2 += (fP1 * dPx + fM1 * dMx);
        a3 += (fP2 * dPy * dPx + fM2 * dMy * dMx);
        a4 += (fP2 * (dPy * dPy - dPx * dPx) + fM2 * (dMy * dMy - dMx * dMx));
    }
    a1 = wave_sum(a1);
    a2 = wave_sum(a2);
    a3 = wave_sum(a3);
    a4 = wave_sum(a4);
    if (lane_id() == 0) taylor[l] = make_double4(-a1 * k1_2Pi, a2 * k1_2Pi, a3 * k1_Pi, a4 * k1_2Pi);
}

// k_far_taylor over a fixed grid, the leaf count read from the device.
__global__ __launch_bounds__(kBlock) void k_far_taylor_gs(TreeDev T, const int32_t* nleaf_dev, double4* taylor) {
    const int nl = *nleaf_dev;
    const int waves = gridDim.x * kWavesPerBlock;
    for (int l = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; l < nl; l += waves) {
        const int L = T.leaf_node[l];
        const double cx = T.nx[L], cy = T.ny[L];
        double a1 = 0, a2 = 0, a3 = 0, a4 = 0;
        for (int64_t e = T.far_off[l] + lane_id(); e < T.far_off[l + 1]; e += kWave) {
            const int F = T.far_list[e];
            const double dPx = cx - T.cmp_x[F], dPy = cy - T.cmp_y[F];
            const double dMx = cx - T.cmm_x[F], dMy = cy - T.cmm_y[F];
            const double aP = dPx * dPx + dPy * dPy, aM = dMx * dMx + dMy * dMy;
            const double fP1 = T.cmp_g[F] / aP, fM1 = T.cmm_g[F] / aM;
            const double fP2 = T.cmp_g[F] / (aP * aP), fM2 = T.cmm_g[F] / (aM * aM);
            a1 += (fP1 * dPy + fM1 * dMy);
            a2 += (fP1 * dPx + fM1 * dMx);
            a3 += (fP2 * dPy * dPx + fM2 * dMy * dMx);
            a4 += (fP2 * (dPy * dPy - dPx * dPx) + fM2 * (dMy * dMy - dMx * dMx));
        }
        a1 = wave_sum(a1);
        a2 = wave_sum(a2);
        a3 = wave_sum(a3);
        a4 = wave_sum(a4);
        if (lane_id() == 0) taylor[l] = make_double4(-a1 * k1_2Pi, a2 * k1_2Pi, a3 * k1_Pi, a4 * k1_2Pi);
    }
}

// Only what the near-field kernel reads (keeps its scalar registers free for
// the target coordinates).
struct NearArgs {
    const int32_t* leaf_node;
    const int32_t* v0;
    const int32_t* v1;
    const double* nx;
    const double* ny;
    const int64_t* near_off;
    const int32_t* near_list;
    int nleaf;
};

// Near-field sums of one chunk of NT targets (base .. base+nt-1, nt <= NT;
// slots beyond nt hold far-away dummies whose sums are discarded), with no
// per-target branches in the pair loop.
// tab: this wave's LDS table of source indices (kNearTab entries).
template <int NT, bool KAHAN>
__device__ inline void near_chunk(const NearArgs& A, const double* __restrict__ x, const double* __restrict__ y,
                                  const float2* __restrict__ src_ge, double2* __restrict__ near_out, int base, int nt,
                                  double cx, double cy, int64_t e_beg, int64_t e_end, int* tab) {
    const int lane = lane_id();
    float my_tx = 1e15f, my_ty = 1e15f;
    if (lane < nt) {
        my_tx = (float)(x[base + lane] - cx);
        my_ty = (float)(y[base + lane] - cy);
    }
    float tx[NT], ty[NT];
#pragma unroll
    for (int k = 0; k < NT; k++) {
        tx[k] = readlane_f(my_tx, k);
        ty[k] = readlane_f(my_ty, k);
    }
    float sx[NT], sy[NT], qx[NT], qy[NT];
#pragma unroll
    for (int k = 0; k < NT; k++) sx[k] = sy[k] = qx[k] = qy[k] = 0.f;

    // software pipeline: the near-list entry of the next chunk and the
    // sources of the next batch are loaded while the current batch computes
    int M_next = (e_beg + lane < e_end) ? A.near_list[e_beg + lane] : -1;
    for (int64_t eb = e_beg; eb < e_end; eb += kWave) {
        const int M = M_next;
        M_next = (eb + kWave + lane < e_end) ? A.near_list[eb + kWave + lane] : -1;
        int s0 = 0, cnt = 0;
        if (M >= 0) {
            s0 = A.v0[M];
            cnt = A.v1[M] - s0;
        }
        const int incl = wave_incl_scan(cnt);
        const int total = __shfl(incl, kWave - 1, kWave);
        // source position -> global index: each lane writes its entry's run
        // into the table (the previous batch's reads are done: one wave
        // issues its LDS operations in order)
        const bool use_tab = total <= kNearTab;   // uniform
        if (use_tab) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            for (int t = 0; t < cnt; t++) tab[incl - cnt + t] = s0 + t;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        auto fetch = [&](int sb, float& sxr, float& syr, float& gj, float& e2) {
            const int sp = sb + lane;
            int j;
            if (use_tab) {
                j = sp < total ? tab[sp] : 0;
            } else {
                // entry index: number of entries whose inclusive prefix <= sp
                int pos = 0;
#pragma unroll
                for (int step = 32; step > 0; step >>= 1) {
                    const int v = __shfl(incl, pos + step - 1, kWave);
                    if (v <= sp) pos += step;
                }
                pos = min(pos, kWave - 1);
                const int ex = __shfl(incl, pos, kWave) - __shfl(cnt, pos, kWave);
                j = __shfl(s0, pos, kWave) + (sp - ex);
            }
            sxr = 0.f;
            syr = 0.f;
            gj = 0.f;
            e2 = 1.f;
            if (sp < total) {
                const float2 ge = src_ge[j];
                gj = ge.x;
                e2 = ge.y;
                sxr = (float)(x[j] - cx);
                syr = (float)(y[j] - cy);
            }
        };
        if (KAHAN) {
            // block-compensated sums: the terms go by FMA into the low-order
            // accumulators qx/qy, which are folded into the running sums
            // sx/sy with a compensated (Fast2Sum) step every eight terms per
            // lane; qx/qy keep the rounding error of each fold, so every
            // rounding is relative to at most eight terms, never to the
            // running sum.  Two sets of prefetched sources are used in turn
            // (no copies between steps); past the batch the fetch returns
            // zero-circulation sources.
            float sa0, ta0, ga0, ea0, sb0, tb0, gb0, eb0;
            float sa1, ta1, ga1, ea1, sb1, tb1, gb1, eb1;
            auto pairs = [&](float sa, float ta, float ga, float ea, float sbx, float tb, float gb, float eb2) {
#pragma unroll
                for (int k = 0; k < NT; k++) {
                    const float dxa = tx[k] - sa, dya = ty[k] - ta;
                    const float fa = ga * __builtin_amdgcn_rcpf(__builtin_fmaf(dxa, dxa, __builtin_fmaf(dya, dya, ea)));
                    const float dxb = tx[k] - sbx, dyb = ty[k] - tb;
                    const float fb = gb * __builtin_amdgcn_rcpf(__builtin_fmaf(dxb, dxb, __builtin_fmaf(dyb, dyb, eb2)));
                    qx[k] = __builtin_fmaf(-dyb, fb, __builtin_fmaf(-dya, fa, qx[k]));
                    qy[k] = __builtin_fmaf(dxb, fb, __builtin_fmaf(dxa, fa, qy[k]));
                }
            };
            auto fold = [&]() {
#pragma unroll
                for (int k = 0; k < NT; k++) {
                    const float tsx = sx[k] + qx[k];
                    qx[k] = qx[k] - (tsx - sx[k]);
                    sx[k] = tsx;
                    const float tsy = sy[k] + qy[k];
                    qy[k] = qy[k] - (tsy - sy[k]);
                    sy[k] = tsy;
                }
            };
            if (total > 0) {
                fetch(0, sa0, ta0, ga0, ea0);
                fetch(kWave, sb0, tb0, gb0, eb0);
            }
            for (int sb = 0; sb < total; sb += 4 * kWave) {
                fetch(sb + 2 * kWave, sa1, ta1, ga1, ea1);
                fetch(sb + 3 * kWave, sb1, tb1, gb1, eb1);
                pairs(sa0, ta0, ga0, ea0, sb0, tb0, gb0, eb0);
                if (sb + 2 * kWave >= total) break;
                fetch(sb + 4 * kWave, sa0, ta0, ga0, ea0);
                fetch(sb + 5 * kWave, sb0, tb0, gb0, eb0);
                pairs(sa1, ta1, ga1, ea1, sb1, tb1, gb1, eb1);
                if (sb & (4 * kWave)) fold();   // every other step: eight terms per lane
            }
            fold();
        } else {
            float s0x, s0y, g0, e0, s1x, s1y, g1, e1;
            auto terms = [&](float sxr, float syr, float gj, float e2) {
#pragma unroll
                for (int k = 0; k < NT; k++) {
                    const float dx = tx[k] - sxr;
                    const float dy = ty[k] - syr;
                    const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, e2));
                    const float f = gj * __builtin_amdgcn_rcpf(r2);
                    sx[k] = __builtin_fmaf(-dy, f, sx[k]);
                    sy[k] = __builtin_fmaf(dx, f, sy[k]);
                }
            };
            if (total > 0) fetch(0, s0x, s0y, g0, e0);
            for (int sb = 0; sb < total; sb += 2 * kWave) {
                fetch(sb + kWave, s1x, s1y, g1, e1);
                terms(s0x, s0y, g0, e0);
                if (sb + kWave >= total) break;
                fetch(sb + 2 * kWave, s0x, s0y, g0, e0);
                terms(s1x, s1y, g1, e1);
            }
        }
    }
    // Totals over the 64 lanes of the P values of each component (P = NT
    // rounded up to a power of two) by recursive halving: at lane offset o a
    // lane keeps the half of its values picked by (lane & o) and adds the
    // partner's copy of that half, so each exchange moves half the remaining
    // values.  Every value is summed in the same order as a butterfly over
    // offsets 32, 16, ..., 1.  Lane l ends with the total of target slot
    // idx, shared by the lanes that differ only in the low bits.
    constexpr int P = NT <= 2 ? 2 : NT <= 4 ? 4 : NT <= 8 ? 8 : 16;
    constexpr int kLow = kWave / P - 1;   // lanes sharing a total: the lowest writes it
#pragma unroll
    for (int comp = 0; comp < 2; comp++) {
        double v[P];
#pragma unroll
        for (int k = 0; k < P; k++) {
            const float s_ = comp == 0 ? sx[k < NT ? k : 0] : sy[k < NT ? k : 0];
            const float q_ = comp == 0 ? qx[k < NT ? k : 0] : qy[k < NT ? k : 0];
            v[k] = k < NT ? (KAHAN ? (double)s_ + (double)q_ : (double)s_) : 0.;
        }
        int idx = 0, o = kWave / 2;
#pragma unroll
        for (int h = P / 2; h >= 1; h >>= 1, o >>= 1) {
            const bool hi = (lane & o) != 0;
#pragma unroll
            for (int j = 0; j < h; j++) {
                const double keep = hi ? v[j + h] : v[j];
                const double send = hi ? v[j] : v[j + h];
                v[j] = keep + __shfl_xor(send, o, kWave);
            }
            if (hi) idx += h;
        }
        for (; o > 0; o >>= 1) v[0] += __shfl_xor(v[0], o, kWave);
        if ((lane & kLow) == 0 && idx < nt) reinterpret_cast<double*>(near_out)[2 * (base + idx) + comp] = v[0];
    }
}

// Near-field work items: one per leaf and chunk of kTargets targets,
// {leaf, first target}; q[0] counts them (the order of the items does not
// matter: each writes only its own targets' sums).
__global__ void k_near_items(NearArgs A, const int32_t* nleaf_dev, int2* items, int32_t* q) {
    const int nl = *nleaf_dev;
    for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < nl; l += gridDim.x * blockDim.x) {
        const int L = A.leaf_node[l];
        const int v0 = A.v0[L], v1 = A.v1[L];
        const int k = (v1 - v0 + kTargets - 1) / kTargets;
        if (k == 0) continue;
        const int b = atomicAdd(&q[0], k);
        for (int c = 0; c < k; c++) items[b + c] = make_int2(l, v0 + c * kTargets);
    }
}

__global__ void k_fill_i32q(int32_t* a, int n) {
    if (threadIdx.x < n) a[threadIdx.x] = 0;
}

template <bool KAHAN>
__global__ __launch_bounds__(kBlock) void k_near_persist(NearArgs A, const int2* __restrict__ items, int32_t* q,
                                                         const double* __restrict__ x, const double* __restrict__ y,
                                                         const float2* __restrict__ src_ge,
                                                         double2* __restrict__ near_out) {
    __shared__ int s_tab[kWavesPerBlock][kNearTab];
    const int w = threadIdx.x / kWave;
    const int total = q[0];
    for (;;) {
        int it = 0;
        if (lane_id() == 0) it = atomicAdd(&q[1], 1);
        it = __shfl(it, 0, kWave);
        if (it >= total) break;   // every wave reaches this: the queue only grows past total
        const int2 item = items[it];
        const int l = item.x, base = item.y;
        const int L = A.leaf_node[l];
        const int nt = min(kTargets, A.v1[L] - base);
        const double cx = A.nx[L], cy = A.ny[L];
        const int64_t e_beg = A.near_off[l], e_end = A.near_off[l + 1];
        switch ((nt + 1) / 2) {
            case 1: near_chunk<2, KAHAN>(A, x, y, src_ge, near_out, base, nt, cx, cy, e_beg, e_end, s_tab[w]); break;
            case 2: near_chunk<4, KAHAN>(A, x, y, src_ge, near_out, base, nt, cx, cy, e_beg, e_end, s_tab[w]); break;
            case 3: near_chunk<6, KAHAN>(A, x, y, src_ge, near_out, base, nt, cx, cy, e_beg, e_end, s_tab[w]); break;
            case 4: near_chunk<8, KAHAN>(A, x, y, src_ge, near_out, base, nt, cx, cy, e_beg, e_end, s_tab[w]); break;
            case 5: near_chunk<10, KAHAN>(A, x, y, src_ge, near_out, base, nt, cx, cy, e_beg, e_end, s_tab[w]); break;
            case 6: near_chunk<12, KAHAN>(A, x, y, src_ge, near_out, base, nt, cx, cy, e_beg, e_end, s_tab[w]); break;
            case 7: near_chunk<14, KAHAN>(A, x, y, src_ge, near_out, base, nt, cx, cy, e_beg, e_end, s_tab[w]); break;
            default: near_chunk<16, KAHAN>(A, x, y, src_ge, near_out, base, nt, cx, cy, e_beg, e_end, s_tab[w]); break;
        }
    }
}

template <bool KAHAN>
__global__ __launch_bounds__(kBlock) void k_near(NearArgs A, const double* __restrict__ x,
                                                 const double* __restrict__ y, const float2* __restrict__ src_ge,
                                                 double2* __restrict__ near_out) {
    __shared__ int s_tab[kWavesPerBlock][kNearTab];
    const int w = threadIdx.x / kWave;
    const int l = xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock + w;
    if (l >= A.nleaf) return;
    const int L = A.leaf_node[l];
    const int v0 = A.v0[L], v1 = A.v1[L];
    const double cx = A.nx[L], cy = A.ny[L];
    const int64_t e_beg = A.near_off[l], e_end = A.near_off[l + 1];
    for (int base = v0; base < v1; base += kTargets) {
        const int nt = min(kTargets, v1 - base);
        // target slots in steps of two: few dummy targets per chunk
        switch ((nt + 1) / 2) {
            case 1: near_chunk<2, KAHAN>(A, x, y, src_ge, near_out, base, nt, cx, cy, e_beg, e_end, s_tab[w]); break;
            case 2: near_chunk<4, KAHAN>(A, x, y, src_ge, near_out, base, nt, cx, cy, e_beg, e_end, s_tab[w]); break;
            case 3: near_chunk<6, KAHAN>(A, x, y, src_ge, near_out, base, nt, cx, cy, e_beg, e_end, s_tab[w]); break;
            case 4: near_chunk<8, KAHAN>(A, x, y, src_ge, near_out, base, nt, cx, cy, e_beg, e_end, s_tab[w]); break;
            case 5: near_chunk<10, KAHAN>(A, x, y, src_ge, near_out, base, nt, cx, cy, e_beg, e_end, s_tab[w]); break;
            case 6: near_chunk<12, KAHAN>(A, x, y, src_ge, near_out, base, nt, cx, cy, e_beg, e_end, s_tab[w]); break;
            case 7: near_chunk<14, KAHAN>(A, x, y, src_ge, near_out, base, nt, cx, cy, e_beg, e_end, s_tab[w]); break;
            default: near_chunk<16, KAHAN>(A, x, y, src_ge, near_out, base, nt, cx, cy, e_beg, e_end, s_tab[w]); break;
        }
    }
}

// v += U_inf; += near * 1/(2 pi); += sinks (none); += bodies (none);
// += (T1, T2); += (T3 dx + T4 dy, T4 dx - T3 dy)   (MConvectiveFast.cpp:76-86)
__global__ void k_conv_finish(const int32_t* vleaf, const double* nx, const double* ny, const int32_t* leaf_node,
                              const double4* taylor, const double* x, const double* y, const double* g,
                              const double2* near_in, double* vx, double* vy, int64_t n, double inf_x, double inf_y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || g[i] == 0) return;
    const int l = vleaf[i];
    const int L = leaf_node[l];
    const double4 T = taylor[l];
    const double2 nr = near_in[i];
    double ax = vx[i] + inf_x, ay = vy[i] + inf_y;
    ax += nr.x * k1_2Pi;
    ay += nr.y * k1_2Pi;
    ax += 0.0;
    ay += 0.0;
    ax += 0.0;
    ay += 0.0;
    ax += T.x;
    ay += T.y;
    const double dxl = x[i] - nx[L], dyl = y[i] - ny[L];
    ax += T.z * dxl + T.w * dyl;
    ay += T.w * dxl + (-T.z) * dyl;
    vx[i] = ax;
    vy[i] = ay;
}

// k_conv_finish over a fixed grid, the vortex count read from the device.
__global__ void k_conv_finish_gs(const int32_t* vleaf, const double* nx, const double* ny, const int32_t* leaf_node,
                                 const double4* taylor, const double* x, const double* y, const double* g,
                                 const double2* near_in, double* vx, double* vy, const int64_t* n_dev, double inf_x,
                                 double inf_y) {
    const int64_t n = *n_dev;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (g[i] == 0) continue;
        const int l = vleaf[i];
        const int L = leaf_node[l];
        const double4 T = taylor[l];
        const double2 nr = near_in[i];
        double ax = vx[i] + inf_x, ay = vy[i] + inf_y;
        ax += nr.x * k1_2Pi;
        ay += nr.y * k1_2Pi;
        ax += 0.0;
        ay += 0.0;
        ax += 0.0;
        ay += 0.0;
        ax += T.x;
        ay += T.y;
        const double dxl = x[i] - nx[L], dyl = y[i] - ny[L];
        ax += T.z * dxl + T.w * dyl;
        ay += T.w * dxl + (-T.z) * dyl;
        vx[i] = ax;
        vy[i] = ay;
    }
}

// MConvectiveFast::velocity for arbitrary points, FP64 (one wave per point).
__global__ void k_velocity_points(TreeDev T, const double* x, const double* y, const double* g, const double* ie,
                                  const double* pts, int64_t n, double inf_x, double inf_y, double* out) {
    const int64_t p = (int64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    if (p >= n) return;
    const double px = pts[2 * p], py = pts[2 * p + 1];
    int node = 0;   // stree::findNode
    while (T.ch1[node] >= 0) {
        const int c = T.ch1[node];
        if (T.nh[node] < T.nw[node])
            node = (px < T.nx[node]) ? c : c + 1;
        else
            node = (py < T.ny[node]) ? c : c + 1;
    }
    const int l = T.leaf_rank[node];
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
    if (lane_id() == 0) {
        double rx = 0, ry = 0;
        rx += nx_ * k1_2Pi;
        ry += ny_ * k1_2Pi;
        rx += fx * k1_2Pi;
        ry += fy * k1_2Pi;
        rx += 0.0;
        ry += 0.0;
        rx += inf_x;
        ry += inf_y;
        out[2 * p] = rx;
        out[2 * p + 1] = ry;
    }
}

// FP64 direct sum of the induced velocity, compensated; sources staged in
// LDS, one thread per target.
constexpr int kDsBlock = 256;
// Sources [lo, n).
__global__ __launch_bounds__(kDsBlock) void k_direct_sum(const double* x, const double* y, const double* g,
                                                         const double* ie, int64_t lo, int64_t n,
                                                         const int64_t* targets, int64_t nt, double* out) {
    __shared__ double sx[kDsBlock], sy[kDsBlock], sg[kDsBlock], se[kDsBlock];
    const int64_t t = (int64_t)blockIdx.x * kDsBlock + threadIdx.x;
    const int64_t i = (t < nt) ? (targets ? targets[t] : t) : 0;
    const double px = (t < nt) ? x[i] : 0, py = (t < nt) ? y[i] : 0;
    double ax = 0, ay = 0, cxs = 0, cys = 0;
    for (int64_t b = lo; b < n; b += kDsBlock) {
        const int64_t j = b + threadIdx.x;
        if (j < n && g[j] != 0) {
            const double e1 = 1. / ie[j];
            sx[threadIdx.x] = x[j];
            sy[threadIdx.x] = y[j];
            sg[threadIdx.x] = g[j];
            se[threadIdx.x] = e1 * e1;
        } else {
            sx[threadIdx.x] = 0;
            sy[threadIdx.x] = 0;
            sg[threadIdx.x] = 0;
            se[threadIdx.x] = 1;
        }
        __syncthreads();
        const int m = (int)std::min<int64_t>(kDsBlock, n - b);
        for (int q = 0; q < m; q++) {
            const double dx = px - sx[q], dy = py - sy[q];
            const double f = sg[q] / (dx * dx + dy * dy + se[q]);
            const double tx = -dy * f - cxs;
            const double ux = ax + tx;
            cxs = (ux - ax) - tx;
            ax = ux;
            const double ty = dx * f - cys;
            const double uy = ay + ty;
            cys = (uy - ay) - ty;
            ay = uy;
        }
        __syncthreads();
    }
    if (t < nt) {
        out[2 * t] = ax * k1_2Pi;
        out[2 * t + 1] = ay * k1_2Pi;
    }
}

}  // namespace

namespace {

constexpr int kPersistWaves[] = {8, 16, 32};   // persistent grid: waves per compute unit

// Persistent-grid blocks for waves_per_cu waves on every compute unit.
int persist_blocks(const Ctx& c, int waves_per_cu) {
    return std::max(1, c.n_cu * waves_per_cu / kWavesPerBlock);
}

// The near field in the form `choice` names: 0 one wave per leaf, k > 0 the
// persistent grid with k waves per compute unit (items built first).
void launch_near(Ctx& c, int choice, int kahan, hipStream_t s, double2* near_out) {
    const TreeDev T = tree_dev(c.t);
    const int nl = T.nleaf;
    NearArgs A{T.leaf_node, T.v0, T.v1, T.nx, T.ny, T.near_off, T.near_list, nl};
    if (choice == 0) {
        const int grid = (nl + kWavesPerBlock - 1) / kWavesPerBlock;
        if (kahan)
            hipLaunchKernelGGL(k_near<true>, dim3(grid), dim3(kBlock), 0, s, A, c.x.p, c.y.p, c.src_ge.p, near_out);
        else
            hipLaunchKernelGGL(k_near<false>, dim3(grid), dim3(kBlock), 0, s, A, c.x.p, c.y.p, c.src_ge.p, near_out);
        return;
    }
    const int32_t* nleaf_dev = c.t.counters.p + 1;
    hipLaunchKernelGGL(k_fill_i32q, dim3(1), dim3(64), 0, s, c.near_q.p, 2);
    hipLaunchKernelGGL(k_near_items, dim3(nblocks(nl)), dim3(kBlock), 0, s, A, nleaf_dev, c.near_items.p, c.near_q.p);
    const int g = persist_blocks(c, choice);
    if (kahan)
        hipLaunchKernelGGL(k_near_persist<true>, dim3(g), dim3(kBlock), 0, s, A, (const int2*)c.near_items.p,
                           c.near_q.p, c.x.p, c.y.p, c.src_ge.p, near_out);
    else
        hipLaunchKernelGGL(k_near_persist<false>, dim3(g), dim3(kBlock), 0, s, A, (const int2*)c.near_items.p,
                           c.near_q.p, c.x.p, c.y.p, c.src_ge.p, near_out);
}

// Workload class of the near field: simulations (one or several), leaves and
// targets per leaf in powers of two.
uint64_t near_key(const Ctx& c, int kahan, bool beside) {
    const int64_t nl = std::max<int64_t>(1, c.t.nleaves_h);
    const int lg_nl = 63 - __builtin_clzll((unsigned long long)nl);
    const int lg_per = 63 - __builtin_clzll((unsigned long long)std::max<int64_t>(1, c.n / nl));
    return ((uint64_t)beside << 40) | ((uint64_t)(c.nsim > 1) << 32) | ((uint64_t)kahan << 24) |
           ((uint64_t)lg_nl << 8) | (uint64_t)lg_per;
}

// The form of the near field for this workload: the one measured fastest the
// first time the class was seen (each candidate once, HIP events on s), or
// one wave per leaf with the autotuner off.  With `beside` (the near field
// shares the GPU with the far-field kernels or the diffusive) only the
// persistent forms are candidates: the candidates are timed alone, where one
// wave per leaf may tie with them, but beside other kernels the persistent
// grid keeps its pace and one wave per leaf falls behind (1.6 against 3.4 ms
// at 1M).
int near_choice(Ctx& c, int kahan, hipStream_t s, double2* near_out, bool beside) {
    if (!c.autotune) return c.near_form;
    const uint64_t key = near_key(c, kahan, beside);
    auto it = c.tune_near.find(key);
    if (it != c.tune_near.end()) return it->second;
    int best = 0;
    float best_ms = 0;
    const int cands[] = {kPersistWaves[0], kPersistWaves[1], kPersistWaves[2], 0};
    for (int k = 0; k < (beside ? 3 : 4); k++) {
        VVHD_HIP(hipEventRecord(c.tune_ev[0], s));
        launch_near(c, cands[k], kahan, s, near_out);
        VVHD_HIP(hipEventRecord(c.tune_ev[1], s));
        VVHD_HIP(hipEventSynchronize(c.tune_ev[1]));
        float ms = 0;
        VVHD_HIP(hipEventElapsedTime(&ms, c.tune_ev[0], c.tune_ev[1]));
        if (k == 0 || ms < best_ms) {
            best = cands[k];
            best_ms = ms;
        }
    }
    c.tune_near[key] = best;
    return best;
}

void grow_near_items(Ctx& c) {
    c.near_items.grow((size_t)(c.n / kTargets + c.t.nleaves_h + 2));
    c.near_q.grow(4);
}

}  // namespace

// Near-field sums of every vortex into near_out (n + 1 entries; the plain and
// the learned path share it).
void convective_near(Ctx& c, int kahan, hipStream_t s, double2* near_out, bool beside) {
    const TreeDev T = tree_dev(c.t);
    if (T.nleaf == 0) return;
    grow_near_items(c);
    const int choice = near_choice(c, kahan, s, near_out, beside);
    c.near_last = choice;
    VVHD_HIP(hipEventRecord(c.ev[12], s));
    launch_near(c, choice, kahan, s, near_out);
    VVHD_HIP(hipEventRecord(c.ev[13], s));
    c.phase_recorded[6] = true;
}

namespace {

// The convective phase as one graph: queue reset, work items, Taylor terms,
// persistent near field, finish -- every size read from the device.  A graph
// is kept per set of buffers and arguments (every pointer its kernels read
// or write, the free stream and the form), since the flow-move swaps the
// vortex arrays with their double buffers every step and buffers move when
// they grow; at most four are kept.
void convective_graph(Ctx& c, double inf_x, double inf_y, int kahan, int choice, double4* taylor, double2* near_out) {
    const TreeDev T = tree_dev(c.t);
    const void* ps[] = {c.x.p,       c.y.p,      c.g.p,       c.vx.p,       c.vy.p,      c.src_ge.p,  taylor,
                        near_out,    c.near_items.p, c.near_q.p, c.t.counters.p, c.sim_off.p + c.nsim, T.leaf_node,
                        T.v0,        T.v1,       T.nx,        T.ny,         T.near_off,  T.near_list, T.far_off,
                        T.far_list,  T.cmp_x,    T.cmp_y,     T.cmp_g,      T.cmm_x,     T.cmm_y,     T.cmm_g,
                        T.vleaf};
    const double fs[2] = {inf_x, inf_y};
    const int is[3] = {kahan, choice, c.n_cu};
    std::vector<uint8_t> key(sizeof(ps) + sizeof(fs) + sizeof(is));
    std::memcpy(key.data(), ps, sizeof(ps));
    std::memcpy(key.data() + sizeof(ps), fs, sizeof(fs));
    std::memcpy(key.data() + sizeof(ps) + sizeof(fs), is, sizeof(is));
    hipGraphExec_t exec = nullptr;
    for (auto& e : c.conv_graphs)
        if (e.first == key) exec = (hipGraphExec_t)e.second;
    if (!exec) {
        VVHD_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
        const int g = persist_blocks(c, choice);
        const int32_t* nleaf_dev = c.t.counters.p + 1;
        hipLaunchKernelGGL(k_far_taylor_gs, dim3(g), dim3(kBlock), 0, c.stream, T, nleaf_dev, taylor);
        launch_near(c, choice, kahan, c.stream, near_out);
        hipLaunchKernelGGL(k_conv_finish_gs, dim3(g), dim3(kBlock), 0, c.stream, T.vleaf, T.nx, T.ny, T.leaf_node,
                           (const double4*)taylor, c.x.p, c.y.p, c.g.p, (const double2*)near_out, c.vx.p, c.vy.p,
                           (const int64_t*)(c.sim_off.p + c.nsim), inf_x, inf_y);
        hipGraph_t graph = nullptr;
        VVHD_HIP(hipStreamEndCapture(c.stream, &graph));
        VVHD_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        VVHD_HIP(hipGraphDestroy(graph));
        if (c.conv_graphs.size() >= 4) {
            (void)hipGraphExecDestroy((hipGraphExec_t)c.conv_graphs.front().second);
            c.conv_graphs.erase(c.conv_graphs.begin());
        }
        c.conv_graphs.emplace_back(key, (void*)exec);
        c.graph_captures++;
    }
    VVHD_HIP(hipEventRecord(c.ev[12], c.stream));
    VVHD_HIP(hipGraphLaunch(exec, c.stream));
    VVHD_HIP(hipEventRecord(c.ev[13], c.stream));
    c.phase_recorded[6] = true;
    c.graph_launches++;
}

}  // namespace

// process_all_lists with the per-leaf Taylor terms in taylor (nleaf + 1) and
// the per-vortex near sums in near_out (n + 1): every entry is written before
// it is read, so neither needs clearing.
// noalloc: the allocation-free caller -- one wave per leaf, no work-item
// buffer, no tuning, no graph.
static void convective_into(Ctx& c, double inf_x, double inf_y, int kahan, double4* taylor, double2* near_out,
                            bool noalloc = false) {
    const TreeDev T = tree_dev(c.t);
    const int nl = T.nleaf;
    if (!noalloc) grow_near_items(c);
    if (!noalloc && c.use_graph && !(c.heat.n || c.streak.n)) {
        const int choice = near_choice(c, kahan, c.stream, near_out, c.overlap);
        c.near_last = choice;
        if (choice > 0) {
            convective_graph(c, inf_x, inf_y, kahan, choice, taylor, near_out);
            VVHD_HIP(hipGetLastError());
            return;
        }
    }
    const int grid = (nl + kWavesPerBlock - 1) / kWavesPerBlock;
    hipLaunchKernelGGL(k_far_taylor, dim3(grid), dim3(kBlock), 0, c.stream, T, taylor);
    if (noalloc) {
        VVHD_HIP(hipEventRecord(c.ev[12], c.stream));
        launch_near(c, 0, kahan, c.stream, near_out);
        VVHD_HIP(hipEventRecord(c.ev[13], c.stream));
        c.phase_recorded[6] = true;
    } else {
        convective_near(c, kahan, c.stream, near_out, c.overlap);
    }
    hipLaunchKernelGGL(k_conv_finish, dim3(nblocks(c.n)), dim3(kBlock), 0, c.stream, T.vleaf, T.nx, T.ny, T.leaf_node,
                       taylor, c.x.p, c.y.p, c.g.p, near_out, c.vx.p, c.vy.p, c.n, inf_x, inf_y);
    VVHD_HIP(hipGetLastError());
    if (c.heat.n || c.streak.n) passive_convective(c, taylor, inf_x, inf_y);
}

// The per-leaf Taylor terms alone (the heat and streak targets of the learned
// path use the reference's far field).
void far_taylor(Ctx& c, double4* taylor) {
    const TreeDev T = tree_dev(c.t);
    const int nl = T.nleaf;
    if (nl == 0) return;
    const int grid = (nl + kWavesPerBlock - 1) / kWavesPerBlock;
    hipLaunchKernelGGL(k_far_taylor, dim3(grid), dim3(kBlock), 0, c.stream, T, taylor);
}

void convective(Ctx& c, double inf_x, double inf_y, int kahan) {
    if (c.t.nleaves_h == 0) return;
    c.taylor.grow(c.t.nleaves_h + 1);
    c.near_v.grow(c.n + 1);
    convective_into(c, inf_x, inf_y, kahan, c.taylor.p, c.near_v.p);
}

// ---------------------------------------------- allocation-free convective --
//
// Scratch of convective_noalloc: the per-leaf Taylor terms followed by the
// per-vortex near-field sums, in a device array of fixed size instead of
// buffers grown on demand.  In host code `scratch` names a host-side stand-in,
// not device memory: its device address is looked up once per device
// (noalloc_init, called by vvhd_create with the device current) and only that
// address is handed to kernels.  Every context on a device shares the array,
// so calls that overlap on one device race; each call writes every entry it
// reads, so nothing carries over from the previous call.
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
}  // namespace

void noalloc_init(int device) {
    VVHD_CHECK(device >= 0 && device < kMaxDevices, VVHD_EUNSUPPORTED,
               "device index " + std::to_string(device) + " beyond the " + std::to_string(kMaxDevices) +
                   " devices the allocation-free scratch table holds");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    VVHD_CHECK(e == hipSuccess && p, VVHD_EHIP,
               std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = p;
}

size_t convective_noalloc_bytes(int64_t n, int64_t nleaf) {
    return align256((size_t)(nleaf + 1) * sizeof(double4)) + align256((size_t)(n + 1) * sizeof(double2));
}

void convective_noalloc(Ctx& c, double inf_x, double inf_y, int kahan) {
    if (c.t.nleaves_h == 0) return;
    const size_t need = convective_noalloc_bytes(c.n, c.t.nleaves_h);
    VVHD_CHECK(need <= kNoallocScratchBytes, VVHD_EINVAL,
               "convective_noalloc: " + std::to_string(c.n) + " vortices in " + std::to_string(c.t.nleaves_h) +
                   " leaves need " + std::to_string(need) + " bytes of scratch, more than its " +
                   std::to_string(kNoallocScratchBytes) + " bytes");
    unsigned char* base;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        base = (unsigned char*)g_scratch_ptr[c.device];
    }
    VVHD_CHECK(base, VVHD_ESTATE, "convective_noalloc: scratch of this device not set up");
    double4* taylor = (double4*)base;
    double2* near_out = (double2*)(base + align256((size_t)(c.t.nleaves_h + 1) * sizeof(double4)));
    convective_into(c, inf_x, inf_y, kahan, taylor, near_out, true);
}

void velocity_points(Ctx& c, const double* xy, int64_t n, double inf_x, double inf_y, double* v_out) {
    if (n == 0) return;
    const TreeDev T = tree_dev(c.t);
    c.tmp.grow((size_t)n * 32);
    double* dpts = (double*)c.tmp.p;
    double* dout = dpts + 2 * n;
    VVHD_HIP(hipMemcpyAsync(dpts, xy, (size_t)n * 16, hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(k_velocity_points, dim3((unsigned)((n + kWavesPerBlock - 1) / kWavesPerBlock)), dim3(kBlock), 0,
                       c.stream, T, c.x.p, c.y.p, c.g.p, c.ie.p, dpts, n, inf_x, inf_y, dout);
    VVHD_HIP(hipMemcpyAsync(v_out, dout, (size_t)n * 16, hipMemcpyDeviceToHost, c.stream));
    VVHD_HIP(hipStreamSynchronize(c.stream));
}

void direct_sum(Ctx& c, const int64_t* targets, int64_t nt, double* v_out) {
    if (nt == 0) return;
    if (c.nsim > 1) {   // each target sums over its own simulation
        std::vector<int64_t> all;
        if (!targets) {
            all.resize(nt);
            for (int64_t k = 0; k < nt; k++) all[k] = k;
            targets = all.data();
        }
        std::vector<std::vector<int64_t>> pos(c.nsim);
        for (int64_t k = 0; k < nt; k++) {
            VVHD_CHECK(targets[k] >= 0 && targets[k] < c.n, VVHD_EINVAL, "direct_sum: target out of range");
            const int s = (int)(std::upper_bound(c.sim_off_h.begin(), c.sim_off_h.end(), targets[k]) -
                                c.sim_off_h.begin()) - 1;
            pos[s].push_back(k);
        }
        c.lscratch0.grow((size_t)nt + 1);
        c.tmp.grow((size_t)nt * 16);
        double* dout = (double*)c.tmp.p;
        std::vector<double> buf;
        std::vector<int64_t> tg;
        for (int s = 0; s < c.nsim; s++) {
            const int64_t m = (int64_t)pos[s].size();
            if (!m) continue;
            tg.resize(m);
            for (int64_t k = 0; k < m; k++) tg[k] = targets[pos[s][k]];
            VVHD_HIP(hipMemcpyAsync(c.lscratch0.p, tg.data(), (size_t)m * 8, hipMemcpyHostToDevice, c.stream));
            hipLaunchKernelGGL(k_direct_sum, dim3((unsigned)((m + kDsBlock - 1) / kDsBlock)), dim3(kDsBlock), 0,
                               c.stream, c.x.p, c.y.p, c.g.p, c.ie.p, c.sim_off_h[s], c.sim_off_h[s + 1],
                               (const int64_t*)c.lscratch0.p, m, dout);
            buf.resize(2 * m);
            VVHD_HIP(hipMemcpyAsync(buf.data(), dout, (size_t)m * 16, hipMemcpyDeviceToHost, c.stream));
            VVHD_HIP(hipStreamSynchronize(c.stream));
            for (int64_t k = 0; k < m; k++) {
                v_out[2 * pos[s][k]] = buf[2 * k];
                v_out[2 * pos[s][k] + 1] = buf[2 * k + 1];
            }
        }
        return;
    }
    c.lscratch0.grow((size_t)nt + 1);
    c.tmp.grow((size_t)nt * 16);
    if (targets) VVHD_HIP(hipMemcpyAsync(c.lscratch0.p, targets, (size_t)nt * 8, hipMemcpyHostToDevice, c.stream));
    double* dout = (double*)c.tmp.p;
    hipLaunchKernelGGL(k_direct_sum, dim3((unsigned)((nt + kDsBlock - 1) / kDsBlock)), dim3(kDsBlock), 0, c.stream,
                       c.x.p, c.y.p, c.g.p, c.ie.p, (int64_t)0, c.n, targets ? c.lscratch0.p : nullptr, nt, dout);
    VVHD_HIP(hipMemcpyAsync(v_out, dout, (size_t)nt * 16, hipMemcpyDeviceToHost, c.stream));
    VVHD_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace vvhd
