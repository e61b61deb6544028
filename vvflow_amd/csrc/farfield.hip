// farfield.hip -- learned far-field correction of the convective velocity.
//
// The plain path (convective.hip) evaluates the far field of a leaf as the
// reference does: a first-order Taylor expansion about the leaf centre of
// the per-sign monopoles computed when the tree was built
// (MConvectiveFast.cpp:48-86, TSortedTree.cpp:150-197).  The learned path
// replaces it, for every (target, far node) pair, by
//   * the exact field at the target of the node's per-sign monopoles, with
//     the moments recomputed after the epsilon pass (merges move and remove
//     circulation after the tree is built) -- the point-query form of the
//     reference's far field (MConvectiveFast::velocity / far_nodes_influence,
//     :20-34, :139-151);
//   * plus a correction predicted by a small MLP from the node's moments and
//     the target offset: 12 inputs, two hidden layers of 64 SiLU units, 2
//     outputs, with FP32 accumulation on the matrix cores (training: bf16
//     weights and activations on v_mfma_f32_32x32x16_bf16; inference: bf16
//     features and layer-1 weights on v_mfma_f32_16x16x32_bf16, fp16 hidden
//     values and layer-2/3 weights on v_mfma_f32_16x16x32_f16).  The network runs per
//     (leaf, far node) at the leaf centre with forward-mode tangents, and
//     each target takes the first-order expansion of the correction about
//     the centre (k_far_learned, k_finish_learned).
// The MLP's target is the residual between the direct Lamb-kernel sum over
// the node's vortices (FP64) and the monopole field, expressed in the frame
// of the target direction and scaled by S d^2 / (2 pi r^3) (S = |G+| + |G-|,
// d^2 = h^2 + w^2 + smoothing radius^2, r = target distance), so that inputs
// and outputs are O(1) and scale-free.  The leading part of that residual is
// the nodes' second moments and the blob smoothing; the network also learns
// the higher-order remainder.  A pair whose moments are degenerate (no
// circulation, zero size) falls back to the monopole alone.
//
// Inputs (before standardisation by the model's mean / deviation):
//   0,1   G+ / S, G- / S
//   2..5  per-sign dipole about the node centre, G (z_s - c) / (S d), rotated
//   6,7   sum of per-sign second moments about the own centroids,
//         (Q+ + Q-) / (S d^2), rotated twice
//   8,9   per-sign smoothing moment sum(g eps^2) / (S d^2)
//   10    d / r
//   11    (h - w) / (h + w)
// "rotated" = multiplied by conj(u), u the unit vector from the node centre
// to the target.  The correction is (y0 + i y1) * ysd * S d^2/(2 pi r^3) * u.
//
// MFMA data flow of one wave (training: 32 samples per tile on 32x32x16;
// inference: 16 columns per tile on 16x16x32, the same scheme): every layer is
// computed transposed, H^T = W X^T, so the hidden units are the M dimension
// and the samples the N dimension.  The accumulator of one layer (lane holds
// sample l%32, units 8j + 4(l/32) + i) feeds the next layer's B operand
// without a transpose: the K order of the next layer is permuted to the
// order the lanes already hold, and the weights are packed in that order.
// Training (fused forward + backward per tile) stages the activations and
// deltas through LDS for the weight-gradient products, whose reduction runs
// over the samples.

#include "api.hpp"

#include <cstring>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

namespace vvhd {
namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr double kPi = 3.14159265358979323846;
constexpr double k1_2Pi = 1. / (2. * kPi);

// The inference weights are folded so that the SiLU exponent needs no scaling
// and the features go in centred but not scaled (L = log2 e): layer 1 takes
// W1' = -L W1 diag(1/sd) and b1' = -L b1 on x - mu, so its outputs are
// z' = -L z and 2^z' = e^-z; the hidden values -L h feed layer 2 unchanged
// (W2' = W2, b2' = -L b2) and layer 3 takes W3' = -W3 / L.  The tangent
// chains (raw feature tangents) carry the same -L factor.  Training uses the
// plain parameters.
constexpr float kLog2e = 1.44269504088896341f;
constexpr float kLn2 = 0.693147180559945309f;   // 1 / L

// parameter layout (also documented in vvhd_gpu.h)
constexpr int kFeat = 12, kHid = 64, kOut = 2;
constexpr int oW1 = 0;                    // [64][12]
constexpr int oB1 = oW1 + kHid * kFeat;   // [64]
constexpr int oW2 = oB1 + kHid;           // [64][64]  (out, in)
constexpr int oB2 = oW2 + kHid * kHid;    // [64]
constexpr int oW3 = oB2 + kHid;           // [2][64]
constexpr int oB3 = oW3 + kOut * kHid;    // [2]
constexpr int kTrain = oB3 + kOut;        // trainable parameters
constexpr int oMu = kTrain;               // [12] feature means
constexpr int oSd = oMu + kFeat;          // [12] feature deviations
constexpr int oYsd = oSd + kFeat;         // [2]  output scales
constexpr int kParams = oYsd + kOut;
static_assert(kParams == VVHD_FAR_PARAMS, "parameter layout");

// bf16 fragment sets: 64 lanes x 8 elements each
constexpr int kSet = 64 * 8;
constexpr int fA1 = 0;               // 2 sets: W1, M blocks of 32 hidden units, K = 16 features
constexpr int fA2 = fA1 + 2 * kSet;  // 8 sets: W2, (M block, K step)
constexpr int fA3 = fA2 + 8 * kSet;  // 4 sets: W3 (rows 0,1 real), K steps
constexpr int fA3T = fA3 + 4 * kSet; // 2 sets: W3^T, M blocks of hidden units, K = outputs
constexpr int fA2T = fA3T + 2 * kSet;// 8 sets: W2^T, (M block of input units, K step over output units)
// inference fragments for v_mfma_f32_16x16x32_bf16 (M blocks of 16 units, K steps of 32)
constexpr int fB1 = fA2T + 8 * kSet; // 4 sets: W1, M blocks, K = 32 feature slots
constexpr int fB2 = fB1 + 4 * kSet;  // 8 sets: W2, (M block b, K step st) = 2 b + st
constexpr int fB3 = fB2 + 8 * kSet;  // 2 sets: W3 (rows 0, 1 real), K steps
constexpr int kFragElems = fB3 + 2 * kSet;

constexpr int kMom = 12;   // gp, gm, zpx, zpy, zmx, zmy, qpr, qpi, qmr, qmi, ep, em
constexpr int kPartStride = kTrain + 2;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));

// Eight floats to bf16 (round to nearest even), converted in pairs, and the
// whole vector zeroed when keep is false (one select per packed word).
__device__ inline bf16x8 pack_bf16x8(const float (&v)[8], bool keep) {
    u32x4 w;
    for (int q = 0; q < 4; q++) {
        const bf16x2 p = __builtin_convertvector((f32x2){v[2 * q], v[2 * q + 1]}, bf16x2);
        const uint32_t u = __builtin_bit_cast(uint32_t, p);
        w[q] = keep ? u : 0u;
    }
    return __builtin_bit_cast(bf16x8, w);
}

inline int nblocks(int64_t n, int b = kBlock) { return (int)std::max<int64_t>(1, (n + b - 1) / b); }

__device__ inline f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ inline f32x4 mfma16h(f16x8 a, f16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// 16x16x32 layout: A row l%16, K 8(l/16)+e; B column l%16, K 8(l/16)+e;
// C column l%16, rows 4(l/16)+i.  Hidden unit held at (K step st, lane
// group g, element e) when a layer's output blocks feed the next layer.
__host__ __device__ inline int chain16(int st, int g, int e) { return 16 * (2 * st + (e >> 2)) + 4 * g + (e & 3); }

__device__ inline f32x16 mfma32(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// Round to nearest even, as the host-side model code does.
__host__ __device__ inline uint16_t bf16_bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline __bf16 to_bf16(float f) {
    const uint16_t b = bf16_bits(f);
    __bf16 r;
    __builtin_memcpy(&r, &b, 2);
    return r;
}
__device__ inline float sigmoid_f(float z) { return __builtin_amdgcn_rcpf(1.f + __expf(-z)); }
__device__ inline float silu_f(float z) { return z * sigmoid_f(z); }
__device__ inline float dsilu_f(float z) {
    const float s = sigmoid_f(z);
    return s + z * s * (1.f - s);
}

// Hidden unit held at (K step st, lane group g, element e) by the chained
// layout: block st/2 of the previous accumulator, accumulator index 8(st%2)+e.
__host__ __device__ inline int chain_unit(int st, int g, int e) {
    return 32 * (st >> 1) + 8 * (2 * (st & 1) + (e >> 2)) + 4 * g + (e & 3);
}

// ------------------------------------------------------------ node moments

struct MomView {
    double* m;
    size_t st;
    __device__ double& operator()(int k, int id) const { return m[(size_t)k * st + id]; }
};

// Per-sign moments of one node from its vortices (a leaf).
__device__ inline void leaf_moments(int id, const int32_t* v0, const int32_t* v1, const double* nx, const double* ny,
                                    const double* x, const double* y, const double* g, const double* ie, MomView M) {
#pragma clang fp contract(off)
    for (int s = 0; s < 2; s++) {
        double G = 0, zx, zy, qr = 0, qi = 0, E = 0;
        double sx = 0, sy = 0;
        for (int k = v0[id]; k < v1[id]; k++) {
            const double gk = g[k];
            if (s == 0 ? !(gk > 0) : !(gk < 0)) continue;
            G += gk;
            sx += gk * x[k];
            sy += gk * y[k];
            const double ek = 1. / ie[k];
            E += gk * (ek * ek);
        }
        if (G != 0) {
            zx = sx / G;
            zy = sy / G;
            for (int k = v0[id]; k < v1[id]; k++) {
                const double gk = g[k];
                if (s == 0 ? !(gk > 0) : !(gk < 0)) continue;
                const double dx = x[k] - zx, dy = y[k] - zy;
                qr += gk * (dx * dx - dy * dy);
                qi += gk * (2 * dx * dy);
            }
        } else {
            zx = nx[id];
            zy = ny[id];
        }
        M(s, id) = G;
        M(2 + 2 * s, id) = zx;
        M(3 + 2 * s, id) = zy;
        M(6 + 2 * s, id) = qr;
        M(7 + 2 * s, id) = qi;
        M(10 + s, id) = E;
    }
}

// Every leaf from its vortices, in one launch.
__global__ void k_moments_leaves(const int32_t* leaf_node, int nleaf, const int32_t* v0, const int32_t* v1,
                                 const double* nx, const double* ny, const double* x, const double* y, const double* g,
                                 const double* ie, MomView M) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nleaf) return;
    leaf_moments(leaf_node[l], v0, v1, nx, ny, x, y, g, ie, M);
}

// The internal nodes of one tree level from their two children (parallel-axis
// combination of the second moments); leaves are done beforehand.
__global__ void k_moments_level(const int32_t* ch1, const int32_t* split, const double* nx, const double* ny,
                                MomView M, int lb, int le) {
#pragma clang fp contract(off)
    const int id = lb + blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= le || split[id] != 2) return;
    for (int s = 0; s < 2; s++) {
        double G, zx, zy, qr = 0, qi = 0, E;
        const int a = ch1[id], b = a + 1;
        const double Ga = M(s, a), Gb = M(s, b);
        G = Ga + Gb;
        E = M(10 + s, a) + M(10 + s, b);
        if (G != 0) {
            const double ax = M(2 + 2 * s, a), ay = M(3 + 2 * s, a), bx = M(2 + 2 * s, b), by = M(3 + 2 * s, b);
            zx = (Ga * ax + Gb * bx) / G;
            zy = (Ga * ay + Gb * by) / G;
            const double dax = ax - zx, day = ay - zy, dbx = bx - zx, dby = by - zy;
            qr = M(6 + 2 * s, a) + M(6 + 2 * s, b) + Ga * (dax * dax - day * day) + Gb * (dbx * dbx - dby * dby);
            qi = M(7 + 2 * s, a) + M(7 + 2 * s, b) + Ga * (2 * dax * day) + Gb * (2 * dbx * dby);
        } else {
            zx = nx[id];
            zy = ny[id];
        }
        M(s, id) = G;
        M(2 + 2 * s, id) = zx;
        M(3 + 2 * s, id) = zy;
        M(6 + 2 * s, id) = qr;
        M(7 + 2 * s, id) = qi;
        M(10 + s, id) = E;
    }
}

// -------------------------------------------------------------- features

struct PairFeat {
    float f[kFeat];
    double ux, uy, scale;
};

// Features of the pair (target t, node F); false when the node's moments are
// degenerate (the pair then takes the monopole field alone).
__device__ bool far_features(const MomView M, int F, double cx, double cy, double h, double w, double tx, double ty,
                             PairFeat& P) {
#pragma clang fp contract(off)
    const double gp = M(0, F), gm = M(1, F);
    const double S = fabs(gp) + fabs(gm);
    const double dx = tx - cx, dy = ty - cy;
    const double r = sqrt(dx * dx + dy * dy);
    if (!(S > 0) || !(r > 0)) return false;
    const double ep = M(10, F), em = M(11, F);
    const double d2 = h * h + w * w + (ep - em) / S;
    if (!(d2 > 0) || !isfinite(d2)) return false;
    const double ux = dx / r, uy = dy / r;
    const double sd = sqrt(d2);
    const double inv1 = 1. / (S * sd), inv2 = 1. / (S * d2);
    double ax = M(2, F) - cx, ay = M(3, F) - cy;
    P.f[2] = (float)(gp * (ax * ux + ay * uy) * inv1);
    P.f[3] = (float)(gp * (ay * ux - ax * uy) * inv1);
    ax = M(4, F) - cx;
    ay = M(5, F) - cy;
    P.f[4] = (float)(gm * (ax * ux + ay * uy) * inv1);
    P.f[5] = (float)(gm * (ay * ux - ax * uy) * inv1);
    const double qr = M(6, F) + M(8, F), qi = M(7, F) + M(9, F);
    const double c2r = ux * ux - uy * uy, c2i = -2 * ux * uy;
    P.f[6] = (float)((qr * c2r - qi * c2i) * inv2);
    P.f[7] = (float)((qr * c2i + qi * c2r) * inv2);
    P.f[0] = (float)(gp / S);
    P.f[1] = (float)(gm / S);
    P.f[8] = (float)(ep * inv2);
    P.f[9] = (float)(em * inv2);
    P.f[10] = (float)(sd / r);
    P.f[11] = (float)((h - w) / fmax(h + w, 1e-300));
    P.ux = ux;
    P.uy = uy;
    P.scale = S * d2 / (2 * kPi * r * r * r);
    return true;
}

// ------------------------------------------------------------------- MLP

struct Weights {
    bf16x8 a1[2], a2[8], a3[4];
};

__device__ inline bf16x8 load_frag(const uint16_t* frag, int set, int lane) {
    const uint4 v = reinterpret_cast<const uint4*>(frag)[set * 64 + lane];
    bf16x8 r;
    __builtin_memcpy(&r, &v, 16);
    return r;
}

__device__ inline void load_weights(const uint16_t* frag, Weights& W) {
    const int l = lane_id();
    for (int b = 0; b < 2; b++) W.a1[b] = load_frag(frag, fA1 / kSet + b, l);
    for (int b = 0; b < 8; b++) W.a2[b] = load_frag(frag, fA2 / kSet + b, l);
    for (int b = 0; b < 4; b++) W.a3[b] = load_frag(frag, fA3 / kSet + b, l);
}

struct Fwd {
    f32x16 c1[2], c2[2], c3;
    bf16x8 h1[4], h2[4];
};

// Forward pass of one tile; x holds this lane's 8 inputs (K slots 8g..8g+7)
// of sample l%32.  Outputs are c3[0], c3[1] in lanes 0..31.
__device__ inline void mlp_forward(const Weights& W, const float* P, bf16x8 x, Fwd& s) {
    const int g = lane_id() >> 5;
    for (int b = 0; b < 2; b++) {
        for (int j = 0; j < 4; j++)
            for (int i = 0; i < 4; i++) s.c1[b][4 * j + i] = P[oB1 + 32 * b + 8 * j + 4 * g + i];
        s.c1[b] = mfma32(W.a1[b], x, s.c1[b]);
    }
    for (int st = 0; st < 4; st++)
        for (int e = 0; e < 8; e++) s.h1[st][e] = to_bf16(silu_f(s.c1[st >> 1][8 * (st & 1) + e]));
    for (int b = 0; b < 2; b++) {
        for (int j = 0; j < 4; j++)
            for (int i = 0; i < 4; i++) s.c2[b][4 * j + i] = P[oB2 + 32 * b + 8 * j + 4 * g + i];
        for (int st = 0; st < 4; st++) s.c2[b] = mfma32(W.a2[4 * b + st], s.h1[st], s.c2[b]);
    }
    for (int st = 0; st < 4; st++)
        for (int e = 0; e < 8; e++) s.h2[st][e] = to_bf16(silu_f(s.c2[st >> 1][8 * (st & 1) + e]));
    for (int j = 0; j < 4; j++)
        for (int i = 0; i < 4; i++) {
            const int row = 8 * j + 4 * g + i;
            s.c3[4 * j + i] = row < kOut ? P[oB3 + row] : 0.f;
        }
    for (int st = 0; st < 4; st++) s.c3 = mfma32(W.a3[st], s.h2[st], s.c3);
}

// Standardised bf16 inputs of this lane's K slots (8g .. 8g+7).
__device__ inline bf16x8 input_frag(const float* f, bool ok, const float* P) {
    const bool hi = (lane_id() >> 5) != 0;
    bf16x8 x;
    for (int e = 0; e < 8; e++) {   // constant indices keep f in registers
        const float fv = hi ? (e + 8 < kFeat ? f[(e + 8) % kFeat] : 0.f) : f[e];
        const int k = hi ? e + 8 : e;
        float v = 0.f;
        if (ok && k < kFeat) v = (fv - P[oMu + k]) / P[oSd + k];
        x[e] = to_bf16(v);
    }
    return x;
}

// FP32 master parameters -> bf16 fragments in the layouts above.
__global__ void k_pack_model(const float* P, uint16_t* frag) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= kFragElems) return;
    int set, rem;
    float v = 0.f;
    auto decode = [&](int base) {
        set = (t - base) / kSet;
        rem = (t - base) % kSet;
    };
    if (t < fA2) {
        decode(fA1);
    } else if (t < fA3) {
        decode(fA2);
    } else if (t < fA3T) {
        decode(fA3);
    } else if (t < fA2T) {
        decode(fA3T);
    } else if (t < fB1) {
        decode(fA2T);
    } else if (t < fB2) {
        decode(fB1);
    } else if (t < fB3) {
        decode(fB2);
    } else {
        decode(fB3);
    }
    const int lane = rem / 8, e = rem % 8, m = lane % 32, g = lane / 32;
    if (t >= fB1) {   // 16x16x32 inference sets
        const int m16 = lane % 16, g16 = lane / 16;
        if (t < fB2) {   // W1' block set: rows 16 set + m16, K = feature 8 g16 + e
            const int k = 8 * g16 + e;
            if (k < kFeat) v = (-kLog2e * P[oW1 + (16 * set + m16) * kFeat + k]) * (1.f / P[oSd + k]);
        } else if (t < fB3) {   // W2: block b = set / 2, K step st = set % 2
            const int b = set / 2, st = set % 2;
            v = P[oW2 + (16 * b + m16) * kHid + chain16(st, g16, e)];
        } else {   // W3' rows 0, 1, K step set
            if (m16 < kOut) v = P[oW3 + m16 * kHid + chain16(set, g16, e)] * -kLn2;
        }
        // layer 1 takes bf16 features; layers 2 and 3 take the fp16 hidden values
        frag[t] = t < fB2 ? bf16_bits(v) : __builtin_bit_cast(uint16_t, (_Float16)v);
        return;
    }
    if (t < fA2) {   // W1: A[m = unit][k = feature]
        const int k = 8 * g + e;
        if (k < kFeat) v = P[oW1 + (32 * set + m) * kFeat + k];
    } else if (t < fA3) {   // W2: block b = set / 4, K step st = set % 4
        const int b = set / 4, st = set % 4;
        v = P[oW2 + (32 * b + m) * kHid + chain_unit(st, g, e)];
    } else if (t < fA3T) {   // W3 (rows 0, 1)
        if (m < kOut) v = P[oW3 + m * kHid + chain_unit(set, g, e)];
    } else if (t < fA2T) {   // W3^T: A[m = hidden unit][k = output]
        const int k = 8 * g + e;
        if (k < kOut) v = P[oW3 + k * kHid + 32 * set + m];
    } else {   // W2^T: A[m = input unit][k = output unit in chained order]
        const int b = set / 4, st = set % 4;
        v = P[oW2 + chain_unit(st, g, e) * kHid + 32 * b + m];
    }
    frag[t] = bf16_bits(v);
}

// -------------------------------------------------------------- inference

// Per-node record of the inference path, built once per step from the
// moments: everything in the features that does not depend on the target,
// in FP32 (the features are rounded to bf16 afterwards), plus the monopoles
// relative to the node centre.
struct FarNode {
    double cx, cy;
    float f0, f1;           // G+/S, G-/S
    float dpx, dpy;         // G+ (z+ - c) / (S d), unrotated
    float dmx, dmy;         // G- (z- - c) / (S d)
    float qqx, qqy;         // (Q+ + Q-) / (S d^2)
    float e8, e9;           // E+ / (S d^2), E- / (S d^2)
    float sd, asp;          // d, (h - w) / (h + w)
    float scl, ok;          // S d^2 / (2 pi); 1 when the features are defined
    float gpk, gmk;         // G+ / (2 pi), G- / (2 pi)
    float zpx, zpy, zmx, zmy;   // z+ - c, z- - c
};
static_assert(sizeof(FarNode) == 96, "FarNode layout");

__global__ void k_far_nodes(TreeDev T, MomView M, int nnodes, FarNode* out) {
#pragma clang fp contract(off)
    const int F = blockIdx.x * blockDim.x + threadIdx.x;
    if (F >= nnodes) return;
    FarNode r;
    const double cx = T.nx[F], cy = T.ny[F], h = T.nh[F], w = T.nw[F];
    const double gp = M(0, F), gm = M(1, F);
    const double S = fabs(gp) + fabs(gm);
    const double ep = M(10, F), em = M(11, F);
    const double d2 = S > 0 ? h * h + w * w + (ep - em) / S : 0.;
    const bool ok = S > 0 && d2 > 0 && isfinite(d2);
    const double sd = ok ? sqrt(d2) : 1., inv1 = ok ? 1. / (S * sd) : 0., inv2 = ok ? 1. / (S * d2) : 0.;
    r.cx = cx;
    r.cy = cy;
    r.f0 = ok ? (float)(gp / S) : 0.f;
    r.f1 = ok ? (float)(gm / S) : 0.f;
    r.dpx = (float)(gp * (M(2, F) - cx) * inv1);
    r.dpy = (float)(gp * (M(3, F) - cy) * inv1);
    r.dmx = (float)(gm * (M(4, F) - cx) * inv1);
    r.dmy = (float)(gm * (M(5, F) - cy) * inv1);
    r.qqx = (float)((M(6, F) + M(8, F)) * inv2);
    r.qqy = (float)((M(7, F) + M(9, F)) * inv2);
    r.e8 = (float)(ep * inv2);
    r.e9 = (float)(em * inv2);
    r.sd = (float)sd;
    r.asp = (float)((h - w) / fmax(h + w, 1e-300));
    r.scl = ok ? (float)(S * d2 / (2 * kPi)) : 0.f;
    r.ok = ok ? 1.f : 0.f;
    r.gpk = (float)(gp * k1_2Pi);
    r.gmk = (float)(gm * k1_2Pi);
    r.zpx = (float)(M(2, F) - cx);
    r.zpy = (float)(M(3, F) - cy);
    r.zmx = (float)(M(4, F) - cx);
    r.zmy = (float)(M(5, F) - cy);
    out[F] = r;
}

// A zero the compiler cannot see through, renewed at every call.  Adding it
// to the LDS indices of the weight fragments and biases makes every tile
// read them from LDS again instead of keeping all 14 fragments and 32 biases
// in registers across the far-list loop: that frees about 90 VGPRs per lane,
// and four waves per SIMD instead of two hide the latency of the MFMA ->
// SiLU -> MFMA chains (2.58 against 2.84 ms at the 1M wake).  The LDS
// traffic this adds (22 x 1 KiB per tile and wave) is a quarter of the
// LDS bandwidth.
__device__ inline int lds_offset_zero() {
    int o = 0;
    asm volatile("" : "+v"(o));
    return o;
}

// Inference forward of one tile of 16 columns with two forward-mode
// tangents, on the matrix cores, with the folded weights: the bf16 features
// and their tangents tx, ty pass through W1 (v_mfma_f32_16x16x32_bf16);
// each SiLU scales the tangents by its derivative s (1 + z (1 - s)),
// s = sigmoid(z).  The hidden values and tangents are fp16 and feed layers
// 2 and 3 (v_mfma_f32_16x16x32_f16, fp16 weights): an activation pair takes
// its exponentials and reciprocals in FP32, then the product, the
// derivative and the tangent products as packed fp16 pairs, whose results
// are already the next layer's operand.  A layer's four output blocks (16
// units each) feed the next layer's two K steps without a transpose
// (chain16).  Outputs (value, x tangent, y tangent) in lanes 0..15.
// The weight fragments come from LDS (sw: W1, W2, W3 sets, 64 lanes each).
//
// The SiLU of z' = -L z: sigmoid(z) = 1 / (1 + 2^z'), h' = z' s = -L h, and
// the derivative s + h (1 - s) = s + h' (s ln2 - ln2).
__device__ inline void silu_pair(float z0, float z1, f16x2& H, f16x2& D) {
    const float s0 = __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(z0));
    const float s1 = __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(z1));
    const f16x2 Z = __builtin_convertvector((f32x2){z0, z1}, f16x2);
    const f16x2 S = __builtin_convertvector((f32x2){s0, s1}, f16x2);
    const f16x2 ln2 = {(_Float16)kLn2, (_Float16)kLn2};
    H = Z * S;
    D = __builtin_elementwise_fma(H, __builtin_elementwise_fma(S, ln2, -ln2), S);
}
__device__ inline f16x2 tangent_pair(f16x2 D, float c0, float c1) {
    return D * __builtin_convertvector((f32x2){c0, c1}, f16x2);
}

__device__ inline void mlp_infer_jvp16(const uint4* sw, const float* sb1, const float* sb2, float b30, float b31,
                                       bf16x8 x, bf16x8 tx, bf16x8 ty, float (&y)[2], float (&yx)[2],
                                       float (&yy)[2]) {
    const int o = lds_offset_zero();
    const int ln = lane_id() + o, g = lane_id() >> 4;
    auto Wb = [&](int set) { return __builtin_bit_cast(bf16x8, sw[set * kWave + ln]); };
    auto Wh = [&](int set) { return __builtin_bit_cast(f16x8, sw[set * kWave + ln]); };
    u32x4 h[2], hx[2], hy[2];   // fp16 pairs: word q of K step st holds elements 2q, 2q + 1
    for (int b = 0; b < 4; b++) {
        f32x4 c, cx = {}, cy = {};
        for (int i = 0; i < 4; i++) c[i] = sb1[16 * b + 4 * g + i + o];
        const bf16x8 a = Wb((fB1 - fB1) / kSet + b);
        c = mfma16(a, x, c);
        cx = mfma16(a, tx, cx);
        cy = mfma16(a, ty, cy);
        for (int i = 0; i < 4; i += 2) {
            f16x2 H, D;
            silu_pair(c[i], c[i + 1], H, D);
            const int st = b >> 1, q = 2 * (b & 1) + i / 2;
            h[st][q] = __builtin_bit_cast(uint32_t, H);
            hx[st][q] = __builtin_bit_cast(uint32_t, tangent_pair(D, cx[i], cx[i + 1]));
            hy[st][q] = __builtin_bit_cast(uint32_t, tangent_pair(D, cy[i], cy[i + 1]));
        }
    }
    u32x4 h2[2], h2x[2], h2y[2];
    for (int b = 0; b < 4; b++) {
        f32x4 c, cx = {}, cy = {};
        for (int i = 0; i < 4; i++) c[i] = sb2[16 * b + 4 * g + i + o];
        for (int st = 0; st < 2; st++) {
            const f16x8 a = Wh((fB2 - fB1) / kSet + 2 * b + st);
            c = mfma16h(a, __builtin_bit_cast(f16x8, h[st]), c);
            cx = mfma16h(a, __builtin_bit_cast(f16x8, hx[st]), cx);
            cy = mfma16h(a, __builtin_bit_cast(f16x8, hy[st]), cy);
        }
        for (int i = 0; i < 4; i += 2) {
            f16x2 H, D;
            silu_pair(c[i], c[i + 1], H, D);
            const int st = b >> 1, q = 2 * (b & 1) + i / 2;
            h2[st][q] = __builtin_bit_cast(uint32_t, H);
            h2x[st][q] = __builtin_bit_cast(uint32_t, tangent_pair(D, cx[i], cx[i + 1]));
            h2y[st][q] = __builtin_bit_cast(uint32_t, tangent_pair(D, cy[i], cy[i + 1]));
        }
    }
    f32x4 c = {}, cx = {}, cy = {};
    c[0] = g == 0 ? b30 : 0.f;
    c[1] = g == 0 ? b31 : 0.f;
    for (int st = 0; st < 2; st++) {
        const f16x8 a = Wh((fB3 - fB1) / kSet + st);
        c = mfma16h(a, __builtin_bit_cast(f16x8, h2[st]), c);
        cx = mfma16h(a, __builtin_bit_cast(f16x8, h2x[st]), cx);
        cy = mfma16h(a, __builtin_bit_cast(f16x8, h2y[st]), cy);
    }
    y[0] = c[0];
    y[1] = c[1];
    yx[0] = cx[0];
    yx[1] = cx[1];
    yy[0] = cy[0];
    yy[1] = cy[1];
}

// The value chain alone (no tangents), same tiles, fragments and rounding.
__device__ inline void mlp_infer_val16(const uint4* sw, const float* sb1, const float* sb2, float b30, float b31,
                                       bf16x8 x, float (&y)[2]) {
    const int o = lds_offset_zero();
    const int ln = lane_id() + o, g = lane_id() >> 4;
    auto Wb = [&](int set) { return __builtin_bit_cast(bf16x8, sw[set * kWave + ln]); };
    auto Wh = [&](int set) { return __builtin_bit_cast(f16x8, sw[set * kWave + ln]); };
    u32x4 h[2];
    for (int b = 0; b < 4; b++) {
        f32x4 c;
        for (int i = 0; i < 4; i++) c[i] = sb1[16 * b + 4 * g + i + o];
        c = mfma16(Wb((fB1 - fB1) / kSet + b), x, c);
        for (int i = 0; i < 4; i += 2) {
            f16x2 H, D;
            silu_pair(c[i], c[i + 1], H, D);
            h[b >> 1][2 * (b & 1) + i / 2] = __builtin_bit_cast(uint32_t, H);
        }
    }
    u32x4 h2[2];
    for (int b = 0; b < 4; b++) {
        f32x4 c;
        for (int i = 0; i < 4; i++) c[i] = sb2[16 * b + 4 * g + i + o];
        for (int st = 0; st < 2; st++)
            c = mfma16h(Wh((fB2 - fB1) / kSet + 2 * b + st), __builtin_bit_cast(f16x8, h[st]), c);
        for (int i = 0; i < 4; i += 2) {
            f16x2 H, D;
            silu_pair(c[i], c[i + 1], H, D);
            h2[b >> 1][2 * (b & 1) + i / 2] = __builtin_bit_cast(uint32_t, H);
        }
    }
    f32x4 c = {};
    c[0] = g == 0 ? b30 : 0.f;
    c[1] = g == 0 ? b31 : 0.f;
    for (int st = 0; st < 2; st++) c = mfma16h(Wh((fB3 - fB1) / kSet + st), __builtin_bit_cast(f16x8, h2[st]), c);
    y[0] = c[0];
    y[1] = c[1];
}

// Network part of the learned far field, one wave per leaf: for every far
// node of the leaf, the correction at the leaf centre c and its gradient
// with respect to the target position (forward-mode tangents through the
// network, and the derivatives of the features, of the rotation back to
// the target frame and of the scale S d^2 / (2 pi r^3)), summed over the far
// list.  k_finish_learned gives each target t the first-order expansion
// C + (t - c) . grad C, so the network runs once per (leaf, far node)
// instead of once per (target, far node); the correction varies slowly over
// a leaf, which lies far from the node.
// Feature derivatives: with theta the angle of u, d theta / dp = (-uy, ux)/r;
// the rotated moments turn with theta (d(f2 + i f3)/d theta = f3 - i f2, the
// quadrupole twice as fast), d/r scales as 1/r, the rest is constant.  So
// the feature tangents along x and y are thx d1 + (ux/r) d2 and
// thy d1 + (uy/r) d2 with two directions fixed per node: d1, the rotated
// moments' derivative in theta, and d2, -d/r at input 10.  The network's
// tangents run along d1 and d2 (no products per tile), and the expansion
// combines them.
// A tile holds 16 far nodes (column l%16); lane groups 0 and 1 (lanes
// 16g..16g+15) build inputs 8g..8g+7 and their tangents, groups 2 and 3 the
// zero padding of the K step.  Output per leaf: 6 doubles
// (u, v, du/dx, dv/dx, du/dy, dv/dy); zeros for a leaf without far nodes or
// without a target of non-zero circulation.
// kJvp: the gradient takes the network's own derivative (forward-mode
// tangents, three chains); otherwise the network output is held fixed over
// the leaf and the gradient keeps the rotation and 1/r^3 terms (one chain).
template <bool kJvp>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_far_learned(
    TreeDev T, const FarNode* __restrict__ nodes, const double* __restrict__ g, const uint16_t* frag, const float* P,
    double* __restrict__ leaf_corr) {
    constexpr int kInfSets = (kFragElems - fB1) / kSet;   // W1, W2, W3 inference sets
    __shared__ float sb1[kHid], sb2[kHid], smu[16];
    __shared__ uint4 sw[kInfSets * kWave];
    for (int k = threadIdx.x; k < kHid; k += blockDim.x) {
        sb1[k] = -kLog2e * P[oB1 + k];   // b1' = -L b1, b2' = -L b2
        sb2[k] = -kLog2e * P[oB2 + k];
        if (k < 16) smu[k] = k < kFeat ? P[oMu + k] : 0.f;
    }
    for (int k = threadIdx.x; k < kInfSets * kWave; k += blockDim.x)
        sw[k] = reinterpret_cast<const uint4*>(frag + fB1)[k];
    __syncthreads();
    const int w = threadIdx.x / kWave;
    const int l = xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock + w;
    if (l >= T.nleaf) return;
    const int lane = lane_id(), grp = lane >> 4, n = lane & 15;
    const int L = T.leaf_node[l];
    const int v0 = T.v0[L], v1 = T.v1[L];
    const int64_t fb = T.far_off[l], fe = T.far_off[l + 1];
    double* out = leaf_corr + 6 * (size_t)l;
    bool target = false;
    for (int i = v0 + lane; i < v1; i += kWave) target |= g[i] != 0;
    if (fb == fe || __ballot(target) == 0) {
        if (lane < 6) out[lane] = 0.;
        return;
    }
    const float b30 = P[oB3], b31 = P[oB3 + 1];
    const float ysd0 = P[oYsd], ysd1 = P[oYsd + 1];
    const double px = T.nx[L], py = T.ny[L];
    float acc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, cmp[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};   // Kahan sums
    float mu8[8];   // the means of this lane's eight inputs
    for (int q = 0; q < 8; q++) mu8[q] = smu[(8 * grp + q) & 15];
    // outputs of one tile held by one lane group until they are expanded
    struct Pending {
        float y0, y1, yx0, yx1, yy0, yy1, ux, uy, rinv, scl;
        bool ok;
    } pend{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, false};
    int tc = 0;   // tiles done (uniform)
    // the correction of the held tile in the target frame and its gradient,
    // added to this lane's Kahan sums
    auto expand = [&]() {
        if (pend.ok) {
            const float ux = pend.ux, uy = pend.uy, rinv = pend.rinv;
            const float thx = -uy * rinv, thy = ux * rinv;
            const float sc = pend.scl * rinv * rinv * rinv;
            const float a = pend.y0 * ysd0, b = pend.y1 * ysd1;
            const float U = (a * ux - b * uy) * sc, V = (a * uy + b * ux) * sc;
            const float kx = -3.f * ux * rinv, ky = -3.f * uy * rinv;
            float t[6];
            t[0] = U;
            t[1] = V;
            t[2] = kx * U - thx * V;
            t[3] = kx * V + thx * U;
            t[4] = ky * U - thy * V;
            t[5] = ky * V + thy * U;
            if constexpr (kJvp) {
                // the feature tangents along x and y are thx d1 + (ux / r) d2 and
                // thy d1 + (uy / r) d2, so the output's are the same
                // combinations of its tangents along d1 and d2
                const float urx = ux * rinv, ury = uy * rinv;
                const float gx0 = __builtin_fmaf(thx, pend.yx0, urx * pend.yy0);
                const float gx1 = __builtin_fmaf(thx, pend.yx1, urx * pend.yy1);
                const float gy0 = __builtin_fmaf(thy, pend.yx0, ury * pend.yy0);
                const float gy1 = __builtin_fmaf(thy, pend.yx1, ury * pend.yy1);
                const float ax = gx0 * ysd0, bx = gx1 * ysd1, ay = gy0 * ysd0, by = gy1 * ysd1;
                t[2] += (ax * ux - bx * uy) * sc;
                t[3] += (ax * uy + bx * ux) * sc;
                t[4] += (ay * ux - by * uy) * sc;
                t[5] += (ay * uy + by * ux) * sc;
            }
            for (int q = 0; q < 6; q++) {
                const float yk = t[q] - cmp[q], tk = acc[q] + yk;
                cmp[q] = (tk - acc[q]) - yk;
                acc[q] = tk;
            }
        }
        pend.ok = false;
    };
    // one tile: 16 far nodes (column n)
    auto tile = [&](const FarNode& nd, bool valid) {
        const float dx = (float)(px - nd.cx), dy = (float)(py - nd.cy);
        const float rinv = __builtin_amdgcn_rsqf(dx * dx + dy * dy);
        const float ux = dx * rinv, uy = dy * rinv;
        const bool ok = valid && nd.ok != 0.f && grp < 2;
        float f[8], fx[8], fy[8];
        if (grp == 0) {
            const float c2r = ux * ux - uy * uy, c2i = -2.f * ux * uy;
            f[0] = nd.f0;
            f[1] = nd.f1;
            f[2] = nd.dpx * ux + nd.dpy * uy;
            f[3] = nd.dpy * ux - nd.dpx * uy;
            f[4] = nd.dmx * ux + nd.dmy * uy;
            f[5] = nd.dmy * ux - nd.dmx * uy;
            f[6] = nd.qqx * c2r - nd.qqy * c2i;
            f[7] = nd.qqx * c2i + nd.qqy * c2r;
            if constexpr (kJvp) {
                // direction 1: the rotated moments' derivative in theta
                const float rot[8] = {0.f, 0.f, f[3], -f[2], f[5], -f[4], f[7] + f[7], -(f[6] + f[6])};
                for (int q = 0; q < 8; q++) {
                    fx[q] = rot[q];
                    fy[q] = 0.f;
                }
            }
        } else {
            f[0] = nd.e8;
            f[1] = nd.e9;
            f[2] = nd.sd * rinv;
            f[3] = nd.asp;
            f[4] = f[5] = f[6] = f[7] = 0.f;
            if constexpr (kJvp) {
                // direction 2: d/r's derivative in r, times r
                for (int q = 0; q < 8; q++) fx[q] = fy[q] = 0.f;
                fy[2] = -f[2];
            }
        }
        // centred features; the 1/sd scale is in the folded layer-1 weights
        for (int q = 0; q < 8; q++) f[q] -= mu8[q];
        const bf16x8 xin = pack_bf16x8(f, ok);
        bf16x8 txin, tyin;
        if constexpr (kJvp) {
            txin = pack_bf16x8(fx, ok);
            tyin = pack_bf16x8(fy, ok);
        }
        float y[2], yx[2] = {0.f, 0.f}, yy[2] = {0.f, 0.f};
        if constexpr (kJvp)
            mlp_infer_jvp16(sw, sb1, sb2, b30, b31, xin, txin, tyin, y, yx, yy);
        else
            mlp_infer_val16(sw, sb1, sb2, b30, b31, xin, y);
        // lane group tc % 4 keeps this tile's outputs (from lanes 0..15) and
        // geometry; every fourth tile all 64 lanes expand four tiles at once
        const int k = tc & 3;
        const float o0 = __shfl(y[0], n, kWave), o1 = __shfl(y[1], n, kWave);
        float ox0 = 0.f, ox1 = 0.f, oy0 = 0.f, oy1 = 0.f;
        if constexpr (kJvp) {
            ox0 = __shfl(yx[0], n, kWave);
            ox1 = __shfl(yx[1], n, kWave);
            oy0 = __shfl(yy[0], n, kWave);
            oy1 = __shfl(yy[1], n, kWave);
        }
        if (grp == k) {
            pend.y0 = o0;
            pend.y1 = o1;
            pend.yx0 = ox0;
            pend.yx1 = ox1;
            pend.yy0 = oy0;
            pend.yy1 = oy1;
            pend.ux = ux;
            pend.uy = uy;
            pend.rinv = rinv;
            pend.scl = nd.scl;
            pend.ok = valid && nd.ok != 0.f;
        }
        tc++;
        if (k == 3) expand();
    };
    // tiles of 16 consecutive far-list entries; the next tile's far-list
    // index loads meanwhile (the records themselves load at the tile start:
    // a second record set would cost 24 VGPRs, and the other waves of the
    // SIMD cover the latency).  Past the list end a lane reads entry fb.
    int idx = T.far_list[(fb + n < fe) ? fb + n : fb];
    for (int64_t f0 = fb; f0 < fe; f0 += 16) {
        const int64_t e0 = f0 + n, e1 = f0 + 16 + n;
        const FarNode nd = nodes[idx];
        idx = T.far_list[(e1 < fe) ? e1 : fb];
        tile(nd, e0 < fe);
    }
    expand();   // tiles still held (fewer than four since the last expansion)
    // sums over the 64 lanes (four tiles of 16 columns each), FP64, fixed order
    for (int q = 0; q < 6; q++) {
        double v = (double)acc[q] - (double)cmp[q];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
        if (lane == 0) out[q] = v;
    }
}

// Monopole part of the learned far field of every vortex: the exact field at
// the vortex of the far nodes' per-sign monopoles (FP32 terms from FP64
// differences, block-compensated sums, lane partials combined in FP64); the leaf's network
// correction is added by k_finish_learned.  One wave per leaf; targets in chunks of up to 16 padded to a power of two, the far
// list split over 64 / npad lanes per target; the monopoles of 64 far nodes
// at a time are staged in LDS.
__global__ __launch_bounds__(kBlock) void k_far_mono(TreeDev T, const FarNode* __restrict__ nodes,
                                                     const double* __restrict__ x, const double* __restrict__ y,
                                                     double2* far_out) {
    // far-node monopoles of the current batch of 64 far-list entries
    __shared__ double s_cx[kWavesPerBlock][kWave], s_cy[kWavesPerBlock][kWave];
    __shared__ float s_m[kWavesPerBlock][6][kWave];   // gpk, gmk, zpx, zpy, zmx, zmy
    const int w = threadIdx.x / kWave;
    const int l = xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock + w;
    if (l >= T.nleaf) return;
    const int lane = lane_id();
    const int L = T.leaf_node[l];
    const int v0 = T.v0[L], v1 = T.v1[L];
    const int64_t fb = T.far_off[l], fe = T.far_off[l + 1];
    double* scx = s_cx[w];
    double* scy = s_cy[w];
    float(*sm)[kWave] = s_m[w];
    auto wave_sync = [] {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    for (int base = v0; base < v1; base += 16) {
        const int ntc = min(16, v1 - base);
        const int npad = ntc <= 1 ? 1 : ntc <= 2 ? 2 : ntc <= 4 ? 4 : ntc <= 8 ? 8 : 16;
        const int nper = kWave / npad;
        const int k = lane % npad, fo = lane / npad;
        const bool tvalid = k < ntc;
        const double tx = tvalid ? x[base + k] : 0., ty = tvalid ? y[base + k] : 0.;
        // block-compensated sums: the terms go by FMA into the low-order
        // parts qu, qv, folded into su, sv (Fast2Sum) every eight far nodes
        // of a lane, so every rounding is relative to a few terms
        float su = 0.f, qu = 0.f, sv = 0.f, qv = 0.f;
        int nterm = 0;
        for (int64_t eb = fb; eb < fe; eb += kWave) {
            if (eb + lane < fe) {
                const FarNode& nd = nodes[T.far_list[eb + lane]];
                scx[lane] = nd.cx;
                scy[lane] = nd.cy;
                sm[0][lane] = nd.gpk;
                sm[1][lane] = nd.gmk;
                sm[2][lane] = nd.zpx;
                sm[3][lane] = nd.zpy;
                sm[4][lane] = nd.zmx;
                sm[5][lane] = nd.zmy;
            }
            wave_sync();
            const int nb = (int)min<int64_t>(kWave, fe - eb);
            if (tvalid)
                for (int q = fo; q < nb; q += nper) {
                    if ((++nterm & 7) == 0) {
                        float t = su + qu;
                        qu = qu - (t - su);
                        su = t;
                        t = sv + qv;
                        qv = qv - (t - sv);
                        sv = t;
                    }
                    const float dx = (float)(tx - scx[q]), dy = (float)(ty - scy[q]);
                    const float gp = sm[0][q], gm = sm[1][q];
                    if (gp != 0.f) {
                        const float ex = dx - sm[2][q], ey = dy - sm[3][q];
                        const float fac = gp * __builtin_amdgcn_rcpf(ex * ex + ey * ey);
                        qu = __builtin_fmaf(-ey, fac, qu);
                        qv = __builtin_fmaf(ex, fac, qv);
                    }
                    if (gm != 0.f) {
                        const float ex = dx - sm[4][q], ey = dy - sm[5][q];
                        const float fac = gm * __builtin_amdgcn_rcpf(ex * ex + ey * ey);
                        qu = __builtin_fmaf(-ey, fac, qu);
                        qv = __builtin_fmaf(ex, fac, qv);
                    }
                }
            wave_sync();
        }
        double du = (double)su + (double)qu, dv = (double)sv + (double)qv;
        for (int o = npad; o < kWave; o <<= 1) {
            du += __shfl_xor(du, o, kWave);
            dv += __shfl_xor(dv, o, kWave);
        }
        if (fo == 0 && tvalid) far_out[base + k] = make_double2(du, dv);
    }
}

// v += U_inf; += near / (2 pi); += far (learned path): the far monopoles at
// the vortex plus its leaf's network correction expanded to first order
// about the leaf centre, C + (t - c) . grad C.
__global__ void k_finish_learned(TreeDev T, const double* x, const double* y, const double* g,
                                 const double2* near_in, const double2* far_in, const double* leaf_corr, double* vx,
                                 double* vy, int64_t n, double inf_x, double inf_y) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || g[i] == 0) return;
    const int l = T.vleaf[i];
    const int L = T.leaf_node[l];
    const double* C = leaf_corr + 6 * (size_t)l;
    const double ox = x[i] - T.nx[L], oy = y[i] - T.ny[L];
    double du = far_in[i].x, dv = far_in[i].y;
    du += C[0] + ox * C[2] + oy * C[4];
    dv += C[1] + ox * C[3] + oy * C[5];
    double ax = vx[i] + inf_x, ay = vy[i] + inf_y;
    ax += near_in[i].x * k1_2Pi;
    ay += near_in[i].y * k1_2Pi;
    ax += du;
    ay += dv;
    vx[i] = ax;
    vy[i] = ay;
}

// MLP outputs for given raw features (one tile of 32 per wave).
__global__ __launch_bounds__(kBlock) void k_forward_feats(const float* X, int64_t n, const uint16_t* frag, const float* P,
                                                          float* Y) {
    const int w = threadIdx.x / kWave;
    const int64_t tile = (int64_t)blockIdx.x * kWavesPerBlock + w;
    if (tile * 32 >= n) return;
    const int lane = lane_id(), g = lane >> 5, nn = lane & 31;
    const int64_t j = tile * 32 + nn;
    const bool valid = j < n;
    Weights W;
    load_weights(frag, W);
    float f[kFeat];
    for (int k = 0; k < kFeat; k++) f[k] = valid ? X[j * kFeat + k] : 0.f;
    const bf16x8 xin = input_frag(f, valid, P);
    Fwd s;
    mlp_forward(W, P, xin, s);
    if (g == 0 && valid) {
        Y[2 * j] = s.c3[0];
        Y[2 * j + 1] = s.c3[1];
    }
}

// ---------------------------------------------------------------- training

__device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Fused forward + backward of one tile of 32 samples per wave (block = one
// wave).  Sample j of slot n: idx == nullptr -> j = tile*32 + n over the
// data (full batch), else j = idx-hash of (seed, tile*32 + n) mod ndata.
// Writes the tile's gradient and loss into part[tile].
__global__ __launch_bounds__(64) void k_train_tile(const float* X, const float* Y, const float* Wt, int64_t ndata,
                                                   int64_t nsamples, int hashed, uint64_t seed, float inv_b,
                                                   const uint16_t* frag, const float* P, float* part) {
    __shared__ uint16_t Tz2[kHid][32], Th1[kHid][32], Tz1[kHid][32], Th2[kHid][32], Tx[16][32], Tdy[2][32];
    const int64_t tile = blockIdx.x;
    const int lane = threadIdx.x, g = lane >> 5, n = lane & 31;
    const int64_t slot = tile * 32 + n;
    int64_t j = -1;
    if (slot < nsamples) j = hashed ? (int64_t)(mix64(seed ^ (uint64_t)slot * 0xD1B54A32D192ED03ull) % (uint64_t)ndata)
                                    : slot;
    const bool valid = j >= 0 && j < ndata;
    Weights W;
    load_weights(frag, W);
    float f[kFeat];
    for (int k = 0; k < kFeat; k++) f[k] = valid ? X[j * kFeat + k] : 0.f;
    const bf16x8 xin = input_frag(f, valid, P);
    Fwd s;
    mlp_forward(W, P, xin, s);
    // loss and output delta (lanes 0..31 hold the outputs)
    float dy0 = 0.f, dy1 = 0.f, loss = 0.f;
    if (g == 0 && valid) {
        const float wgt = Wt[j];
        const float r0 = s.c3[0] - Y[2 * j] / P[oYsd], r1 = s.c3[1] - Y[2 * j + 1] / P[oYsd + 1];
        loss = wgt * (r0 * r0 + r1 * r1) * inv_b;
        dy0 = 2.f * wgt * r0 * inv_b;
        dy1 = 2.f * wgt * r1 * inv_b;
    }
    bf16x8 dyf;
    for (int e = 0; e < 8; e++) dyf[e] = to_bf16(g == 0 && e == 0 ? dy0 : (g == 0 && e == 1 ? dy1 : 0.f));
    // dZ2 = (W3^T dY) * silu'(Z2)
    f32x16 dz2[2];
    for (int b = 0; b < 2; b++) {
        f32x16 c = {};
        c = mfma32(load_frag(frag, fA3T / kSet + b, lane), dyf, c);
        for (int q = 0; q < 16; q++) dz2[b][q] = c[q] * dsilu_f(s.c2[b][q]);
    }
    bf16x8 z2b[4];
    for (int st = 0; st < 4; st++)
        for (int e = 0; e < 8; e++) z2b[st][e] = to_bf16(dz2[st >> 1][8 * (st & 1) + e]);
    // dZ1 = (W2^T dZ2) * silu'(Z1)
    f32x16 dz1[2];
    for (int b = 0; b < 2; b++) {
        f32x16 c = {};
        for (int st = 0; st < 4; st++) c = mfma32(load_frag(frag, fA2T / kSet + 4 * b + st, lane), z2b[st], c);
        for (int q = 0; q < 16; q++) dz1[b][q] = c[q] * dsilu_f(s.c1[b][q]);
    }
    // stage bf16 operands [unit][sample] for the weight gradients
    for (int b = 0; b < 2; b++)
        for (int j4 = 0; j4 < 4; j4++)
            for (int i = 0; i < 4; i++) {
                const int u = 32 * b + 8 * j4 + 4 * g + i, q = 4 * j4 + i;
                Tz2[u][n] = bf16_bits(dz2[b][q]);
                Tz1[u][n] = bf16_bits(dz1[b][q]);
                Th1[u][n] = bf16_bits(silu_f(s.c1[b][q]));
                Th2[u][n] = bf16_bits(silu_f(s.c2[b][q]));
            }
    for (int e = 0; e < 8; e++) {
        uint16_t b16;
        const __bf16 v = xin[e];
        __builtin_memcpy(&b16, &v, 2);
        Tx[8 * g + e][n] = b16;
    }
    if (g == 0) {
        Tdy[0][n] = bf16_bits(dy0);
        Tdy[1][n] = bf16_bits(dy1);
    }
    __syncthreads();
    float* out = part + tile * kPartStride;
    auto frag_lds = [&](const uint16_t* row0, int stride, int nrows, int row, int t) {
        // A or B fragment: 8 samples 16t + 8g .. +8 of one row (zero past nrows)
        bf16x8 r;
        for (int e = 0; e < 8; e++) {
            const uint16_t b16 = row < nrows ? row0[row * stride + 16 * t + 8 * g + e] : (uint16_t)0;
            __bf16 v;
            __builtin_memcpy(&v, &b16, 2);
            r[e] = v;
        }
        return r;
    };
    // dW2[out][in]
    for (int bo = 0; bo < 2; bo++)
        for (int bi = 0; bi < 2; bi++) {
            f32x16 c = {};
            for (int t = 0; t < 2; t++)
                c = mfma32(frag_lds(&Tz2[0][0], 32, kHid, 32 * bo + n, t), frag_lds(&Th1[0][0], 32, kHid, 32 * bi + n, t),
                           c);
            for (int j4 = 0; j4 < 4; j4++)
                for (int i = 0; i < 4; i++) {
                    const int o = 32 * bo + 8 * j4 + 4 * g + i, in = 32 * bi + n;
                    out[oW2 + o * kHid + in] = c[4 * j4 + i];
                }
        }
    // dW1[out][feature]
    for (int bo = 0; bo < 2; bo++) {
        f32x16 c = {};
        for (int t = 0; t < 2; t++)
            c = mfma32(frag_lds(&Tz1[0][0], 32, kHid, 32 * bo + n, t), frag_lds(&Tx[0][0], 32, kFeat, n, t), c);
        if (n < kFeat)
            for (int j4 = 0; j4 < 4; j4++)
                for (int i = 0; i < 4; i++) out[oW1 + (32 * bo + 8 * j4 + 4 * g + i) * kFeat + n] = c[4 * j4 + i];
    }
    // dW3[k][in]
    for (int bi = 0; bi < 2; bi++) {
        f32x16 c = {};
        for (int t = 0; t < 2; t++)
            c = mfma32(frag_lds(&Tdy[0][0], 32, kOut, n, t), frag_lds(&Th2[0][0], 32, kHid, 32 * bi + n, t), c);
        if (g == 0)
            for (int i = 0; i < kOut; i++) out[oW3 + i * kHid + 32 * bi + n] = c[i];
    }
    // bias gradients: sums over the 32 samples (lanes of one group)
    for (int b = 0; b < 2; b++)
        for (int q = 0; q < 16; q++) {
            float a1 = dz1[b][q], a2 = dz2[b][q];
            for (int o = 1; o < 32; o <<= 1) {
                a1 += __shfl_xor(a1, o, kWave);
                a2 += __shfl_xor(a2, o, kWave);
            }
            if (n == 0) {
                const int u = 32 * b + 8 * (q >> 2) + 4 * g + (q & 3);
                out[oB1 + u] = a1;
                out[oB2 + u] = a2;
            }
        }
    float b0 = dy0, b1 = dy1, ls = loss;
    for (int o = 1; o < 64; o <<= 1) {
        b0 += __shfl_xor(b0, o, kWave);
        b1 += __shfl_xor(b1, o, kWave);
        ls += __shfl_xor(ls, o, kWave);
    }
    if (lane == 0) {
        out[oB3] = b0;
        out[oB3 + 1] = b1;
        out[kTrain] = ls;
    }
}

// Sum of the tile partials in tile order (deterministic); optional Adam step.
__global__ void k_reduce_adam(const float* part, int ntiles, float* grad, float* P, float* m, float* v, int update,
                              float lr, float bc1, float bc2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > kTrain) return;
    float s = 0.f;
    for (int t = 0; t < ntiles; t++) s += part[(size_t)t * kPartStride + p];
    grad[p] = s;
    if (!update || p == kTrain) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float mm = b1 * m[p] + (1.f - b1) * s;
    const float vv = b2 * v[p] + (1.f - b2) * s * s;
    m[p] = mm;
    v[p] = vv;
    P[p] -= lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
}

// ----------------------------------------------------------- training data

// Far entries: sampled targets of the entry's leaf (hash below threshold).
__global__ void k_pair_count(TreeDev T, const double* g, int64_t nfar, uint64_t seed, uint32_t thresh, int32_t* cnt,
                             const int64_t* off, int32_t* pt, int32_t* pf) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nfar) return;
    int lo = 0, hi = T.nleaf;   // leaf l with far_off[l] <= e < far_off[l + 1]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (T.far_off[mid] <= e)
            lo = mid;
        else
            hi = mid;
    }
    const int L = T.leaf_node[lo];
    const int F = T.far_list[e];
    int c = 0;
    for (int i = T.v0[L]; i < T.v1[L]; i++) {
        if (g[i] == 0) continue;
        const uint64_t h = mix64(seed ^ ((uint64_t)i * 0xD1B54A32D192ED03ull) ^ ((uint64_t)F << 32));
        if ((uint32_t)(h >> 32) < thresh) {
            if (off) {
                pt[off[e] + c] = i;
                pf[off[e] + c] = F;
            }
            c++;
        }
    }
    if (!off) cnt[e] = c;
}

// Labels of sampled pairs, one wave per pair: FP64 direct Lamb sum over the
// node's vortices minus the monopole field, in the model's normalised frame.
__global__ __launch_bounds__(kBlock) void k_pair_label(TreeDev T, MomView M, const double* x, const double* y,
                                                       const double* g, const double* ie, const int32_t* pt,
                                                       const int32_t* pf, int64_t np, float* X, float* Y, float* Wsc) {
#pragma clang fp contract(off)
    const int64_t p = (int64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    if (p >= np) return;
    const int i = pt[p], F = pf[p];
    const double tx = x[i], ty = y[i];
    double su = 0, sv = 0;
    for (int j = T.v0[F] + lane_id(); j < T.v1[F]; j += kWave) {
        const double gj = g[j];
        if (gj == 0) continue;
        const double dx = tx - x[j], dy = ty - y[j];
        const double ej = 1. / ie[j];
        const double fac = gj / (dx * dx + dy * dy + ej * ej);
        su += -dy * fac;
        sv += dx * fac;
    }
    su = wave_sum(su) * k1_2Pi;
    sv = wave_sum(sv) * k1_2Pi;
    if (lane_id() != 0) return;
    double bu = 0, bv = 0;
    for (int s = 0; s < 2; s++) {
        const double G = M(s, F);
        if (G == 0) continue;
        const double dx = tx - M(2 + 2 * s, F), dy = ty - M(3 + 2 * s, F);
        const double fac = G / (dx * dx + dy * dy);
        bu += -dy * fac * k1_2Pi;
        bv += dx * fac * k1_2Pi;
    }
    PairFeat pfe;
    const bool ok = far_features(M, F, T.nx[F], T.ny[F], T.nh[F], T.nw[F], tx, ty, pfe);
    const double ru = su - bu, rv = sv - bv;
    for (int k = 0; k < kFeat; k++) X[p * kFeat + k] = ok ? pfe.f[k] : 0.f;
    if (ok) {
        Y[2 * p] = (float)((ru * pfe.ux + rv * pfe.uy) / pfe.scale);
        Y[2 * p + 1] = (float)((rv * pfe.ux - ru * pfe.uy) / pfe.scale);
        Wsc[p] = (float)pfe.scale;
    } else {
        Y[2 * p] = Y[2 * p + 1] = 0.f;
        Wsc[p] = 0.f;
    }
}

MomView moments(Ctx& c) {
    TreeArrays& t = c.t;
    const size_t st = std::max<size_t>(t.nnodes, 1);
    if (c.fm_stride < st) {
        c.fm.release();
        c.fm.grow(kMom * st);
        c.fm_stride = st;
    }
    MomView M{c.fm.p, c.fm_stride};
    if (t.nleaves_h > 0)
        hipLaunchKernelGGL(k_moments_leaves, dim3(nblocks(t.nleaves_h)), dim3(kBlock), 0, c.stream, t.leaf_node.p,
                           t.nleaves_h, t.v0.p, t.v1.p, t.x.p, t.y.p, c.x.p, c.y.p, c.g.p, c.ie.p, M);
    // internal nodes level by level from the deepest one that has any (the
    // last level holds leaves only)
    for (int d = (int)t.levels.size() - 2; d >= 0; d--) {
        const int lb = t.levels[d].first, le = t.levels[d].second;
        hipLaunchKernelGGL(k_moments_level, dim3(nblocks(le - lb)), dim3(kBlock), 0, c.stream, t.ch1.p, t.split.p,
                           t.x.p, t.y.p, M, lb, le);
    }
    return M;
}

void require_model(Ctx& c) { VVHD_CHECK(c.mlp_loaded, VVHD_ESTATE, "learned far field: no model loaded"); }

void pack_model(Ctx& c) {
    c.mlp_frags.grow(kFragElems);
    hipLaunchKernelGGL(k_pack_model, dim3(nblocks(kFragElems)), dim3(kBlock), 0, c.stream, c.mlp_params.p,
                       c.mlp_frags.p);
}

// Runs the tile kernel over nsamples and reduces (optionally with an Adam step).
float train_pass(Ctx& c, const float* X, const float* Y, const float* Wt, int64_t ndata, int64_t nsamples, int hashed,
                 uint64_t seed, float inv_b, int update, float lr) {
    const int64_t ntiles = (nsamples + 31) / 32;
    VVHD_CHECK(ntiles <= (1 << 20), VVHD_EINVAL, "learned far field: batch too large");
    c.grad_part.grow((size_t)ntiles * kPartStride);
    c.grad_out.grow(kPartStride);
    hipLaunchKernelGGL(k_train_tile, dim3((unsigned)ntiles), dim3(64), 0, c.stream, X, Y, Wt, ndata, nsamples, hashed,
                       seed, inv_b, c.mlp_frags.p, c.mlp_params.p, c.grad_part.p);
    float bc1 = 1.f, bc2 = 1.f;
    if (update) {
        c.adam_t++;
        bc1 = 1.f - std::pow(0.9f, (float)c.adam_t);
        bc2 = 1.f - std::pow(0.999f, (float)c.adam_t);
    }
    hipLaunchKernelGGL(k_reduce_adam, dim3(nblocks(kTrain + 1)), dim3(kBlock), 0, c.stream, c.grad_part.p, (int)ntiles,
                       c.grad_out.p, c.mlp_params.p, c.adam_m.p, c.adam_v.p, update, lr, bc1, bc2);
    if (update) pack_model(c);
    return 0.f;
}

}  // namespace

void farfield_release(Ctx& c) {
    c.fm.release();
    c.fm_stride = 0;
    c.far_v.release();
    c.far_nodes.release();
    c.far_leaf.release();
    c.mlp_params.release();
    c.mlp_frags.release();
    c.mlp_loaded = false;
    for (auto* b : {&c.tr_x, &c.tr_y, &c.tr_w, &c.adam_m, &c.adam_v, &c.grad_part, &c.grad_out}) b->release();
    c.pair_off.release();
    for (auto* b : {&c.pair_cnt, &c.pair_t, &c.pair_f}) b->release();
}

}  // namespace vvhd

using namespace vvhd;

extern "C" {

int vvhd_far_model_set(vvhd_ctx* ctx, const float* params, int32_t n) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    Ctx& c = ctx->c;
    VVHD_CHECK(params && n == kParams, VVHD_EINVAL, "far_model_set: expected VVHD_FAR_PARAMS parameters");
    for (int k = 0; k < kFeat; k++)
        VVHD_CHECK(params[oSd + k] > 0 && std::isfinite(params[oSd + k]), VVHD_EINVAL,
                   "far_model_set: feature deviations must be positive");
    VVHD_HIP(hipSetDevice(c.device));
    c.mlp_params.grow(kParams);
    c.adam_m.grow(kTrain + 1);
    c.adam_v.grow(kTrain + 1);
    VVHD_HIP(hipMemcpyAsync(c.mlp_params.p, params, kParams * 4, hipMemcpyHostToDevice, c.stream));
    VVHD_HIP(hipMemsetAsync(c.adam_m.p, 0, (kTrain + 1) * 4, c.stream));
    VVHD_HIP(hipMemsetAsync(c.adam_v.p, 0, (kTrain + 1) * 4, c.stream));
    c.adam_t = 0;
    pack_model(c);
    VVHD_HIP(hipStreamSynchronize(c.stream));
    c.mlp_loaded = true;
    API_END(ctx)
}

int vvhd_far_model_get(vvhd_ctx* ctx, float* params, int32_t n) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    Ctx& c = ctx->c;
    VVHD_CHECK(params && n == kParams, VVHD_EINVAL, "far_model_get: expected VVHD_FAR_PARAMS parameters");
    require_model(c);
    VVHD_HIP(hipSetDevice(c.device));
    VVHD_HIP(hipMemcpyAsync(params, c.mlp_params.p, kParams * 4, hipMemcpyDeviceToHost, c.stream));
    VVHD_HIP(hipStreamSynchronize(c.stream));
    API_END(ctx)
}

int vvhd_far_moments(vvhd_ctx* ctx, int32_t* nnodes, double* out) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    Ctx& c = ctx->c;
    VVHD_CHECK(nnodes, VVHD_EINVAL, "far_moments: null count");
    VVHD_CHECK(c.t.built, VVHD_ESTATE, "far_moments: tree is not built");
    VVHD_HIP(hipSetDevice(c.device));
    *nnodes = c.t.nnodes;
    if (out) {
        MomView M = moments(c);
        for (int k = 0; k < kMom; k++)
            VVHD_HIP(hipMemcpyAsync(out + (size_t)k * c.t.nnodes, M.m + (size_t)k * M.st, (size_t)c.t.nnodes * 8,
                                    hipMemcpyDeviceToHost, c.stream));
        VVHD_HIP(hipStreamSynchronize(c.stream));
    }
    API_END(ctx)
}

int vvhd_far_sample(vvhd_ctx* ctx, double rate, uint64_t seed, int64_t cap, int64_t* npairs, float* X, float* Y,
                    float* scale, int32_t* target, int32_t* node) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    Ctx& c = ctx->c;
    VVHD_CHECK(npairs && rate >= 0 && rate <= 1, VVHD_EINVAL, "far_sample: bad arguments");
    VVHD_CHECK(c.t.built, VVHD_ESTATE, "far_sample: tree is not built");
    VVHD_HIP(hipSetDevice(c.device));
    TreeArrays& t = c.t;
    const TreeDev T = tree_dev(t);
    const int64_t nfar = t.nfar;
    const uint32_t thresh = (uint32_t)std::min(4294967295.0, rate * 4294967296.0);
    c.pair_cnt.grow(nfar + 1);
    c.pair_off.grow(nfar + 1);
    VVHD_HIP(hipMemsetAsync(c.pair_cnt.p + nfar, 0, 4, c.stream));
    hipLaunchKernelGGL(k_pair_count, dim3(nblocks(nfar)), dim3(kBlock), 0, c.stream, T, c.g.p, nfar, seed, thresh,
                       c.pair_cnt.p, (const int64_t*)nullptr, (int32_t*)nullptr, (int32_t*)nullptr);
    size_t bytes = 0;
    VVHD_HIP(rocprim::exclusive_scan(nullptr, bytes, c.pair_cnt.p, c.pair_off.p, (int64_t)0, (size_t)nfar + 1,
                                     rocprim::plus<int64_t>(), c.stream));
    c.tmp.grow(bytes + 16);
    VVHD_HIP(rocprim::exclusive_scan(c.tmp.p, bytes, c.pair_cnt.p, c.pair_off.p, (int64_t)0, (size_t)nfar + 1,
                                     rocprim::plus<int64_t>(), c.stream));
    int64_t np = 0;
    VVHD_HIP(hipMemcpyAsync(&np, c.pair_off.p + nfar, 8, hipMemcpyDeviceToHost, c.stream));
    VVHD_HIP(hipStreamSynchronize(c.stream));
    *npairs = np;
    if (!X) return VVHD_OK;
    VVHD_CHECK(cap >= np && Y && scale, VVHD_EINVAL, "far_sample: output capacity below the pair count");
    c.pair_t.grow(np + 1);
    c.pair_f.grow(np + 1);
    hipLaunchKernelGGL(k_pair_count, dim3(nblocks(nfar)), dim3(kBlock), 0, c.stream, T, c.g.p, nfar, seed, thresh,
                       c.pair_cnt.p, (const int64_t*)c.pair_off.p, c.pair_t.p, c.pair_f.p);
    MomView M = moments(c);
    DBuf<float> dX, dY, dW;
    dX.grow(np * kFeat + 1);
    dY.grow(2 * np + 1);
    dW.grow(np + 1);
    hipLaunchKernelGGL(k_pair_label, dim3(nblocks(np, kWavesPerBlock)), dim3(kBlock), 0, c.stream, T, M, c.x.p, c.y.p,
                       c.g.p, c.ie.p, c.pair_t.p, c.pair_f.p, np, dX.p, dY.p, dW.p);
    if (np) {
        VVHD_HIP(hipMemcpyAsync(X, dX.p, (size_t)np * kFeat * 4, hipMemcpyDeviceToHost, c.stream));
        VVHD_HIP(hipMemcpyAsync(Y, dY.p, (size_t)np * 2 * 4, hipMemcpyDeviceToHost, c.stream));
        VVHD_HIP(hipMemcpyAsync(scale, dW.p, (size_t)np * 4, hipMemcpyDeviceToHost, c.stream));
        if (target) VVHD_HIP(hipMemcpyAsync(target, c.pair_t.p, (size_t)np * 4, hipMemcpyDeviceToHost, c.stream));
        if (node) VVHD_HIP(hipMemcpyAsync(node, c.pair_f.p, (size_t)np * 4, hipMemcpyDeviceToHost, c.stream));
    }
    VVHD_HIP(hipStreamSynchronize(c.stream));
    dX.release();
    dY.release();
    dW.release();
    API_END(ctx)
}

int vvhd_far_forward(vvhd_ctx* ctx, const float* X, int64_t n, float* Y) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    Ctx& c = ctx->c;
    VVHD_CHECK(n >= 0 && (n == 0 || (X && Y)), VVHD_EINVAL, "far_forward: bad arguments");
    require_model(c);
    VVHD_HIP(hipSetDevice(c.device));
    if (n == 0) return VVHD_OK;
    DBuf<float> dX, dY;
    dX.grow(n * kFeat);
    dY.grow(2 * n);
    VVHD_HIP(hipMemcpyAsync(dX.p, X, (size_t)n * kFeat * 4, hipMemcpyHostToDevice, c.stream));
    const int64_t tiles = (n + 31) / 32;
    hipLaunchKernelGGL(k_forward_feats, dim3(nblocks(tiles, kWavesPerBlock)), dim3(kBlock), 0, c.stream, dX.p, n,
                       c.mlp_frags.p, c.mlp_params.p, dY.p);
    VVHD_HIP(hipMemcpyAsync(Y, dY.p, (size_t)n * 2 * 4, hipMemcpyDeviceToHost, c.stream));
    VVHD_HIP(hipStreamSynchronize(c.stream));
    dX.release();
    dY.release();
    API_END(ctx)
}

int vvhd_far_gradients(vvhd_ctx* ctx, const float* X, const float* Y, const float* W, int64_t n, float* grad,
                       float* loss) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    Ctx& c = ctx->c;
    VVHD_CHECK(n > 0 && X && Y && W && grad, VVHD_EINVAL, "far_gradients: bad arguments");
    require_model(c);
    VVHD_HIP(hipSetDevice(c.device));
    DBuf<float> dX, dY, dW;
    dX.grow(n * kFeat);
    dY.grow(2 * n);
    dW.grow(n);
    VVHD_HIP(hipMemcpyAsync(dX.p, X, (size_t)n * kFeat * 4, hipMemcpyHostToDevice, c.stream));
    VVHD_HIP(hipMemcpyAsync(dY.p, Y, (size_t)n * 2 * 4, hipMemcpyHostToDevice, c.stream));
    VVHD_HIP(hipMemcpyAsync(dW.p, W, (size_t)n * 4, hipMemcpyHostToDevice, c.stream));
    train_pass(c, dX.p, dY.p, dW.p, n, n, 0, 0, 1.f / (float)n, 0, 0.f);
    std::vector<float> h(kTrain + 1);
    VVHD_HIP(hipMemcpyAsync(h.data(), c.grad_out.p, (kTrain + 1) * 4, hipMemcpyDeviceToHost, c.stream));
    VVHD_HIP(hipStreamSynchronize(c.stream));
    std::memcpy(grad, h.data(), kTrain * 4);
    if (loss) *loss = h[kTrain];
    dX.release();
    dY.release();
    dW.release();
    API_END(ctx)
}

int vvhd_far_train_data(vvhd_ctx* ctx, const float* X, const float* Y, const float* W, int64_t n) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    Ctx& c = ctx->c;
    VVHD_CHECK(n > 0 && X && Y && W, VVHD_EINVAL, "far_train_data: bad arguments");
    VVHD_HIP(hipSetDevice(c.device));
    c.tr_x.grow(n * kFeat);
    c.tr_y.grow(2 * n);
    c.tr_w.grow(n);
    VVHD_HIP(hipMemcpyAsync(c.tr_x.p, X, (size_t)n * kFeat * 4, hipMemcpyHostToDevice, c.stream));
    VVHD_HIP(hipMemcpyAsync(c.tr_y.p, Y, (size_t)n * 2 * 4, hipMemcpyHostToDevice, c.stream));
    VVHD_HIP(hipMemcpyAsync(c.tr_w.p, W, (size_t)n * 4, hipMemcpyHostToDevice, c.stream));
    VVHD_HIP(hipStreamSynchronize(c.stream));
    c.tr_n = n;
    API_END(ctx)
}

int vvhd_far_train(vvhd_ctx* ctx, int32_t steps, int32_t batch, const float* lr, uint64_t seed, float* loss) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    Ctx& c = ctx->c;
    VVHD_CHECK(steps >= 0 && batch > 0 && batch % 32 == 0 && lr, VVHD_EINVAL,
               "far_train: steps >= 0, batch a positive multiple of 32, lr per step");
    VVHD_CHECK(c.tr_n > 0, VVHD_ESTATE, "far_train: no training data");
    require_model(c);
    VVHD_HIP(hipSetDevice(c.device));
    std::vector<float> lsum(steps, 0.f);
    for (int s = 0; s < steps; s++) {
        train_pass(c, c.tr_x.p, c.tr_y.p, c.tr_w.p, c.tr_n, batch, 1, seed + (uint64_t)s * 0x9E3779B97F4A7C15ull,
                   1.f / (float)batch, 1, lr[s]);
        if (loss) VVHD_HIP(hipMemcpyAsync(&lsum[s], c.grad_out.p + kTrain, 4, hipMemcpyDeviceToHost, c.stream));
    }
    VVHD_HIP(hipStreamSynchronize(c.stream));
    if (loss) std::memcpy(loss, lsum.data(), (size_t)steps * 4);
    API_END(ctx)
}

int vvhd_convective_learned(vvhd_ctx* ctx, double inf_x, double inf_y, int kahan) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    Ctx& c = ctx->c;
    VVHD_CHECK(c.t.built, VVHD_ESTATE, "convective_learned: tree is not built");
    require_model(c);
    VVHD_HIP(hipSetDevice(c.device));
    VVHD_HIP(hipEventRecord(c.ev_pre_conv, c.stream));
    c.conv_call = c.api_calls;
    VVHD_HIP(hipEventRecord(c.ev[4], c.stream));
    // the near field runs on its own stream beside the far field
    const hipStream_t ns = c.overlap ? c.stream3 : c.stream;
    if (c.overlap) VVHD_HIP(hipStreamWaitEvent(c.stream3, c.ev_pre_conv, 0));
    c.near_v.grow(c.n + 1);
    convective_near(c, kahan, ns, c.near_v.p, c.overlap);
    const TreeDev T = tree_dev(c.t);
    if (T.nleaf > 0) {
        MomView M = moments(c);
        c.far_v.grow(c.n + 1);
        c.far_nodes.grow((size_t)c.t.nnodes * sizeof(FarNode));
        FarNode* nodes = reinterpret_cast<FarNode*>(c.far_nodes.p);
        c.far_leaf.grow(6 * (size_t)T.nleaf);
        hipLaunchKernelGGL(k_far_nodes, dim3(nblocks(c.t.nnodes)), dim3(kBlock), 0, c.stream, T, M, c.t.nnodes, nodes);
        const int grid = (T.nleaf + kWavesPerBlock - 1) / kWavesPerBlock;
        // the far monopoles at the vortices need only the node records: after
        // the near field on the side stream, beside the network (one stream:
        // before the network, outside the network's own phase events)
        if (!c.overlap)
            hipLaunchKernelGGL(k_far_mono, dim3(grid), dim3(kBlock), 0, c.stream, T, (const FarNode*)nodes, c.x.p,
                               c.y.p, c.far_v.p);
        VVHD_HIP(hipEventRecord(c.ev[14], c.stream));
        if (c.overlap) {
            VVHD_HIP(hipStreamWaitEvent(c.stream3, c.ev[14], 0));
            hipLaunchKernelGGL(k_far_mono, dim3(grid), dim3(kBlock), 0, c.stream3, T, (const FarNode*)nodes, c.x.p,
                               c.y.p, c.far_v.p);
            VVHD_HIP(hipEventRecord(c.ev_side, c.stream3));
        }
        if (c.far_net_grad)
            hipLaunchKernelGGL(k_far_learned<true>, dim3(grid), dim3(kBlock), 0, c.stream, T, (const FarNode*)nodes,
                               c.g.p, c.mlp_frags.p, c.mlp_params.p, c.far_leaf.p);
        else
            hipLaunchKernelGGL(k_far_learned<false>, dim3(grid), dim3(kBlock), 0, c.stream, T, (const FarNode*)nodes,
                               c.g.p, c.mlp_frags.p, c.mlp_params.p, c.far_leaf.p);
        VVHD_HIP(hipEventRecord(c.ev[15], c.stream));
        c.phase_recorded[7] = true;
        if (c.overlap) VVHD_HIP(hipStreamWaitEvent(c.stream, c.ev_side, 0));
        hipLaunchKernelGGL(k_finish_learned, dim3(nblocks(c.n)), dim3(kBlock), 0, c.stream, T, c.x.p, c.y.p, c.g.p,
                           c.near_v.p, c.far_v.p, c.far_leaf.p, c.vx.p, c.vy.p, c.n, inf_x, inf_y);
        if (c.heat.n || c.streak.n) {   // heat and streak targets: the reference's Taylor far field
            c.taylor.grow(T.nleaf + 1);
            far_taylor(c, c.taylor.p);
            passive_convective(c, c.taylor.p, inf_x, inf_y);
        }
    }
    VVHD_HIP(hipEventRecord(c.ev[5], c.stream));
    c.phase_recorded[2] = true;
    VVHD_HIP(hipGetLastError());
    API_END(ctx)
}

}  // extern "C"
