// export.hip -- read-back of the built tree (leaf table, near/far lists,
// point location) and the algorithmic work counts of its interaction lists.

#include "api.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace vvhd {
namespace {

constexpr int kBlock = 256;
inline int nblocks(int64_t n, int b = kBlock) { return (int)std::max<int64_t>(1, (n + b - 1) / b); }

template <typename T>
std::vector<T> d2h(Ctx& c, const T* p, size_t n) {
    std::vector<T> v(n);
    if (n) VVHD_HIP(hipMemcpyAsync(v.data(), p, n * sizeof(T), hipMemcpyDeviceToHost, c.stream));
    return v;
}

__global__ void k_find(TreeDev T, const double* pts, int64_t n, int32_t* out) {
#pragma clang fp contract(off)
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double px = pts[2 * p], py = pts[2 * p + 1];
    int node = 0;
    while (T.ch1[node] >= 0) {
        const int c = T.ch1[node];
        if (T.nh[node] < T.nw[node])
            node = (px < T.nx[node]) ? c : c + 1;
        else
            node = (py < T.ny[node]) ? c : c + 1;
    }
    out[p] = T.leaf_rank[node];
}

// Per leaf: near sources (all vortices of the near leaves), far terms.
__global__ void k_work(TreeDev T, const double* g, unsigned long long* acc) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= T.nleaf) return;
    const int L = T.leaf_node[l];
    int nt = 0;
    for (int i = T.v0[L]; i < T.v1[L]; i++) nt += (g[i] != 0);
    long long src = 0;
    for (int64_t e = T.near_off[l]; e < T.near_off[l + 1]; e++) {
        const int M = T.near_list[e];
        src += T.v1[M] - T.v0[M];
    }
    atomicAdd(&acc[0], (unsigned long long)(src * nt));
    atomicAdd(&acc[1], (unsigned long long)((T.far_off[l + 1] - T.far_off[l]) * (nt ? 1 : 0)));
    atomicAdd(&acc[2], (unsigned long long)((T.far_off[l + 1] - T.far_off[l]) * (T.v1[L] - T.v0[L])));
}

}  // namespace
}  // namespace vvhd

using namespace vvhd;

extern "C" {

int vvhd_tree_leaves(vvhd_ctx* ctx, int32_t* nleaves, double* geom, int32_t* ints, uint64_t* path) {
    if (!ctx) return VVHD_EINVAL;
    API_BEGIN
    Ctx& c = ctx->c;
    TreeArrays& t = c.t;
    VVHD_CHECK(t.built, VVHD_ESTATE, "tree_leaves: tree is not built");
    const int nl = t.nleaves_h, nn = t.nnodes;
    if (nleaves) *nleaves = nl;
    if (!geom && !ints && !path) return VVHD_OK;
    auto leaf = d2h(c, t.leaf_node.p, nl);
    auto x = d2h(c, t.x.p, nn), y = d2h(c, t.y.p, nn), h = d2h(c, t.h.p, nn), w = d2h(c, t.w.p, nn);
    auto v0 = d2h(c, t.v0.p, nn), v1 = d2h(c, t.v1.p, nn), dp = d2h(c, t.depth.p, nn), ns = d2h(c, t.nseg.p, nn);
    auto pa = d2h(c, t.path.p, nn);
    sync(c);
    for (int l = 0; l < nl; l++) {
        const int L = leaf[l];
        if (geom) {
            geom[4 * l] = x[L];
            geom[4 * l + 1] = y[L];
            geom[4 * l + 2] = h[L];
            geom[4 * l + 3] = w[L];
        }
        if (ints) {
            ints[4 * l] = v0[L];
            ints[4 * l + 1] = v1[L];
            ints[4 * l + 2] = dp[L];
            ints[4 * l + 3] = ns[L];
        }
        if (path) path[l] = pa[L];
    }
    API_END(ctx)
}

int vvhd_tree_nodes(vvhd_ctx* ctx, int32_t* nnodes, double* geom, int32_t* ints, uint64_t* path, double* cm) {
    if (!ctx) return VVHD_EINVAL;
    API_BEGIN
    Ctx& c = ctx->c;
    TreeArrays& t = c.t;
    VVHD_CHECK(t.built, VVHD_ESTATE, "tree_nodes: tree is not built");
    const int nn = t.nnodes;
    if (nnodes) *nnodes = nn;
    if (!geom && !ints && !path && !cm) return VVHD_OK;
    auto x = d2h(c, t.x.p, nn), y = d2h(c, t.y.p, nn), h = d2h(c, t.h.p, nn), w = d2h(c, t.w.p, nn);
    auto v0 = d2h(c, t.v0.p, nn), v1 = d2h(c, t.v1.p, nn), ch = d2h(c, t.ch1.p, nn), dp = d2h(c, t.depth.p, nn);
    auto sp = d2h(c, t.split.p, nn);
    auto pa = d2h(c, t.path.p, nn);
    std::vector<double> cmv[6];
    if (cm) {
        const double* src[6] = {t.cmp_x.p, t.cmp_y.p, t.cmp_g.p, t.cmm_x.p, t.cmm_y.p, t.cmm_g.p};
        for (int q = 0; q < 6; q++) cmv[q] = d2h(c, src[q], nn);
    }
    sync(c);
    for (int k = 0; k < nn; k++) {
        if (geom) {
            geom[4 * k] = x[k];
            geom[4 * k + 1] = y[k];
            geom[4 * k + 2] = h[k];
            geom[4 * k + 3] = w[k];
        }
        if (ints) {
            ints[4 * k] = v0[k];
            ints[4 * k + 1] = v1[k];
            ints[4 * k + 2] = sp[k] == 2 ? ch[k] : -1;
            ints[4 * k + 3] = dp[k];
        }
        if (path) path[k] = pa[k];
        if (cm)
            for (int q = 0; q < 6; q++) cm[6 * k + q] = cmv[q][k];
    }
    API_END(ctx)
}

int vvhd_tree_lists(vvhd_ctx* ctx, int far, int64_t* count, int64_t* offsets, int32_t* depth, uint64_t* path,
                    double* cm) {
    if (!ctx) return VVHD_EINVAL;
    API_BEGIN
    Ctx& c = ctx->c;
    TreeArrays& t = c.t;
    VVHD_CHECK(t.built, VVHD_ESTATE, "tree_lists: tree is not built");
    const int nl = t.nleaves_h, nn = t.nnodes;
    const int64_t tot = far ? t.nfar : t.nnear;
    if (count) *count = tot;
    if (!offsets && !depth && !path && !cm) return VVHD_OK;
    auto off = d2h(c, far ? t.far_off.p : t.near_off.p, (size_t)nl + 1);
    auto lst = d2h(c, far ? t.far_list.p : t.near_list.p, (size_t)tot);
    auto dp = d2h(c, t.depth.p, nn);
    auto pa = d2h(c, t.path.p, nn);
    std::vector<double> px, py, pg, mx, my, mg;
    if (cm) {
        px = d2h(c, t.cmp_x.p, nn);
        py = d2h(c, t.cmp_y.p, nn);
        pg = d2h(c, t.cmp_g.p, nn);
        mx = d2h(c, t.cmm_x.p, nn);
        my = d2h(c, t.cmm_y.p, nn);
        mg = d2h(c, t.cmm_g.p, nn);
    }
    sync(c);
    if (offsets) std::memcpy(offsets, off.data(), ((size_t)nl + 1) * 8);
    for (int64_t e = 0; e < tot; e++) {
        const int N = lst[e];
        if (depth) depth[e] = dp[N];
        if (path) path[e] = pa[N];
        if (cm) {
            cm[6 * e] = px[N];
            cm[6 * e + 1] = py[N];
            cm[6 * e + 2] = pg[N];
            cm[6 * e + 3] = mx[N];
            cm[6 * e + 4] = my[N];
            cm[6 * e + 5] = mg[N];
        }
    }
    API_END(ctx)
}

int vvhd_tree_find(vvhd_ctx* ctx, const double* xy, int64_t n, int32_t* leaf_out) {
    if (!ctx) return VVHD_EINVAL;
    API_BEGIN
    REQUIRE_ONE_SIM(ctx->c, "tree_find");
    Ctx& c = ctx->c;
    VVHD_CHECK(c.t.built, VVHD_ESTATE, "TTree::findNode(): tree is not built");
    VVHD_CHECK(n >= 0 && (n == 0 || (xy && leaf_out)), VVHD_EINVAL, "tree_find: bad arguments");
    if (n == 0) return VVHD_OK;
    c.tmp.grow((size_t)n * 20);
    double* dp = (double*)c.tmp.p;
    int32_t* dout = (int32_t*)(dp + 2 * n);
    VVHD_HIP(hipMemcpyAsync(dp, xy, (size_t)n * 16, hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(k_find, dim3(nblocks(n)), dim3(kBlock), 0, c.stream, tree_dev(c.t), dp, n, dout);
    VVHD_HIP(hipMemcpyAsync(leaf_out, dout, (size_t)n * 4, hipMemcpyDeviceToHost, c.stream));
    sync(c);
    API_END(ctx)
}

int vvhd_work_counts(vvhd_ctx* ctx, int64_t* out7) {
    if (!ctx) return VVHD_EINVAL;
    API_BEGIN
    Ctx& c = ctx->c;
    VVHD_CHECK(out7, VVHD_EINVAL, "work_counts: null output");
    VVHD_CHECK(c.t.built, VVHD_ESTATE, "work_counts: tree is not built");
    c.lscratch0.grow(8);
    VVHD_HIP(hipMemsetAsync(c.lscratch0.p, 0, 64, c.stream));
    hipLaunchKernelGGL(k_work, dim3(nblocks(c.t.nleaves_h)), dim3(kBlock), 0, c.stream, tree_dev(c.t), c.g.p,
                       (unsigned long long*)c.lscratch0.p);
    int64_t acc[8];
    VVHD_HIP(hipMemcpyAsync(acc, c.lscratch0.p, 64, hipMemcpyDeviceToHost, c.stream));
    sync(c);
    out7[0] = acc[0];
    out7[1] = acc[1];
    out7[2] = acc[2];
    out7[3] = (c.far_net_grad ? 3 : 1) * acc[1];
    out7[4] = (int64_t)c.t.levels.size();
    out7[5] = c.t.nleaves_h;
    out7[6] = c.t.nnodes;
    API_END(ctx)
}

}  // extern "C"
