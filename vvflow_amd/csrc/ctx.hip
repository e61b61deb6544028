// ctx.hip -- device context and the C ABI entry points (include/vvhd_gpu.h).
//
// Every entry point converts exceptions into a status code plus a message
// (vvhd_last_error); nothing throws across the ABI.

#include "api.hpp"

#include <cmath>
#include <cstring>

namespace vvhd {
namespace {

constexpr int kBlock = 256;
inline int nblocks(int64_t n, int b = kBlock) { return (int)std::max<int64_t>(1, (n + b - 1) / b); }

__global__ void k_aos_to_soa(const double* a, int64_t n, double* x, double* y, double* g, double* vx, double* vy,
                             double* ie) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* o = a + 6 * i;
    x[i] = o[0];
    y[i] = o[1];
    g[i] = o[2];
    vx[i] = o[3];
    vy[i] = o[4];
    ie[i] = o[5];
}

__global__ void k_soa_to_aos(double* a, int64_t n, const double* x, const double* y, const double* g, const double* vx,
                             const double* vy, const double* ie) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double* o = a + 6 * i;
    o[0] = x[i];
    o[1] = y[i];
    o[2] = g[i];
    o[3] = vx[i];
    o[4] = vy[i];
    o[5] = ie[i];
}

// FP64 reductions of Space::gsum / integral / hydrodynamic_momentum / gmax
// (TSpace.hpp:85-117): per-block partial sums, then one block.
__global__ void k_integrals_partial(const double* x, const double* y, const double* g, int64_t n, double* part) {
    __shared__ double sh[5][kBlock];
    double gs = 0, in = 0, mx = 0, my = 0, gm = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double gi = g[i];
        gs += gi;
        in += gi * (x[i] * x[i] + y[i] * y[i]);
        mx += gi * x[i];
        my += gi * y[i];
        if (fabs(gi) > fabs(gm)) gm = gi;
    }
    const int t = threadIdx.x;
    sh[0][t] = gs;
    sh[1][t] = in;
    sh[2][t] = mx;
    sh[3][t] = my;
    sh[4][t] = gm;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if (t < o) {
            for (int q = 0; q < 4; q++) sh[q][t] += sh[q][t + o];
            if (fabs(sh[4][t + o]) > fabs(sh[4][t])) sh[4][t] = sh[4][t + o];
        }
        __syncthreads();
    }
    if (t == 0)
        for (int q = 0; q < 5; q++) part[5 * blockIdx.x + q] = sh[q][0];
}

__global__ void k_integrals_final(const double* part, int nb, double* out) {
    if (threadIdx.x != 0) return;
    double r[5] = {0, 0, 0, 0, 0};
    for (int b = 0; b < nb; b++) {
        for (int q = 0; q < 4; q++) r[q] += part[5 * b + q];
        if (fabs(part[5 * b + 4]) > fabs(r[4])) r[4] = part[5 * b + 4];
    }
    for (int q = 0; q < 5; q++) out[q] = r[q];
}

}  // namespace

void reserve_vortices(Ctx& c, int64_t n, bool preserve);

void sim_sync(Ctx& c) {
    if (c.nsim == 1) {
        c.sim_off_h = {0, c.n};
        c.sim_body_h = {0, (int32_t)c.bodies_h.size()};
        c.sim_re_h = {c.re};
    }
    const int ns = c.nsim, nseg = (int)c.segs_h.size();
    c.sim_off.grow(ns + 1);
    c.sim_body.grow(ns + 1);
    c.sim_re.grow(ns);
    c.seg_root.grow(std::max(nseg, 1));
    std::vector<int32_t> root(nseg, 0);
    for (int s = 0; s < ns; s++)
        for (int b = c.sim_body_h[s]; b < c.sim_body_h[s + 1]; b++)
            for (int k = 0; k < c.bodies_h[b].nseg; k++) root[c.bodies_h[b].first_seg + k] = s;
    VVHD_HIP(hipMemcpyAsync(c.sim_off.p, c.sim_off_h.data(), (ns + 1) * 8, hipMemcpyHostToDevice, c.stream));
    VVHD_HIP(hipMemcpyAsync(c.sim_body.p, c.sim_body_h.data(), (ns + 1) * 4, hipMemcpyHostToDevice, c.stream));
    VVHD_HIP(hipMemcpyAsync(c.sim_re.p, c.sim_re_h.data(), ns * 8, hipMemcpyHostToDevice, c.stream));
    if (nseg) VVHD_HIP(hipMemcpyAsync(c.seg_root.p, root.data(), (size_t)nseg * 4, hipMemcpyHostToDevice, c.stream));
    VVHD_HIP(hipStreamSynchronize(c.stream));
}

void ensure_vortex_capacity(Ctx& c, int64_t n) { reserve_vortices(c, n, false); }

// Grows every per-vortex array to at least n entries, keeping the first c.n
// entries of the master arrays when preserve is set.
void reserve_vortices(Ctx& c, int64_t n, bool preserve) {
    if ((size_t)n <= c.x.cap) return;
    const size_t cap = std::max<size_t>((size_t)n, c.x.cap + c.x.cap / 2);
    auto regrow = [&](DBuf<double>& b, bool keep) {
        DBuf<double> nb;
        nb.grow(cap);
        if (keep && c.n > 0) VVHD_HIP(hipMemcpyAsync(nb.p, b.p, (size_t)c.n * 8, hipMemcpyDeviceToDevice, c.stream));
        VVHD_HIP(hipStreamSynchronize(c.stream));
        b.release();
        b = nb;
    };
    for (auto* b : {&c.x, &c.y, &c.g, &c.vx, &c.vy, &c.ie}) regrow(*b, preserve);
    for (auto* b : {&c.x2, &c.y2, &c.g2, &c.vx2, &c.vy2, &c.ie2}) regrow(*b, false);
    c.id.release();
    c.id2.release();
    c.src_ge.release();
    c.id.grow(cap);
    c.id2.grow(cap);
    c.src_ge.grow(cap);
}

void sync(Ctx& c) { VVHD_HIP(hipStreamSynchronize(c.stream)); }

}  // namespace vvhd

using namespace vvhd;

extern "C" {

int vvhd_create(int device, vvhd_ctx** out) {
    if (!out) return VVHD_EINVAL;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return VVHD_EHIP;
    if (device < 0 || device >= count) return VVHD_EINVAL;
    vvhd_ctx* ctx = new vvhd_ctx();
    ctx->c.device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return VVHD_EHIP;
    }
    for (auto& e : ctx->c.ev)
        if (hipEventCreate(&e) != hipSuccess) {
            delete ctx;
            return VVHD_EHIP;
        }
    for (auto& e : ctx->c.lev_ev)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            delete ctx;
            return VVHD_EHIP;
        }
    for (auto& e : ctx->c.tune_ev)
        if (hipEventCreate(&e) != hipSuccess) {
            delete ctx;
            return VVHD_EHIP;
        }
    {
        int ncu = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
            ctx->c.n_cu = ncu;
    }
    if (hipStreamCreateWithFlags(&ctx->c.stream2, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->c.stream3, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->c.ev_pre_conv, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->c.ev_side, hipEventDisableTiming) != hipSuccess) {
        delete ctx;
        return VVHD_EHIP;
    }
    if (hipHostMalloc((void**)&ctx->c.lev_counters, Ctx::kLevSlots * 8 * sizeof(int32_t), hipHostMallocDefault) !=
            hipSuccess ||
        hipHostMalloc((void**)&ctx->c.fm_counts, 4 * sizeof(int32_t), hipHostMallocDefault) != hipSuccess) {
        delete ctx;
        return VVHD_EHIP;
    }
    // device address of the allocation-free entry point's scratch array,
    // looked up once per device while the device is current
    try {
        noalloc_init(device);
    } catch (const Error& e) {
        vvhd_destroy(ctx);
        return e.code;
    }
    *out = ctx;
    return VVHD_OK;
}

void vvhd_destroy(vvhd_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->c.device);
    (void)hipStreamSynchronize(ctx->c.stream);
    Ctx& c = ctx->c;
    for (auto* b : {&c.x, &c.y, &c.g, &c.vx, &c.vy, &c.ie, &c.x2, &c.y2, &c.g2, &c.vx2, &c.vy2, &c.ie2, &c.snap,
                    &c.d1, &c.d2, &c.leaf_crit, &c.leaf_restr, &c.fric})
        b->release();
    for (auto* b : {&c.id, &c.id2, &c.iscratch0, &c.iscratch1, &c.iscratch2, &c.iscratch3, &c.nn1, &c.nn2, &c.ecount,
                    &c.rnear_list})
        b->release();
    for (auto* b : {&c.flag_list, &c.rnear_cnt, &c.leaf_stamp, &c.seg_root, &c.sim_body}) b->release();
    c.sim_off.release();
    c.sim_re.release();
    c.leaf_maxd2.release();
    c.leaf_mark.release();
    c.rep_state.release();
    c.near_v.release();
    c.taylor.release();
    c.src_ge.release();
    c.lscratch0.release();
    c.rnear_off.release();
    c.tmp.release();
    c.eflag.release();
    c.segs.release();
    c.body_range.release();
    bodies_release(c);
    farfield_release(c);
    TreeArrays& t = c.t;
    for (auto* b : {&t.x, &t.y, &t.h, &t.w, &t.cmp_x, &t.cmp_y, &t.cmp_g, &t.cmm_x, &t.cmm_y, &t.cmm_g}) b->release();
    for (auto* b : {&t.blx, &t.bly, &t.trx, &t.try_, &t.path}) b->release();
    for (auto* b : {&t.v0, &t.v1, &t.ch1, &t.parent, &t.depth, &t.nseg, &t.split, &t.nleaves, &t.arrive, &t.leaf_rank,
                    &t.counters, &t.leaf_node, &t.near_list, &t.far_list, &t.seg_list, &t.seg_leaf, &t.vleaf,
                    &t.node_L, &t.node_m, &t.leaf_tmp, &t.seg_node, &t.cnt_near, &t.cnt_far, &t.err})
        b->release();
    for (auto* b : {&t.near_off, &t.far_off, &t.seg_off, &t.slot_near, &t.slot_far}) b->release();
    for (auto* b : {&t.groups, &t.leaf_sim, &t.root_first, &t.pre}) b->release();
    t.entries.release();
    t.trav.release();
    t.pbox.release();
    t.cmp_ie.release();
    t.cmm_ie.release();
    for (auto& e : c.ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : c.tune_ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& g : c.conv_graphs) (void)hipGraphExecDestroy((hipGraphExec_t)g.second);
    c.conv_graphs.clear();
    c.near_items.release();
    c.near_q.release();
    if (c.ev_pre_conv) (void)hipEventDestroy(c.ev_pre_conv);
    if (c.ev_side) (void)hipEventDestroy(c.ev_side);
    for (hipStream_t* st : {&c.stream2, &c.stream3})
        if (*st) {
            (void)hipStreamSynchronize(*st);
            (void)hipStreamDestroy(*st);
        }
    c.dvx.release();
    c.dvy.release();
    c.heat.release();
    c.streak.release();
    c.heat_cond.release();
    c.heat_const.release();
    c.heat_layer.release();
    c.hres.release();
    c.hflag.release();
    c.hdirty.release();
    for (auto* b : {&t.h0, &t.h1, &t.s0, &t.s1, &t.seg_keys, &t.seg_ids}) b->release();
    for (auto* b : {&c.fld_a, &c.fld_b, &c.fld_gtmp}) b->release();
    c.fld_map.release();
    for (auto& e : c.lev_ev)
        if (e) (void)hipEventDestroy(e);
    if (c.lev_counters) (void)hipHostFree(c.lev_counters);
    if (c.fm_counts) (void)hipHostFree(c.fm_counts);
    (void)hipStreamDestroy(c.stream);
    delete ctx;
}

const char* vvhd_last_error(const vvhd_ctx* ctx) { return ctx ? ctx->c.err.c_str() : "null context"; }

void* vvhd_stream(vvhd_ctx* ctx) { return ctx ? (void*)ctx->c.stream : nullptr; }

int vvhd_set_params(vvhd_ctx* ctx, double re, double dt, double inf_g, double time) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    VVHD_CHECK(re > 0, VVHD_EINVAL, "set_params: re must be positive");
    ctx->c.re = re;
    ctx->c.dt = dt;
    ctx->c.inf_g = inf_g;
    ctx->c.time = time;
    API_END(ctx)
}

int vvhd_upload_vortices(vvhd_ctx* ctx, const double* objs6, int64_t n) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    Ctx& c = ctx->c;
    VVHD_CHECK(n >= 0 && (n == 0 || objs6), VVHD_EINVAL, "upload_vortices: bad arguments");
    VVHD_HIP(hipSetDevice(c.device));
    ensure_vortex_capacity(c, std::max<int64_t>(n, 1));
    c.n = n;
    c.t.built = false;
    c.nsim = 1;
    if (n) {
        // stage through the x2..ie2 double buffers (6 * n doubles fit in x2+y2+g2+vx2+vy2+ie2 only
        // as separate arrays, so use the snapshot buffer as staging)
        c.snap.grow((size_t)6 * n);
        VVHD_HIP(hipMemcpyAsync(c.snap.p, objs6, (size_t)n * 48, hipMemcpyHostToDevice, c.stream));
        hipLaunchKernelGGL(k_aos_to_soa, dim3(nblocks(n)), dim3(kBlock), 0, c.stream, c.snap.p, n, c.x.p, c.y.p, c.g.p,
                           c.vx.p, c.vy.p, c.ie.p);
        c.snap_n = -1;
        VVHD_HIP(hipStreamSynchronize(c.stream));
    }
    API_END(ctx)
}

int vvhd_download_vortices(vvhd_ctx* ctx, double* objs6, int64_t n) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    Ctx& c = ctx->c;
    VVHD_CHECK(n == c.n, VVHD_EINVAL, "download_vortices: n does not match the vortex count");
    VVHD_CHECK(n == 0 || objs6, VVHD_EINVAL, "download_vortices: null output");
    if (n) {
        c.tmp.grow((size_t)n * 48);
        double* st = (double*)c.tmp.p;
        hipLaunchKernelGGL(k_soa_to_aos, dim3(nblocks(n)), dim3(kBlock), 0, c.stream, st, n, c.x.p, c.y.p, c.g.p,
                           c.vx.p, c.vy.p, c.ie.p);
        VVHD_HIP(hipMemcpyAsync(objs6, st, (size_t)n * 48, hipMemcpyDeviceToHost, c.stream));
        VVHD_HIP(hipStreamSynchronize(c.stream));
    }
    API_END(ctx)
}

int64_t vvhd_num_vortices(vvhd_ctx* ctx) { return ctx ? ctx->c.n : -1; }

int vvhd_snapshot(vvhd_ctx* ctx) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    Ctx& c = ctx->c;
    const int64_t n = c.n;
    c.snap.grow((size_t)6 * std::max<int64_t>(n, 1));
    double* arrs[6] = {c.x.p, c.y.p, c.g.p, c.vx.p, c.vy.p, c.ie.p};
    for (int q = 0; q < 6; q++)
        if (n) VVHD_HIP(hipMemcpyAsync(c.snap.p + q * n, arrs[q], (size_t)n * 8, hipMemcpyDeviceToDevice, c.stream));
    c.snap_n = n;
    API_END(ctx)
}

int vvhd_restore(vvhd_ctx* ctx) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    Ctx& c = ctx->c;
    VVHD_CHECK(c.snap_n >= 0, VVHD_ESTATE, "restore: no snapshot");
    const int64_t n = c.snap_n;
    reserve_vortices(c, std::max<int64_t>(n, 1), false);
    c.n = n;
    double* arrs[6] = {c.x.p, c.y.p, c.g.p, c.vx.p, c.vy.p, c.ie.p};
    for (int q = 0; q < 6; q++)
        if (n) VVHD_HIP(hipMemcpyAsync(arrs[q], c.snap.p + q * n, (size_t)n * 8, hipMemcpyDeviceToDevice, c.stream));
    c.t.built = false;
    API_END(ctx)
}

int vvhd_upload_bodies(vvhd_ctx* ctx, const vvhd_seg* segs, int32_t nseg, const vvhd_body* bodies, int32_t nbody) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    Ctx& c = ctx->c;
    VVHD_CHECK(nseg >= 0 && nbody >= 0 && (nseg == 0 || segs) && (nbody == 0 || bodies), VVHD_EINVAL,
               "upload_bodies: bad arguments");
    int32_t expect = 0;
    for (int b = 0; b < nbody; b++) {
        VVHD_CHECK(bodies[b].first_seg == expect && bodies[b].nseg > 0, VVHD_EINVAL,
                   "upload_bodies: bodies must list consecutive, non-empty segment ranges");
        VVHD_CHECK(bodies[b].special_segment_no >= 0 && bodies[b].special_segment_no < bodies[b].nseg, VVHD_EINVAL,
                   "upload_bodies: special segment out of range");
        expect += bodies[b].nseg;
    }
    VVHD_CHECK(expect == nseg, VVHD_EINVAL, "upload_bodies: segment count mismatch");
    c.segs_h.assign(segs, segs + nseg);
    c.bodies_h.assign(bodies, bodies + nbody);
    c.nsim = 1;
    c.body_range.grow(2 * std::max(nbody, 1));
    if (nbody) {
        std::vector<int32_t> br(2 * nbody);
        for (int b = 0; b < nbody; b++) {
            br[2 * b] = bodies[b].first_seg;
            br[2 * b + 1] = bodies[b].nseg;
        }
        VVHD_HIP(hipMemcpyAsync(c.body_range.p, br.data(), br.size() * 4, hipMemcpyHostToDevice, c.stream));
        VVHD_HIP(hipStreamSynchronize(c.stream));
    }
    c.segs.grow(std::max(nseg, 1));
    c.fric.grow(std::max(nseg, 1));
    if (nseg) {
        VVHD_HIP(hipMemcpyAsync(c.segs.p, segs, (size_t)nseg * sizeof(vvhd_seg), hipMemcpyHostToDevice, c.stream));
        std::vector<double> fr(nseg);
        for (int s = 0; s < nseg; s++) fr[s] = segs[s].fric;
        VVHD_HIP(hipMemcpyAsync(c.fric.p, fr.data(), (size_t)nseg * 8, hipMemcpyHostToDevice, c.stream));
        VVHD_HIP(hipStreamSynchronize(c.stream));
    }
    bodies_prepare(c);
    // heat conditions start at neglect, heat_const at 0, no layer particles
    c.heat_cond_h.assign(nbody, 0);
    c.heat_cond.grow(std::max(nbody, 1));
    c.heat_const.grow(std::max(nseg, 1));
    if (nbody) VVHD_HIP(hipMemsetAsync(c.heat_cond.p, 0, (size_t)nbody * 4, c.stream));
    if (nseg) VVHD_HIP(hipMemsetAsync(c.heat_const.p, 0, (size_t)nseg * 4, c.stream));
    heat_layer_reset(c);
    VVHD_HIP(hipStreamSynchronize(c.stream));
    c.t.built = false;
    API_END(ctx)
}

int vvhd_set_batch(vvhd_ctx* ctx, int32_t nsim, const int64_t* vortex_off, const int32_t* body_off,
                   const double* re) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    Ctx& c = ctx->c;
    VVHD_CHECK(nsim >= 1 && vortex_off && body_off && re, VVHD_EINVAL, "set_batch: bad arguments");
    VVHD_CHECK(vortex_off[0] == 0 && vortex_off[nsim] == c.n, VVHD_EINVAL,
               "set_batch: vortex offsets must run from 0 to the vortex count");
    VVHD_CHECK(body_off[0] == 0 && body_off[nsim] == (int32_t)c.bodies_h.size(), VVHD_EINVAL,
               "set_batch: body offsets must run from 0 to the body count");
    for (int s = 0; s < nsim; s++) {
        VVHD_CHECK(vortex_off[s + 1] >= vortex_off[s] && body_off[s + 1] >= body_off[s], VVHD_EINVAL,
                   "set_batch: offsets must not decrease");
        VVHD_CHECK(re[s] > 0, VVHD_EINVAL, "set_batch: re must be positive");
    }
    c.nsim = nsim;
    c.sim_off_h.assign(vortex_off, vortex_off + nsim + 1);
    c.sim_body_h.assign(body_off, body_off + nsim + 1);
    c.sim_re_h.assign(re, re + nsim);
    c.t.built = false;
    API_END(ctx)
}

int vvhd_set_tuning(vvhd_ctx* ctx, int autotune, int near_form) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    VVHD_CHECK(near_form == 0 || near_form == 8 || near_form == 16 || near_form == 32, VVHD_EINVAL,
               "set_tuning: near_form must be 0 (one wave per leaf) or 8, 16, 32 (persistent waves per CU)");
    ctx->c.autotune = autotune != 0;
    ctx->c.near_form = near_form;
    ctx->c.tune_near.clear();
    API_END(ctx)
}

int vvhd_set_graph(vvhd_ctx* ctx, int enable) {
    REQUIRE_CTX(ctx);
    ctx->c.use_graph = enable != 0;
    return VVHD_OK;
}

int vvhd_tuning(vvhd_ctx* ctx, int64_t* out4) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    VVHD_CHECK(out4, VVHD_EINVAL, "tuning: null output");
    out4[0] = ctx->c.near_last;
    out4[1] = (int64_t)ctx->c.tune_near.size();
    out4[2] = ctx->c.graph_captures;
    out4[3] = ctx->c.graph_launches;
    API_END(ctx)
}

int vvhd_batch_offsets(vvhd_ctx* ctx, int32_t* nsim, int64_t* vortex_off) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    Ctx& c = ctx->c;
    VVHD_CHECK(nsim, VVHD_EINVAL, "batch_offsets: null count");
    *nsim = c.nsim;
    if (vortex_off) {
        if (c.nsim == 1) {
            vortex_off[0] = 0;
            vortex_off[1] = c.n;
        } else {
            std::memcpy(vortex_off, c.sim_off_h.data(), (size_t)(c.nsim + 1) * 8);
        }
    }
    API_END(ctx)
}

int vvhd_download_bodies(vvhd_ctx* ctx, vvhd_seg* segs, int32_t nseg, vvhd_body* bodies, int32_t nbody) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    Ctx& c = ctx->c;
    VVHD_CHECK(nseg == (int32_t)c.segs_h.size() && nbody == (int32_t)c.bodies_h.size(), VVHD_EINVAL,
               "download_bodies: counts do not match");
    if (nseg) {
        std::vector<double> fr(nseg);
        VVHD_HIP(hipMemcpyAsync(fr.data(), c.fric.p, (size_t)nseg * 8, hipMemcpyDeviceToHost, c.stream));
        VVHD_HIP(hipStreamSynchronize(c.stream));
        for (int s = 0; s < nseg; s++) c.segs_h[s].fric = fr[s];
        std::memcpy(segs, c.segs_h.data(), (size_t)nseg * sizeof(vvhd_seg));
    }
    if (nbody) std::memcpy(bodies, c.bodies_h.data(), (size_t)nbody * sizeof(vvhd_body));
    API_END(ctx)
}

int vvhd_tree_build(vvhd_ctx* ctx, int far_criteria, double min_node_size, double max_node_size, int32_t* perm_out) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    Ctx& c = ctx->c;
    VVHD_HIP(hipSetDevice(c.device));
    VVHD_HIP(hipEventRecord(c.ev[0], c.stream));
    tree_build(c, far_criteria, min_node_size, max_node_size, perm_out);
    VVHD_HIP(hipEventRecord(c.ev[1], c.stream));
    c.phase_recorded[0] = true;
    if (perm_out) sync(c);
    API_END(ctx)
}

int vvhd_tree_destroy(vvhd_ctx* ctx) {
    REQUIRE_CTX(ctx);
    ctx->c.t.built = false;
    return VVHD_OK;
}

int vvhd_epsilon(vvhd_ctx* ctx, int merge, int32_t* nn1, int32_t* nn2, int32_t* merged) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    Ctx& c = ctx->c;
    VVHD_CHECK(c.t.built, VVHD_ESTATE, "MEpsilonFast::CalcEpsilonFast(): tree is not built");
    VVHD_HIP(hipEventRecord(c.ev[2], c.stream));
    epsilon(c, merge, nn1, nn2, merged);
    VVHD_HIP(hipEventRecord(c.ev[3], c.stream));
    c.phase_recorded[1] = true;
    if (nn1 || nn2 || merged) sync(c);
    API_END(ctx)
}

int vvhd_convective(vvhd_ctx* ctx, double inf_x, double inf_y, int kahan) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    Ctx& c = ctx->c;
    VVHD_CHECK(c.t.built, VVHD_ESTATE, "MConvectiveFast::process_all_lists(): tree is not built");
    VVHD_HIP(hipEventRecord(c.ev_pre_conv, c.stream));
    c.conv_call = c.api_calls;
    VVHD_HIP(hipEventRecord(c.ev[4], c.stream));
    convective(c, inf_x, inf_y, kahan);
    VVHD_HIP(hipEventRecord(c.ev[5], c.stream));
    c.phase_recorded[2] = true;
    API_END(ctx)
}

int vvhd_convective_noalloc(vvhd_ctx* ctx, double inf_x, double inf_y, int kahan) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    Ctx& c = ctx->c;
    VVHD_CHECK(c.t.built, VVHD_ESTATE, "MConvectiveFast::process_all_lists(): tree is not built");
    VVHD_HIP(hipSetDevice(c.device));
    VVHD_HIP(hipEventRecord(c.ev_pre_conv, c.stream));
    c.conv_call = c.api_calls;
    VVHD_HIP(hipEventRecord(c.ev[4], c.stream));
    convective_noalloc(c, inf_x, inf_y, kahan);
    VVHD_HIP(hipEventRecord(c.ev[5], c.stream));
    c.phase_recorded[2] = true;
    API_END(ctx)
}

int vvhd_velocity_points(vvhd_ctx* ctx, const double* xy, int64_t n, double inf_x, double inf_y, double* v_out) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    REQUIRE_ONE_SIM(ctx->c, "velocity_points");
    Ctx& c = ctx->c;
    VVHD_CHECK(c.t.built, VVHD_ESTATE, "TTree::findNode(): tree is not built");
    VVHD_CHECK(n >= 0 && (n == 0 || (xy && v_out)), VVHD_EINVAL, "velocity_points: bad arguments");
    velocity_points(c, xy, n, inf_x, inf_y, v_out);
    API_END(ctx)
}

int vvhd_diffusive(vvhd_ctx* ctx) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    Ctx& c = ctx->c;
    VVHD_CHECK(c.t.built, VVHD_ESTATE, "MDiffusiveFast::process_vort_list(): tree is not built");
    // right after a convective call the state it started from is the
    // diffusive's input too: run beside the convective kernels
    const bool overlap = c.overlap && c.conv_call + 1 == c.api_calls && c.segs_h.empty() && c.t.nleaves_h > 0;
    if (overlap) {
        VVHD_HIP(hipStreamWaitEvent(c.stream2, c.ev_pre_conv, 0));
        VVHD_HIP(hipEventRecord(c.ev[6], c.stream2));
    } else {
        VVHD_HIP(hipEventRecord(c.ev[6], c.stream));
    }
    diffusive(c, overlap);
    c.phase_recorded[3] = true;
    API_END(ctx)
}

int vvhd_integrals(vvhd_ctx* ctx, double* out5) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    REQUIRE_ONE_SIM(ctx->c, "integrals");
    Ctx& c = ctx->c;
    VVHD_CHECK(out5, VVHD_EINVAL, "integrals: null output");
    const int nb = 256;
    c.tmp.grow((size_t)(5 * nb + 5) * 8);
    double* part = (double*)c.tmp.p;
    hipLaunchKernelGGL(k_integrals_partial, dim3(nb), dim3(kBlock), 0, c.stream, c.x.p, c.y.p, c.g.p, c.n, part);
    hipLaunchKernelGGL(k_integrals_final, dim3(1), dim3(64), 0, c.stream, part, nb, part + 5 * nb);
    VVHD_HIP(hipMemcpyAsync(out5, part + 5 * nb, 40, hipMemcpyDeviceToHost, c.stream));
    sync(c);
    API_END(ctx)
}

int vvhd_direct_sum(vvhd_ctx* ctx, const int64_t* targets, int64_t nt, double* v_out) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    Ctx& c = ctx->c;
    VVHD_CHECK(nt >= 0 && (nt == 0 || v_out), VVHD_EINVAL, "direct_sum: bad arguments");
    direct_sum(c, targets, nt, v_out);
    API_END(ctx)
}

int vvhd_set_far_gradient(vvhd_ctx* ctx, int network) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    ctx->c.far_net_grad = network != 0;
    API_END(ctx)
}

int vvhd_set_overlap(vvhd_ctx* ctx, int enable) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    ctx->c.overlap = enable != 0;
    API_END(ctx)
}

int vvhd_phase_times(vvhd_ctx* ctx, float* ms6) {
    REQUIRE_CTX(ctx);
    API_BEGIN
    Ctx& c = ctx->c;
    VVHD_CHECK(ms6, VVHD_EINVAL, "phase_times: null output");
    sync(c);
    for (int p = 0; p < 8; p++) {
        float ms = -1;
        if (c.phase_recorded[p] && hipEventElapsedTime(&ms, c.ev[2 * p], c.ev[2 * p + 1]) != hipSuccess) {
            ms = -1;
            (void)hipGetLastError();
        }
        ms6[p] = ms;
    }
    API_END(ctx)
}

}  // extern "C"
