// modules.cpp -- the libvvhd module classes (include/vvhd_modules.hpp) over
// the C ABI of libvvhd_gpu.  Host code only: every method is one or two ABI
// calls plus the host-mirror bookkeeping of Space.

#include "vvhd_modules.hpp"

#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "vvhd_hdf.h"

namespace {

bool vortex_dtype_ok() { return sizeof(TObj) == 48; }

}  // namespace

// ---------------------------------------------------------------- Space --

Space::Space(int device) {
    if (vvhd_create(device, &ctx_) != VVHD_OK || !ctx_)
        throw std::runtime_error("Space: no usable GPU (vvhd_create failed)");
    (void)vortex_dtype_ok();
}

Space::~Space() {
    if (ctx_) vvhd_destroy(ctx_);
}

void Space::check(int rc, bool state_error) const {
    if (rc == VVHD_OK) return;
    const std::string msg = vvhd_last_error(ctx_);
    if (rc == VVHD_ESTATE || rc == VVHD_EINVAL || state_error) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

void Space::set_params() {
    check(vvhd_set_params(ctx_, re, dt, inf_g, time));
    check(vvhd_set_pr(ctx_, pr));
}

void Space::push() {
    std::vector<vvhd_seg> segs;
    std::vector<vvhd_body> bodies;
    for (const auto& bp : BodyList) {
        const TBody& b = *bp;
        vvhd_body r;
        std::memset(&r, 0, sizeof(r));
        r.first_seg = (int32_t)segs.size();
        r.nseg = (int32_t)b.alist.size();
        r.special_segment_no = b.special_segment_no;
        r.kutta = b.kutta;
        const TVec3D* v3[9] = {&b.holder, &b.dpos, &b.speed_slae, &b.speed_slae_prev, &b.force_hydro,
                               &b.force_holder, &b.friction, &b.friction_prev, &b.fdt_dead};
        double* dst[9] = {r.holder, r.dpos, r.speed_slae, r.speed_slae_prev, r.force_hydro,
                          r.force_holder, r.friction, r.friction_prev, r.fdt_dead};
        for (int k = 0; k < 9; k++) {
            dst[k][0] = v3[k]->r.x;
            dst[k][1] = v3[k]->r.y;
            dst[k][2] = v3[k]->o;
        }
        r.g_dead = b.g_dead;
        r.nusselt = b.nusselt;
        for (const TAtt& a : b.alist) {
            vvhd_seg s;
            std::memset(&s, 0, sizeof(s));
            s.x = a.r.x;
            s.y = a.r.y;
            s.g = a.g;
            s.inv_eps = a._1_eps;
            s.cx = a.corner.x;
            s.cy = a.corner.y;
            s.dlx = a.dl.x;
            s.dly = a.dl.y;
            s.gsum = a.gsum;
            s.hsum = a.hsum;
            s.fric = a.fric;
            s.Cp = a.Cp;
            s.Fr = a.Fr;
            s.Nu = a.Nu;
            s.slip = a.slip;
            s.eq_no = a.eq_no;
            segs.push_back(s);
        }
        bodies.push_back(r);
    }
    nseg_ = (int32_t)segs.size();
    set_params();
    check(vvhd_upload_bodies(ctx_, segs.data(), nseg_, bodies.data(), (int32_t)bodies.size()));
    check(vvhd_upload_vortices(ctx_, reinterpret_cast<const double*>(VortexList.data()),
                               (int64_t)VortexList.size()));
    // heat conditions, heat constants and layer particles of the bodies
    std::vector<int32_t> cond;
    std::vector<float> hconst;
    std::vector<int64_t> layer;
    for (const auto& bp : BodyList) {
        cond.push_back((int32_t)bp->heat_condition);
        for (const TAtt& a : bp->alist) {
            hconst.push_back(a.heat_const);
            layer.push_back(a.heat_layer_obj_no);
        }
    }
    check(vvhd_set_heat(ctx_, cond.data(), hconst.data()));
    check(vvhd_set_heat_layer(ctx_, layer.data(), nseg_));
    check(vvhd_set_pr(ctx_, pr));
    const std::vector<TObj>* lists[3] = {&HeatList, &StreakList, &StreakSourceList};
    const int ids[3] = {VVHD_LIST_HEAT, VVHD_LIST_STREAK, VVHD_LIST_STREAK_SOURCE};
    for (int q = 0; q < 3; q++)
        check(vvhd_upload_list(ctx_, ids[q], reinterpret_cast<const double*>(lists[q]->data()),
                               (int64_t)lists[q]->size()));
    pushed_ = true;
    pull_bodies();
}

void Space::ensure_pushed() {
    if (!pushed_) push();
}

void Space::pull_bodies() {
    const int32_t nb = (int32_t)BodyList.size();
    if (!nb) return;
    std::vector<vvhd_seg> segs(nseg_);
    std::vector<vvhd_body> bodies(nb);
    std::vector<int64_t> layer(nseg_);
    check(vvhd_download_bodies(ctx_, segs.data(), nseg_, bodies.data(), nb));
    check(vvhd_heat_layer(ctx_, layer.data(), nseg_));
    for (int32_t bi = 0; bi < nb; bi++) {
        TBody& b = *BodyList[bi];
        const vvhd_body& r = bodies[bi];
        TVec3D* v3[9] = {&b.holder, &b.dpos, &b.speed_slae, &b.speed_slae_prev, &b.force_hydro,
                         &b.force_holder, &b.friction, &b.friction_prev, &b.fdt_dead};
        const double* src[9] = {r.holder, r.dpos, r.speed_slae, r.speed_slae_prev, r.force_hydro,
                                r.force_holder, r.friction, r.friction_prev, r.fdt_dead};
        for (int k = 0; k < 9; k++) {
            v3[k]->r = TVec(src[k][0], src[k][1]);
            v3[k]->o = src[k][2];
        }
        b.g_dead = r.g_dead;
        b.nusselt = r.nusselt;
        b.slen_ = r.slen;
        b.area_ = r.area;
        b.cofm_ = TVec(r.cofm[0], r.cofm[1]);
        b.moi_cofm_ = r.moi_cofm;
        for (int32_t k = 0; k < r.nseg; k++) {
            const vvhd_seg& s = segs[r.first_seg + k];
            TAtt& a = b.alist[k];
            a.r = TVec(s.x, s.y);
            a.g = s.g;
            a._1_eps = s.inv_eps;
            a.corner = TVec(s.cx, s.cy);
            a.dl = TVec(s.dlx, s.dly);
            a.gsum = s.gsum;
            a.hsum = s.hsum;
            a.fric = s.fric;
            a.Cp = s.Cp;
            a.Fr = s.Fr;
            a.Nu = s.Nu;
            a.slip = s.slip;
            a.eq_no = s.eq_no;
            a.heat_layer_obj_no = layer[r.first_seg + k];
        }
    }
}

void Space::pull_lists() {
    std::vector<TObj>* lists[2] = {&HeatList, &StreakList};
    const int ids[2] = {VVHD_LIST_HEAT, VVHD_LIST_STREAK};
    for (int q = 0; q < 2; q++) {
        const int64_t n = vvhd_list_size(ctx_, ids[q]);
        lists[q]->resize((size_t)n);
        check(vvhd_download_list(ctx_, ids[q], reinterpret_cast<double*>(lists[q]->data()), n));
    }
}

void Space::pull() {
    ensure_pushed();
    const int64_t n = vvhd_num_vortices(ctx_);
    VortexList.resize((size_t)n);
    check(vvhd_download_vortices(ctx_, reinterpret_cast<double*>(VortexList.data()), n));
    pull_lists();
    pull_bodies();
}

int64_t Space::vortex_count() const { return pushed_ ? vvhd_num_vortices(ctx_) : (int64_t)VortexList.size(); }

void Space::calc_forces() {
    ensure_pushed();
    set_params();
    check(vvhd_calc_forces(ctx_));
    pull_bodies();
}

void Space::zero_forces() {
    ensure_pushed();
    check(vvhd_zero_forces(ctx_));
    pull_bodies();
}

// Space::average_segment_length (TSpace.hpp:119-129): the first body's
// contour length over its segment count minus one.
double Space::average_segment_length() const {
    if (BodyList.empty()) return 0;
    const TBody& b = *BodyList.front();
    const size_t m = b.alist.size();
    if (m <= 1) return 0;
    double len = b.get_slen();
    if (!(len > 0)) {   // before the first push: from the corners
        len = 0;
        for (size_t k = 0; k < m; k++) len += (b.alist[(k + 1) % m].corner - b.alist[k].corner).abs();
    }
    return len / (double)(m - 1);
}

double Space::gsum() {
    ensure_pushed();
    double out[5];
    check(vvhd_integrals(ctx_, out));
    return out[0];
}

double Space::gmax() {
    ensure_pushed();
    double out[5];
    check(vvhd_integrals(ctx_, out));
    return out[4];
}

double Space::integral() {
    ensure_pushed();
    double out[5];
    check(vvhd_integrals(ctx_, out));
    return out[1];
}

TVec Space::hydrodynamic_momentum() {
    ensure_pushed();
    double out[5];
    check(vvhd_integrals(ctx_, out));
    return TVec(out[2], out[3]);
}

// ------------------------------------------------------------ snapshots --

namespace {

// libvvhd_hdf.so beside this library, opened on first use.
struct HdfLib {
    decltype(&vvhd_hdf_save) save = nullptr;
    decltype(&vvhd_hdf_sizes) sizes = nullptr;
    decltype(&vvhd_hdf_load) load = nullptr;
    decltype(&vvhd_hdf_error) error = nullptr;
    decltype(&vvhd_stepdata_open) step_open = nullptr;
    decltype(&vvhd_stepdata_write) step_write = nullptr;
    decltype(&vvhd_stepdata_flush) step_flush = nullptr;
    decltype(&vvhd_stepdata_close) step_close = nullptr;
};

const HdfLib& hdf_lib() {
    static HdfLib L;
    static bool done = false;
    if (done) return L;
    Dl_info info;
    std::string dir;
    if (dladdr(reinterpret_cast<void*>(&hdf_lib), &info) && info.dli_fname) {
        dir = info.dli_fname;
        dir = dir.substr(0, dir.find_last_of('/') + 1);
    }
    void* h = dlopen((dir + "libvvhd_hdf.so").c_str(), RTLD_NOW | RTLD_LOCAL);
    if (!h) throw std::runtime_error(std::string("libvvhd_hdf.so: ") + dlerror());
    L.save = reinterpret_cast<decltype(&vvhd_hdf_save)>(dlsym(h, "vvhd_hdf_save"));
    L.sizes = reinterpret_cast<decltype(&vvhd_hdf_sizes)>(dlsym(h, "vvhd_hdf_sizes"));
    L.load = reinterpret_cast<decltype(&vvhd_hdf_load)>(dlsym(h, "vvhd_hdf_load"));
    L.error = reinterpret_cast<decltype(&vvhd_hdf_error)>(dlsym(h, "vvhd_hdf_error"));
    L.step_open = reinterpret_cast<decltype(&vvhd_stepdata_open)>(dlsym(h, "vvhd_stepdata_open"));
    L.step_write = reinterpret_cast<decltype(&vvhd_stepdata_write)>(dlsym(h, "vvhd_stepdata_write"));
    L.step_flush = reinterpret_cast<decltype(&vvhd_stepdata_flush)>(dlsym(h, "vvhd_stepdata_flush"));
    L.step_close = reinterpret_cast<decltype(&vvhd_stepdata_close)>(dlsym(h, "vvhd_stepdata_close"));
    if (!L.save || !L.sizes || !L.load || !L.error || !L.step_open || !L.step_write || !L.step_flush || !L.step_close)
        throw std::runtime_error("libvvhd_hdf.so: missing entry points");
    done = true;
    return L;
}

void copy_str(char* dst, const std::string& src) { std::snprintf(dst, VVHD_HDF_STR, "%s", src.c_str()); }

}  // namespace

void Space::save(const char* path) {
    pull();
    const HdfLib& H = hdf_lib();
    vvhd_hdf_space sp;
    std::memset(&sp, 0, sizeof(sp));
    copy_str(sp.caption, caption);
    const TTime* tt[5] = {&time, &dt, &dt_save, &dt_streak, &dt_profile};
    double* tv[5] = {&sp.time, &sp.dt, &sp.dt_save, &sp.dt_streak, &sp.dt_profile};
    for (int k = 0; k < 5; k++) {
        *tv[k] = double(*tt[k]);
        sp.frac_value[k] = tt[k]->value;
        sp.frac_scale[k] = tt[k]->timescale;
    }
    sp.re = re;
    sp.pr = pr;
    sp.inf_g = inf_g;
    sp.finish = finish;
    sp.inf_marker[0] = inf_marker.x;
    sp.inf_marker[1] = inf_marker.y;
    sp.gravity[0] = gravity.x;
    sp.gravity[1] = gravity.y;
    copy_str(sp.inf_speed_x, inf_vx);
    copy_str(sp.inf_speed_y, inf_vy);
    const std::vector<TObj> empty;
    const std::vector<TObj>* lists[VVHD_HDF_LISTS] = {&VortexList, &HeatList, &StreakList, &StreakSourceList, &empty};
    const double* ptrs[VVHD_HDF_LISTS];
    int64_t counts[VVHD_HDF_LISTS];
    for (int k = 0; k < VVHD_HDF_LISTS; k++) {
        ptrs[k] = reinterpret_cast<const double*>(lists[k]->data());
        counts[k] = (int64_t)lists[k]->size();
    }
    std::vector<vvhd_hdf_body> bodies;
    std::vector<double> seg4;
    std::vector<uint32_t> slip;
    for (const auto& bp : BodyList) {
        const TBody& b = *bp;
        vvhd_hdf_body r;
        std::memset(&r, 0, sizeof(r));
        const TVec3D* v3[5] = {&b.holder, &b.dpos, &b.speed_slae, &b.speed_slae_prev, &b.force_hydro};
        double* d3[5] = {r.holder, r.dpos, r.speed_slae, r.speed_slae_prev, r.force_hydro};
        for (int k = 0; k < 5; k++) {
            d3[k][0] = v3[k]->r.x;
            d3[k][1] = v3[k]->r.y;
            d3[k][2] = v3[k]->o;
        }
        const TVec3D* w3[3] = {&b.force_holder, &b.friction_prev, &b.fdt_dead};
        double* e3[3] = {r.force_holder, r.friction_prev, r.fdt_dead};
        for (int k = 0; k < 3; k++) {
            e3[k][0] = w3[k]->r.x;
            e3[k][1] = w3[k]->r.y;
            e3[k][2] = w3[k]->o;
        }
        for (int k = 0; k < 3; k++) {   // rigid: infinite springs (TBody.cpp:24-30)
            r.kspring[k] = -1;
            r.damping[k] = 0;
        }
        copy_str(r.speed_x, "0");
        copy_str(r.speed_y, "0");
        copy_str(r.speed_o, "0");
        r.density = 1;
        r.g_dead = b.g_dead;
        r.area = b.get_area();
        r.cofm[0] = b.get_cofm().x;
        r.cofm[1] = b.get_cofm().y;
        r.moi_cofm = b.get_moi_cofm();
        const TVec ax = b.get_axis();
        r.moi_axis = b.get_moi_cofm() + (ax - b.get_cofm()).abs2() * b.get_area();
        r.special_segment_no = b.special_segment_no;
        r.boundary_condition = b.kutta ? 1 : 0;
        r.heat_condition = (uint32_t)b.heat_condition;
        r.heat_const = b.alist.empty() ? 0 : (double)b.alist.front().heat_const;
        r.nseg = (int32_t)b.alist.size();
        for (const TAtt& a : b.alist) {
            seg4.insert(seg4.end(), {a.corner.x, a.corner.y, a.g, a.gsum});
            slip.push_back(a.slip);
        }
        bodies.push_back(r);
    }
    if (H.save(path, &sp, ptrs, counts, (int32_t)bodies.size(), bodies.data(), seg4.data(), slip.data()) != 0)
        throw std::runtime_error(std::string("Space::save: ") + H.error());
}

void Space::load(const char* path) {
    const HdfLib& H = hdf_lib();
    int64_t counts[VVHD_HDF_LISTS];
    int32_t nb = 0, ns = 0;
    if (H.sizes(path, counts, &nb, &ns) != 0) throw std::runtime_error(std::string("Space::load: ") + H.error());
    std::vector<TObj> src(counts[4]);
    std::vector<TObj>* lists[VVHD_HDF_LISTS] = {&VortexList, &HeatList, &StreakList, &StreakSourceList, &src};
    double* ptrs[VVHD_HDF_LISTS];
    for (int k = 0; k < VVHD_HDF_LISTS; k++) {
        lists[k]->assign((size_t)counts[k], TObj());
        ptrs[k] = reinterpret_cast<double*>(lists[k]->data());
    }
    std::vector<vvhd_hdf_body> bodies(std::max(nb, 1));
    std::vector<double> seg4(4 * (size_t)std::max(ns, 1));
    std::vector<uint32_t> slip(std::max(ns, 1));
    vvhd_hdf_space sp;
    if (H.load(path, &sp, ptrs, bodies.data(), seg4.data(), slip.data()) != 0)
        throw std::runtime_error(std::string("Space::load: ") + H.error());
    caption = sp.caption;
    TTime* tt[5] = {&time, &dt, &dt_save, &dt_streak, &dt_profile};
    for (int k = 0; k < 5; k++) *tt[k] = TTime(sp.frac_value[k], sp.frac_scale[k]);
    re = sp.re;
    pr = sp.pr;
    inf_g = sp.inf_g;
    finish = sp.finish;
    inf_marker = TVec(sp.inf_marker[0], sp.inf_marker[1]);
    gravity = TVec(sp.gravity[0], sp.gravity[1]);
    inf_vx = sp.inf_speed_x;
    inf_vy = sp.inf_speed_y;
    char* end = nullptr;
    const double ux = std::strtod(sp.inf_speed_x, &end);
    if (!(end && *end == 0 && end != sp.inf_speed_x)) throw std::invalid_argument("Space::load: inf_speed_x is not a constant");
    const double uy = std::strtod(sp.inf_speed_y, &end);
    if (!(end && *end == 0 && end != sp.inf_speed_y)) throw std::invalid_argument("Space::load: inf_speed_y is not a constant");
    inf_vel = TVec(ux, uy);
    BodyList.clear();
    size_t s = 0;
    for (int32_t b = 0; b < nb; b++) {
        const vvhd_hdf_body& r = bodies[b];
        auto body = std::make_shared<TBody>();
        for (int32_t k = 0; k < r.nseg; k++, s++) {
            body->alist.emplace_back(seg4[4 * s], seg4[4 * s + 1], slip[s] != 0);
            TAtt& a = body->alist.back();
            a.g = seg4[4 * s + 2];
            a.gsum = seg4[4 * s + 3];
            a.heat_const = (float)r.heat_const;
        }
        const double* d3[5] = {r.holder, r.dpos, r.speed_slae, r.speed_slae_prev, r.force_hydro};
        TVec3D* v3[5] = {&body->holder, &body->dpos, &body->speed_slae, &body->speed_slae_prev, &body->force_hydro};
        for (int k = 0; k < 5; k++) *v3[k] = TVec3D{TVec(d3[k][0], d3[k][1]), d3[k][2]};
        body->force_holder = TVec3D{TVec(r.force_holder[0], r.force_holder[1]), r.force_holder[2]};
        body->friction_prev = TVec3D{TVec(r.friction_prev[0], r.friction_prev[1]), r.friction_prev[2]};
        body->fdt_dead = TVec3D{TVec(r.fdt_dead[0], r.fdt_dead[1]), r.fdt_dead[2]};
        body->g_dead = r.g_dead;
        body->special_segment_no = r.special_segment_no;
        body->kutta = r.boundary_condition == 1;
        body->heat_condition = static_cast<hc_t>(r.heat_condition);
        BodyList.push_back(body);
    }
    push();   // segments and body geometry are rebuilt from the corners on the device
}

// ---------------------------------------------------------------- stree --

stree::stree(Space* sS, int sFarCriteria, double sMinNodeSize, double sMaxNodeSize)
    : S(sS), farCriteria_(sFarCriteria), minNodeSize(sMinNodeSize), maxNodeSize(sMaxNodeSize) {}

void stree::build(bool IncludeVortexes, bool IncludeBody, bool IncludeHeat) {
    if (!IncludeVortexes || !IncludeBody || (!IncludeHeat && vvhd_list_size(S->ctx_, VVHD_LIST_HEAT) > 0))
        throw std::invalid_argument("stree::build: partial trees are not supported by the device path");
    if (built_) {
        fprintf(stderr, "Tree is already built\n");   // TSortedTree.cpp:234
        return;
    }
    S->ensure_pushed();
    S->set_params();
    S->check(vvhd_tree_build(S->ctx_, farCriteria_, minNodeSize, maxNodeSize, nullptr));
    built_ = true;
    leaves_valid_ = false;
}

void stree::destroy() {
    S->check(vvhd_tree_destroy(S->ctx_));
    built_ = false;
    leaves_valid_ = false;
}

// Views of every node with its monopoles, the leaves in DFS order and their
// near/far lists, the nodes of a list found by (depth, path code).
void stree::load_leaves() const {
    if (leaves_valid_) return;
    if (!built_) throw std::invalid_argument("stree::getBottomNodes(): tree is not built");
    vvhd_ctx* ctx = S->ctx_;
    int32_t nn = 0, nl = 0;
    S->check(vvhd_tree_nodes(ctx, &nn, nullptr, nullptr, nullptr, nullptr));
    std::vector<double> geom(4 * (size_t)nn), cm(6 * (size_t)nn);
    std::vector<int32_t> ints(4 * (size_t)nn);
    std::vector<uint64_t> path(nn);
    S->check(vvhd_tree_nodes(ctx, &nn, geom.data(), ints.data(), path.data(), cm.data()));
    const bool mirror = (int64_t)S->VortexList.size() == vvhd_num_vortices(ctx);
    nodes_.assign(nn, TSortedNode());
    std::unordered_map<uint64_t, int32_t> by_key;   // path code with a leading bit at its depth
    by_key.reserve((size_t)nn * 2);
    auto key = [](int depth, uint64_t p) { return (depth < 64 ? (1ull << depth) : 0ull) | p; };
    for (int32_t k = 0; k < nn; k++) {
        TSortedNode& t = nodes_[k];
        t.x = geom[4 * k];
        t.y = geom[4 * k + 1];
        t.h = geom[4 * k + 2];
        t.w = geom[4 * k + 3];
        t.v0 = ints[4 * k];
        t.v1 = ints[4 * k + 1];
        t.i = ints[4 * k + 3];
        t.j = (int)(uint32_t)path[k];   // snode::j wraps at 32 bits
        t.CMp = TObj(cm[6 * k], cm[6 * k + 1], cm[6 * k + 2]);
        t.CMm = TObj(cm[6 * k + 3], cm[6 * k + 4], cm[6 * k + 5]);
        const int c = ints[4 * k + 2];
        t.ch1 = c >= 0 ? &nodes_[c] : nullptr;
        t.ch2 = c >= 0 ? &nodes_[c + 1] : nullptr;
        t.tree = this;
        if (mirror && t.v1 > t.v0) {
            t.vRange.first = S->VortexList.data() + t.v0;
            t.vRange.last = S->VortexList.data() + t.v1;
        }
        by_key[key(t.i, path[k])] = k;
    }
    S->check(vvhd_tree_leaves(ctx, &nl, nullptr, nullptr, nullptr));
    std::vector<int32_t> lints(4 * (size_t)nl);
    std::vector<uint64_t> lpath(nl);
    S->check(vvhd_tree_leaves(ctx, &nl, nullptr, lints.data(), lpath.data()));
    std::vector<int32_t> hs(4 * (size_t)std::max(nl, 1));
    if (nl) S->check(vvhd_tree_leaf_lists(ctx, hs.data()));
    const bool hmirror = (int64_t)S->HeatList.size() == vvhd_list_size(ctx, VVHD_LIST_HEAT);
    const bool smirror = (int64_t)S->StreakList.size() == vvhd_list_size(ctx, VVHD_LIST_STREAK);
    leaf_ptrs_.resize(nl);
    near_.assign(nl, {});
    far_.assign(nl, {});
    for (int32_t l = 0; l < nl; l++) {
        TSortedNode* t = &nodes_[by_key.at(key(lints[4 * l + 2], lpath[l]))];
        t->leaf = l;
        t->h0 = hs[4 * l];
        t->h1 = hs[4 * l + 1];
        t->s0 = hs[4 * l + 2];
        t->s1 = hs[4 * l + 3];
        if (hmirror && t->h1 > t->h0) {
            t->hRange.first = S->HeatList.data() + t->h0;
            t->hRange.last = S->HeatList.data() + t->h1;
        }
        if (smirror && t->s1 > t->s0) {
            t->sRange.first = S->StreakList.data() + t->s0;
            t->sRange.last = S->StreakList.data() + t->s1;
        }
        t->NearNodes = &near_[l];
        t->FarNodes = &far_[l];
        leaf_ptrs_[l] = t;
    }
    for (int far = 0; far < 2; far++) {
        int64_t cnt = 0;
        S->check(vvhd_tree_lists(ctx, far, &cnt, nullptr, nullptr, nullptr, nullptr));
        std::vector<int64_t> off((size_t)nl + 1);
        std::vector<int32_t> dep((size_t)std::max<int64_t>(cnt, 1));
        std::vector<uint64_t> pth((size_t)std::max<int64_t>(cnt, 1));
        S->check(vvhd_tree_lists(ctx, far, &cnt, off.data(), dep.data(), pth.data(), nullptr));
        auto& lists = far ? far_ : near_;
        for (int32_t l = 0; l < nl; l++) {
            lists[l].reserve((size_t)(off[l + 1] - off[l]));
            for (int64_t e = off[l]; e < off[l + 1]; e++) lists[l].push_back(&nodes_[by_key.at(key(dep[e], pth[e]))]);
        }
    }
    leaves_valid_ = true;
}

const std::vector<TSortedNode*>& stree::getBottomNodes() const {
    load_leaves();
    return leaf_ptrs_;
}

const TSortedNode* stree::findNode(TVec p) const {
    if (!built_) throw std::invalid_argument("TTree::findNode(): tree is not built");   // TSortedTree.cpp:287
    load_leaves();
    int32_t leaf = -1;
    const double xy[2] = {p.x, p.y};
    S->check(vvhd_tree_find(S->ctx_, xy, 1, &leaf));
    return leaf >= 0 ? leaf_ptrs_[leaf] : nullptr;
}

void stree::printBottomNodes(FILE* f, bool PrintDepth) const {
    load_leaves();
    for (const TSortedNode* tp : leaf_ptrs_) {
        const TSortedNode& t = *tp;
        if (PrintDepth)
            fprintf(f, "%g\t%g\t%g\t%g\t%d\n", t.x, t.y, t.w, t.h, t.i);
        else
            fprintf(f, "%g\t%g\t%g\t%g\n", t.x, t.y, t.w, t.h);
    }
}

// ------------------------------------------------------- MConvectiveFast --

MConvectiveFast::MConvectiveFast(Space* S_, const TSortedTree* tree_) : S(S_), tree(tree_) {}

void MConvectiveFast::process_all_lists() {
    S->set_params();
    const TVec u = S->inf_speed();
    if (learned_)
        S->check(vvhd_convective_learned(S->ctx_, u.x, u.y, kahan_));
    else
        S->check(vvhd_convective(S->ctx_, u.x, u.y, kahan_));
}

TVec MConvectiveFast::velocity(TVec p) const {
    const TVec u = S->inf_speed();
    double xy[2] = {p.x, p.y}, v[2];
    S->check(vvhd_velocity_points(S->ctx_, xy, 1, u.x, u.y, v));
    return TVec(v[0], v[1]);
}

void MConvectiveFast::calc_circulation(const void** collision) {
    S->set_params();
    const TVec u = S->inf_speed();
    S->check(vvhd_calc_circulation(S->ctx_, u.x, u.y));
    if (collision) *collision = nullptr;   // collisions are not modelled
    S->pull_bodies();
}

void MConvectiveFast::set_learned(const std::vector<float>& params) {
    S->check(vvhd_far_model_set(S->ctx_, params.data(), (int32_t)params.size()));
    learned_ = true;
}

void MConvectiveFast::set_learned_gradient(bool network) {
    S->check(vvhd_set_far_gradient(S->ctx_, network ? 1 : 0));
}

// ---------------------------------------------------------- MEpsilonFast --

void MEpsilonFast::CalcEpsilonFast(bool merge) {
    (void)Tree;
    int32_t m = 0;
    S->check(vvhd_epsilon(S->ctx_, merge ? 1 : 0, nullptr, nullptr, &m));
    merged_ = m;
}

double MEpsilonFast::near_query(const TSortedNode& node, TVec p, bool h2) {
    if (!node.tree || node.leaf < 0)
        throw std::invalid_argument("MEpsilonFast::eps2h/h2: the node is not a leaf of a built tree");
    Space* S = node.tree->space();
    const double xy[2] = {p.x, p.y};
    const int32_t leaf = node.leaf;
    double out = 0;
    S->check(h2 ? vvhd_h2(S->ctx_, xy, &leaf, 1, &out) : vvhd_eps2h(S->ctx_, xy, &leaf, 1, &out));
    return out;
}

double MEpsilonFast::eps2h(const TSortedNode& node, TVec p) { return near_query(node, p, false); }
double MEpsilonFast::h2(const TSortedNode& node, TVec p) { return near_query(node, p, true); }

// -------------------------------------------------------- MDiffusiveFast --

void MDiffusiveFast::process_vort_list() {
    (void)tree;
    S->set_params();
    S->check(vvhd_diffusive(S->ctx_));
}

void MDiffusiveFast::process_heat_list() {
    S->set_params();
    S->check(vvhd_diffusive_heat(S->ctx_));
}

// ------------------------------------------------------------- MFlowmove --

MFlowmove::MFlowmove(Space* S_, double remove_eps_) : S(S_) {
    if (remove_eps_ != 1E-10) throw std::invalid_argument("MFlowmove: the device path removes |g| < 1e-10 only");
}

void MFlowmove::move_and_clean(bool remove, const void** collision, size_t* cleaned_v) {
    S->set_params();
    int64_t cl = 0;
    S->check(vvhd_move_and_clean(S->ctx_, remove ? 1 : 0, &cl));
    if (cleaned_v) *cleaned_v = (size_t)cl;
    if (collision) *collision = nullptr;
    S->pull_bodies();
}

void MFlowmove::vortex_shed() {
    S->check(vvhd_vortex_shed(S->ctx_));
    S->pull_bodies();
}

void MFlowmove::heat_shed() {
    S->ensure_pushed();
    S->check(vvhd_heat_shed(S->ctx_));
    S->pull_bodies();
}

void MFlowmove::heat_crop(double scale) {
    S->ensure_pushed();
    S->check(vvhd_heat_crop(S->ctx_, scale));
}

// MFlowmove::streak_shed (MFlowmove.cpp:284-291): when the time is a whole
// multiple of dt_streak (exact, on the fractions).
void MFlowmove::streak_shed() {
    S->ensure_pushed();
    if (!S->time.divisibleBy(S->dt_streak)) return;
    S->check(vvhd_streak_shed(S->ctx_));
}

// ---------------------------------------------------------------- XField --

XField::XField(float xmin_, float ymin_, float dxdy_, int xres_, int yres_)
    : xmin(xmin_), ymin(ymin_), dxdy(dxdy_), xres(xres_), yres(yres_) {
    if (xres <= 1) throw std::invalid_argument("XField(): xres must be > 1");
    if (yres <= 1) throw std::invalid_argument("XField(): yres must be > 1");
    if (!std::isfinite(dxdy)) throw std::invalid_argument("XField(): dxdy must be finite");
    if (dxdy <= 0) throw std::invalid_argument("XField(): dxdy must be positive");
    map_.assign((size_t)xres * yres, 0.f);
}

float XField::min() const {
    if (!evaluated) throw std::invalid_argument("XField::min(): not evaluated");
    return *std::min_element(map_.begin(), map_.end());
}

float XField::max() const {
    if (!evaluated) throw std::invalid_argument("XField::max(): not evaluated");
    return *std::max_element(map_.begin(), map_.end());
}

// XField::percentile (XField.cpp:95-121): over the non-zero values.
float XField::percentile(float p) const {
    if (!evaluated) throw std::invalid_argument("XField::percentile(): not evaluated");
    if (!(p > 0 && p < 1)) throw std::invalid_argument("XField::percentile(p): p must be in range (0, 1)");
    const bool low = p < 0.5f;
    if (!low) p = 1 - p;
    std::vector<float> v;
    for (float x : map_)
        if (x != 0) v.push_back(x);
    const size_t N = (size_t)std::floor(v.size() * p);
    if (low)
        std::partial_sort(v.begin(), v.begin() + N, v.end(), [](float a, float b) { return a < b; });
    else
        std::partial_sort(v.begin(), v.begin() + N, v.end(), [](float a, float b) { return a > b; });
    return v[N - 1];
}

// XField's binary matrix (XField.cpp:124-152): N = xres, the x coordinates,
// then per row its y and the row's values, all float.
std::ostream& operator<<(std::ostream& os, const XField& field) {
    if (!field.evaluated) throw std::invalid_argument("XField::operator<<: not evaluated");
    const float N = (float)field.xres;
    os.write(reinterpret_cast<const char*>(&N), sizeof(float));
    for (int xi = 0; xi < field.xres; xi++) {
        const float x = field.xmin + field.dxdy * xi;
        os.write(reinterpret_cast<const char*>(&x), sizeof(float));
    }
    for (int yj = 0; yj < field.yres; yj++) {
        const float y = field.ymin + field.dxdy * yj;
        os.write(reinterpret_cast<const char*>(&y), sizeof(float));
        os.write(reinterpret_cast<const char*>(field.map_.data() + (size_t)yj * field.xres), sizeof(float) * field.xres);
    }
    return os;
}

vvhd_grid XField::grid() const {
    vvhd_grid g;
    g.xmin = xmin;
    g.ymin = ymin;
    g.dxdy = dxdy;
    g.xres = xres;
    g.yres = yres;
    return g;
}

// The evaluators' copy of the Space (the reference's evaluators hold a Space
// by value): parameters, bodies and the vortex list as the device has them.
void XField::copy_space(const Space& S, Space& copy) const {
    Space& src = const_cast<Space&>(S);
    if (src.pushed_) src.pull();
    copy.re = S.re;
    copy.pr = S.pr;
    copy.inf_g = S.inf_g;
    copy.inf_vel = S.inf_vel;
    copy.time = S.time;
    copy.dt = S.dt;
    copy.finish = S.finish;
    copy.BodyList.clear();
    for (const auto& b : S.BodyList) copy.BodyList.push_back(std::make_shared<TBody>(*b));
    copy.VortexList = S.VortexList;
    copy.push();
}

XVelocity::XVelocity(const Space& S_, double xmin, double ymin, double dxdy, int xres, int yres)
    : XField((float)xmin, (float)ymin, (float)dxdy, xres, yres), S(S_) {}

// XVelocity::evaluate (XVelocity.cpp:24-73)
void XVelocity::evaluate() {
    TVec ref(0, 0);
    switch (ref_frame) {
        case 'o': break;
        case 'f': ref = S.inf_speed(); break;
        case 'b': ref = S.BodyList.at(0)->speed_slae.r; break;
        default: throw std::invalid_argument("XVelocity(): bad ref_frame");
    }
    if (mode != 'x' && mode != 'y') throw std::invalid_argument("XPressure(): bad mode");
    if (evaluated) return;
    Space copy(0);
    copy_space(S, copy);
    const double dl = copy.average_segment_length();
    TSortedTree tree(&copy, 8, dl * 20, 0.1);
    tree.build();
    MEpsilonFast eps(&copy, &tree);
    eps.CalcEpsilonFast(false);
    const vvhd_grid g = grid();
    const TVec u = copy.inf_speed();
    copy.check(vvhd_field(copy.ctx_, mode == 'x' ? VVHD_FIELD_VX : VVHD_FIELD_VY, &g, 1.0, dl, u.x, u.y, 1, ref.x,
                          ref.y, map_.data()));
    evaluated = true;
}

XVorticity::XVorticity(const Space& S_, double xmin, double ymin, double dxdy, int xres, int yres)
    : XField((float)xmin, (float)ymin, (float)dxdy, xres, yres), S(S_) {}

// XVorticity::evaluate (XVorticity.cpp:26-72)
void XVorticity::evaluate() {
    if (eps_mult <= 0) throw std::invalid_argument("XVorticity(): eps_mult must be positive");
    if (evaluated) return;
    Space copy(0);
    copy_space(S, copy);
    const double dl = copy.average_segment_length();
    MFlowmove flowmove(&copy);
    flowmove.vortex_shed();
    TSortedTree tree(&copy, 8, dl * 20);
    tree.build();
    const vvhd_grid g = grid();
    copy.check(vvhd_field(copy.ctx_, VVHD_FIELD_VORTICITY, &g, eps_mult, dl, 0, 0, 0, 0, 0, map_.data()));
    evaluated = true;
}

XStreamfunction::XStreamfunction(const Space& S_, double xmin, double ymin, double dxdy, int xres, int yres)
    : XField((float)xmin, (float)ymin, (float)dxdy, xres, yres), S(S_) {}

// XStreamfunction::evaluate (XStreamfunction.cpp:47-118)
void XStreamfunction::evaluate() {
    if (eps_mult <= 0) throw std::invalid_argument("XStreamfunction(): eps_mult must be positive");
    TVec ref(0, 0);
    switch (ref_frame) {
        case 'o': break;
        case 'f': ref = S.inf_speed(); break;
        case 'b': ref = S.BodyList.at(0)->speed_slae.r; break;
        default: throw std::invalid_argument("XStreamfunction(): bad ref_frame");
    }
    if (evaluated) return;
    Space copy(0);
    copy_space(S, copy);
    const double dl = copy.average_segment_length();
    const double min_node = dl > 0 ? dl * 10 : 0;
    const double max_node = dl > 0 ? dl * 20 : std::numeric_limits<double>::infinity();
    TSortedTree tree(&copy, 8, min_node, max_node);
    tree.build();
    const vvhd_grid g = grid();
    const TVec u = copy.inf_speed();
    copy.check(vvhd_field(copy.ctx_, VVHD_FIELD_STREAMFUNCTION, &g, eps_mult, dl, u.x, u.y, 1, ref.x, ref.y,
                          map_.data()));
    evaluated = true;
}

XPressure::XPressure(const Space& S_, double xmin, double ymin, double dxdy, int xres, int yres)
    : XField((float)xmin, (float)ymin, (float)dxdy, xres, yres), S(S_) {}

// XPressure::evaluate (XPressure.cpp:32-100)
void XPressure::evaluate() {
    if (eps_mult <= 0) throw std::invalid_argument("XPressure(): eps_mult must be positive");
    TVec ref(0, 0);
    switch (ref_frame) {
        case 's': break;
        case 'o': break;
        case 'f': ref = S.inf_speed(); break;
        case 'b': ref = S.BodyList.at(0)->speed_slae.r; break;
        default: throw std::invalid_argument("XPressure(): bad ref_frame");
    }
    if (evaluated) return;
    Space copy(0);
    copy_space(S, copy);
    const double dl = copy.average_segment_length();
    MFlowmove flowmove(&copy);
    flowmove.vortex_shed();
    TSortedTree tree(&copy, 8, dl * 20, 0.1);
    tree.build();
    MEpsilonFast eps(&copy, &tree);
    eps.CalcEpsilonFast(false);
    MConvectiveFast convective(&copy, &tree);
    convective.process_all_lists();
    MDiffusiveFast diffusive(&copy, &tree);
    diffusive.process_vort_list();
    const vvhd_grid g = grid();
    const TVec u = copy.inf_speed();
    copy.check(vvhd_field(copy.ctx_, VVHD_FIELD_PRESSURE, &g, eps_mult, dl, u.x, u.y, ref_frame != 's', ref.x,
                          ref.y, map_.data()));
    evaluated = true;
}

// --------------------------------------------------------------- sensors --

sensors::sensors(Space* S_, MConvectiveFast* conv_, const char* sensors_file, const char* output)
    : S(S_), conv(conv_), fout(nullptr), slist() {
    loadFile(sensors_file);
    if (slist.size()) fout = fopen(output, "a");
}

sensors::~sensors() {
    if (fout) fclose(fout);
    fout = nullptr;
}

void sensors::loadFile(const char* file) {
    if (!file) return;
    FILE* fin = fopen(file, "r");
    if (!fin) {
        fprintf(stderr, "No file called '%s'\n", file);
        return;
    }
    TVec vec(0, 0);
    while (fscanf(fin, "%lf %lf", &vec.x, &vec.y) == 2) slist.push_back(vec);
    fclose(fin);
}

// sensors::output (sensors.cpp:58-69): MConvectiveFast::velocity at every
// probe, all probes in one device call.
void sensors::output() {
    if (!fout) return;
    (void)conv;
    std::vector<double> xy(2 * slist.size()), v(2 * slist.size());
    for (size_t k = 0; k < slist.size(); k++) {
        xy[2 * k] = slist[k].x;
        xy[2 * k + 1] = slist[k].y;
    }
    const TVec u = S->inf_speed();
    S->check(vvhd_velocity_points(S->ctx_, xy.data(), (int64_t)slist.size(), u.x, u.y, v.data()));
    fprintf(fout, "%lg", double(S->time));
    for (size_t k = 0; k < slist.size(); k++) fprintf(fout, " \t%lg \t%lg", v[2 * k], v[2 * k + 1]);
    fprintf(fout, "\n");
    fflush(fout);
}

// ------------------------------------------------------------ step data --

Stepdata::Stepdata(Space* S_, const char* fname, bool profile) : S(S_), b_save_profile(profile) {
    const HdfLib& H = hdf_lib();
    std::vector<int32_t> nseg;
    for (const auto& b : S->BodyList) nseg.push_back((int32_t)b->alist.size());
    h_ = H.step_open(fname, S->caption.c_str(), S->time, (int32_t)nseg.size(), nseg.data(), profile ? 1 : 0);
    if (!h_) throw std::runtime_error(std::string("Stepdata: ") + H.error());
}

Stepdata::~Stepdata() {
    if (h_) hdf_lib().step_close(h_);
}

void Stepdata::flush() { hdf_lib().step_flush(h_); }

void Stepdata::write() {
    const HdfLib& H = hdf_lib();
    std::vector<vvhd_stepdata_body> rows(S->BodyList.size());
    std::vector<double> pressure, friction;
    for (size_t bi = 0; bi < S->BodyList.size(); bi++) {
        TBody& b = *S->BodyList[bi];
        const TVec3D* src[6] = {&b.force_hydro, &b.force_holder, &b.friction, &b.holder, &b.dpos, &b.speed_slae};
        double* dst[6] = {rows[bi].force_hydro, rows[bi].force_holder, rows[bi].friction,
                          rows[bi].holder,      rows[bi].dpos,         rows[bi].speed_slae};
        for (int k = 0; k < 6; k++) {
            dst[k][0] = src[k]->r.x;
            dst[k][1] = src[k]->r.y;
            dst[k][2] = src[k]->o;
        }
        if (b_save_profile)
            for (TAtt& a : b.alist) {
                pressure.push_back(a.Cp / S->dt);
                friction.push_back(a.Fr / S->dt);
                a.Cp = a.Fr = a.Nu = 0;
            }
    }
    if (H.step_write(h_, S->time, rows.data(), b_save_profile ? pressure.data() : nullptr,
                     b_save_profile ? friction.data() : nullptr) != 0)
        throw std::runtime_error(std::string("Stepdata::write: ") + H.error());
    if (b_save_profile && !S->BodyList.empty()) S->check(vvhd_zero_profiles(S->ctx_));
}
