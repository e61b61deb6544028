// modules.cpp -- the libvvhd module classes (include/vvhd_modules.hpp) over
// the C ABI of libvvhd_gpu.  Host code only: every method is one or two ABI
// calls plus the host-mirror bookkeeping of Space.

#include "vvhd_modules.hpp"

#include <cstring>
#include <stdexcept>

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

void Space::set_params() { check(vvhd_set_params(ctx_, re, dt, inf_g, time)); }

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
    check(vvhd_download_bodies(ctx_, segs.data(), nseg_, bodies.data(), nb));
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
        }
    }
}

void Space::pull() {
    ensure_pushed();
    const int64_t n = vvhd_num_vortices(ctx_);
    VortexList.resize((size_t)n);
    check(vvhd_download_vortices(ctx_, reinterpret_cast<double*>(VortexList.data()), n));
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

double Space::average_segment_length() const {
    double len = 0;
    size_t n = 0;
    for (const auto& b : BodyList) {
        const size_t m = b->alist.size();
        for (size_t k = 0; k < m; k++) len += (b->alist[(k + 1) % m].corner - b->alist[k].corner).abs();
        n += m;
    }
    return n ? len / (double)n : 0;
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

// ---------------------------------------------------------------- stree --

stree::stree(Space* sS, int sFarCriteria, double sMinNodeSize, double sMaxNodeSize)
    : S(sS), farCriteria_(sFarCriteria), minNodeSize(sMinNodeSize), maxNodeSize(sMaxNodeSize) {}

void stree::build(bool IncludeVortexes, bool IncludeBody, bool IncludeHeat) {
    (void)IncludeHeat;   // no heat list
    if (!IncludeVortexes || !IncludeBody)
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

void stree::load_leaves() const {
    if (leaves_valid_) return;
    if (!built_) throw std::invalid_argument("stree::getBottomNodes(): tree is not built");
    int32_t nl = 0;
    S->check(vvhd_tree_leaves(S->ctx_, &nl, nullptr, nullptr, nullptr));
    std::vector<double> geom(4 * (size_t)nl);
    std::vector<int32_t> ints(4 * (size_t)nl);
    std::vector<uint64_t> path(nl);
    S->check(vvhd_tree_leaves(S->ctx_, &nl, geom.data(), ints.data(), path.data()));
    leaves_.assign(nl, TSortedNode());
    leaf_ptrs_.resize(nl);
    for (int32_t l = 0; l < nl; l++) {
        TSortedNode& t = leaves_[l];
        t.x = geom[4 * l];
        t.y = geom[4 * l + 1];
        t.h = geom[4 * l + 2];
        t.w = geom[4 * l + 3];
        t.v0 = ints[4 * l];
        t.v1 = ints[4 * l + 1];
        t.i = ints[4 * l + 2];
        t.j = (int)(uint32_t)path[l];   // snode::j wraps at 32 bits
        leaf_ptrs_[l] = &t;
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
    return leaf >= 0 ? &leaves_[leaf] : nullptr;
}

void stree::printBottomNodes(FILE* f, bool PrintDepth) const {
    load_leaves();
    for (const TSortedNode& t : leaves_) {
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

// -------------------------------------------------------- MDiffusiveFast --

void MDiffusiveFast::process_vort_list() {
    (void)tree;
    S->set_params();
    S->check(vvhd_diffusive(S->ctx_));
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
