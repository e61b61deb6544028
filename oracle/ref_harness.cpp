// ref_harness.cpp -- drives the reference's own tree, epsilon and diffusive
// translation units (vvflow/vvflow libvvhd/src/TSortedTree.cpp,
// MEpsilonFast.cpp, MDiffusiveFast.cpp, compiled unmodified from
// /root/reference by oracle/Makefile) so tests can compare the FP64
// restatement with the reference bit for bit.
//
// TEST INFRASTRUCTURE ONLY (built into oracle/_ref/, never shipped).
//
// Those three translation units need nothing but libstdc++/libm.  The
// reference's Space and TBody classes, however, also own Lua-backed TEval
// members whose constructors live in TEval.cpp (Lua is not in this image).
// The tree/epsilon/diffusive code never touches those members, so the
// harness sets up the Space and TBody storage member by member (the
// std::vector lists, re) and never constructs, reads or destroys a TEval.
// No stand-in for Lua, TEval, LAPACK or any other missing dependency exists.
// Segment geometry (r, dl, _1_eps) is supplied by the caller.

#include "MDiffusiveFast.hpp"
#include "MEpsilonFast.hpp"
#include "TSortedTree.hpp"

#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

namespace {

struct SegIn {   // same layout as vvo_seg_state (oracle/vvhd_oracle.h)
    double x, y, g, inv_eps, cx, cy, dlx, dly, gsum, hsum, fric, Cp, Fr, Nu;
    uint32_t slip, eq_no;
};

struct Harness {
    alignas(Space) unsigned char space_mem[sizeof(Space)];
    Space* S;
    std::vector<unsigned char*> body_mem;
    stree* tree = nullptr;
    int merged = 0;
};

void init_space(Harness* h) {
    h->S = reinterpret_cast<Space*>(h->space_mem);
    Space* S = h->S;
    new (&S->BodyList) std::vector<std::shared_ptr<TBody>>();
    new (&S->VortexList) std::vector<TObj>();
    new (&S->HeatList) std::vector<TObj>();
    new (&S->SourceList) std::vector<TObj>();
    new (&S->StreakSourceList) std::vector<TObj>();
    new (&S->StreakList) std::vector<TObj>();
    S->re = 0;
    S->pr = 0;
}

}  // namespace

extern "C" {

void* vvr_create(void) {
    Harness* h = new Harness();
    init_space(h);
    return h;
}

void vvr_destroy(void* p) {
    Harness* h = static_cast<Harness*>(p);
    if (h->tree) {
        h->tree->destroy();
        delete h->tree;
    }
    Space* S = h->S;
    for (auto& b : S->BodyList) b->alist.~vector();
    S->BodyList.~vector();
    S->VortexList.~vector();
    S->HeatList.~vector();
    S->SourceList.~vector();
    S->StreakSourceList.~vector();
    S->StreakList.~vector();
    for (unsigned char* m : h->body_mem) delete[] m;
    delete h;
}

void vvr_set_re(void* p, double re) { static_cast<Harness*>(p)->S->re = re; }
void vvr_set_pr(void* p, double pr) { static_cast<Harness*>(p)->S->pr = pr; }

static std::vector<TObj>* list_of(Space* S, int list) {
    switch (list) {
        case 0: return &S->VortexList;
        case 1: return &S->HeatList;
        case 2: return &S->StreakList;
        default: return nullptr;
    }
}

// list: 0 VortexList, 1 HeatList, 2 StreakList; records of 6 doubles.
int vvr_set_list(void* p, int list, const double* o6, int64_t n) {
    std::vector<TObj>* L = list_of(static_cast<Harness*>(p)->S, list);
    if (!L) return 1;
    L->clear();
    for (int64_t i = 0; i < n; i++) {
        TObj o(o6[6 * i], o6[6 * i + 1], o6[6 * i + 2]);
        o.v = TVec(o6[6 * i + 3], o6[6 * i + 4]);
        o._1_eps = o6[6 * i + 5];
        L->push_back(o);
    }
    return 0;
}

int64_t vvr_list_size(void* p, int list) {
    std::vector<TObj>* L = list_of(static_cast<Harness*>(p)->S, list);
    return L ? (int64_t)L->size() : -1;
}

int vvr_get_list(void* p, int list, double* o6) {
    std::vector<TObj>* L = list_of(static_cast<Harness*>(p)->S, list);
    if (!L) return 1;
    for (size_t i = 0; i < L->size(); i++) {
        const TObj& o = (*L)[i];
        o6[6 * i] = o.r.x;
        o6[6 * i + 1] = o.r.y;
        o6[6 * i + 2] = o.g;
        o6[6 * i + 3] = o.v.x;
        o6[6 * i + 4] = o.v.y;
        o6[6 * i + 5] = o._1_eps;
    }
    return 0;
}

void vvr_set_vortices(void* p, const double* o6, int64_t n) {
    Space* S = static_cast<Harness*>(p)->S;
    S->VortexList.clear();
    for (int64_t i = 0; i < n; i++) {
        TObj o(o6[6 * i], o6[6 * i + 1], o6[6 * i + 2]);
        o.v = TVec(o6[6 * i + 3], o6[6 * i + 4]);
        o._1_eps = o6[6 * i + 5];
        S->VortexList.push_back(o);
    }
}

void vvr_get_vortices(void* p, double* o6) {
    Space* S = static_cast<Harness*>(p)->S;
    for (size_t i = 0; i < S->VortexList.size(); i++) {
        const TObj& o = S->VortexList[i];
        o6[6 * i] = o.r.x;
        o6[6 * i + 1] = o.r.y;
        o6[6 * i + 2] = o.g;
        o6[6 * i + 3] = o.v.x;
        o6[6 * i + 4] = o.v.y;
        o6[6 * i + 5] = o._1_eps;
    }
}

void vvr_add_body(void* p, const SegIn* segs, int32_t n) {
    Harness* h = static_cast<Harness*>(p);
    unsigned char* mem = new unsigned char[sizeof(TBody) + alignof(TBody)];
    h->body_mem.push_back(mem);
    void* aligned = mem;
    size_t space = sizeof(TBody) + alignof(TBody);
    std::align(alignof(TBody), sizeof(TBody), aligned, space);
    TBody* b = static_cast<TBody*>(aligned);
    new (&b->alist) std::vector<TAtt>();
    for (int32_t k = 0; k < n; k++) {
        TAtt a(segs[k].cx, segs[k].cy, segs[k].slip != 0);
        a.r = TVec(segs[k].x, segs[k].y);
        a.g = segs[k].g;
        a._1_eps = segs[k].inv_eps;
        a.dl = TVec(segs[k].dlx, segs[k].dly);
        a.gsum = segs[k].gsum;
        a.hsum = segs[k].hsum;
        a.fric = segs[k].fric;
        a.eq_no = segs[k].eq_no;
        b->alist.push_back(a);
    }
    h->S->BodyList.push_back(std::shared_ptr<TBody>(b, [](TBody*) {}));
}

void vvr_get_fric(void* p, double* out) {
    Space* S = static_cast<Harness*>(p)->S;
    size_t k = 0;
    for (auto& b : S->BodyList)
        for (auto& a : b->alist) out[k++] = a.fric;
}

void vvr_tree_build(void* p, int far_criteria, double min_node, double max_node) {
    Harness* h = static_cast<Harness*>(p);
    if (h->tree) {
        h->tree->destroy();
        delete h->tree;
    }
    h->tree = new stree(h->S, far_criteria, min_node, max_node);
    h->tree->build();
}

int32_t vvr_num_leaves(void* p) {
    return (int32_t) static_cast<Harness*>(p)->tree->getBottomNodes().size();
}

static int32_t seg_index(Space* S, const TObj* ptr) {
    int32_t base = 0;
    for (auto& b : S->BodyList) {
        const TAtt* a0 = b->alist.data();
        if (ptr >= a0 && ptr < a0 + b->alist.size()) return base + (int32_t)(static_cast<const TAtt*>(ptr) - a0);
        base += (int32_t)b->alist.size();
    }
    return -1;
}

// Per leaf (DFS order): geom {x, y, h, w}, ints {v0, v1, depth, path_code,
// nseg, nnear, nfar}.
void vvr_leaves(void* p, double* geom, int32_t* ints) {
    Harness* h = static_cast<Harness*>(p);
    const TObj* base = h->S->VortexList.data();
    const auto& bn = h->tree->getBottomNodes();
    for (size_t l = 0; l < bn.size(); l++) {
        const snode* n = bn[l];
        geom[4 * l + 0] = n->x;
        geom[4 * l + 1] = n->y;
        geom[4 * l + 2] = n->h;
        geom[4 * l + 3] = n->w;
        ints[7 * l + 0] = n->vRange.first ? (int32_t)(n->vRange.first - base) : 0;
        ints[7 * l + 1] = n->vRange.last ? (int32_t)(n->vRange.last - base) : 0;
        ints[7 * l + 2] = n->i;
        ints[7 * l + 3] = n->j;
        ints[7 * l + 4] = (int32_t)n->bllist.size();
        ints[7 * l + 5] = (int32_t)n->NearNodes->size();
        ints[7 * l + 6] = (int32_t)n->FarNodes->size();
    }
}

// Per leaf (DFS order): hRange and sRange as list indices {h0, h1, s0, s1}.
void vvr_leaves_hs(void* p, int32_t* out) {
    Harness* h = static_cast<Harness*>(p);
    const TObj* hb = h->S->HeatList.data();
    const TObj* sb = h->S->StreakList.data();
    const auto& bn = h->tree->getBottomNodes();
    for (size_t l = 0; l < bn.size(); l++) {
        const snode* n = bn[l];
        out[4 * l + 0] = n->hRange.first ? (int32_t)(n->hRange.first - hb) : 0;
        out[4 * l + 1] = n->hRange.last ? (int32_t)(n->hRange.last - hb) : 0;
        out[4 * l + 2] = n->sRange.first ? (int32_t)(n->sRange.first - sb) : 0;
        out[4 * l + 3] = n->sRange.last ? (int32_t)(n->sRange.last - sb) : 0;
    }
}

// Flattened lists in leaf DFS order.  near/far entries: {depth, path_code};
// segs: global segment index.
void vvr_lists(void* p, int32_t* near_ij, int32_t* far_ij, int32_t* segs) {
    Harness* h = static_cast<Harness*>(p);
    const auto& bn = h->tree->getBottomNodes();
    size_t a = 0, b = 0, c = 0;
    for (const snode* n : bn) {
        for (const snode* m : *n->NearNodes) {
            near_ij[2 * a] = m->i;
            near_ij[2 * a + 1] = m->j;
            a++;
        }
        for (const snode* m : *n->FarNodes) {
            far_ij[2 * b] = m->i;
            far_ij[2 * b + 1] = m->j;
            b++;
        }
        for (const TObj* s : n->bllist) segs[c++] = seg_index(h->S, s);
    }
}

// Far-node monopoles in the same flattened order: {CMp.x, CMp.y, CMp.g,
// CMp._1_eps, CMm.x, CMm.y, CMm.g, CMm._1_eps}.
void vvr_far_cm(void* p, double* out) {
    Harness* h = static_cast<Harness*>(p);
    size_t b = 0;
    for (const snode* n : h->tree->getBottomNodes())
        for (const snode* m : *n->FarNodes) {
            double* o = out + 8 * b++;
            o[0] = m->CMp.r.x;
            o[1] = m->CMp.r.y;
            o[2] = m->CMp.g;
            o[3] = m->CMp._1_eps;
            o[4] = m->CMm.r.x;
            o[5] = m->CMm.r.y;
            o[6] = m->CMm.g;
            o[7] = m->CMm._1_eps;
        }
}

int32_t vvr_epsilon(void* p, int merge) {
    Harness* h = static_cast<Harness*>(p);
    MEpsilonFast e(h->S, h->tree);
    e.CalcEpsilonFast(merge != 0);
    return e.Merged();
}

void vvr_diffusive(void* p) {
    Harness* h = static_cast<Harness*>(p);
    MDiffusiveFast d(h->S, h->tree);
    d.process_vort_list();
}

void vvr_diffusive_heat(void* p) {
    Harness* h = static_cast<Harness*>(p);
    MDiffusiveFast d(h->S, h->tree);
    d.process_heat_list();
}

}  // extern "C"
