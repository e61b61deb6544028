/*
 * vvhd_gpu.h -- C ABI of the MI355X (gfx950) implementation of libvvhd's
 * per-step particle hot path: tree build, epsilon search with vortex merging,
 * convective (Biot-Savart) velocity and vortex-domain diffusive velocity,
 * plus flow-move, the FP64 boundary-condition SLAE and the FP64 integrals.
 *
 * Boundary.  The reference exposes this path as C++ classes over a borrowed
 * Space* (vvflow/vvflow libvvhd/headers/TSortedTree.hpp:60-92,
 * MEpsilonFast.hpp:7-31, MConvectiveFast.hpp:9-27, MDiffusiveFast.hpp:7-18,
 * MFlowmove.hpp:5-19).  This header is the flat C layer beneath a
 * replacement of those classes (vvflow_amd/csrc/shim): plain pointers and
 * sizes, an opaque device context, int status codes, no exceptions.
 *
 * Conventions.
 *  - Every call returns VVHD_OK (0) or a VVHD_E* code; vvhd_last_error()
 *    then holds a message.  Nothing throws across the ABI.
 *  - Host arrays of vortices use the reference's TObj record: 6 doubles
 *    {r.x, r.y, g, v.x, v.y, _1_eps} per vortex (TObj.hpp:13-16).  On the
 *    device the context keeps them structure-of-arrays (FP64 master copy).
 *  - All device work of one context is ordered on one HIP stream
 *    (vvhd_stream).  Calls block until their results are on the host only
 *    when they return host data.
 *  - One host thread per context.
 */
#ifndef VVHD_GPU_H
#define VVHD_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    VVHD_OK = 0,
    VVHD_EINVAL = 1,       /* invalid argument                               */
    VVHD_ESTATE = 2,       /* call out of order (e.g. tree not built)        */
    VVHD_EHIP = 3,         /* HIP runtime error                              */
    VVHD_ENOMEM = 4,       /* device allocation failed                       */
    VVHD_ESINGULAR = 5,    /* singular boundary-condition matrix             */
    VVHD_EUNSUPPORTED = 6  /* input outside what this implementation covers  */
};

typedef struct vvhd_ctx vvhd_ctx;

/* One body segment (the TAtt fields the path reads and writes,
 * libvvhd/headers/TBody.hpp:17-55).  Same layout for upload and download. */
typedef struct vvhd_seg {
    double x, y;          /* centre r                                  */
    double g;             /* bound circulation (SLAE solution)         */
    double inv_eps;       /* _1_eps = 3/|dl|                           */
    double cx, cy;        /* corner                                    */
    double dlx, dly;      /* segment vector dl                         */
    double gsum, hsum, fric, Cp, Fr, Nu;
    uint32_t slip, eq_no;
} vvhd_seg;

/* One body (TBody fields used by the SLAE, flow-move and force integrals,
 * libvvhd/headers/TBody.hpp:67-158).  Bodies are rigid and static
 * (kspring = inf, prescribed speed 0): the configurations of this path. */
typedef struct vvhd_body {
    double holder[3], dpos[3], speed_slae[3], speed_slae_prev[3];
    double force_hydro[3], force_holder[3], friction[3], friction_prev[3];
    double fdt_dead[3];
    double g_dead, nusselt;
    double slen, area, cofm[2], moi_cofm;
    int32_t first_seg, nseg;
    int32_t special_segment_no, kutta;
} vvhd_body;

/* ---- context -------------------------------------------------------- */
int  vvhd_create(int device, vvhd_ctx** out);
void vvhd_destroy(vvhd_ctx* ctx);
const char* vvhd_last_error(const vvhd_ctx* ctx);
/* The hipStream_t all work of this context is ordered on. */
void* vvhd_stream(vvhd_ctx* ctx);
/* Space::re, Space::dt, Space::inf_g, Space::time (TSpace.hpp:24-36). */
int  vvhd_set_params(vvhd_ctx* ctx, double re, double dt, double inf_g, double time);

/* ---- particles (Space::VortexList, TSpace.hpp:39) -------------------- */
int     vvhd_upload_vortices(vvhd_ctx* ctx, const double* objs6, int64_t n);
int     vvhd_download_vortices(vvhd_ctx* ctx, double* objs6, int64_t n);
int64_t vvhd_num_vortices(vvhd_ctx* ctx);
/* Device-side copy of the vortex list and its restore (benchmarks start
 * every repeat from the same input because merges mutate it). */
int     vvhd_snapshot(vvhd_ctx* ctx);
int     vvhd_restore(vvhd_ctx* ctx);

/* ---- heat and streak lists (Space::HeatList, StreakList,
 * StreakSourceList, TSpace.hpp:40-43) ---------------------------------------
 * The tree build partitions the heat and streak lists with the vortices
 * (hRange / sRange, TSortedTree.cpp:68-69); vvhd_epsilon also runs epsh on
 * the heat list (MEpsilonFast.cpp:53-60, 175-212; its merges count in
 * *merged), vvhd_convective and vvhd_convective_learned also move the heat
 * (g != 0) and streak targets (MConvectiveFast.cpp:87-110; the learned path
 * gives them the reference's Taylor far field), vvhd_move_and_clean also
 * moves and cleans both lists (MFlowmove.cpp:107-198).  One simulation per
 * context: with vvhd_set_batch in force the tree build rejects them. */
enum { VVHD_LIST_VORTEX = 0, VVHD_LIST_HEAT = 1, VVHD_LIST_STREAK = 2, VVHD_LIST_STREAK_SOURCE = 3 };
int     vvhd_upload_list(vvhd_ctx* ctx, int list, const double* objs6, int64_t n);
int     vvhd_download_list(vvhd_ctx* ctx, int list, double* objs6, int64_t n);
int64_t vvhd_list_size(vvhd_ctx* ctx, int list);
/* Space::pr (TSpace.hpp:25), used by process_heat_list and calc_forces. */
int     vvhd_set_pr(vvhd_ctx* ctx, double pr);
/* TBody::heat_condition per body (hc_t, TBody.hpp:15) and TAtt::heat_const
 * per segment (TBody.hpp:27) of the uploaded bodies; both start at zero
 * (neglect) when bodies are uploaded. */
enum { VVHD_HC_NEGLECT = 0, VVHD_HC_ISOLATE = 1, VVHD_HC_CONST_T = 2, VVHD_HC_CONST_W = 3 };
int     vvhd_set_heat(vvhd_ctx* ctx, const int32_t* heat_condition, const float* heat_const);
/* TAtt::heat_layer_obj_no per segment (TBody.hpp:26): -1 after an upload
 * and after zero_forces (TSpace.cpp:131), set by move_and_clean. */
int     vvhd_heat_layer(vvhd_ctx* ctx, int64_t* obj_no, int32_t nseg);
int     vvhd_set_heat_layer(vvhd_ctx* ctx, const int64_t* obj_no, int32_t nseg);
/* MDiffusiveFast::process_heat_list (MDiffusiveFast.cpp:50-89). */
int     vvhd_diffusive_heat(vvhd_ctx* ctx);
/* MFlowmove::heat_shed, heat_crop(scale), streak_shed (MFlowmove.cpp:
 * 237-291).  streak_shed appends the sources unconditionally: the caller
 * tests Space::time.divisibleBy(dt_streak). */
int     vvhd_heat_shed(vvhd_ctx* ctx);
int     vvhd_heat_crop(vvhd_ctx* ctx, double scale);
int     vvhd_streak_shed(vvhd_ctx* ctx);

/* ---- bodies (Space::BodyList, TSpace.hpp:38) ------------------------- */
int vvhd_upload_bodies(vvhd_ctx* ctx, const vvhd_seg* segs, int32_t nseg,
                       const vvhd_body* bodies, int32_t nbody);
int vvhd_download_bodies(vvhd_ctx* ctx, vvhd_seg* segs, int32_t nseg,
                         vvhd_body* bodies, int32_t nbody);

/* ---- independent simulations in one context (no counterpart in the
 * reference: one Space per process there) --------------------------------
 * Declares the uploaded vortex and body lists as the concatenation of nsim
 * independent simulations: simulation s owns vortices [vortex_off[s],
 * vortex_off[s+1]) (vortex_off[0] = 0, vortex_off[nsim] = vortex count),
 * bodies [body_off[s], body_off[s+1]) and Reynolds number re[s].  The tree
 * build then grows one tree per simulation (a forest, one root each), and
 * epsilon, convective (plain and learned) and diffusive advance every
 * simulation in the same launches; no list, merge or sum crosses
 * simulations, and the direct sum of a target runs over its own simulation.
 * calc_circulation solves one boundary system per simulation (all of one
 * order: strided-batched LU), vortex_shed appends each simulation's new
 * vortices after its own, move_and_clean removes each vortex only against
 * its own simulation's bodies, and calc_forces uses each simulation's
 * Reynolds number; the ranges then move (vvhd_batch_offsets).  Holds until
 * the next vortex or body upload (back to one simulation).  integrals,
 * velocity_points, tree_find, the point queries, the grid fields and the
 * heat and streak lists, whose reference counterparts act on the whole
 * Space, return VVHD_EUNSUPPORTED meanwhile. */
int vvhd_set_batch(vvhd_ctx* ctx, int32_t nsim, const int64_t* vortex_off, const int32_t* body_off,
                   const double* re);
/* The simulations' current vortex ranges (nsim + 1 offsets; vortex_off may
 * be NULL to get the count). */
int vvhd_batch_offsets(vvhd_ctx* ctx, int32_t* nsim, int64_t* vortex_off);

/* ---- launch forms (no counterpart in the reference) ------------------
 * The near-field kernel runs one wave per leaf (form 0) or as a persistent
 * grid of 8, 16 or 32 waves per compute unit taking {leaf, 16 targets} work
 * items from a device queue; every form gives the same bits.  With autotune
 * set (the default) the form is the fastest of the four measured on the
 * first call of each workload class (one or several simulations, leaves and
 * targets per leaf in powers of two); otherwise near_form.  With the graph
 * set, vvhd_convective (plain path, no heat or streak particles) runs as one
 * captured HIP graph when the form is persistent: its kernels read every
 * size from device memory, and the graph is captured once per set of
 * buffers and replayed.  vvhd_tuning: {last form, workload classes tuned,
 * graphs captured, graph launches}. */
int vvhd_set_tuning(vvhd_ctx* ctx, int autotune, int near_form);
int vvhd_set_graph(vvhd_ctx* ctx, int enable);
int vvhd_tuning(vvhd_ctx* ctx, int64_t* out4);

/* ---- tree: stree (TSortedTree.hpp:60-92, TSortedTree.cpp:232-303) ----
 * Builds the adaptive binary tree over the vortex list and the body
 * segments, reordering the vortex list in place exactly as the reference's
 * Hoare partitions do, then the per-sign monopoles and the near/far
 * interaction lists of every leaf.  perm_out (optional, host, n entries)
 * receives for every new position the vortex's index before the build. */
int vvhd_tree_build(vvhd_ctx* ctx, int far_criteria, double min_node_size,
                    double max_node_size, int32_t* perm_out);
int vvhd_tree_destroy(vvhd_ctx* ctx);
/* Leaf table in DFS order (getBottomNodes):
 *   geom: 4 doubles per leaf {x, y, h, w}; ints: 4 int32 {v0, v1, depth,
 *   nseg}; path: uint64 path code (bit per level, 1 = second child).
 * Any pointer may be NULL.  Returns the leaf count through nleaves. */
int vvhd_tree_leaves(vvhd_ctx* ctx, int32_t* nleaves, double* geom,
                     int32_t* ints, uint64_t* path);
/* Near (far = 0) or far (far = 1) lists of all leaves, flattened in leaf DFS
 * order; each entry is a node given by {depth, path code} plus, for far
 * entries, its monopoles cm = {CMp.x, CMp.y, CMp.g, CMm.x, CMm.y, CMm.g}.
 * Call with entries == NULL to get the total count. */
int vvhd_tree_lists(vvhd_ctx* ctx, int far, int64_t* count, int64_t* offsets,
                    int32_t* depth, uint64_t* path, double* cm);
/* hRange and sRange of every leaf (DFS order) as heat / streak list
 * indices: 4 int32 per leaf {h0, h1, s0, s1}. */
int vvhd_tree_leaf_lists(vvhd_ctx* ctx, int32_t* ranges4);
/* stree::findNode for n points: leaf index (DFS order) or -1. */
int vvhd_tree_find(vvhd_ctx* ctx, const double* xy, int64_t n, int32_t* leaf_out);
/* Every node of the built tree in device numbering (inspection):
 * geom {x, y, h, w}, ints {v0, v1, first child or -1, depth}, path code, and
 * the per-sign monopoles cm = {CMp.x, CMp.y, CMp.g, CMm.x, CMm.y, CMm.g}
 * (snode::CMp/CMm, TSortedTree.hpp:30).  Any pointer may be NULL. */
int vvhd_tree_nodes(vvhd_ctx* ctx, int32_t* nnodes, double* geom, int32_t* ints, uint64_t* path, double* cm);

/* ---- epsilon: MEpsilonFast::CalcEpsilonFast (MEpsilonFast.cpp:11-63) --
 * Sets _1_eps of every vortex with g != 0 from its two nearest neighbours
 * in the near leaves, merging vortices with the reference's criterion and in
 * the reference's order when merge != 0.  nn1/nn2 (optional, host, n
 * entries, tree order) receive the neighbour indices of the final search of
 * each target (-1 where none).  merged (optional) receives Merged(). */
int vvhd_epsilon(vvhd_ctx* ctx, int merge, int32_t* nn1, int32_t* nn2, int32_t* merged);

/* ---- convective: MConvectiveFast::process_all_lists (:36-114) ---------
 * v += U_inf + near-field Lamb-kernel sum (FP32 pair sums, compensated when
 * kahan != 0: terms into low-order parts folded into the running sums every
 * eight terms, in leaf-local coordinates) + far-field Taylor terms
 * (FP64).  inf_x/inf_y: Space::inf_speed() evaluated by the caller. */
int vvhd_convective(vvhd_ctx* ctx, double inf_x, double inf_y, int kahan);
/* Allocation-free variant of vvhd_convective (no counterpart in the
 * reference): for callers that cannot allocate device memory -- code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc.  Same
 * kernels and results (bit for bit) as vvhd_convective, but its scratch (the
 * per-leaf Taylor terms and the per-vortex near-field sums) lives in one
 * fixed device array of VVHD_NOALLOC_SCRATCH_BYTES per device instead of
 * buffers grown on demand, and it calls no hipMalloc/hipFree.  Needs
 * 32 (leaves + 1) + 16 (vortices + 1) bytes, each part rounded up to 256
 * (about 20 MB for a 1M-vortex wake); a larger input returns VVHD_EINVAL.  All
 * contexts on a device share that array: calls that overlap on one device
 * race.  The tree must be built (and the vortex list uploaded) beforehand, as
 * for vvhd_convective. */
#define VVHD_NOALLOC_SCRATCH_BYTES (64ll << 20)
int vvhd_convective_noalloc(vvhd_ctx* ctx, double inf_x, double inf_y, int kahan);
/* MConvectiveFast::velocity(TVec) (:20-34) for n points (host xy, host out). */
int vvhd_velocity_points(vvhd_ctx* ctx, const double* xy, int64_t n,
                         double inf_x, double inf_y, double* v_out);

/* ---- learned far-field correction (no counterpart in the reference; it
 * replaces the far-field part of MConvectiveFast::process_all_lists,
 * :48-86, and keeps its near field and call order) -----------------------
 * Model: an MLP 12 -> 64 -> 64 -> 2 with SiLU, VVHD_FAR_PARAMS floats in this
 * order: W1[64][12], b1[64], W2[64][64] (out, in), b2[64], W3[2][64], b3[2]
 * (the VVHD_FAR_TRAINABLE trainable ones), then the input means mu[12], the
 * input deviations sd[12] and the output scales ysd[2].  Inference and
 * training use bf16 weights and activations with FP32 accumulation. */
#define VVHD_FAR_TRAINABLE 5122
#define VVHD_FAR_PARAMS 5148
int vvhd_far_model_set(vvhd_ctx* ctx, const float* params, int32_t n);
int vvhd_far_model_get(vvhd_ctx* ctx, float* params, int32_t n);
/* Convective velocity with the learned far field: the same near field as
 * vvhd_convective, and for every (vortex, far node) pair the field of the
 * node's per-sign monopoles (moments taken after the epsilon pass) plus the
 * MLP's correction.  Requires a model (VVHD_ESTATE otherwise). */
int vvhd_convective_learned(vvhd_ctx* ctx, double inf_x, double inf_y, int kahan);
/* Per-node moments after the epsilon pass of the built tree: *nnodes, and
 * when out != NULL 12 arrays of nnodes doubles: G+, G-, z+ (x, y), z- (x, y),
 * Q+ (re, im), Q- (re, im), sum(g eps^2) for + and -. */
int vvhd_far_moments(vvhd_ctx* ctx, int32_t* nnodes, double* out);
/* Training pairs of the built tree: every (vortex with g != 0, far node of
 * its leaf) pair kept with probability rate (hash of seed, vortex, node).
 * *npairs receives the count; with X != NULL (cap >= count) also the raw
 * features X[12], the normalised FP64 residual labels Y[2], the scale
 * S d^2 / (2 pi r^3) (0 for a degenerate pair) and, optionally, the vortex
 * (tree order) and node of each pair. */
int vvhd_far_sample(vvhd_ctx* ctx, double rate, uint64_t seed, int64_t cap, int64_t* npairs,
                    float* X, float* Y, float* scale, int32_t* target, int32_t* node);
/* Raw network outputs (before ysd) for n raw feature rows. */
int vvhd_far_forward(vvhd_ctx* ctx, const float* X, int64_t n, float* Y);
/* Gradient of the loss sum_j W_j |net(X_j) - Y_j / ysd|^2 / n with respect
 * to the VVHD_FAR_TRAINABLE parameters (same kernels as training). */
int vvhd_far_gradients(vvhd_ctx* ctx, const float* X, const float* Y, const float* W, int64_t n,
                       float* grad, float* loss);
/* Training set kept on the device, then `steps` Adam steps on batches of
 * `batch` samples drawn with replacement (hash of seed and step); lr holds
 * one learning rate per step; loss (optional) the batch loss per step. */
int vvhd_far_train_data(vvhd_ctx* ctx, const float* X, const float* Y, const float* W, int64_t n);
int vvhd_far_train(vvhd_ctx* ctx, int32_t steps, int32_t batch, const float* lr, uint64_t seed,
                   float* loss);

/* ---- diffusive: MDiffusiveFast::process_vort_list (:8-48) ------------ */
int vvhd_diffusive(vvhd_ctx* ctx);

/* ---- flow-move: MFlowmove::move_and_clean / vortex_shed (:11-235) ---- */
int vvhd_move_and_clean(vvhd_ctx* ctx, int remove, int64_t* cleaned);
int vvhd_vortex_shed(vvhd_ctx* ctx);

/* ---- boundary-condition SLAE: MConvectiveFast::calc_circulation -------
 * (:235-273, fill :882-1017) with Matrix::solveUsingInverseMatrix
 * (TMatrix.cpp:166-259): FP64 fill and LU solve on the device; writes the
 * bound circulations, speed_slae, force_hydro and force_holder. */
int vvhd_calc_circulation(vvhd_ctx* ctx, double inf_x, double inf_y);
/* The system filled by the last vvhd_calc_circulation: order into *n; when
 * non-NULL, the matrix row-major into A (n*n) and the right-hand side into
 * rhs (n).  Rows and columns follow Space::EnumerateBodies (TSpace.cpp:
 * 189-203): per body its segments, then speed x/y/o, hydro x/y/o and holder
 * x/y/o.  (Inspection of MConvectiveFast::fillSlaeMatrix, :882-1017.) */
int vvhd_slae_system(vvhd_ctx* ctx, int32_t* n, double* A, double* rhs);

/* ---- FP64 integrals: Space::calc_forces / zero_forces (TSpace.cpp:94-143)
 * and Space::gsum / integral / hydrodynamic_momentum / gmax
 * (TSpace.hpp:85-117); out5 = {gsum, integral, momentum.x, momentum.y, gmax}. */
int vvhd_calc_forces(vvhd_ctx* ctx);
int vvhd_zero_forces(vvhd_ctx* ctx);
int vvhd_integrals(vvhd_ctx* ctx, double* out5);
/* Stepdata::write's reset of every segment's Cp, Fr and Nu accumulators
 * once the profiles are written (MStepdata.cpp:166-176). */
int vvhd_zero_profiles(vvhd_ctx* ctx);

/* ---- point queries and grid fields -----------------------------------
 * MEpsilonFast::eps2h / h2 (MEpsilonFast.cpp:66-107, MEpsilonFast.hpp:15-18)
 * for n points (host xy; out n doubles) over the near list of the leaf given
 * by leaf[k] (DFS index) or, with leaf == NULL, of the leaf stree::findNode
 * gives; with xy == NULL and n == vortex count, for every vortex (tree order)
 * with its own leaf.  eps2h: the second smallest non-zero squared distance to
 * the vortices of the near leaves (the smallest when only one, the lowest
 * double when none); h2: the smallest squared distance to the body segments
 * of the near leaves (+inf when none). */
int vvhd_eps2h(vvhd_ctx* ctx, const double* xy, const int32_t* leaf, int64_t n, double* out);
int vvhd_h2(vvhd_ctx* ctx, const double* xy, const int32_t* leaf, int64_t n, double* out);

/* A field on the grid p(xi, yj) = (xmin + dxdy xi, ymin + dxdy yj), xi < xres,
 * yj < yres, written row by row (yj major) as floats, as XField::map
 * (XField.hpp:56-63). */
typedef struct vvhd_grid {
    double xmin, ymin, dxdy;
    int32_t xres, yres;
} vvhd_grid;
enum {
    VVHD_FIELD_VX = 0,             /* XVelocity mode 'x' (XVelocity.cpp:24-73)          */
    VVHD_FIELD_VY = 1,             /* XVelocity mode 'y'                                */
    VVHD_FIELD_VORTICITY = 2,      /* XVorticity (XVorticity.cpp:26-94)                 */
    VVHD_FIELD_STREAMFUNCTION = 3, /* XStreamfunction (XStreamfunction.cpp:47-118, 160-205) */
    VVHD_FIELD_PRESSURE = 4        /* XPressure (XPressure.cpp:32-146)                  */
};
/* Evaluates one field over the built tree of this context.  The reference's
 * evaluators prepare a copy of the Space first; the caller does the same here
 * in a context of its own (the evaluator classes of vvhd_modules.hpp do):
 *  - VX / VY: tree (8, 20 dl, 0.1), epsilon without merging; map = velocity(p)
 *    minus (ref_x, ref_y); inf = U.
 *  - VORTICITY: vortex_shed, tree (8, 20 dl); per-vortex weights from eps2h
 *    and eps_mult are computed here.
 *  - STREAMFUNCTION: tree (8, 10 dl, 20 dl) (0 and +inf without bodies);
 *    weights from eps2h here; inf - ref enters as p . rotl(U - U_ref).
 *  - PRESSURE: vortex_shed, tree (8, 20 dl, 0.1), epsilon without merging,
 *    convective and diffusive (the vortex velocities); with_ref = 0 is
 *    ref_frame 's'.
 * dl: Space::average_segment_length().  The context's source list is empty
 * (no sources), so the streamfunction has no gaps. */
int vvhd_field(vvhd_ctx* ctx, int kind, const vvhd_grid* grid, double eps_mult, double dl, double inf_x,
               double inf_y, int with_ref, double ref_x, double ref_y, float* map);

/* ---- measurement (no counterpart in the reference) -------------------
 * FP64 direct sum of the induced Lamb-kernel velocity over all vortices for
 * nt targets (tree order indices; NULL = all), into host v_out (2*nt). */
int vvhd_direct_sum(vvhd_ctx* ctx, const int64_t* targets, int64_t nt, double* v_out);
/* Device time in ms of the phases of the last calls, measured with HIP
 * events on the context stream: {tree, epsilon, convective, diffusive,
 * move_and_clean, slae, convective near-field kernel alone, learned
 * far-field kernel alone}; -1 for a phase not run yet. */
int vvhd_phase_times(vvhd_ctx* ctx, float* ms8);
/* Algorithmic work counted on the interaction lists of the current tree:
 * {near convective pairs, far (leaf, node) terms, (vortex, far node) pairs,
 *  learned far-field network passes (value and two tangents per far
 *  (leaf, node) term), tree levels (depths the build split at), leaves,
 *  nodes}; far terms and network passes count leaves with a target of
 *  non-zero circulation. */
int vvhd_work_counts(vvhd_ctx* ctx, int64_t* out7);
/* Concurrency of the step's kernels (default on): the learned near field
 * runs beside its far field, and a diffusive call made right after a
 * convective call runs beside the convective kernels; results are the same
 * bit for bit either way.  Off, every kernel runs alone on the context
 * stream, which is how a single kernel's own rate is measured. */
int vvhd_set_overlap(vvhd_ctx* ctx, int enable);
/* Gradient of the learned far-field correction used for its first-order
 * expansion over a leaf (no counterpart in the reference): network = 1
 * (default) differentiates through the network (forward-mode tangents, three
 * MFMA chains); network = 0 holds the network output fixed over the leaf and
 * keeps the rotation and 1/r^3 terms (one chain, faster, larger error). */
int vvhd_set_far_gradient(vvhd_ctx* ctx, int network);

/* ---- synthetic inputs (host) -------------------------------------------
 * Karman-street wake: 400 Gaussian clusters (sigma 0.05) in two staggered
 * rows at y = +-0.2, SplitMix64 seeded, Box-Muller normals; objs6 receives
 * n TObj records. */
int vvhd_gen_wake(int64_t n, uint64_t seed, double* objs6);
/* Cylinder corners as gen_cylinder{R, N} makes them (gen_body.cpp:74-88). */
int vvhd_gen_cylinder(double radius, int32_t n, double* corners_xy);

#ifdef __cplusplus
}
#endif
#endif
