/*
 * vvhd_modules.hpp -- the libvvhd module classes over the device context.
 *
 * The classes keep the names and the public signatures of the reference's
 * modules, so that its time loop (utils/vvflow/vvflow.cpp:214-262) is
 * written against them call for call.  Where they differ: the free stream is
 * set as numbers (no TEval), and a Space names its device.  The classes:
 *   stree / TSortedTree   libvvhd/headers/TSortedTree.hpp:60-92
 *   MConvectiveFast       libvvhd/headers/MConvectiveFast.hpp:9-27
 *   MEpsilonFast          libvvhd/headers/MEpsilonFast.hpp:7-31
 *   MDiffusiveFast        libvvhd/headers/MDiffusiveFast.hpp:7-18
 *   MFlowmove             libvvhd/headers/MFlowmove.hpp:5-19
 *   Space (the members the hot path reads and writes), TObj, TAtt, TBody,
 *   TVec                  libvvhd/headers/TSpace.hpp, TObj.hpp, TBody.hpp
 *   XField, XVelocity, XVorticity, XStreamfunction, XPressure
 *                         libvvhd/headers/XField.hpp:18-63, XVelocity.hpp,
 *                         XVorticity.hpp, XStreamfunction.hpp, XPressure.hpp
 *   sensors               utils/vvflow/sensors.cpp:8-70
 *   Stepdata              libvvhd/headers/MStepdata.hpp:7-39
 *
 * The vortex list and the bodies live on the device between calls; the host
 * containers of Space are a mirror that is synchronised explicitly:
 *   Space::push()  host -> device, after the caller edited VortexList or
 *                  BodyList (stree::build pushes once by itself if nothing
 *                  has been pushed yet);
 *   Space::pull()  device -> host, before the caller reads them (saving,
 *                  sensors, scripting).
 * Errors of the device library are thrown as std::invalid_argument for
 * calls out of order (the reference's own convention for an unbuilt tree)
 * and std::runtime_error otherwise.  Bodies are rigid (no springs or
 * collisions).  The heat and streak lists (HeatList,
 * StreakList, StreakSourceList) are carried like the vortex list.
 */
#ifndef VVHD_MODULES_HPP
#define VVHD_MODULES_HPP

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <limits>
#include <memory>
#include <numeric>
#include <ostream>
#include <string>
#include <unordered_map>
#include <vector>

#include "vvhd_gpu.h"

// Simulation time as an exact fraction value / timescale (TTime.hpp): time
// steps add without rounding, and "every dt_save" is exact divisibility.
class TTime {
   public:
    int32_t value = 0;
    uint32_t timescale = 1;

    TTime() = default;
    TTime(int32_t v, uint32_t ts) : value(v), timescale(ts) {}

    static TTime makeWithSeconds(double seconds, uint32_t preferredTimescale) {
        return TTime((int32_t)(seconds * preferredTimescale), preferredTimescale);
    }
    // The smallest power-of-ten timescale (below 0x19999999) at which the
    // truncated value gives `seconds` back, then reduced by tens while the
    // value stays a multiple of ten.
    static TTime makeWithSecondsDecimal(double seconds) {
        uint32_t ts = 1;
        while ((double)(int32_t)(seconds * ts) / (double)ts != seconds && ts < 0x19999999u) ts *= 10;
        while (ts > 1 && (int32_t)(seconds * ts) % 10 == 0) ts /= 10;
        return TTime((int32_t)(seconds * ts), ts);
    }
    static TTime add(TTime a, TTime b) {
        const uint32_t ts = lcm(a.timescale, b.timescale);
        return TTime(a.value * (int32_t)(ts / a.timescale) + b.value * (int32_t)(ts / b.timescale), ts);
    }
    // value / timescale is a whole multiple of the divisor (never for a zero
    // divisor)
    bool divisibleBy(TTime d) const {
        if (!d.value || !d.timescale) return false;
        const uint32_t ts = lcm(timescale, d.timescale);
        return ((int64_t)value * ts / timescale) % ((int64_t)d.value * ts / d.timescale) == 0;
    }
    operator double() const { return double(value) / double(timescale); }
    operator std::string() const { return std::to_string(value) + "/" + std::to_string(timescale); }

   private:
    // least common multiple; 1 when either is zero, and the first factor
    // again when the multiple does not fit 32 bits
    static uint32_t lcm(uint32_t x, uint32_t y) {
        if (x == y) return x;
        if (!x || !y) return 1;
        const uint64_t l = (uint64_t)(x / std::gcd(x, y)) * y;
        return l > 0xffffffffu ? x : (uint32_t)l;
    }
};

struct TVec {
    double x = 0, y = 0;
    TVec() = default;
    TVec(double x_, double y_) : x(x_), y(y_) {}
    TVec operator+(TVec o) const { return {x + o.x, y + o.y}; }
    TVec operator-(TVec o) const { return {x - o.x, y - o.y}; }
    TVec operator*(double c) const { return {x * c, y * c}; }
    double operator*(TVec o) const { return x * o.x + y * o.y; }   // scalar product, as TVec.hpp
    TVec& operator+=(TVec o) {
        x += o.x;
        y += o.y;
        return *this;
    }
    double abs2() const { return x * x + y * y; }
    double abs() const { return std::sqrt(abs2()); }
};

struct TVec3D {
    TVec r;
    double o = 0;
};

class TObj {
   public:
    TVec r;
    double g = 0;
    TVec v;
    double _1_eps = 0;
    TObj() = default;
    TObj(double rx, double ry, double g_) : r(rx, ry), g(g_) {}
    TObj(TVec r_, double g_) : r(r_), g(g_) {}
};
static_assert(sizeof(TObj) == 6 * sizeof(double), "TObj is the 48-byte record of the device ABI");

// Heat boundary condition of a body (TBody.hpp:15).
enum class hc_t { neglect, isolate, const_t, const_w };

class TAtt : public TObj {
   public:
    TVec corner;
    int64_t heat_layer_obj_no = -1;   // index of its layer particle in Space::HeatList, or -1
    float heat_const = 0;
    uint32_t slip = 0;
    TVec dl;
    double gsum = 0, hsum = 0, fric = 0, Cp = 0, Fr = 0, Nu = 0;
    uint32_t eq_no = 0;
    TAtt(double x, double y, bool slip_ = false) : corner(x, y), slip(slip_) {}
    TAtt(TVec c, bool slip_ = false) : TAtt(c.x, c.y, slip_) {}
};

class TBody {
   public:
    std::vector<TAtt> alist;
    TVec3D holder, dpos, speed_slae, speed_slae_prev;
    TVec3D force_hydro, force_holder, friction, friction_prev, fdt_dead;
    double g_dead = 0, nusselt = 0;
    int special_segment_no = 0;
    bool kutta = false;
    hc_t heat_condition = hc_t::neglect;
    // geometry, computed on the device (TBody::doFillProperties)
    double get_slen() const { return slen_; }
    double get_area() const { return area_; }
    TVec get_cofm() const { return cofm_; }
    double get_moi_cofm() const { return moi_cofm_; }
    TVec get_axis() const { return holder.r + dpos.r; }

   private:
    friend class Space;
    double slen_ = 0, area_ = 0, moi_cofm_ = 0;
    TVec cofm_;
};

class Space {
   public:
    explicit Space(int device = 0);
    ~Space();
    Space(const Space&) = delete;
    Space& operator=(const Space&) = delete;

    double re = std::numeric_limits<double>::infinity();
    double pr = 0;
    double inf_g = 0;
    TVec inf_vel{0, 0};   // U_inf (the reference evaluates inf_vx/inf_vy expressions)
    TTime time;                        // current simulation time
    TTime dt{1, 1};                    // time step
    TTime dt_save, dt_streak, dt_profile;   // periods of saving, streak shedding, profiles
    double finish = 0;                 // time to finish the simulation
    std::string caption;
    std::string inf_vx = "1", inf_vy = "0";   // the free stream's expressions, kept for snapshots
    TVec inf_marker{0, 0};
    TVec gravity{0, 0};
    std::vector<std::shared_ptr<TBody>> BodyList;
    std::vector<TObj> VortexList;
    std::vector<TObj> HeatList;
    std::vector<TObj> StreakSourceList;
    std::vector<TObj> StreakList;

    TVec inf_speed() const { return inf_vel; }
    void calc_forces();
    void zero_forces();
    void finish_step() { time = TTime::add(time, dt); }
    double average_segment_length() const;
    // sums over the vortex list on the device (TSpace.hpp:85-117)
    double gsum();
    double gmax();
    double integral();
    TVec hydrodynamic_momentum();

    // HDF5 snapshots (TSpace.hpp: save / load, the format of
    // TSpace_save_hdf.cpp and TSpace_load_hdf.cpp) through libvvhd_hdf.so:
    // save pulls the device state first; load replaces the lists and bodies
    // (every vortex with v = 0 and _1_eps = 0, as the reference's reader
    // leaves them) and pushes them to the device.
    void save(const char* path);
    void load(const char* path);

    // device synchronisation (no counterpart in the reference)
    void push();
    void pull();
    vvhd_ctx* device() const { return ctx_; }
    int64_t vortex_count() const;

   private:
    friend class stree;
    friend class MConvectiveFast;
    friend class MEpsilonFast;
    friend class MDiffusiveFast;
    friend class MFlowmove;
    friend class XField;
    friend class XVelocity;
    friend class XVorticity;
    friend class XStreamfunction;
    friend class XPressure;
    friend class sensors;
    friend class Stepdata;
    vvhd_ctx* ctx_ = nullptr;
    bool pushed_ = false;
    int32_t nseg_ = 0;
    void ensure_pushed();
    void set_params();
    void pull_bodies();
    void pull_lists();
    void check(int rc, bool state_error = false) const;
};

// Host view of one node of the device tree (TSortedTree.hpp:20-57): centre
// and size, depth i, path code j, per-sign monopoles CMp/CMm, the vortex
// range as indices [v0, v1) into Space::VortexList (and as TObj pointers in
// vRange while the host mirror holds the device's list), a leaf's heat and
// streak ranges [h0, h1), [s0, s1) (hRange, sRange likewise), and for a leaf
// its near (leaves) and far (any level) interaction lists in the
// reference's order.  Views of every node and the lists are loaded from the device when
// the tree is first asked for its leaves.
class stree;
class TSortedNode {
   public:
    struct range {
        TObj *first = nullptr, *last = nullptr;
        size_t size() const { return (size_t)(last - first); }
    };
    double x = 0, y = 0, h = 0, w = 0;
    int i = 0, j = 0;
    TObj CMp, CMm;
    int64_t v0 = 0, v1 = 0;
    range vRange;
    int64_t h0 = 0, h1 = 0, s0 = 0, s1 = 0;
    range hRange, sRange;
    std::vector<TSortedNode*>* NearNodes = nullptr;
    std::vector<TSortedNode*>* FarNodes = nullptr;
    TSortedNode *ch1 = nullptr, *ch2 = nullptr;
    // leaf index in DFS order (-1 for an internal node) and the owning tree
    int leaf = -1;
    const stree* tree = nullptr;
};

class stree {
   public:
    stree(Space* sS, int sFarCriteria, double sMinNodeSize,
          double sMaxNodeSize = std::numeric_limits<double>::max());
    stree() = delete;
    stree(const stree&) = delete;
    stree& operator=(const stree&) = delete;

    void build(bool IncludeVortexes = true, bool IncludeBody = true, bool IncludeHeat = true);
    void destroy();
    const std::vector<TSortedNode*>& getBottomNodes() const;
    const TSortedNode* findNode(TVec p) const;
    void printBottomNodes(FILE* f, bool PrintDepth = false) const;

    int farCriteria() const { return farCriteria_; }

    Space* space() const { return S; }

   private:
    Space* S;
    int farCriteria_;
    double minNodeSize, maxNodeSize;
    bool built_ = false;
    mutable std::vector<TSortedNode> nodes_;              // every node, device numbering
    mutable std::vector<TSortedNode*> leaf_ptrs_;         // leaves in DFS order
    mutable std::vector<std::vector<TSortedNode*>> near_, far_;
    mutable bool leaves_valid_ = false;
    void load_leaves() const;
};
typedef stree TSortedTree;

class MConvectiveFast {
   public:
    MConvectiveFast() = delete;
    MConvectiveFast(Space* S, const TSortedTree* tree);
    MConvectiveFast(const MConvectiveFast&) = delete;
    MConvectiveFast& operator=(const MConvectiveFast&) = delete;
    void process_all_lists();
    TVec velocity(TVec p) const;
    void calc_circulation(const void** collision);

    // Learned far field (no counterpart in the reference): with a model
    // set, process_all_lists uses it; clear_learned() restores the plain path.
    void set_learned(const std::vector<float>& params);
    void clear_learned() { learned_ = false; }
    // Gradient of the learned correction over a leaf: the network's own
    // forward-mode derivative (true, default) or the geometric terms only.
    void set_learned_gradient(bool network);
    // Pair sums: Kahan-compensated FP32 (default) or plain FP32.
    void set_kahan(bool k) { kahan_ = k; }

   private:
    Space* S;
    const TSortedTree* tree;
    bool learned_ = false, kahan_ = true;
};

class MEpsilonFast {
   public:
    MEpsilonFast(Space* S_, const TSortedTree* Tree_) : S(S_), Tree(Tree_), merged_(0) {}
    void CalcEpsilonFast(bool merge);
    int Merged() { return merged_; }
    // Squared distances over the near list of a leaf (MEpsilonFast.cpp:66-107),
    // evaluated on the device of the node's tree.
    static double eps2h(const TSortedNode& node, TVec p);
    static double h2(const TSortedNode& node, TVec p);

   private:
    static double near_query(const TSortedNode& node, TVec p, bool h2);

   private:
    Space* S;
    const TSortedTree* Tree;
    int merged_;
};

class MDiffusiveFast {
   public:
    MDiffusiveFast(Space* S_, const TSortedTree* tree_) : S(S_), tree(tree_) {}
    void process_vort_list();
    void process_heat_list();

   private:
    Space* S;
    const TSortedTree* tree;
};

class MFlowmove {
   public:
    MFlowmove(Space* S_, double remove_eps_ = 1E-10);
    void move_and_clean(bool remove, const void** collision, size_t* cleaned_v = nullptr);
    void vortex_shed();
    void streak_shed();
    void heat_shed();
    void heat_crop(double scale = 16);

   private:
    Space* S;
};

// --------------------------------------------------------- step data --
// The per-step time series (MStepdata.hpp:7-39, MStepdata.cpp:10-290) in the
// reference's HDF5 layout, through libvvhd_hdf.so: every write() appends the
// time and each body's force_hydro, force_holder, friction, holder, dpos and
// speed_slae; with b_save_profile also every segment's Cp / dt and Fr / dt,
// after which the segments' Cp, Fr and Nu accumulators are reset (on the
// device too).  An existing file is continued from S->time (its later rows
// are dropped).
struct vvhd_stepdata;
class Stepdata {
   public:
    Stepdata(Space* S, const char* fname, bool b_save_profile);
    Stepdata(const Stepdata&) = delete;
    Stepdata& operator=(const Stepdata&) = delete;
    ~Stepdata();
    void write();
    void flush();

   private:
    Space* S;
    bool b_save_profile;
    vvhd_stepdata* h_ = nullptr;
};

// ------------------------------------------------------------ fields --
// Grid evaluators (XField.hpp:18-63 and its subclasses): each works on a copy
// of the Space in a device context of its own, prepared as the reference's
// evaluate() prepares its copy, then fills map on the device (vvhd_field).
class XField {
   public:
    XField() = delete;
    XField(const XField&) = delete;
    XField& operator=(const XField&) = delete;
    XField(float xmin, float ymin, float dxdy, int xres, int yres);
    virtual ~XField() = default;
    const float& at(int xi, int yj) const { return map_[(size_t)yj * xres + xi]; }
    float min() const;
    float max() const;
    float percentile(float p) const;
    friend std::ostream& operator<<(std::ostream& os, const XField& field);

   protected:
    float xmin, ymin, dxdy;
    int xres, yres;
    std::vector<float> map_;
    bool evaluated = false;
    // the copy: device context, host lists and the average segment length
    void copy_space(const Space& S, Space& copy) const;
    vvhd_grid grid() const;
};

class XVelocity : public XField {
   public:
    XVelocity(const Space& S, double xmin, double ymin, double dxdy, int xres, int yres);
    char mode = 0;        // 'x' or 'y'
    char ref_frame = 0;   // 'o', 'f' or 'b'
    void evaluate();

   private:
    const Space& S;
};

class XVorticity : public XField {
   public:
    XVorticity(const Space& S, double xmin, double ymin, double dxdy, int xres, int yres);
    double eps_mult = 0;
    void evaluate();

   private:
    const Space& S;
};

class XStreamfunction : public XField {
   public:
    XStreamfunction(const Space& S, double xmin, double ymin, double dxdy, int xres, int yres);
    double eps_mult = 0;
    char ref_frame = 0;   // 'o', 'f' or 'b'
    void evaluate();

   private:
    const Space& S;
};

class XPressure : public XField {
   public:
    XPressure(const Space& S, double xmin, double ymin, double dxdy, int xres, int yres);
    double eps_mult = 0;
    char ref_frame = 0;   // 's', 'o', 'f' or 'b'
    void evaluate();

   private:
    const Space& S;
};

// Velocity probes (utils/vvflow/sensors.cpp:8-70): points read from a file
// ("x y" per line); output() appends "time \tu \tv ..." with the
// convective velocity at every point (one device call for all points).
class sensors {
   public:
    sensors() = delete;
    sensors(Space* S, MConvectiveFast* conv, const char* sensors_file, const char* output);
    sensors(const sensors&) = delete;
    sensors& operator=(const sensors&) = delete;
    ~sensors();
    void loadFile(const char* file);
    void output();

   private:
    Space* S;
    MConvectiveFast* conv;
    FILE* fout;
    std::vector<TVec> slist;
};

#endif
