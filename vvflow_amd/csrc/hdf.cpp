// hdf.cpp -- the reference's HDF5 snapshot format (include/vvhd_hdf.h) over
// libhdf5 1.10.  Host code only; the layout follows Space::save_hdf /
// load_hdf (TSpace_save_hdf.cpp:11-221, TSpace_load_hdf.cpp:11-228) and the
// attribute types of vvhdf.cpp / vvhdf_h5t.cpp.

#include "vvhd_hdf.h"

#include <hdf5.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

struct H5Fail : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void check(herr_t e, const char* what, const std::string& name) {
    if (e < 0) throw H5Fail(std::string(what) + " (" + name + ") failed");
}
hid_t checkid(hid_t id, const char* what, const std::string& name) {
    if (id < 0) throw H5Fail(std::string(what) + " (" + name + ") failed");
    return id;
}

// ---------------------------------------------------------------- types --
// The committed attribute types (vvhdf_h5t.cpp): vec_t {x, y},
// vec3d_t {x, y, o}, fraction_t {value int32, timescale uint32}, string_t
// (variable-length C string).
struct Types {
    hid_t vec = -1, vec3 = -1, frac = -1, str = -1;
    Types() {
        vec = H5Tcreate(H5T_COMPOUND, 16);
        H5Tinsert(vec, "x", 0, H5T_NATIVE_DOUBLE);
        H5Tinsert(vec, "y", 8, H5T_NATIVE_DOUBLE);
        vec3 = H5Tcreate(H5T_COMPOUND, 24);
        H5Tinsert(vec3, "x", 0, H5T_NATIVE_DOUBLE);
        H5Tinsert(vec3, "y", 8, H5T_NATIVE_DOUBLE);
        H5Tinsert(vec3, "o", 16, H5T_NATIVE_DOUBLE);
        frac = H5Tcreate(H5T_COMPOUND, 8);
        H5Tinsert(frac, "value", 0, H5T_STD_I32LE);
        H5Tinsert(frac, "timescale", 4, H5T_STD_U32LE);
        str = H5Tcopy(H5T_C_S1);
        H5Tset_size(str, H5T_VARIABLE);
    }
    ~Types() {
        for (hid_t t : {vec, vec3, frac, str})
            if (t >= 0) H5Tclose(t);
    }
    void commit(hid_t fid) {
        const std::pair<hid_t, const char*> all[] = {{vec, "vec_t"}, {vec3, "vec3d_t"}, {str, "string_t"},
                                                     {frac, "fraction_t"}};
        for (const auto& t : all)
            if (H5Tcommitted(t.first) <= 0)
                check(H5Tcommit2(fid, t.second, t.first, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "H5Tcommit",
                      t.second);
    }
};

struct Frac {
    int32_t value;
    uint32_t timescale;
};

// A double as the nearest fraction value / timescale with timescale below
// 2^31 (continued fractions), exact for the decimal time steps of a run.
Frac to_frac(double t) {
    if (!std::isfinite(t) || std::fabs(t) >= 2147483647.0) throw H5Fail("time value out of the fraction range");
    const double a = std::fabs(t);
    long long p0 = 0, q0 = 1, p1 = 1, q1 = 0;
    double x = a;
    for (int it = 0; it < 64; it++) {
        const double fl = std::floor(x);
        const long long ai = (long long)fl;
        const long long p2 = ai * p1 + p0, q2 = ai * q1 + q0;
        if (q2 > 2147483647LL || p2 > 2147483647LL) break;
        p0 = p1;
        q0 = q1;
        p1 = p2;
        q1 = q2;
        if ((double)p1 / (double)q1 == a) break;
        const double r = x - fl;
        if (r == 0) break;
        x = 1.0 / r;
    }
    Frac f;
    f.value = (int32_t)(t < 0 ? -p1 : p1);
    f.timescale = (uint32_t)(q1 ? q1 : 1);
    return f;
}

hid_t scalar_space() {
    static thread_local hid_t s = H5Screate(H5S_SCALAR);
    return s;
}

void attr_write(hid_t hid, const char* name, hid_t type, const void* val) {
    const hid_t aid = checkid(H5Acreate2(hid, name, type, scalar_space(), H5P_DEFAULT, H5P_DEFAULT), "H5Acreate", name);
    check(H5Awrite(aid, type, val), "H5Awrite", name);
    check(H5Aclose(aid), "H5Aclose", name);
}
void attr_double(hid_t h, const char* n, double v) { attr_write(h, n, H5T_NATIVE_DOUBLE, &v); }
void attr_u32(hid_t h, const char* n, uint32_t v) { attr_write(h, n, H5T_NATIVE_UINT32, &v); }
void attr_i32(hid_t h, const char* n, int32_t v) { attr_write(h, n, H5T_NATIVE_INT32, &v); }
// Empty strings are not written (vvhdf.cpp: h5a_write<std::string const&>).
void attr_str(hid_t h, const char* n, const char* v, const Types& T) {
    if (!v || !*v) return;
    attr_write(h, n, T.str, &v);
}

bool attr_read(hid_t hid, const char* name, hid_t type, void* out) {
    if (H5Aexists(hid, name) <= 0) return false;
    const hid_t aid = checkid(H5Aopen(hid, name, H5P_DEFAULT), "H5Aopen", name);
    check(H5Aread(aid, type, out), "H5Aread", name);
    check(H5Aclose(aid), "H5Aclose", name);
    return true;
}
double read_double(hid_t h, const char* n) {
    double v = 0;
    attr_read(h, n, H5T_NATIVE_DOUBLE, &v);
    return v;
}
void read_str(hid_t h, const char* n, char* out, const Types& T) {
    out[0] = 0;
    char* p = nullptr;
    if (!attr_read(h, n, T.str, &p) || !p) return;
    std::snprintf(out, VVHD_HDF_STR, "%s", p);
    H5free_memory(p);
}

const char* kListNames[VVHD_HDF_LISTS] = {"vort", "heat", "ink", "ink_source", "source"};

// A TObj list as N x 3 {x, y, g}: the memory is the TObj array seen as
// 2N x 3 doubles, every other row selected (TSpace_save_hdf.cpp:11-52).
void write_list(hid_t fid, const char* name, const double* objs6, int64_t n) {
    if (n <= 0) return;
    const hsize_t N = (hsize_t)n;
    hsize_t dims1[2] = {N, 3}, dims2[2] = {2 * N, 3}, chunk[2] = {std::min<hsize_t>(512, N), 3};
    const hid_t prop = H5Pcreate(H5P_DATASET_CREATE);
    H5Pset_chunk(prop, 2, chunk);
    H5Pset_deflate(prop, 9);
    const hid_t sf = checkid(H5Screate_simple(2, dims1, dims1), "H5Screate_simple", name);
    const hid_t sm = checkid(H5Screate_simple(2, dims2, dims2), "H5Screate_simple", name);
    hsize_t start[2] = {0, 0}, stride[2] = {2, 1};
    check(H5Sselect_hyperslab(sm, H5S_SELECT_SET, start, stride, dims1, nullptr), "H5Sselect_hyperslab", name);
    const hid_t d = checkid(H5Dcreate2(fid, name, H5T_NATIVE_DOUBLE, sf, H5P_DEFAULT, prop, H5P_DEFAULT), "H5Dcreate",
                            name);
    check(H5Dwrite(d, H5T_NATIVE_DOUBLE, sm, sf, H5P_DEFAULT, objs6), "H5Dwrite", name);
    check(H5Dclose(d), "H5Dclose", name);
    H5Sclose(sf);
    H5Sclose(sm);
    H5Pclose(prop);
}

int64_t list_rows(hid_t fid, const char* name) {
    if (H5Lexists(fid, name, H5P_DEFAULT) <= 0) return 0;
    const hid_t d = checkid(H5Dopen2(fid, name, H5P_DEFAULT), "H5Dopen", name);
    const hid_t sf = H5Dget_space(d);
    hsize_t dims[2] = {0, 0};
    H5Sget_simple_extent_dims(sf, dims, nullptr);
    H5Sclose(sf);
    H5Dclose(d);
    return (int64_t)dims[0];
}

// The strided read leaves v and _1_eps of every record zero
// (TSpace_load_hdf.cpp:11-45).
void read_list(hid_t fid, const char* name, double* objs6, int64_t n) {
    if (n <= 0) return;
    std::memset(objs6, 0, (size_t)n * 48);
    const hid_t d = checkid(H5Dopen2(fid, name, H5P_DEFAULT), "H5Dopen", name);
    const hid_t sf = H5Dget_space(d);
    hsize_t dims1[2] = {(hsize_t)n, 3}, dims2[2] = {2 * (hsize_t)n, 3};
    const hid_t sm = H5Screate_simple(2, dims2, dims2);
    hsize_t start[2] = {0, 0}, stride[2] = {2, 1};
    check(H5Sselect_hyperslab(sm, H5S_SELECT_SET, start, stride, dims1, nullptr), "H5Sselect_hyperslab", name);
    check(H5Dread(d, H5T_NATIVE_DOUBLE, sm, sf, H5P_DEFAULT, objs6), "H5Dread", name);
    H5Sclose(sm);
    H5Sclose(sf);
    H5Dclose(d);
}

void write_body(hid_t fid, const char* name, const vvhd_hdf_body& b, const double* seg4, const uint32_t* slip,
                const Types& T) {
    const hsize_t N = (hsize_t)b.nseg;
    hsize_t dims[2] = {N, 4}, chunk[2] = {std::min<hsize_t>(512, N), 4};
    const hid_t prop = H5Pcreate(H5P_DATASET_CREATE);
    H5Pset_chunk(prop, 2, chunk);
    H5Pset_deflate(prop, 9);
    const hid_t sf = checkid(H5Screate_simple(2, dims, dims), "H5Screate_simple", name);
    const hid_t d = checkid(H5Dcreate2(fid, name, H5T_NATIVE_DOUBLE, sf, H5P_DEFAULT, prop, H5P_DEFAULT), "H5Dcreate",
                            name);
    const uint32_t s0 = slip ? slip[0] : 0;
    bool same = true;
    for (hsize_t i = 0; slip && i < N; i++) same = same && slip[i] == s0;
    if (!same) {
        const hid_t ss = H5Screate_simple(1, dims, dims);
        const hid_t aid = checkid(H5Acreate2(d, "slip", H5T_NATIVE_UINT32, ss, H5P_DEFAULT, H5P_DEFAULT), "H5Acreate",
                                  name);
        check(H5Awrite(aid, H5T_NATIVE_UINT32, slip), "H5Awrite", name);
        H5Aclose(aid);
        H5Sclose(ss);
    } else {
        attr_u32(d, "general_slip", s0);
    }
    attr_double(d, "general_heat_const", b.heat_const);
    attr_u32(d, "simplified_dataset", 2);
    attr_str(d, "label", b.label, T);
    attr_str(d, "root_body", b.root_body, T);
    attr_write(d, "holder_position", T.vec3, b.holder);
    attr_write(d, "delta_position", T.vec3, b.dpos);
    attr_str(d, "speed_x", b.speed_x, T);
    attr_str(d, "speed_y", b.speed_y, T);
    attr_str(d, "speed_o", b.speed_o, T);
    attr_str(d, "force_o", b.force_o, T);
    attr_write(d, "speed_slae", T.vec3, b.speed_slae);
    attr_write(d, "speed_slae_prev", T.vec3, b.speed_slae_prev);
    attr_write(d, "spring_const", T.vec3, b.kspring);
    attr_write(d, "spring_damping", T.vec3, b.damping);
    attr_double(d, "density", b.density);
    attr_write(d, "force_hydro", T.vec3, b.force_hydro);
    attr_write(d, "force_holder", T.vec3, b.force_holder);
    attr_write(d, "friction_prev", T.vec3, b.friction_prev);
    attr_write(d, "fdt_dead", T.vec3, b.fdt_dead);
    attr_double(d, "g_dead", b.g_dead);
    attr_write(d, "collision_min", T.vec3, b.collision_min);
    attr_write(d, "collision_max", T.vec3, b.collision_max);
    attr_double(d, "bounce", b.bounce);
    attr_double(d, "area", b.area);
    attr_write(d, "cofm", T.vec, b.cofm);
    attr_double(d, "moi_cofm", b.moi_cofm);
    attr_double(d, "moi_axis", b.moi_axis);
    attr_i32(d, "special_segment_no", b.special_segment_no);
    attr_u32(d, "boundary_condition", b.boundary_condition);
    attr_u32(d, "heat_condition", b.heat_condition);
    check(H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, sf, H5P_DEFAULT, seg4), "H5Dwrite", name);
    check(H5Dclose(d), "H5Dclose", name);
    H5Sclose(sf);
    H5Pclose(prop);
}

herr_t collect_bodies(hid_t, const char* name, const H5L_info_t*, void* op) {
    if (std::strncmp(name, "body", 4) == 0) static_cast<std::vector<std::string>*>(op)->push_back(name);
    return 0;
}

std::vector<std::string> body_names(hid_t fid) {
    std::vector<std::string> names;
    check(H5Literate(fid, H5_INDEX_NAME, H5_ITER_NATIVE, nullptr, collect_bodies, &names), "H5Literate", "bodies");
    std::sort(names.begin(), names.end());   // the reference keeps them in a std::map
    return names;
}

struct File {
    hid_t id = -1;
    ~File() {
        if (id >= 0) H5Fclose(id);
    }
};

template <typename F>
int guarded(F f) {
    try {
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
        f();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// ----------------------------------------------------------- step data --

// A float dataset of rows x cols with unlimited rows (chunks of 128 rows,
// deflate 9, NaN fill); an existing one is cut or grown to rows
// (MStepdata.cpp:255-289).
hid_t step_dataset(hid_t loc, const char* name, hsize_t rows, hsize_t cols) {
    if (H5Lexists(loc, name, H5P_DEFAULT) > 0) {
        const hid_t d = checkid(H5Dopen2(loc, name, H5P_DEFAULT), "H5Dopen", name);
        hsize_t dims[2] = {rows, cols};
        check(H5Dset_extent(d, dims), "H5Dset_extent", name);
        return d;
    }
    static const float nanf_ = std::nanf("");
    const hid_t prop = H5Pcreate(H5P_DATASET_CREATE);
    hsize_t chunk[2] = {128, cols};
    H5Pset_chunk(prop, 2, chunk);
    H5Pset_deflate(prop, 9);
    H5Pset_fill_value(prop, H5T_NATIVE_FLOAT, &nanf_);
    H5Pset_fill_time(prop, H5D_FILL_TIME_ALLOC);
    hsize_t cur[2] = {rows, cols}, mx[2] = {H5S_UNLIMITED, cols};
    const hid_t sp = checkid(H5Screate_simple(2, cur, mx), "H5Screate_simple", name);
    const hid_t d = checkid(H5Dcreate2(loc, name, H5T_NATIVE_FLOAT, sp, H5P_DEFAULT, prop, H5P_DEFAULT), "H5Dcreate",
                            name);
    H5Sclose(sp);
    H5Pclose(prop);
    return d;
}

// One more row of doubles at the end of a dataset (MStepdata.cpp:196-240).
void step_append(hid_t d, const double* row) {
    if (d < 0) return;
    hid_t sf = checkid(H5Dget_space(d), "H5Dget_space", "append");
    hsize_t dims[2] = {0, 0};
    check(H5Sget_simple_extent_dims(sf, dims, nullptr), "H5Sget_simple_extent_dims", "append");
    H5Sclose(sf);
    hsize_t off[2] = {dims[0], 0}, ext[2] = {1, dims[1]};
    dims[0]++;
    check(H5Dset_extent(d, dims), "H5Dset_extent", "append");
    sf = checkid(H5Dget_space(d), "H5Dget_space", "append");
    check(H5Sselect_hyperslab(sf, H5S_SELECT_SET, off, nullptr, ext, nullptr), "H5Sselect_hyperslab", "append");
    const hid_t sm = checkid(H5Screate_simple(2, ext, nullptr), "H5Screate_simple", "append");
    check(H5Dwrite(d, H5T_NATIVE_DOUBLE, sm, sf, H5P_DEFAULT, row), "H5Dwrite", "append");
    H5Sclose(sm);
    H5Sclose(sf);
}

struct BodySets {
    hid_t force_hydro = -1, force_holder = -1, force_friction = -1, position = -1, spring = -1, speed = -1;
    hid_t pressure = -1, friction = -1;
    int32_t nseg = 0;
};

}  // namespace

struct vvhd_stepdata {
    hid_t file = -1;
    hid_t time = -1;
    bool profile = false;
    std::vector<BodySets> bodies;
    long long last_flush = 0;
    ~vvhd_stepdata() {
        if (time >= 0) H5Dclose(time);
        for (BodySets& b : bodies)
            for (hid_t d : {b.force_hydro, b.force_holder, b.force_friction, b.position, b.spring, b.speed, b.pressure,
                            b.friction})
                if (d >= 0) H5Dclose(d);
        if (file >= 0) H5Fclose(file);
    }
};

extern "C" {

const char* vvhd_hdf_error(void) { return g_err.c_str(); }

int vvhd_hdf_save(const char* path, const vvhd_hdf_space* sp, const double* const* lists, const int64_t* counts,
                  int32_t nbody, const vvhd_hdf_body* bodies, const double* seg4, const uint32_t* slip) {
    return guarded([&] {
        if (!path || !sp || !lists || !counts || nbody < 0 || (nbody && (!bodies || !seg4)))
            throw H5Fail("vvhd_hdf_save: bad arguments");
        File f;
        f.id = checkid(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), "H5Fcreate", path);
        Types T;
        attr_str(f.id, "caption", sp->caption, T);
        const double tv[5] = {sp->time, sp->dt, sp->dt_save, sp->dt_streak, sp->dt_profile};
        Frac fr[5];
        for (int k = 0; k < 5; k++)
            fr[k] = sp->frac_scale[k] ? Frac{sp->frac_value[k], sp->frac_scale[k]} : to_frac(tv[k]);
        const char* fn[5] = {"time", "dt", "dt_save", "dt_streak", "dt_profile"};
        for (int k = 0; k < 5; k++) attr_write(f.id, fn[k], T.frac, &fr[k]);
        attr_double(f.id, "re", sp->re);
        attr_double(f.id, "pr", sp->pr);
        attr_write(f.id, "inf_marker", T.vec, sp->inf_marker);
        attr_str(f.id, "inf_speed_x", sp->inf_speed_x, T);
        attr_str(f.id, "inf_speed_y", sp->inf_speed_y, T);
        attr_double(f.id, "inf_circulation", sp->inf_g);
        attr_write(f.id, "gravity", T.vec, sp->gravity);
        attr_double(f.id, "time_to_finish", sp->finish);
        attr_str(f.id, "git_rev", "vvflow_amd", T);
        attr_str(f.id, "git_info", "MI355X device path", T);
        attr_str(f.id, "git_diff", " ", T);
        std::time_t rt = std::time(nullptr);
        std::string ts = std::ctime(&rt);
        if (!ts.empty() && ts.back() == '\n') ts.pop_back();
        attr_str(f.id, "time_local", ts.c_str(), T);
        for (int k = 0; k < VVHD_HDF_LISTS; k++) write_list(f.id, kListNames[k], lists[k], counts[k]);
        int64_t s = 0;
        for (int b = 0; b < nbody; b++) {
            char name[32];
            std::snprintf(name, sizeof(name), "body%02d", b);
            write_body(f.id, name, bodies[b], seg4 + 4 * s, slip ? slip + s : nullptr, T);
            s += bodies[b].nseg;
        }
        T.commit(f.id);
    });
}

int vvhd_hdf_sizes(const char* path, int64_t* counts, int32_t* nbody, int32_t* nseg) {
    return guarded([&] {
        if (!path || !counts || !nbody || !nseg) throw H5Fail("vvhd_hdf_sizes: bad arguments");
        File f;
        f.id = checkid(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), "H5Fopen", path);
        for (int k = 0; k < VVHD_HDF_LISTS; k++) counts[k] = list_rows(f.id, kListNames[k]);
        const auto names = body_names(f.id);
        *nbody = (int32_t)names.size();
        int64_t s = 0;
        for (const auto& n : names) s += list_rows(f.id, n.c_str());
        *nseg = (int32_t)s;
    });
}

int vvhd_hdf_load(const char* path, vvhd_hdf_space* sp, double* const* lists, vvhd_hdf_body* bodies, double* seg4,
                  uint32_t* slip) {
    return guarded([&] {
        if (!path || !sp || !lists) throw H5Fail("vvhd_hdf_load: bad arguments");
        File f;
        f.id = checkid(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), "H5Fopen", path);
        Types T;
        std::memset(sp, 0, sizeof(*sp));
        read_str(f.id, "caption", sp->caption, T);
        const char* fn[5] = {"time", "dt", "dt_save", "dt_streak", "dt_profile"};
        double* tv[5] = {&sp->time, &sp->dt, &sp->dt_save, &sp->dt_streak, &sp->dt_profile};
        for (int k = 0; k < 5; k++) {
            Frac q{0, 1};
            if (!attr_read(f.id, fn[k], T.frac, &q) || q.timescale == 0) q = Frac{0, 1};
            sp->frac_value[k] = q.value;
            sp->frac_scale[k] = q.timescale;
            *tv[k] = (double)q.value / (double)q.timescale;
        }
        sp->re = read_double(f.id, "re");
        sp->pr = read_double(f.id, "pr");
        attr_read(f.id, "inf_marker", T.vec, sp->inf_marker);
        read_str(f.id, "inf_speed_x", sp->inf_speed_x, T);
        read_str(f.id, "inf_speed_y", sp->inf_speed_y, T);
        sp->inf_g = read_double(f.id, "inf_circulation");
        attr_read(f.id, "gravity", T.vec, sp->gravity);
        sp->finish = read_double(f.id, "time_to_finish");
        for (int k = 0; k < VVHD_HDF_LISTS; k++) read_list(f.id, kListNames[k], lists[k], list_rows(f.id, kListNames[k]));
        int64_t s = 0;
        const auto names = body_names(f.id);
        for (size_t b = 0; b < names.size(); b++) {
            vvhd_hdf_body& B = bodies[b];
            std::memset(&B, 0, sizeof(B));
            const char* name = names[b].c_str();
            const hid_t d = checkid(H5Dopen2(f.id, name, H5P_DEFAULT), "H5Dopen", name);
            B.nseg = (int32_t)list_rows(f.id, name);
            read_str(d, "label", B.label, T);
            read_str(d, "root_body", B.root_body, T);
            attr_read(d, "holder_position", T.vec3, B.holder);
            attr_read(d, "delta_position", T.vec3, B.dpos);
            read_str(d, "speed_x", B.speed_x, T);
            read_str(d, "speed_y", B.speed_y, T);
            read_str(d, "speed_o", B.speed_o, T);
            read_str(d, "force_o", B.force_o, T);
            attr_read(d, "speed_slae", T.vec3, B.speed_slae);
            attr_read(d, "speed_slae_prev", T.vec3, B.speed_slae_prev);
            attr_read(d, "spring_const", T.vec3, B.kspring);
            attr_read(d, "spring_damping", T.vec3, B.damping);
            B.density = read_double(d, "density");
            attr_read(d, "force_hydro", T.vec3, B.force_hydro);
            attr_read(d, "force_holder", T.vec3, B.force_holder);
            attr_read(d, "friction_prev", T.vec3, B.friction_prev);
            attr_read(d, "fdt_dead", T.vec3, B.fdt_dead);
            B.g_dead = read_double(d, "g_dead");
            attr_read(d, "collision_min", T.vec3, B.collision_min);
            attr_read(d, "collision_max", T.vec3, B.collision_max);
            B.bounce = read_double(d, "bounce");
            B.area = read_double(d, "area");
            attr_read(d, "cofm", T.vec, B.cofm);
            B.moi_cofm = read_double(d, "moi_cofm");
            B.moi_axis = read_double(d, "moi_axis");
            uint32_t version = 0;
            attr_read(d, "simplified_dataset", H5T_NATIVE_UINT32, &version);
            if (version != 2) throw H5Fail(std::string("vvhd_hdf_load: ") + name + ": dataset version other than 2");
            attr_read(d, "special_segment_no", H5T_NATIVE_INT32, &B.special_segment_no);
            attr_read(d, "boundary_condition", H5T_NATIVE_UINT32, &B.boundary_condition);
            attr_read(d, "heat_condition", H5T_NATIVE_UINT32, &B.heat_condition);
            B.heat_const = read_double(d, "general_heat_const");
            if (seg4) check(H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, seg4 + 4 * s), "H5Dread", name);
            if (slip) {
                if (H5Aexists(d, "slip") > 0) {
                    const hid_t aid = H5Aopen(d, "slip", H5P_DEFAULT);
                    check(H5Aread(aid, H5T_NATIVE_UINT32, slip + s), "H5Aread", "slip");
                    H5Aclose(aid);
                } else {
                    uint32_t g = 0;
                    attr_read(d, "general_slip", H5T_NATIVE_UINT32, &g);
                    for (int32_t k = 0; k < B.nseg; k++) slip[s + k] = g;
                }
            }
            H5Dclose(d);
            s += B.nseg;
        }
    });
}

vvhd_stepdata* vvhd_stepdata_open(const char* path, const char* caption, double time, int32_t nbody,
                                  const int32_t* nseg, int32_t save_profile) {
    vvhd_stepdata* sd = new vvhd_stepdata;
    const int rc = guarded([&] {
        if (!path || nbody < 0 || (nbody && !nseg)) throw H5Fail("vvhd_stepdata_open: bad arguments");
        sd->profile = save_profile != 0;
        hsize_t rows = 0;
        if (H5Fis_hdf5(path) <= 0) {
            sd->file = checkid(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), "H5Fcreate", path);
            Types T;
            attr_str(sd->file, "caption", caption, T);
            attr_str(sd->file, "git_rev", "vvflow_amd", T);
            attr_str(sd->file, "git_info", "MI355X device path", T);
            attr_str(sd->file, "git_diff", " ", T);
            check(H5Tcommit2(sd->file, "string_t", T.str, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "H5Tcommit",
                  "string_t");
        } else {
            sd->file = checkid(H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT), "H5Fopen", path);
            if (H5Lexists(sd->file, "time", H5P_DEFAULT) > 0) {
                // keep the rows written before this time (MStepdata.cpp:54-75)
                const hid_t d = checkid(H5Dopen2(sd->file, "time", H5P_DEFAULT), "H5Dopen", "time");
                const hid_t sp = H5Dget_space(d);
                hsize_t dims[2] = {0, 0};
                H5Sget_simple_extent_dims(sp, dims, nullptr);
                std::vector<float> t((size_t)(dims[0] * dims[1]) + 1);
                check(H5Dread(d, H5T_NATIVE_FLOAT, sp, sp, H5P_DEFAULT, t.data()), "H5Dread", "time");
                while (rows < dims[0] && t[(size_t)(rows * dims[1])] < (float)time) rows++;
                H5Sclose(sp);
                H5Dclose(d);
            }
        }
        sd->time = step_dataset(sd->file, "time", rows, 1);
        sd->bodies.resize((size_t)nbody);
        for (int32_t b = 0; b < nbody; b++) {
            char name[32];
            std::snprintf(name, sizeof(name), "body%02d", b);
            const hid_t g = H5Lexists(sd->file, name, H5P_DEFAULT) > 0
                                ? checkid(H5Gopen2(sd->file, name, H5P_DEFAULT), "H5Gopen", name)
                                : checkid(H5Gcreate2(sd->file, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                                          "H5Gcreate", name);
            BodySets& B = sd->bodies[(size_t)b];
            B.nseg = nseg[b];
            B.force_hydro = step_dataset(g, "force_hydro", rows, 3);
            B.force_holder = step_dataset(g, "force_holder", rows, 3);
            B.force_friction = step_dataset(g, "force_friction", rows, 3);
            B.position = step_dataset(g, "holder_position", rows, 3);
            B.spring = step_dataset(g, "delta_position", rows, 3);
            B.speed = step_dataset(g, "speed_slae", rows, 3);
            const hsize_t ns = (hsize_t)std::max<int32_t>(nseg[b], 1);
            if (sd->profile) {
                B.pressure = step_dataset(g, "pressure", rows, ns);
                B.friction = step_dataset(g, "friction", rows, ns);
            } else {
                for (const char* n : {"pressure", "friction"})
                    if (H5Lexists(g, n, H5P_DEFAULT) > 0) H5Dclose(step_dataset(g, n, rows, ns));
            }
            H5Gclose(g);
        }
    });
    if (rc != 0) {
        delete sd;
        return nullptr;
    }
    return sd;
}

int vvhd_stepdata_write(vvhd_stepdata* sd, double time, const vvhd_stepdata_body* bodies, const double* pressure,
                        const double* friction) {
    return guarded([&] {
        if (!sd || (!sd->bodies.empty() && !bodies) || (sd->profile && (!pressure || !friction)))
            throw H5Fail("vvhd_stepdata_write: bad arguments");
        step_append(sd->time, &time);
        size_t s = 0;
        for (size_t b = 0; b < sd->bodies.size(); b++) {
            const BodySets& B = sd->bodies[b];
            const vvhd_stepdata_body& r = bodies[b];
            step_append(B.force_hydro, r.force_hydro);
            step_append(B.force_holder, r.force_holder);
            step_append(B.force_friction, r.friction);
            step_append(B.position, r.holder);
            step_append(B.spring, r.dpos);
            step_append(B.speed, r.speed_slae);
            if (sd->profile && B.nseg > 0) {
                step_append(B.pressure, pressure + s);
                step_append(B.friction, friction + s);
            }
            s += (size_t)B.nseg;
        }
        timespec clk;
        clock_gettime(CLOCK_MONOTONIC, &clk);
        if (clk.tv_sec - sd->last_flush > 20) {
            check(H5Fflush(sd->file, H5F_SCOPE_GLOBAL), "H5Fflush", "stepdata");
            sd->last_flush = clk.tv_sec;
        }
    });
}

int vvhd_stepdata_flush(vvhd_stepdata* sd) {
    if (!sd) return -1;
    sd->last_flush = 0;
    return 0;
}

int vvhd_stepdata_close(vvhd_stepdata* sd) {
    delete sd;
    return 0;
}

}  // extern "C"
