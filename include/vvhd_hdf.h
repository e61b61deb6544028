/*
 * vvhd_hdf.h -- the reference's HDF5 formats, host code in libvvhd_hdf.so
 * over libhdf5: snapshots (Space::save_hdf / load_hdf, vvflow/vvflow
 * libvvhd/src/TSpace_save_hdf.cpp:11-221, TSpace_load_hdf.cpp:11-228,
 * vvhdf.cpp, vvhdf_h5t.cpp) and the per-step time series (Stepdata, below).
 *
 * A file holds:
 *  - root attributes: caption, time, dt, dt_save, dt_streak, dt_profile
 *    (fractions, committed type "fraction_t" {int32 value, uint32
 *    timescale}), re, pr, inf_marker (vec_t {x, y}), inf_speed_x,
 *    inf_speed_y (variable-length strings, "string_t"), inf_circulation,
 *    gravity, time_to_finish, git_rev, git_info, git_diff, time_local;
 *  - datasets vort, heat, ink, ink_source, source: N x 3 doubles
 *    {x, y, g} per TObj (chunked by 512 rows, deflate 9; absent when the
 *    list is empty);
 *  - datasets body00, body01, ...: N x 4 doubles {corner.x, corner.y, g,
 *    gsum} per segment with the body's attributes (simplified_dataset = 2:
 *    general_slip or a slip array, general_heat_const, label, root_body,
 *    holder_position, delta_position, speed_x/y/o, force_o, speed_slae,
 *    speed_slae_prev, spring_const, spring_damping, density, force_hydro,
 *    force_holder, friction_prev, fdt_dead, g_dead, collision_min,
 *    collision_max, bounce, area, cofm, moi_cofm, moi_axis,
 *    special_segment_no, boundary_condition, heat_condition).
 * Times are written as the fractions the caller gives (frac_value /
 * frac_scale, as the reference writes its TTime), or from the doubles in
 * lowest terms.
 * Loading gives every TObj v = 0 and _1_eps = 0, as the reference's
 * strided read does (TSpace_load_hdf.cpp:11-45); the caller rebuilds the
 * segments from the corners (TBody::doUpdateSegments).
 */
#ifndef VVHD_HDF_H
#define VVHD_HDF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { VVHD_HDF_LISTS = 5 };   /* vort, heat, ink, ink_source, source */
enum { VVHD_HDF_STR = 256 };

typedef struct vvhd_hdf_space {
    char caption[VVHD_HDF_STR];
    double time, dt, dt_save, dt_streak, dt_profile;   /* stored as fractions */
    double re, pr, inf_g, finish;
    double inf_marker[2], gravity[2];
    char inf_speed_x[VVHD_HDF_STR], inf_speed_y[VVHD_HDF_STR];
    /* the same five times as exact fractions value / timescale, in the order
     * time, dt, dt_save, dt_streak, dt_profile: written as given when the
     * timescale is non-zero (else the double, in lowest terms), and filled
     * by the loader */
    int32_t frac_value[5];
    uint32_t frac_scale[5];
} vvhd_hdf_space;

typedef struct vvhd_hdf_body {
    char label[VVHD_HDF_STR], root_body[VVHD_HDF_STR];
    char speed_x[VVHD_HDF_STR], speed_y[VVHD_HDF_STR], speed_o[VVHD_HDF_STR], force_o[VVHD_HDF_STR];
    double holder[3], dpos[3], speed_slae[3], speed_slae_prev[3], kspring[3], damping[3];
    double force_hydro[3], force_holder[3], friction_prev[3], fdt_dead[3];
    double collision_min[3], collision_max[3];
    double density, g_dead, bounce, area, cofm[2], moi_cofm, moi_axis;
    double heat_const;
    int32_t special_segment_no;
    uint32_t boundary_condition;   /* bc_t: 0 steady, 1 kutta */
    uint32_t heat_condition;       /* hc_t: 0 neglect, 1 isolate, 2 const_t, 3 const_w */
    int32_t nseg;
} vvhd_hdf_body;

/* Writes a snapshot.  lists[k] holds counts[k] TObj records (6 doubles)
 * of list k; seg4 holds per segment of all bodies in order {corner.x,
 * corner.y, g, gsum}; slip per segment (may be NULL: no slip).  Returns 0,
 * or -1 with a message in vvhd_hdf_error(). */
int vvhd_hdf_save(const char* path, const vvhd_hdf_space* space, const double* const* lists, const int64_t* counts,
                  int32_t nbody, const vvhd_hdf_body* bodies, const double* seg4, const uint32_t* slip);
/* Sizes of a snapshot: the five list lengths, the bodies and the segments. */
int vvhd_hdf_sizes(const char* path, int64_t* counts, int32_t* nbody, int32_t* nseg);
/* Reads a snapshot into buffers of the sizes vvhd_hdf_sizes gave. */
int vvhd_hdf_load(const char* path, vvhd_hdf_space* space, double* const* lists, vvhd_hdf_body* bodies,
                  double* seg4, uint32_t* slip);
const char* vvhd_hdf_error(void);

/* ---- step data: Stepdata (MStepdata.hpp:7-39, MStepdata.cpp:10-290) ----
 * One row per step in stepdata_<caption>.h5: a dataset "time" (rows x 1)
 * and per body a group bodyNN with force_hydro, force_holder,
 * force_friction, holder_position, delta_position, speed_slae (rows x 3
 * {x, y, o}) and, with profiles, pressure and friction (rows x nseg).
 * Datasets are stored as float, chunked by 128 rows, deflate 9, NaN fill,
 * unlimited rows.  A new file gets the attributes caption, git_rev,
 * git_info, git_diff and the committed type string_t.  Opening an existing
 * file keeps the leading rows whose time is below `time` (a restart) and
 * cuts every dataset there; without profiles, existing pressure and
 * friction datasets are cut too.  The file is flushed every 20 s of wall
 * time, and at the next write after vvhd_stepdata_flush. */
typedef struct vvhd_stepdata vvhd_stepdata;
typedef struct vvhd_stepdata_body {
    double force_hydro[3], force_holder[3], friction[3];
    double holder[3], dpos[3], speed_slae[3];
} vvhd_stepdata_body;
/* NULL on failure (message in vvhd_hdf_error()); nseg[b] is body b's
 * segment count. */
vvhd_stepdata* vvhd_stepdata_open(const char* path, const char* caption, double time, int32_t nbody,
                                  const int32_t* nseg, int32_t save_profile);
/* Appends one row: time, every body's record, and with profiles the
 * per-segment pressure and friction of all bodies in order (else NULL). */
int vvhd_stepdata_write(vvhd_stepdata* sd, double time, const vvhd_stepdata_body* bodies, const double* pressure,
                        const double* friction);
int vvhd_stepdata_flush(vvhd_stepdata* sd);
int vvhd_stepdata_close(vvhd_stepdata* sd);

#ifdef __cplusplus
}
#endif

#endif
