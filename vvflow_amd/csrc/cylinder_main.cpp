// cylinder_main.cpp -- the reference's simulate() loop (utils/vvflow/
// vvflow.cpp:177-266) written against the module classes of
// include/vvhd_modules.hpp, for an impulsively started circular cylinder
// (utils/vvflow/gen_body.cpp:74-88 corners, R = 0.5, 500 segments).
//
//   vvflow_cylinder [--steps 400] [--re 600] [--dt 0.005] [--nseg 500]
//                   [--learned model.npy] [--far-criteria 8] [--every 1]
//                   [--sensors points.txt --sensors-out out.txt]
//                   [--fields PREFIX] [--heat COND:HEAT_CONST:PR]
//                   [--streak N:DT_STREAK] [--save PATH] [--load PATH]
//                   [--stepdata PATH [--profile]]
//
// Prints one line per `every` steps: step, time, vortex count, force_hydro
// (x, y, o), friction (x, y, o), step wall time in ms; then a summary line.
// --sensors appends the velocity at the listed points every step, after the
// convective velocities (utils/vvflow/vvflow.cpp:248, sensors::output).  --fields evaluates, after
// the last step, XVelocity ('x', frame 'o'), XVorticity, XStreamfunction
// (frame 'o') and XPressure (frame 's') with eps_mult 1 on the grid
// x in [-1, 3), y in [-1, 1) step 0.05, written with XField's operator<< to
// PREFIX_{velocity,vorticity,streamfunction,pressure}.bin (and the state
// they start from to PREFIX_state.bin), and prints
// MEpsilonFast::eps2h / h2 at three points with the near and far list sizes
// of their leaves.  --heat gives the body a heat condition (0 neglect,
// 1 isolate, 2 const_t, 3 const_w), every segment heat_const and Space::pr;
// --streak puts N streak sources on the segment x = -1, y in [-1, 1], shed
// every DT_STREAK.  The summary then also reports the heat and streak lists.
// --save writes an HDF5 snapshot of the final state (Space::save); --load
// starts from a snapshot instead of an empty cylinder (Space::load).
// --stepdata writes the per-step time series (Stepdata, MStepdata.cpp) after
// the forces of every step, as the reference's loop does (vvflow.cpp:235-237),
// with the segments' pressure and friction profiles under --profile.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "vvhd_modules.hpp"

namespace {

// numpy .npy (version 1.x / 2.x, little-endian float32, one dimension)
std::vector<float> load_npy_f32(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    char magic[6];
    f.read(magic, 6);
    if (std::memcmp(magic, "\x93NUMPY", 6) != 0) throw std::runtime_error(path + ": not a .npy file");
    unsigned char ver[2];
    f.read(reinterpret_cast<char*>(ver), 2);
    uint32_t hlen = 0;
    if (ver[0] == 1) {
        unsigned char b[2];
        f.read(reinterpret_cast<char*>(b), 2);
        hlen = b[0] | (b[1] << 8);
    } else {
        unsigned char b[4];
        f.read(reinterpret_cast<char*>(b), 4);
        hlen = b[0] | (b[1] << 8) | (b[2] << 16) | ((uint32_t)b[3] << 24);
    }
    std::string header(hlen, ' ');
    f.read(&header[0], hlen);
    if (header.find("'<f4'") == std::string::npos || header.find("'fortran_order': False") == std::string::npos)
        throw std::runtime_error(path + ": expected a little-endian float32 array");
    std::vector<float> out;
    float v;
    while (f.read(reinterpret_cast<char*>(&v), 4)) out.push_back(v);
    return out;
}

}  // namespace

int main(int argc, char** argv) {
    int steps = 400, nseg = 500, fc = 8, every = 1;
    double re = 600, dt = 0.005, radius = 0.5;
    std::string learned, sensors_in, sensors_out, fields, save_path, load_path, stepdata_path;
    int heat_cond = 0, nstreak = 0;
    bool profile = false;
    double heat_const = 0, pr = 0, dt_streak = 0;
    for (int k = 1; k < argc; k++) {
        const std::string a = argv[k];
        auto next = [&]() -> const char* {
            if (k + 1 >= argc) throw std::invalid_argument("missing value for " + a);
            return argv[++k];
        };
        if (a == "--steps") steps = std::atoi(next());
        else if (a == "--re") re = std::atof(next());
        else if (a == "--dt") dt = std::atof(next());
        else if (a == "--nseg") nseg = std::atoi(next());
        else if (a == "--far-criteria") fc = std::atoi(next());
        else if (a == "--learned") learned = next();
        else if (a == "--every") every = std::atoi(next());
        else if (a == "--sensors") sensors_in = next();
        else if (a == "--sensors-out") sensors_out = next();
        else if (a == "--fields") fields = next();
        else if (a == "--save") save_path = next();
        else if (a == "--load") load_path = next();
        else if (a == "--stepdata") stepdata_path = next();
        else if (a == "--profile") profile = true;
        else if (a == "--heat") {
            if (std::sscanf(next(), "%d:%lf:%lf", &heat_cond, &heat_const, &pr) != 3)
                throw std::invalid_argument("--heat expects COND:HEAT_CONST:PR");
        } else if (a == "--streak") {
            if (std::sscanf(next(), "%d:%lf", &nstreak, &dt_streak) != 2)
                throw std::invalid_argument("--streak expects N:DT_STREAK");
        }
        else {
            fprintf(stderr, "unknown option %s\n", a.c_str());
            return 2;
        }
    }
    try {
        Space S(0);
        if (!load_path.empty()) {
            S.load(load_path.c_str());
        } else {
            S.re = re;
            S.dt = TTime::makeWithSecondsDecimal(dt);
            S.inf_vel = TVec(1, 0);
            S.caption = "cylinder";
            auto body = std::make_shared<TBody>();
            std::vector<double> xy(2 * (size_t)nseg);
            if (vvhd_gen_cylinder(radius, nseg, xy.data()) != VVHD_OK) throw std::runtime_error("gen_cylinder failed");
            for (int k = 0; k < nseg; k++) body->alist.emplace_back(xy[2 * k], xy[2 * k + 1]);
            body->heat_condition = static_cast<hc_t>(heat_cond);
            for (TAtt& a : body->alist) a.heat_const = (float)heat_const;
            S.BodyList.push_back(body);
            S.pr = pr;
            S.dt_streak = TTime::makeWithSecondsDecimal(dt_streak);
            for (int k = 0; k < nstreak; k++)
                S.StreakSourceList.emplace_back(-1.0, nstreak > 1 ? -1.0 + 2.0 * k / (nstreak - 1) : 0.0, 0.0);
            S.push();
        }

        const double dl = S.average_segment_length();
        const double min_node_size = dl > 0 ? dl * 5 : 0;
        const double max_node_size = dl > 0 ? dl * 100 : std::numeric_limits<double>::max();
        TSortedTree tr(&S, fc, min_node_size, max_node_size);
        MConvectiveFast convective(&S, &tr);
        MEpsilonFast epsilon(&S, &tr);
        MDiffusiveFast diffusive(&S, &tr);
        MFlowmove flowmove(&S);
        if (!learned.empty()) convective.set_learned(load_npy_f32(learned));
        const bool is_viscous = std::isfinite(S.re);
        std::unique_ptr<sensors> probes;
        if (!sensors_in.empty())
            probes.reset(new sensors(&S, &convective, sensors_in.c_str(),
                                     sensors_out.empty() ? "sensors.txt" : sensors_out.c_str()));
        std::unique_ptr<Stepdata> stepdata;
        if (!stepdata_path.empty()) stepdata.reset(new Stepdata(&S, stepdata_path.c_str(), profile));

        const void* collision = nullptr;
        double total_ms = 0;
        for (int step = 0; step < steps; step++) {
            const auto t0 = std::chrono::steady_clock::now();
            if (S.BodyList.size()) {
                tr.build();
                convective.calc_circulation(&collision);
                tr.destroy();
            }
            flowmove.heat_shed();
            flowmove.vortex_shed();
            flowmove.streak_shed();
            S.calc_forces();
            const TBody& b = *S.BodyList[0];
            const TVec3D fh = b.force_hydro, fr = b.friction;
            if (stepdata) stepdata->write();
            S.zero_forces();
            tr.build();
            epsilon.CalcEpsilonFast(is_viscous);
            convective.process_all_lists();
            if (probes) probes->output();   // vvflow.cpp:248, on this step's tree
            if (is_viscous) {
                diffusive.process_vort_list();
                diffusive.process_heat_list();
            }
            tr.destroy();
            flowmove.move_and_clean(true, &collision);
            flowmove.heat_crop();
            S.finish_step();
            const double ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            total_ms += ms;
            if (every > 0 && (step + 1) % every == 0)
                printf("%d %.6f %lld %.10e %.10e %.10e %.10e %.10e %.10e %.3f\n", step + 1, double(S.time),
                       (long long)S.vortex_count(), fh.r.x, fh.r.y, fh.o, fr.r.x, fr.r.y, fr.o, ms);
        }
        printf("# steps %d vortices %lld total %.1f ms mean %.3f ms/step\n", steps, (long long)S.vortex_count(),
               total_ms, total_ms / std::max(1, steps));
        if (!save_path.empty()) S.save(save_path.c_str());
        if (heat_cond || nstreak) {
            S.pull();
            double hg = 0;
            for (const TObj& o : S.HeatList) hg += o.g;
            double hsum = 0;
            for (const TAtt& a : S.BodyList[0]->alist) hsum += a.hsum;
            printf("# heat %zu %.17g hsum %.17g streak %zu nusselt %.17g\n", S.HeatList.size(), hg, hsum,
                   S.StreakList.size(), S.BodyList[0]->nusselt);
        }
        if (!fields.empty()) {
            // the state the evaluators copy: vortices (6 doubles each), then
            // per segment its vvhd_seg fields as 16 doubles, then the body's
            // holder, dpos, speed_slae, speed_slae_prev, friction_prev and
            // fdt_dead triples and g_dead
            S.pull();
            std::ofstream st(fields + "_state.bin", std::ios::binary);
            const int64_t nv = (int64_t)S.VortexList.size();
            st.write(reinterpret_cast<const char*>(&nv), 8);
            st.write(reinterpret_cast<const char*>(S.VortexList.data()), nv * (int64_t)sizeof(TObj));
            const TBody& bd = *S.BodyList[0];
            const int64_t ns = (int64_t)bd.alist.size();
            st.write(reinterpret_cast<const char*>(&ns), 8);
            for (const TAtt& a : bd.alist) {
                const double rec[16] = {a.r.x,  a.r.y,  a.g,  a._1_eps, a.corner.x, a.corner.y,     a.dl.x,
                                        a.dl.y, a.gsum, a.hsum, a.fric, a.Cp,      a.Fr,       a.Nu,
                                        (double)a.slip, (double)a.eq_no};
                st.write(reinterpret_cast<const char*>(rec), sizeof(rec));
            }
            const TVec3D* v3[6] = {&bd.holder, &bd.dpos, &bd.speed_slae, &bd.speed_slae_prev, &bd.friction_prev,
                                   &bd.fdt_dead};
            for (const TVec3D* v : v3) {
                const double t[3] = {v->r.x, v->r.y, v->o};
                st.write(reinterpret_cast<const char*>(t), sizeof(t));
            }
            st.write(reinterpret_cast<const char*>(&bd.g_dead), 8);
            st.close();
            XVelocity fv(S, -1, -1, 0.05, 80, 40);
            fv.mode = 'x';
            fv.ref_frame = 'o';
            XVorticity fw(S, -1, -1, 0.05, 80, 40);
            fw.eps_mult = 1;
            XStreamfunction fs(S, -1, -1, 0.05, 80, 40);
            fs.eps_mult = 1;
            fs.ref_frame = 'o';
            XPressure fp(S, -1, -1, 0.05, 80, 40);
            fp.eps_mult = 1;
            fp.ref_frame = 's';
            XField* all[4] = {&fv, &fw, &fs, &fp};
            const char* names[4] = {"velocity", "vorticity", "streamfunction", "pressure"};
            for (int k = 0; k < 4; k++) {
                if (k == 0) fv.evaluate();
                if (k == 1) fw.evaluate();
                if (k == 2) fs.evaluate();
                if (k == 3) fp.evaluate();
                std::ofstream out(fields + "_" + names[k] + ".bin", std::ios::binary);
                out << *all[k];
                printf("# field %s min %.9g max %.9g\n", names[k], all[k]->min(), all[k]->max());
            }
            tr.build();
            const TVec probe[3] = {TVec(0.7, 0.1), TVec(1.5, -0.2), TVec(0.0, 0.55)};
            for (const TVec& p : probe) {
                const TSortedNode* nd = tr.findNode(p);
                printf("# node %.17g %.17g eps2h %.17g h2 %.17g near %zu far %zu\n", p.x, p.y,
                       MEpsilonFast::eps2h(*nd, p), MEpsilonFast::h2(*nd, p), nd->NearNodes->size(),
                       nd->FarNodes->size());
            }
            printf("# integrals gsum %.17g integral %.17g momentum %.17g %.17g\n", S.gsum(), S.integral(),
                   S.hydrodynamic_momentum().x, S.hydrodynamic_momentum().y);
            tr.destroy();
        }
    } catch (const std::exception& e) {
        fprintf(stderr, "vvflow_cylinder: %s\n", e.what());
        return 1;
    }
    return 0;
}
