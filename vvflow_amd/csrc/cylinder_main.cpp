// cylinder_main.cpp -- the reference's simulate() loop (utils/vvflow/
// vvflow.cpp:177-266) written against the module classes of
// include/vvhd_modules.hpp, for an impulsively started circular cylinder
// (utils/vvflow/gen_body.cpp:74-88 corners, R = 0.5, 500 segments).
//
//   vvflow_cylinder [--steps 400] [--re 600] [--dt 0.005] [--nseg 500]
//                   [--learned model.npy] [--far-criteria 8] [--every 1]
//
// Prints one line per `every` steps: step, time, vortex count, force_hydro
// (x, y, o), friction (x, y, o), step wall time in ms; then a summary line.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
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
    std::string learned;
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
        else {
            fprintf(stderr, "unknown option %s\n", a.c_str());
            return 2;
        }
    }
    try {
        Space S(0);
        S.re = re;
        S.dt = dt;
        S.inf_vel = TVec(1, 0);
        auto body = std::make_shared<TBody>();
        std::vector<double> xy(2 * (size_t)nseg);
        if (vvhd_gen_cylinder(radius, nseg, xy.data()) != VVHD_OK) throw std::runtime_error("gen_cylinder failed");
        for (int k = 0; k < nseg; k++) body->alist.emplace_back(xy[2 * k], xy[2 * k + 1]);
        S.BodyList.push_back(body);
        S.push();

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
            S.zero_forces();
            tr.build();
            epsilon.CalcEpsilonFast(is_viscous);
            convective.process_all_lists();
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
                printf("%d %.6f %lld %.10e %.10e %.10e %.10e %.10e %.10e %.3f\n", step + 1, S.time,
                       (long long)S.vortex_count(), fh.r.x, fh.r.y, fh.o, fr.r.x, fr.r.y, fr.o, ms);
        }
        printf("# steps %d vortices %lld total %.1f ms mean %.3f ms/step\n", steps, (long long)S.vortex_count(),
               total_ms, total_ms / std::max(1, steps));
    } catch (const std::exception& e) {
        fprintf(stderr, "vvflow_cylinder: %s\n", e.what());
        return 1;
    }
    return 0;
}
