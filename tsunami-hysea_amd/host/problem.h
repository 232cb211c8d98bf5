// problem.h -- parameter file, grid input, non-dimensionalisation, geometry and
// gauges of one run: the host setup of the reference (src/Problema.cxx:176-477).
#pragma once
#include <string>
#include <vector>

#include "../../include/thysea.h"

namespace thyhost {

constexpr double EARTH_RADIUS = 6378136.6;  // src/Constantes.hxx:12
constexpr int MAX_FAULTS = 5000;            // src/Constantes.hxx:13
constexpr int SPONGE_SIZE = 4;              // src/Constantes.hxx:14
constexpr double SEA_LEVEL = 0.0;           // src/Constantes.hxx:15
constexpr int SEA_SURFACE_FROM_FILE = 0, OKADA_STANDARD = 1;
constexpr int NO_CROP = 0, CROP_RELATIVE = 1, CROP_ABSOLUTE = 2;

struct Problem {
    std::string name, prefix, bathy_path;
    int okada_flag = OKADA_STANDARD;
    std::vector<thy_fault> faults;   // sorted by time
    int crop_flag = NO_CROP;
    double crop_value = 0.0;         // non-dimensional when absolute
    int nx = 0, ny = 0;
    std::vector<double> lon, lat;
    std::vector<double> state1, state2;  // (h, H) and (qx, qy) pairs, non-dimensional
    double hmin = 0.0;
    std::vector<double> area, cosphi, ancho, alto;
    double borde_sup = 1, borde_inf = 1, borde_izq = 1, borde_der = 1;
    int sp_sup = 0, sp_inf = 0, sp_izq = 0, sp_der = 0;
    double t_end = 0, t_save = -1, t_series = -1;  // non-dimensional
    double cfl = 0.5, mf0 = 0, vmax = 0, eps_h = 0, hpos = 0, cvis = 0, dif_at = 0;
    double L = 1e5, H = 1e3, Q = 0, T = 0;
    int read_points = 0;
    std::vector<double> plon, plat;
    std::vector<int> ppos;
    bool periodic() const;
    thy_grid grid() const;
    thy_params params(int precision, int track_arrival) const;
};

// Returns 0 on success; on error prints the reference's message to stderr and returns 1.
int load_problem(const std::string &param_file, Problem &p);
void print_problem(const Problem &p, const std::string &version);

}  // namespace thyhost
