// problem.cpp -- host setup of one run (mirrors src/Problema.cxx:11-533).
#include "problem.h"

#include <sys/stat.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <fstream>
#include <iostream>
#include <sstream>

#include "ncclassic.h"

namespace thyhost {

namespace {

bool file_exists(const std::string &f) {
    struct stat st;
    return stat(f.c_str(), &st) == 0;
}

// One datum per line, skipping blank lines and lines whose first non-blank
// character is '#' (src/Problema.cxx:150-170).  Returns false at end of file.
bool next_line(std::ifstream &f, std::string &line) {
    std::string l;
    while (std::getline(f, l)) {
        size_t p = l.find_first_not_of(" \t\r\n");
        if (p == std::string::npos) continue;
        l.erase(0, p);
        if (l[0] == '#') continue;
        line = l;
        return true;
    }
    line.clear();
    return false;
}

template <typename T>
void next_datum(std::ifstream &f, T &v) {
    std::string l;
    next_line(f, l);
    std::istringstream iss(l);
    iss >> v;  // a missing line leaves v unchanged
}

// nearest grid cell of a point (src/Problema.cxx:54-95)
int nearest_cell(const std::vector<double> &lon, const std::vector<double> &lat, double plon, double plat) {
    const int nx = (int)lon.size(), ny = (int)lat.size();
    if (!(plon >= lon[0])) return -1;
    int i = 0;
    while (i < nx && !(lon[i] >= plon)) i++;
    if (i >= nx) return -1;
    if (std::fabs(plon - lon[std::max(i - 1, 0)]) < std::fabs(plon - lon[i])) i = std::max(i - 1, 0);
    if (!(plat >= lat[0])) return -1;
    int j = 0;
    while (j < ny && !(lat[j] >= plat)) j++;
    if (j >= ny) return -1;
    if (std::fabs(plat - lat[std::max(j - 1, 0)]) < std::fabs(plat - lat[j])) j = std::max(j - 1, 0);
    return j * nx + i;
}

bool read_grid(const std::string &path, std::vector<double> &lon, std::vector<double> &lat,
               std::vector<float> &z, std::string &err) {
    nc::Reader r;
    if (!r.open(path, &err)) return false;
    int vlon = r.find_var("lon"), vlat = r.find_var("lat");
    if (vlon < 0) vlon = r.find_var("x");
    if (vlat < 0) vlat = r.find_var("y");
    int vz = r.find_var("z");
    if (vlon < 0 || vlat < 0 || vz < 0) {
        err = path + ": variables lon/lat (or x/y) and z are required";
        return false;
    }
    if (!r.read_double(vlon, lon, &err) || !r.read_double(vlat, lat, &err) || !r.read_float(vz, z, &err))
        return false;
    if (z.size() != lon.size() * lat.size()) {
        err = path + ": z is not lat x lon";
        return false;
    }
    return true;
}

}  // namespace

bool Problem::periodic() const {
    return std::fabs(borde_izq - 2.0) < DBL_EPSILON && std::fabs(borde_der - 2.0) < DBL_EPSILON;
}

thy_grid Problem::grid() const {
    thy_grid g;
    g.nx = nx;
    g.ny = ny;
    g.lon = lon.data();
    g.lat = lat.data();
    g.area = area.data();
    g.cosphi = cosphi.data();
    g.ancho = ancho.data();
    g.alto = alto.data();
    return g;
}

thy_params Problem::params(int precision, int track_arrival) const {
    thy_params p{};
    p.borde_sup = borde_sup; p.borde_inf = borde_inf; p.borde_izq = borde_izq; p.borde_der = borde_der;
    p.sponge_sup = sp_sup; p.sponge_inf = sp_inf; p.sponge_izq = sp_izq; p.sponge_der = sp_der;
    p.cfl = cfl; p.mf0 = mf0; p.vmax = vmax; p.eps_h = eps_h; p.hpos = hpos; p.cvis = cvis; p.dif_at = dif_at;
    p.hmin = hmin;
    p.sea_level = SEA_LEVEL / H;
    p.L = L; p.H = H; p.Q = Q; p.T = T;
    p.precision = precision;
    p.track_arrival = track_arrival;
    return p;
}

int load_problem(const std::string &fich_ent, Problem &p) {
    std::string dir;
    size_t k = fich_ent.find_last_of('/');
    if (k != std::string::npos) dir = fich_ent.substr(0, k) + "/";
    else if ((k = fich_ent.find_last_of('\\')) != std::string::npos) dir = fich_ent.substr(0, k) + "\\";

    std::ifstream f(fich_ent.c_str());
    std::string topo, est, pts;
    next_datum(f, p.name);
    next_datum(f, topo);
    if (!file_exists(dir + topo)) {
        std::cerr << "Error: File '" << dir + topo << "' not found" << std::endl;
        return 1;
    }
    next_datum(f, p.okada_flag);
    if (p.okada_flag != SEA_SURFACE_FROM_FILE && p.okada_flag != OKADA_STANDARD) {
        std::cerr << "Error: The initialization flag should be " << SEA_SURFACE_FROM_FILE << " or " << OKADA_STANDARD
                  << std::endl;
        return 1;
    }
    if (p.okada_flag == SEA_SURFACE_FROM_FILE) {
        next_datum(f, est);
        if (!file_exists(dir + est)) {
            std::cerr << "Error: File '" << dir + est << "' not found" << std::endl;
            return 1;
        }
    } else {
        int nf = 0;
        next_datum(f, nf);
        if (nf < 1 || nf > MAX_FAULTS) {
            std::cerr << "Error: The number of faults should be between 1 and " << MAX_FAULTS << std::endl;
            return 1;
        }
        p.faults.resize(nf);
        for (int i = 0; i < nf; i++) {
            std::string l;
            next_line(f, l);
            std::istringstream iss(l);
            thy_fault &q = p.faults[i];
            iss >> q.time >> q.lon_c >> q.lat_c >> q.depth_c >> q.length >> q.width >> q.strike >> q.dip >> q.rake >>
                q.slip;
        }
        // selection sort by time, swapping only on a strictly smaller time (src/Problema.cxx:97-126)
        for (int i = 0; i < nf - 1; i++) {
            int m = i;
            for (int j = i + 1; j < nf; j++)
                if (p.faults[j].time < p.faults[m].time) m = j;
            if (m != i) std::swap(p.faults[i], p.faults[m]);
        }
        next_datum(f, p.crop_flag);
        if (p.crop_flag != NO_CROP && p.crop_flag != CROP_RELATIVE && p.crop_flag != CROP_ABSOLUTE) {
            std::cerr << "Error: The cropping flag should be " << NO_CROP << ", " << CROP_RELATIVE << " or "
                      << CROP_ABSOLUTE << std::endl;
            return 1;
        }
        if (p.crop_flag == CROP_RELATIVE || p.crop_flag == CROP_ABSOLUTE) next_datum(f, p.crop_value);
    }
    next_datum(f, p.prefix);
    std::vector<float> z;
    std::string err;
    p.bathy_path = dir + topo;
    if (!read_grid(p.bathy_path, p.lon, p.lat, z, err)) {
        std::cerr << "Error: " << err << std::endl;
        return 1;
    }
    p.nx = (int)p.lon.size();
    p.ny = (int)p.lat.size();
    if (p.nx < 2 || p.ny < 2) {
        std::cerr << "Error: Mesh size too small. The number of rows and columns should be >= 2" << std::endl;
        return 1;
    }
    int levels = 0;
    next_datum(f, levels);
    if (levels != 1) {
        std::cerr << "Error: The number of levels should be 1" << std::endl;
        return 1;
    }
    next_datum(f, p.borde_sup);
    next_datum(f, p.borde_inf);
    next_datum(f, p.borde_izq);
    next_datum(f, p.borde_der);
    auto bad = [](double b) {
        return std::fabs(b - 1.0) > DBL_EPSILON && std::fabs(b - 2.0) > DBL_EPSILON && std::fabs(b + 1.0) > DBL_EPSILON;
    };
    if (bad(p.borde_sup) || bad(p.borde_inf) || bad(p.borde_izq) || bad(p.borde_der)) {
        std::cerr << "Error: The boundary conditions should be -1, 1 or 2" << std::endl;
        return 1;
    }
    if (std::fabs(p.borde_sup - 2.0) < DBL_EPSILON || std::fabs(p.borde_inf - 2.0) < DBL_EPSILON) {
        std::cerr << "Error: Periodic boundary conditions are only supported for left and right borders" << std::endl;
        return 1;
    }
    if ((std::fabs(p.borde_izq - 2.0) < DBL_EPSILON) != (std::fabs(p.borde_der - 2.0) < DBL_EPSILON)) {
        std::cerr << "Error: Both left and right borders should have periodic boundary conditions" << std::endl;
        return 1;
    }
    next_datum(f, p.t_end);
    next_datum(f, p.t_save);
    next_datum(f, p.read_points);
    if (p.read_points == 1) {
        next_datum(f, pts);
        next_datum(f, p.t_series);
        if (!file_exists(dir + pts)) {
            std::cerr << "Error: File '" << dir + pts << "' not found" << std::endl;
            return 1;
        }
    }
    double stability = 0.0;
    next_datum(f, p.cfl);
    next_datum(f, p.eps_h);
    next_datum(f, p.hpos);
    next_datum(f, stability);
    next_datum(f, p.mf0);
    next_datum(f, p.vmax);
    next_datum(f, p.L);
    next_datum(f, p.H);
    p.dif_at = p.eps_h;
    next_datum(f, p.dif_at);
    // non-dimensionalisation (src/Problema.cxx:327-343)
    p.dif_at /= p.H;
    p.Q = std::sqrt(9.81 * std::pow(p.H, 3.0));
    p.T = p.L * p.H / p.Q;
    p.t_end /= p.T;
    p.t_save /= p.T;
    p.t_series /= p.T;
    p.mf0 *= 9.81 * p.mf0 * p.L / std::pow(p.H, 4.0 / 3.0);
    p.vmax /= p.Q / p.H;
    p.eps_h /= p.H;
    p.hpos /= p.H;
    p.cvis = 1.0 - stability;
    const double R = EARTH_RADIUS / p.L;
    if (p.crop_flag == CROP_ABSOLUTE) p.crop_value /= p.H;
    p.sp_izq = std::fabs(p.borde_izq - 1.0) < DBL_EPSILON ? SPONGE_SIZE : 0;
    p.sp_der = std::fabs(p.borde_der - 1.0) < DBL_EPSILON ? SPONGE_SIZE : 0;
    p.sp_sup = std::fabs(p.borde_sup - 1.0) < DBL_EPSILON ? SPONGE_SIZE : 0;
    p.sp_inf = std::fabs(p.borde_inf - 1.0) < DBL_EPSILON ? SPONGE_SIZE : 0;

    const size_t n = (size_t)p.nx * p.ny;
    p.state1.assign(2 * n, 0.0);
    p.state2.assign(2 * n, 0.0);
    p.hmin = 1e30;
    for (size_t c = 0; c < n; c++) {
        double v = -1.0 * (double)z[c];
        v /= p.H;
        p.state1[2 * c + 1] = v;
        if (p.okada_flag == OKADA_STANDARD) p.state1[2 * c] = v > 0.0 ? v : 0.0;
        if (v < p.hmin) p.hmin = v;
    }
    // gauges (src/Problema.cxx:400-429)
    if (p.read_points == 1) {
        std::ifstream fp((dir + pts).c_str());
        int np = 0;
        fp >> np;
        for (int i = 0; i < np; i++) {
            double lo, la;
            fp >> lo >> la;
            bool dup = false;
            for (size_t q = 0; q < p.plon.size(); q++)
                if (std::fabs(p.plon[q] - lo) < 1e-6 && std::fabs(p.plat[q] - la) < 1e-6) dup = true;
            if (!dup) {
                p.plon.push_back(lo);
                p.plat.push_back(la);
                p.ppos.push_back(nearest_cell(p.lon, p.lat, lo, la));
            }
        }
    }
    // geometry (src/Problema.cxx:431-448)
    const double g2r = M_PI / 180.0;
    p.ancho.assign(p.ny + 1, 0.0);
    p.alto.assign(p.ny, 0.0);
    p.area.assign(p.ny, 0.0);
    p.cosphi.assign(p.ny, 0.0);
    const double dlon = std::fabs(p.lon[1] - p.lon[0]);
    double dlat = 0.0;
    for (int i = 0; i < p.ny - 1; i++) {
        dlat = std::fabs(p.lat[i + 1] - p.lat[i]);
        p.ancho[i] = dlon;
        p.alto[i] = dlat;
        p.area[i] = R * dlon * dlat * g2r;
        p.cosphi[i] = std::cos(p.lat[i] * g2r);
    }
    p.ancho[p.ny - 1] = dlon;
    p.alto[p.ny - 1] = p.alto[p.ny - 2];
    p.ancho[p.ny] = dlon;
    p.area[p.ny - 1] = R * dlon * dlat * g2r;
    p.cosphi[p.ny - 1] = std::cos(p.lat[p.ny - 1] * g2r);
    // Hmin shift (src/Problema.cxx:450-456)
    if (p.hmin < 0.0) {
        for (size_t c = 0; c < n; c++) p.state1[2 * c + 1] -= p.hmin;
    } else {
        p.hmin = 0.0;
    }
    // initial sea surface from a file (src/Problema.cxx:458-473)
    if (p.okada_flag == SEA_SURFACE_FROM_FILE) {
        std::vector<double> elon, elat;
        std::vector<float> eta;
        if (!read_grid(dir + est, elon, elat, eta, err) || eta.size() != n) {
            std::cerr << "Error: " << (err.empty() ? "initial state grid does not match the bathymetry" : err)
                      << std::endl;
            return 1;
        }
        for (size_t c = 0; c < n; c++) {
            double v = (double)eta[c] / p.H + p.state1[2 * c + 1] + p.hmin;
            p.state1[2 * c] = std::max(v, 0.0);
        }
    }
    return 0;
}

// src/Problema.cxx:479-533
void print_problem(const Problem &p, const std::string &version) {
    std::cout << "/**********************************************************************" << std::endl;
    std::cout << " Tsunami-HySEA numerical model open source v" << version << std::endl;
    std::cout << " MI355X (gfx950) time-stepping path                                    " << std::endl;
    std::cout << "**********************************************************************/" << std::endl;
    std::cout << std::endl;
    std::cout << "Problem data" << std::endl;
    if (p.okada_flag == SEA_SURFACE_FROM_FILE) {
        std::cout << "Initialization: Sea surface displacement from file" << std::endl;
    } else {
        std::cout << "Initialization: Standard Okada" << std::endl;
        std::cout << "Number of faults: " << p.faults.size() << std::endl;
        if (p.crop_flag == NO_CROP) std::cout << "Crop deformations: No" << std::endl;
        else if (p.crop_flag == CROP_RELATIVE)
            std::cout << "Crop deformations: Yes (relative percentage [0,1]: " << p.crop_value << ")" << std::endl;
        else
            std::cout << "Crop deformations: Yes (absolute threshold: " << p.crop_value * p.H << " m)" << std::endl;
    }
    std::cout << "CFL: " << p.cfl << std::endl;
    std::cout << "Water-bottom friction: " << std::sqrt(p.mf0 * std::pow(p.H, 4.0 / 3.0) / (9.81 * p.L)) << std::endl;
    std::cout << "Maximum allowed velocity of water: " << p.vmax * (p.Q / p.H) << std::endl;
    std::cout << "Epsilon h: " << p.eps_h * p.H << " m" << std::endl;
    std::cout << "Threshold for the 2s+WAF scheme: " << p.hpos * p.H << " m" << std::endl;
    std::cout << "Stability coefficient: " << 1.0 - p.cvis << std::endl;
    std::cout << "Threshold for the arrival times: " << p.dif_at * p.H << " m" << std::endl;
    std::cout << "Simulation time: " << p.t_end * p.T << " sec" << std::endl;
    std::cout << "Saving time of NetCDF files: " << p.t_save * p.T << " sec" << std::endl;
    if (p.read_points == 1) std::cout << "Time series: yes (saving time: " << p.t_series * p.T << " sec)" << std::endl;
    else std::cout << "Time series: no" << std::endl;
    std::cout << "Number of levels: 1" << std::endl;
    std::cout << "Level 0" << std::endl;
    std::cout << "  Volumes: " << p.nx << " x " << p.ny << " = " << p.nx * p.ny << std::endl;
    std::cout << "  Longitude: [" << p.lon[0] << ", " << p.lon[p.nx - 1] << "]" << std::endl;
    std::cout << "  Latitude: [" << p.lat[0] << ", " << p.lat[p.ny - 1] << "]" << std::endl;
    std::cout << std::endl;
}

}  // namespace thyhost
