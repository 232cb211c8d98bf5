// main.cpp -- TsunamiHySEA <dataFile>: the reference's command line
// (src/ShallowWater.cxx:29-166) on the MI355X library.
//
// Reads the reference's parameter file and NetCDF grid and, as the reference's main,
// calls shallowWater (include/thysea_sw.h, libthysea_sw.so), which runs the time loop
// on one gfx950 GPU through the C-ABI (include/thysea.h) and writes <prefix>.nc (when
// NetCDF saving is enabled) and <prefix>_ts.nc (when a points file is given).
#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/thysea.h"
#include "ncclassic.h"
#include "outputs.h"
#include "problem.h"
#include "../../include/thysea_sw.h"

using namespace thyhost;

namespace {

const std::string kVersion = "1.3.0";

void usage(const char *prog) {
    std::cerr << "/**********************************************************************" << std::endl;
    std::cerr << " Tsunami-HySEA numerical model open source v" << kVersion << std::endl;
    std::cerr << " MI355X (gfx950) time-stepping path                                    " << std::endl;
    std::cerr << "**********************************************************************/" << std::endl << std::endl;
    std::cerr << "Use: " << std::endl << prog << " dataFile" << std::endl << std::endl;
    std::cerr << "dataFile format:" << std::endl;
    std::cerr << "  Problem name" << std::endl;
    std::cerr << "  Bathymetry file" << std::endl;
    std::cerr << "  Initialization of states (" << SEA_SURFACE_FROM_FILE << ": Sea surface displacement from file,"
              << std::endl;
    std::cerr << "                            " << OKADA_STANDARD << ": Standard Okada)" << std::endl;
    std::cerr << "  If " << SEA_SURFACE_FROM_FILE << ":" << std::endl;
    std::cerr << "    Initial state file" << std::endl;
    std::cerr << "  Else if " << OKADA_STANDARD << ":" << std::endl;
    std::cerr << "    Number of faults (>= 1)" << std::endl;
    std::cerr << "    For every fault, a line containing:" << std::endl;
    std::cerr << "      Time(sec) Lon_epicenter Lat_epicenter Depth_hypocenter(km) Fault_length(km) Fault_width(km) "
                 "Strike Dip Rake Slip(m)"
              << std::endl;
    std::cerr << "    Crop deformations (" << NO_CROP << ": No, " << CROP_RELATIVE << ": Relative cropping, "
              << CROP_ABSOLUTE << ": Absolute cropping)" << std::endl;
    std::cerr << "    If " << CROP_RELATIVE << ":" << std::endl << "      Percentage [0,1]" << std::endl;
    std::cerr << "    Else if " << CROP_ABSOLUTE << ":" << std::endl << "      Threshold value (m)" << std::endl;
    std::cerr << "  NetCDF file prefix" << std::endl;
    std::cerr << "  Number of levels (should be 1)" << std::endl;
    std::cerr << "  Upper border condition (1: open, -1: wall)" << std::endl;
    std::cerr << "  Lower border condition (1: open, -1: wall)" << std::endl;
    std::cerr << "  Left border condition (1: open, -1: wall, 2: periodic)" << std::endl;
    std::cerr << "  Right border condition (1: open, -1: wall, 2: periodic)" << std::endl;
    std::cerr << "  Simulation time (sec)" << std::endl;
    std::cerr << "  Saving time of NetCDF files (sec) (-1: do not save)" << std::endl;
    std::cerr << "  Read points from file (0: no, 1: yes)" << std::endl;
    std::cerr << "  If 1:" << std::endl << "    File with points" << std::endl;
    std::cerr << "    Saving time of time series (sec)" << std::endl;
    std::cerr << "  CFL" << std::endl << "  Epsilon h (m)" << std::endl;
    std::cerr << "  Threshold for the 2s+WAF scheme (m)" << std::endl << "  Stability coefficient" << std::endl;
    std::cerr << "  Water-bottom friction (Manning coefficient)" << std::endl;
    std::cerr << "  Maximum allowed velocity of water" << std::endl;
    std::cerr << "  L (typical length)" << std::endl << "  H (typical depth)" << std::endl;
    std::cerr << "  Threshold for arrival times (m) (epsilon h if not specified)" << std::endl;
}

}  // namespace

// --check <dataFile> <dump.nc>: parse and set up the problem without a GPU, write the
// non-dimensional inputs to <dump.nc> and the output files of a run with no steps.
int check_mode(const std::string &fich, const std::string &dump) {
    Problem p;
    if (load_problem(fich, p) != 0) return EXIT_FAILURE;
    std::string err;
    nc::Writer w;
    const size_t n = (size_t)p.nx * p.ny;
    if (!w.create(dump, 2, &err)) { std::cerr << err << std::endl; return EXIT_FAILURE; }
    const int dx = w.def_dim("x", p.nx), dy = w.def_dim("y", p.ny), dy1 = w.def_dim("y1", p.ny + 1);
    const int dp = w.def_dim("points", std::max<size_t>(1, p.ppos.size())), df = w.def_dim("faults", std::max<size_t>(1, p.faults.size()));
    const int dk = w.def_dim("ten", 10);
    const int vh = w.def_var("h", nc::NC_DOUBLE, {dy, dx}), vH = w.def_var("H", nc::NC_DOUBLE, {dy, dx});
    const int va = w.def_var("area", nc::NC_DOUBLE, {dy}), vc = w.def_var("cosphi", nc::NC_DOUBLE, {dy});
    const int van = w.def_var("ancho", nc::NC_DOUBLE, {dy1}), val = w.def_var("alto", nc::NC_DOUBLE, {dy});
    const int vp = w.def_var("gauge_pos", nc::NC_INT, {dp}), vf = w.def_var("faults", nc::NC_DOUBLE, {df, dk});
    const double scal[] = {p.hmin, p.cfl, p.mf0, p.vmax, p.eps_h, p.hpos, p.cvis, p.dif_at, p.t_end, p.t_save,
                           p.t_series, p.L, p.H, p.Q, p.T, p.borde_sup, p.borde_inf, p.borde_izq, p.borde_der,
                           (double)p.sp_sup, (double)p.sp_inf, (double)p.sp_izq, (double)p.sp_der,
                           (double)p.okada_flag, (double)p.crop_flag, p.crop_value};
    const char *names[] = {"hmin", "cfl", "mf0", "vmax", "eps_h", "hpos", "cvis", "dif_at", "t_end", "t_save",
                           "t_series", "L", "H_ref", "Q", "T", "borde_sup", "borde_inf", "borde_izq", "borde_der",
                           "sponge_sup", "sponge_inf", "sponge_izq", "sponge_der", "okada_flag", "crop_flag",
                           "crop_value"};
    for (size_t k = 0; k < sizeof(scal) / sizeof(scal[0]); k++) w.put_att(-1, nc::Att::f64(names[k], scal[k]));
    w.put_att(-1, nc::Att::text("name", p.name));
    w.put_att(-1, nc::Att::text("prefix", p.prefix));
    if (!w.enddef(&err)) { std::cerr << err << std::endl; return EXIT_FAILURE; }
    std::vector<double> h(n), H(n);
    for (size_t c = 0; c < n; c++) { h[c] = p.state1[2 * c]; H[c] = p.state1[2 * c + 1]; }
    std::vector<int> pos(p.ppos);
    if (pos.empty()) pos.push_back(-2);
    std::vector<double> fl;
    for (const thy_fault &f : p.faults)
        for (double v : {f.time, f.lon_c, f.lat_c, f.depth_c, f.length, f.width, f.strike, f.dip, f.rake, f.slip})
            fl.push_back(v);
    if (fl.empty()) fl.assign(10, 0.0);
    bool ok = w.put_var(vh, h.data(), &err) && w.put_var(vH, H.data(), &err) && w.put_var(va, p.area.data(), &err) &&
              w.put_var(vc, p.cosphi.data(), &err) && w.put_var(van, p.ancho.data(), &err) &&
              w.put_var(val, p.alto.data(), &err) && w.put_var(vp, pos.data(), &err) && w.put_var(vf, fl.data(), &err) &&
              w.close(&err);
    if (!ok) { std::cerr << err << std::endl; return EXIT_FAILURE; }
    // outputs of a run with no steps: bathymetry and fill values
    if (p.t_save >= 0.0) {
        GridOutput g;
        std::vector<float> z(n);
        for (size_t c = 0; c < n; c++) z[c] = (float)((p.state1[2 * c + 1] + p.hmin) * p.H);
        std::vector<float> fill(n, -9999.0f), arr(n, -1.0f);
        ok = g.create(p.prefix + ".nc", p, kVersion, &err) && g.write_snapshot(0, 0.0, fill.data(), fill.data(), fill.data(), &err) &&
             g.write_final(z.data(), fill.data(), arr.data(), &err) && g.close(&err);
        if (!ok) { std::cerr << err << std::endl; return EXIT_FAILURE; }
    }
    if (p.read_points == 1) {
        SeriesOutput s;
        const size_t ng = p.ppos.size();
        std::vector<float> t1(1, 0.0f), e(ng, 0.0f), b(ng, 0.0f);
        ok = s.create(p.prefix + "_ts.nc", p, kVersion, &err) && s.write_samples(1, t1.data(), e.data(), e.data(), e.data(), &err) &&
             s.write_final(b.data(), e.data(), e.data(), &err) && s.close(&err);
        if (!ok) { std::cerr << err << std::endl; return EXIT_FAILURE; }
    }
    return EXIT_SUCCESS;
}

int main(int argc, char *argv[]) {
    if (argc < 2) {
        usage(argv[0]);
        return EXIT_FAILURE;
    }
    if (std::string(argv[1]) == "--check") {
        if (argc < 4) {
            std::cerr << "Use: " << argv[0] << " --check dataFile dump.nc" << std::endl;
            return EXIT_FAILURE;
        }
        return check_mode(argv[2], argv[3]);
    }
    const std::string fich = argv[1];
    struct stat st;
    if (stat(fich.c_str(), &st) != 0) {
        std::cerr << "Error: File '" << fich << "' not found" << std::endl;
        return EXIT_FAILURE;
    }
    const int probe = thy_probe();
    if (probe == 1) {
        std::cerr << "Error: There is no graphics card" << std::endl;
        return EXIT_FAILURE;
    } else if (probe == 2) {
        std::cerr << "Error: There is no graphics card supporting gfx950" << std::endl;
        return EXIT_FAILURE;
    }
    Problem p;
    if (load_problem(fich, p) != 0) return EXIT_FAILURE;
    print_problem(p, kVersion);
    std::cout << std::scientific;

    // the reference's call into its GPU part (src/ShallowWater.cxx:146-151), with its
    // arrays built as cargarDatosProblema leaves them
    const int nf = (int)p.faults.size();
    std::vector<double> fv[10];
    for (const thy_fault &f : p.faults) {
        const double v[10] = {f.lon_c, f.lat_c, f.depth_c, f.length, f.width, f.strike, f.dip, f.rake, f.slip, f.time};
        for (int k = 0; k < 10; k++) fv[k].push_back(v[k]);
    }
    for (auto &a : fv) a.resize(std::max<size_t>(1, a.size()));
    std::vector<thy_double2> ayc(p.ny);
    for (int j = 0; j < p.ny; j++) ayc[j] = thy_double2{p.area[j], p.cosphi[j]};
    std::vector<double> ancho(p.ancho), alto(p.alto), lon(p.lon), lat(p.lat);
    thy_submalla nivel{ayc.data(), ancho.data(), alto.data(), lon.data(), lat.data()};
    std::vector<int> pos(p.ppos);
    std::vector<double> plon(p.plon), plat(p.plat);
    std::string name = p.name, version = kVersion;
    double tiempo = 0.0;
    const int err = shallowWater(1, p.okada_flag, nf, p.crop_flag, p.crop_value, fv[0].data(), fv[1].data(),
                                 fv[2].data(), fv[3].data(), fv[4].data(), fv[5].data(), fv[6].data(), fv[7].data(),
                                 fv[8].data(), fv[9].data(), reinterpret_cast<thy_double2 *>(p.state1.data()),
                                 reinterpret_cast<thy_double2 *>(p.state2.data()), nivel, p.read_points,
                                 (int)pos.size(), pos.data(), plon.data(), plat.data(), p.nx, p.ny,
                                 (int64_t)p.nx * p.ny, p.hmin, &name[0], p.prefix, p.borde_sup, p.borde_inf,
                                 p.borde_izq, p.borde_der, p.sp_sup, p.sp_inf, p.sp_izq, p.sp_der, p.t_end, p.t_save,
                                 p.t_series, p.cfl, p.mf0, p.vmax, p.eps_h, p.hpos, p.cvis, p.dif_at, p.L, p.H, p.Q,
                                 p.T, &version[0], &tiempo);
    if (err != THY_OK) {
        if (err == THY_ERR_DEVICE_MEMORY) std::cerr << "Error: Not enough GPU memory" << std::endl;
        else if (err == THY_ERR_HOST_MEMORY) std::cerr << "Error: Not enough CPU memory" << std::endl;
        return EXIT_FAILURE;
    }
    std::cout << std::endl << "Runtime: " << tiempo << " sec" << std::endl;
    return EXIT_SUCCESS;
}
