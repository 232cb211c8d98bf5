// outputs.cpp -- NetCDF outputs with the reference's names, dimensions, attributes
// and fill values (src/GPU/netcdfTH.cu:93-544, src/GPU/netcdfSeries.cu:26-373).
#include "outputs.h"

#include <sys/time.h>

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <ctime>

namespace thyhost {

using nc::Att;

namespace {

std::string now_string() {
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    char buf[24];
    strftime(buf, sizeof(buf), "%Y-%m-%d %H:%M:%S", localtime(&tv.tv_sec));
    return buf;
}

std::string border_name(double b) { return std::fabs(b - 1.0) < DBL_EPSILON ? "open" : "wall"; }

// global attributes shared by both files (src/GPU/netcdfTH.cu:319-412)
void run_attributes(nc::Writer &w, const Problem &p, const std::string &title, const std::string &version,
                    bool grid_interval) {
    w.put_att(-1, Att::text("Conventions", "CF-1.0"));
    w.put_att(-1, Att::text("title", title));
    w.put_att(-1, Att::text("Tsunami-HySEA_miniapp_version", version));
    w.put_att(-1, Att::text("creator_name", "EDANYA Group"));
    w.put_att(-1, Att::text("institution", "University of Malaga"));
    w.put_att(-1, Att::text("comments", " "));
    w.put_att(-1, Att::text("references", " "));
    w.put_att(-1, Att::text("history", now_string()));
    w.put_att(-1, Att::text("problem_name", p.name));
    w.put_att(-1, Att::f32("simulation_time", (float)(p.t_end * p.T)));
    if (grid_interval) w.put_att(-1, Att::i32("output_grid_interval", 1));
    w.put_att(-1, Att::text("upper_border", border_name(p.borde_sup)));
    w.put_att(-1, Att::text("lower_border", border_name(p.borde_inf)));
    if (p.periodic()) {
        w.put_att(-1, Att::text("left_border", "periodic"));
        w.put_att(-1, Att::text("right_border", "periodic"));
    } else {
        w.put_att(-1, Att::text("left_border", border_name(p.borde_izq)));
        w.put_att(-1, Att::text("right_border", border_name(p.borde_der)));
    }
    const double mf0_ini = std::sqrt(p.mf0 * std::pow(p.H, 4.0 / 3.0) / (9.81 * p.L));
    w.put_att(-1, Att::f32("CFL", (float)p.cfl));
    w.put_att(-1, Att::f32("epsilon_h", (float)(p.eps_h * p.H)));
    w.put_att(-1, Att::f32("water_bottom_friction", (float)mf0_ini));
    w.put_att(-1, Att::f32("max_velocity_water", (float)(p.vmax * p.Q / p.H)));
    w.put_att(-1, Att::f32("threshold_2swaf", (float)(p.hpos * p.H)));
    w.put_att(-1, Att::f32("stability_coefficient", (float)(1.0 - p.cvis)));
    w.put_att(-1, Att::f32("threshold_arrival_times", (float)(p.dif_at * p.H)));
    if (p.okada_flag == SEA_SURFACE_FROM_FILE) {
        w.put_att(-1, Att::text("initialization_mode", "sea_surface_from_file"));
    } else {
        w.put_att(-1, Att::text("initialization_mode", "okada_standard"));
        w.put_att(-1, Att::i32("num_faults", (int)p.faults.size()));
        for (size_t i = 0; i < p.faults.size(); i++) {
            const thy_fault &f = p.faults[i];
            char name[64], val[512];
            snprintf(name, sizeof(name), "fault_%d", (int)i + 1);
            snprintf(val, sizeof(val),
                     "time: %.4f, lon: %.4f, lat: %.4f, depth: %.4f, length: %.4f, width: %.4f, strike: %.4f, dip: "
                     "%.4f, rake: %.4f, slip: %.4f",
                     f.time, f.lon_c, f.lat_c, f.depth_c, f.length, f.width, f.strike, f.dip, f.rake, f.slip);
            w.put_att(-1, Att::text(name, val));
        }
        if (p.crop_flag == CROP_RELATIVE) {
            w.put_att(-1, Att::text("crop_deformations", "relative"));
            w.put_att(-1, Att::f32("crop_percentage_0-1", (float)p.crop_value));
        } else if (p.crop_flag == CROP_ABSOLUTE) {
            w.put_att(-1, Att::text("crop_deformations", "absolute"));
            w.put_att(-1, Att::f32("crop_threshold", (float)(p.crop_value * p.H)));
        } else {
            w.put_att(-1, Att::text("crop_deformations", "no"));
        }
    }
    w.put_att(-1, Att::text("implementation", "thysea, MI355X (gfx950) HIP time-stepping path, classic NetCDF"));
}

void fill_atts(nc::Writer &w, int v) {
    w.put_att(v, Att::f32("missing_value", -9999.0f));
    w.put_att(v, Att::f32("_FillValue", -9999.0f));
}

}  // namespace

bool GridOutput::create(const std::string &path, const Problem &p, const std::string &version, std::string *err) {
    okada_ = p.okada_flag == OKADA_STANDARD;
    const uint64_t field = (uint64_t)p.nx * p.ny * 4;
    if (!w_.create(path, field * 3 > 0xFFFFFFF0ull ? 5 : 2, err)) return false;
    const int d_lon = w_.def_dim("lon", p.nx), d_lat = w_.def_dim("lat", p.ny);
    const int d_glon = w_.def_dim("grid_lon", p.nx), d_glat = w_.def_dim("grid_lat", p.ny);
    const int d_time = w_.def_dim("time", 0);
    const int v_lon = w_.def_var("lon", nc::NC_DOUBLE, {d_lon});
    const int v_lat = w_.def_var("lat", nc::NC_DOUBLE, {d_lat});
    const int v_glon = w_.def_var("grid_lon", nc::NC_DOUBLE, {d_glon});
    const int v_glat = w_.def_var("grid_lat", nc::NC_DOUBLE, {d_glat});
    int v_bat;
    if (okada_) {
        v_bat = w_.def_var("original_bathy", nc::NC_FLOAT, {d_glat, d_glon});
        v_def_ = w_.def_var("deformed_bathy", nc::NC_FLOAT, {d_glat, d_glon});
    } else {
        v_bat = w_.def_var("bathymetry", nc::NC_FLOAT, {d_glat, d_glon});
    }
    v_time_ = w_.def_var("time", nc::NC_FLOAT, {d_time});
    v_maxh_ = w_.def_var("max_height", nc::NC_FLOAT, {d_lat, d_lon});
    v_arr_ = w_.def_var("arrival_times", nc::NC_FLOAT, {d_lat, d_lon});
    v_eta_ = w_.def_var("eta", nc::NC_FLOAT, {d_time, d_lat, d_lon});
    w_.put_att(v_eta_, Att::text("units", "meters"));
    w_.put_att(v_eta_, Att::text("long_name", "Wave amplitude"));
    v_ux_ = w_.def_var("ux", nc::NC_FLOAT, {d_time, d_lat, d_lon});
    w_.put_att(v_ux_, Att::text("units", "meters/second"));
    w_.put_att(v_ux_, Att::text("long_name", "Velocity of water along longitude"));
    v_uy_ = w_.def_var("uy", nc::NC_FLOAT, {d_time, d_lat, d_lon});
    w_.put_att(v_uy_, Att::text("units", "meters/second"));
    w_.put_att(v_uy_, Att::text("long_name", "Velocity of water along latitude"));
    if (okada_) {
        w_.put_att(v_bat, Att::text("long_name", "Grid original bathymetry"));
        w_.put_att(v_bat, Att::text("standard_name", "original depth"));
        w_.put_att(v_bat, Att::text("units", "meters"));
        fill_atts(w_, v_bat);
        w_.put_att(v_def_, Att::text("long_name", "Grid deformed bathymetry"));
        w_.put_att(v_def_, Att::text("standard_name", "deformed depth"));
        w_.put_att(v_def_, Att::text("units", "meters"));
        fill_atts(w_, v_def_);
    } else {
        w_.put_att(v_bat, Att::text("long_name", "Grid bathymetry"));
        w_.put_att(v_bat, Att::text("standard_name", "depth"));
        w_.put_att(v_bat, Att::text("units", "meters"));
        fill_atts(w_, v_bat);
    }
    w_.put_att(v_glon, Att::text("long_name", "Grid longitude"));
    w_.put_att(v_glon, Att::text("units", "degrees_east"));
    w_.put_att(v_glat, Att::text("long_name", "Grid latitude"));
    w_.put_att(v_glat, Att::text("units", "degrees_north"));
    w_.put_att(v_maxh_, Att::text("long_name", "Maximum wave amplitude"));
    w_.put_att(v_maxh_, Att::text("units", "meters"));
    fill_atts(w_, v_maxh_);
    w_.put_att(v_lon, Att::text("standard_name", "longitude"));
    w_.put_att(v_lon, Att::text("long_name", "longitude"));
    w_.put_att(v_lon, Att::text("units", "degrees_east"));
    w_.put_att(v_lat, Att::text("standard_name", "latitude"));
    w_.put_att(v_lat, Att::text("long_name", "latitude"));
    w_.put_att(v_lat, Att::text("units", "degrees_north"));
    w_.put_att(v_time_, Att::text("long_name", "Time"));
    w_.put_att(v_time_, Att::text("units", "seconds since 1970-01-01"));
    fill_atts(w_, v_eta_);
    fill_atts(w_, v_ux_);
    fill_atts(w_, v_uy_);
    run_attributes(w_, p, "TsunamiHySEA model output", version, true);
    if (!w_.enddef(err)) return false;
    if (!w_.put_var(v_lon, p.lon.data(), err) || !w_.put_var(v_lat, p.lat.data(), err) ||
        !w_.put_var(v_glon, p.lon.data(), err) || !w_.put_var(v_glat, p.lat.data(), err))
        return false;
    // original bathymetry (H + Hmin) * H_ref before any deformation (src/GPU/ShallowWater.cu:321-323)
    std::vector<float> bat((size_t)p.nx * p.ny);
    for (size_t c = 0; c < bat.size(); c++) bat[c] = (float)((p.state1[2 * c + 1] + p.hmin) * p.H);
    if (!w_.put_var(v_bat, bat.data(), err)) return false;
    // fixed-size outputs written at the end start as fill values
    std::vector<float> fill(bat.size(), -9999.0f);
    if (okada_ && !w_.put_var(v_def_, fill.data(), err)) return false;
    return w_.put_var(v_maxh_, fill.data(), err) && w_.put_var(v_arr_, fill.data(), err) && w_.sync(err);
}

bool GridOutput::write_snapshot(int index, double t_seconds, const float *eta, const float *ux, const float *uy,
                                std::string *err) {
    const float t = (float)t_seconds;
    return w_.put_record(v_time_, index, &t, err) && w_.put_record(v_eta_, index, eta, err) &&
           w_.put_record(v_ux_, index, ux, err) && w_.put_record(v_uy_, index, uy, err) && w_.sync(err);
}

bool GridOutput::write_final(const float *deformed_bathy, const float *max_height, const float *arrival,
                             std::string *err) {
    if (okada_ && deformed_bathy && !w_.put_var(v_def_, deformed_bathy, err)) return false;
    return w_.put_var(v_maxh_, max_height, err) && w_.put_var(v_arr_, arrival, err);
}

bool GridOutput::close(std::string *err) { return w_.close(err); }

bool SeriesOutput::create(const std::string &path, const Problem &p, const std::string &version, std::string *err) {
    np_ = (int)p.plon.size();
    if (!w_.create(path, 2, err)) return false;
    const int d_np = w_.def_dim("grid_npoints", np_);
    const int d_time = w_.def_dim("time", 0);
    if (p.okada_flag == OKADA_STANDARD) v_bat_ = w_.def_var("deformed_bathy", nc::NC_FLOAT, {d_np});
    else v_bat_ = w_.def_var("bathymetry", nc::NC_FLOAT, {d_np});
    v_time_ = w_.def_var("time", nc::NC_FLOAT, {d_time});
    const int v_lon = w_.def_var("longitude", nc::NC_DOUBLE, {d_np});
    w_.put_att(v_lon, Att::text("units", "degrees_east"));
    w_.put_att(v_lon, Att::f64("missing_value", -9999.0));
    w_.put_att(v_lon, Att::f64("_FillValue", -9999.0));
    const int v_lat = w_.def_var("latitude", nc::NC_DOUBLE, {d_np});
    w_.put_att(v_lat, Att::text("units", "degrees_north"));
    w_.put_att(v_lat, Att::f64("missing_value", -9999.0));
    w_.put_att(v_lat, Att::f64("_FillValue", -9999.0));
    v_min_ = w_.def_var("min_height", nc::NC_FLOAT, {d_np});
    w_.put_att(v_min_, Att::text("long_name", "Minimum wave amplitude"));
    w_.put_att(v_min_, Att::text("units", "meters"));
    fill_atts(w_, v_min_);
    v_max_ = w_.def_var("max_height", nc::NC_FLOAT, {d_np});
    w_.put_att(v_max_, Att::text("long_name", "Maximum wave amplitude"));
    w_.put_att(v_max_, Att::text("units", "meters"));
    fill_atts(w_, v_max_);
    v_eta_ = w_.def_var("eta", nc::NC_FLOAT, {d_time, d_np});
    w_.put_att(v_eta_, Att::text("units", "meters"));
    w_.put_att(v_eta_, Att::text("long_name", "Wave amplitude"));
    fill_atts(w_, v_eta_);
    v_u_ = w_.def_var("ux", nc::NC_FLOAT, {d_time, d_np});
    w_.put_att(v_u_, Att::text("units", "meters/second"));
    w_.put_att(v_u_, Att::text("long_name", "Velocity of water along longitude"));
    fill_atts(w_, v_u_);
    v_v_ = w_.def_var("uy", nc::NC_FLOAT, {d_time, d_np});
    w_.put_att(v_v_, Att::text("units", "meters/second"));
    w_.put_att(v_v_, Att::text("long_name", "Velocity of water along latitude"));
    fill_atts(w_, v_v_);
    if (p.okada_flag == OKADA_STANDARD) {
        w_.put_att(v_bat_, Att::text("long_name", "Grid deformed bathymetry"));
        w_.put_att(v_bat_, Att::text("standard_name", "deformed depth"));
    } else {
        w_.put_att(v_bat_, Att::text("long_name", "Grid bathymetry"));
        w_.put_att(v_bat_, Att::text("standard_name", "depth"));
    }
    w_.put_att(v_bat_, Att::text("units", "meters"));
    fill_atts(w_, v_bat_);
    w_.put_att(v_time_, Att::text("long_name", "Time"));
    w_.put_att(v_time_, Att::text("units", "seconds since 1970-01-01"));
    run_attributes(w_, p, "Time series output of TsunamiHySEA model", version, false);
    if (!w_.enddef(err)) return false;
    return w_.put_var(v_lon, p.plon.data(), err) && w_.put_var(v_lat, p.plat.data(), err) && w_.sync(err);
}

bool SeriesOutput::write_samples(int n, const float *time, const float *eta, const float *u, const float *v,
                                 std::string *err) {
    for (int s = 0; s < n; s++) {
        if (!w_.put_record(v_time_, s, &time[s], err) || !w_.put_record(v_eta_, s, eta + (size_t)s * np_, err) ||
            !w_.put_record(v_u_, s, u + (size_t)s * np_, err) || !w_.put_record(v_v_, s, v + (size_t)s * np_, err))
            return false;
    }
    return w_.sync(err);
}

bool SeriesOutput::write_final(const float *bathy, const float *eta_min, const float *eta_max, std::string *err) {
    return w_.put_var(v_bat_, bathy, err) && w_.put_var(v_min_, eta_min, err) && w_.put_var(v_max_, eta_max, err);
}

bool SeriesOutput::close(std::string *err) { return w_.close(err); }

}  // namespace thyhost
