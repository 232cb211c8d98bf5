// outputs.h -- the reference's NetCDF outputs, written in the classic format:
//   <prefix>.nc    grid, bathymetry, snapshots of eta/ux/uy, max_height, arrival_times
//                  (src/GPU/netcdfTH.cu:118-544)
//   <prefix>_ts.nc gauge time series (src/GPU/netcdfSeries.cu:51-373)
#pragma once
#include <string>
#include <vector>

#include "ncclassic.h"
#include "problem.h"

namespace thyhost {

class GridOutput {
  public:
    // creates the file and writes lon/lat/grid_lon/grid_lat and the original bathymetry
    bool create(const std::string &path, const Problem &p, const std::string &version, std::string *err);
    bool write_snapshot(int index, double t_seconds, const float *eta, const float *ux, const float *uy,
                        std::string *err);
    bool write_final(const float *deformed_bathy, const float *max_height, const float *arrival, std::string *err);
    bool close(std::string *err);

  private:
    nc::Writer w_;
    int v_time_ = -1, v_eta_ = -1, v_ux_ = -1, v_uy_ = -1, v_maxh_ = -1, v_arr_ = -1, v_def_ = -1;
    bool okada_ = true;
};

class SeriesOutput {
  public:
    bool create(const std::string &path, const Problem &p, const std::string &version, std::string *err);
    bool write_samples(int n, const float *time, const float *eta, const float *u, const float *v, std::string *err);
    bool write_final(const float *bathy, const float *eta_min, const float *eta_max, std::string *err);
    bool close(std::string *err);

  private:
    nc::Writer w_;
    int np_ = 0;
    int v_time_ = -1, v_eta_ = -1, v_u_ = -1, v_v_ = -1, v_min_ = -1, v_max_ = -1, v_bat_ = -1;
};

}  // namespace thyhost
