/*
 * thy_oracle.h -- FP64 CPU restatement of the Tsunami-HySEA v1.3.0 time-stepping
 * path (edge fluxes, cell updates, CFL reduction, metrics, Okada initial condition,
 * output conversions and the driver loop).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker and the CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it; the product path (libthysea.so and the TsunamiHySEA executable)
 * never links or calls it.
 *
 * Every function cites the reference file:line whose semantics it restates.
 * Data model follows the reference: per cell (h, H) and (qx, qy) in double,
 * non-dimensional (src/Problema.cxx:327-343), row-major pos = j*nx + i
 * (src/Problema.cxx:386), per-row geometry (src/Problema.cxx:431-448).
 *
 * Parity pinning (see tests/test_oracle_*.py): lake at rest, mass conservation
 * with walls, mirror symmetry, closed-form initial dt, brute-force min,
 * scalar closed forms for desingularisation / friction / filter / limiter,
 * DC3D at z=0 against an independent Okada (1985) surface formula, and the
 * Gauss-Krueger forward/inverse round trip.
 */
#ifndef THY_ORACLE_H
#define THY_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    int nx, ny;
    /* border codes as in the parameter file: 1 open, -1 wall, 2 periodic
       (src/Problema.cxx:280-305); sup = row 0 side, inf = row ny side,
       izq = column 0 side, der = column nx side. */
    double borde_sup, borde_inf, borde_izq, borde_der;
    int sponge_sup, sponge_inf, sponge_izq, sponge_der; /* src/Problema.cxx:346-349 */
    double cfl, mf0, vmax, eps_h, hpos, cvis, dif_at;   /* non-dimensional */
    double hmin;       /* Hmin shift (src/Problema.cxx:450-456) */
    double sea_level;  /* SEA_LEVEL/H (src/GPU/ShallowWater.cu:205) */
    int quirk_reduction; /* 1: reproduce the reference's partial min on small grids */
} orc_params;

typedef struct orc_sim orc_sim;

/* geometry: area[ny], cosphi[ny], ancho[ny+1] (dlon), alto[ny] (dlat), degrees
   except area (src/Problema.cxx:431-448). */
orc_sim *orc_create(const orc_params *p, const double *area, const double *cosphi,
                    const double *ancho, const double *alto);
void orc_destroy(orc_sim *s);
void orc_set_threads(int n);

/* state1 = (h, H), state2 = (qx, qy), SoA views of the reference's double2 arrays */
void orc_set_state(orc_sim *s, const double *h, const double *H, const double *qx, const double *qy);
void orc_get_state(const orc_sim *s, double *h, double *H, double *qx, double *qy);

/* src/GPU/ShallowWater.cu:305-309 (eta_max, eta_ini, arrival times) */
void orc_init_metrics(orc_sim *s);
/* src/GPU/Metrics.cu:23-42 and :57-77 */
void orc_update_metrics(orc_sim *s, double t);
void orc_get_metrics(const orc_sim *s, double *eta_max, double *t_arr, double *eta_ini);

/* src/GPU/Timestep.cu:11-33: initial dt (min over cells, quirk-aware) */
double orc_initial_dt(orc_sim *s);
/* per-cell dt field from the last reduction input (for brute-force min tests) */
void orc_get_cell_dt(const orc_sim *s, double *dt_cell);
/* src/GPU/Timestep.cu:70-120: one step (mass + momentum); returns next raw dt */
double orc_step(orc_sim *s, double dt);
/* Split halves of one step, for per-pass checks: mass pass writes h* into state1;
   momentum pass writes q and the per-cell dt and returns the reduced next dt. */
void orc_mass_pass(orc_sim *s, double dt);
double orc_momentum_pass(orc_sim *s, double dt);

/* min-reduction semantics of src/GPU/Reduccion_kernel.cu:130-173 over n values.
   quirk=0: exact global minimum.  quirk=1: the reference's launch arithmetic, with
   stages whose thread count matches no switch case skipped (race-free reading). */
double orc_reduce_min(const double *v, int64_t n, int quirk);
float orc_reduce_min_f(const float *v, int64_t n, int quirk);

/* Okada standard source, src/GPU/Deformacion.cu:118-143.  lon0/dlon/lat0/dlat give
   the uniform coordinates lon0 + i*dlon (src/GPU/ShallowWater.cu:215-218).
   Writes U_Z/H into def_out[N] when non-NULL (before crop). */
void orc_okada(orc_sim *s, double lon0, double dlon, double lat0, double dlat,
               double lon_c, double lat_c, double depth_c, double length, double width,
               double strike, double dip, double rake, double slip,
               int crop_flag, double crop_value, double H, double *def_out);

/* Scalar pieces for closed-form tests */
int orc_prtoxy(double *lat, double *lon, double lat0, double lon0, double *x, double *y, int ind);
void orc_dc3d_surface(double alpha, double x, double y, double depth, double dip,
                      double al1, double al2, double aw1, double aw2,
                      double disl1, double disl2, double disl3,
                      double *ux, double *uy, double *uz, int *iret);
double orc_limiter(double p0m, double p0, double p1, double p1p, double S);
double orc_desing(double h, double q, double eps_h);
void orc_friction_filter(double h_star, double qx_old, double qy_old, double *qx, double *qy,
                         double dt, double mf0, double vmax, double eps_h);

/* output conversions, src/GPU/netcdf_kernel.cu:6-126.
   which: 0 eta, 1 ux, 2 uy, 3 deformed bathymetry, 4 max height, 5 arrival times */
void orc_output_field(const orc_sim *s, int which, double H, double Q, double T, float *out);

/* gauge gather + conversion, src/GPU/Volumen_kernel.cu:223-254 and
   src/GPU/ShallowWater.cu:69-90: eta/u/v as float for each position (-1 -> -9999) */
void orc_gauges(const orc_sim *s, const int *pos, int n, double H, double Q,
                float *eta, float *u, float *v);

/* Driver loop of src/GPU/ShallowWater.cu:374-453 with the truncation rules of
   src/GPU/Timestep.cu:35-68.  Times in the run config are non-dimensional except
   fault times (seconds, as defTime).  Faults must be sorted by time. */
typedef struct orc_fault {
    double time, lon_c, lat_c, depth_c, length, width, strike, dip, rake, slip;
} orc_fault;

typedef void (*orc_snapshot_cb)(void *user, int index, double t);

typedef struct orc_run_cfg {
    double t_end, save_interval, series_interval; /* non-dimensional; <0 disables */
    double T, H, Q;                                /* scales */
    int n_faults; const orc_fault *faults;
    int crop_flag; double crop_value;
    double lon0, dlon, lat0, dlat;                 /* uniform Okada coordinates */
    int n_gauges; const int *gauge_pos;
    int max_samples;                               /* capacity of the sample buffers */
    float *g_time, *g_eta, *g_u, *g_v;             /* [max_samples], [max_samples*n_gauges]... */
    float *g_eta_min, *g_eta_max;                  /* [n_gauges] running min/max */
    int64_t max_steps;                             /* 0: unlimited */
    orc_snapshot_cb on_snapshot; void *user;
} orc_run_cfg;

typedef struct orc_run_out {
    int64_t steps; int n_samples; int n_snapshots;
    double t_final, dt_initial, dt_last;
} orc_run_out;

int orc_run(orc_sim *s, const orc_run_cfg *c, orc_run_out *o);

/* read-only geometry/state access used by tests */
double orc_mass(const orc_sim *s);   /* sum h*area*cos(phi) (compensated) */
double orc_time_of(const orc_sim *s);

#ifdef __cplusplus
}
#endif
#endif
