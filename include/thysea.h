/*
 * thysea.h -- C-ABI of the MI355X-native Tsunami-HySEA time-stepping library
 * (libthysea.so, hand-written HIP kernels for gfx950).
 *
 * This is the drop-in boundary for the reference's host/GPU split:
 *   reference: extern "C" int shallowWater(...53 args...)
 *              (src/GPU/ShallowWater.cu:182-190, declared src/ShallowWater.cxx:19-27)
 *   reference: int comprobarSoporteCUDA()  (src/GPU/SoporteCUDA.cu:8)
 * The monolithic call is split into plain-C entry points (no C++ or HIP types in the
 * signatures, plain pointers and sizes).  Inputs are the reference's non-dimensional
 * host arrays (src/Problema.cxx:327-473); the caller owns host buffers, the context
 * owns device memory, the host (not this library) owns files.
 *
 * Return codes follow the reference's int convention (0 ok, 1 device memory, 2 host
 * memory; src/GPU/ShallowWater.cu:275,285,319) extended with the codes below.
 *
 * Threads: calls on different contexts and batches may run on different threads at
 * once, also while a batch records its step sequence (every entry point relaxes the
 * calling thread's stream-capture mode for its duration).  One context, or one batch
 * with the contexts it holds, is used from one thread at a time.
 */
#ifndef THYSEA_H
#define THYSEA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum thy_status {
    THY_OK = 0,
    THY_ERR_DEVICE_MEMORY = 1, /* as shallowWater's return 1 */
    THY_ERR_HOST_MEMORY = 2,   /* as shallowWater's return 2 */
    THY_ERR_INVALID = 3,       /* bad argument or call order */
    THY_ERR_NO_DEVICE = 4,     /* no usable gfx950 device */
    THY_ERR_HIP = 5,           /* a HIP runtime or launch error */
    THY_ERR_BUDGET = 6,        /* batch admission refused by the memory budget */
    THY_ERR_NONFINITE = 7      /* a health check found NaN/Inf in the state */
};

typedef struct thy_ctx thy_ctx;

/* Grid geometry, as tipoDatosSubmalla (src/Constantes.hxx:32-38) built by
   src/Problema.cxx:431-448.  Row j is the j-th latitude of the input file. */
typedef struct thy_grid {
    int nx, ny;
    const double *lon;    /* [nx] degrees */
    const double *lat;    /* [ny] degrees */
    const double *area;   /* [ny] areaYCosPhi.x */
    const double *cosphi; /* [ny] areaYCosPhi.y */
    const double *ancho;  /* [ny+1] anchoVolumenes (dlon) */
    const double *alto;   /* [ny] altoVolumenes (dlat) */
} thy_grid;

/* Scheme parameters, non-dimensional as passed to shallowWater
   (src/GPU/ShallowWater.cu:186-190). */
typedef struct thy_params {
    double borde_sup, borde_inf, borde_izq, borde_der; /* 1 open, -1 wall, 2 periodic */
    int sponge_sup, sponge_inf, sponge_izq, sponge_der;
    double cfl, mf0, vmax, eps_h, hpos, cvis, dif_at;
    double hmin;       /* Hmin shift applied to H */
    double sea_level;  /* SEA_LEVEL / H */
    double L, H, Q, T; /* scales */
    int precision;     /* 0: FP32 state and fluxes; 1: FP64 everywhere */
    int track_arrival; /* 1: arrival times tracked (reference behaviour) */
    int tile_x, tile_y;/* 2D-tile kernel: tile shape, 0: default */
    int kernel;        /* step kernel: 0 row sweep with a deep-water launch for blocks far from
                          coast and borders (default), 1 2D LDS tiles, 2 row sweep with the
                          general faces everywhere */
    int strip_rows;    /* row-sweep kernel: rows per strip, 0: default (24) */
    int general_rows;  /* row-sweep kernel: rows per general strip beside the deep-water
                          launch (must divide strip_rows), 0: default (taller strips when
                          the general strips are many, 4-row strips when they are few) */
} thy_params;

/* One Okada fault, fields as a parameter-file fault line (src/Problema.cxx:128-148). */
typedef struct thy_fault {
    double time, lon_c, lat_c, depth_c, length, width, strike, dip, rake, slip;
} thy_fault;

/* Per-run schedule for thy_run (mirrors the loop of src/GPU/ShallowWater.cu:374-453). */
typedef void (*thy_snapshot_fn)(void *user, int index, double t_nondim, thy_ctx *ctx);
typedef void (*thy_okada_fn)(void *user, int fault_index);
/* Called once per finished iteration (1-based), in order, with the dt chosen for the
   next iteration and the time reached (non-dimensional), as the reference prints them
   (src/GPU/ShallowWater.cu:447-448).  Delivered after each graph replay. */
typedef void (*thy_step_fn)(void *user, int64_t iteration, double dt_next, double t_nondim);

typedef struct thy_run_config {
    double t_end;           /* tiempo_tot (non-dimensional) */
    double save_interval;   /* tiempoGuardarNetCDF (non-dimensional, <0: none) */
    double series_interval; /* tiempoGuardarSeries (non-dimensional, <0: none) */
    int n_faults;           /* faults sorted by time (src/Problema.cxx:97-126) */
    const thy_fault *faults;
    int crop_flag;          /* 0 none, 1 relative, 2 absolute (non-dimensional) */
    double crop_value;
    int n_gauges;           /* gauge cell indices (-1: outside the grid) */
    const int *gauge_pos;
    int64_t max_steps;      /* 0: unlimited */
    int steps_per_sync;     /* steps per captured graph replay (0: default) */
    thy_snapshot_fn on_snapshot; /* called at each NetCDF snapshot time */
    thy_okada_fn on_okada;       /* called before each Okada application */
    thy_step_fn on_step;         /* per-iteration log (may be NULL) */
    void *user;
} thy_run_config;

typedef struct thy_run_result {
    int64_t steps;
    int n_samples;          /* gauge samples recorded */
    int n_snapshots;
    double t_final;         /* non-dimensional */
    double dt_initial;      /* non-dimensional */
    double seconds;         /* wall time of the stepping window (hipEvent) */
} thy_run_result;

/* Output fields (src/GPU/netcdf_kernel.cu:6-126). */
enum thy_field {
    THY_ETA = 0, THY_UX = 1, THY_UY = 2, THY_DEFORMED_BATHY = 3,
    THY_MAX_HEIGHT = 4, THY_ARRIVAL_TIMES = 5
};

/* Device probe.  Replaces comprobarSoporteCUDA (src/GPU/SoporteCUDA.cu:8-29):
   0 usable gfx950 device, 1 no GPU, 2 no gfx950 GPU. */
int thy_probe(void);

/* Allocate the device state of one grid.  Replaces the allocation part of
   shallowWater (src/GPU/ShallowWater.cu:241-303). */
int thy_create(const thy_grid *grid, const thy_params *params, thy_ctx **out);
void thy_destroy(thy_ctx *ctx);
const char *thy_last_error(const thy_ctx *ctx);

/* Upload the reference's state arrays: state1 = (h, H) pairs and state2 = (qx, qy)
   pairs, interleaved as the double2 arrays datosVolumenesNivel_1/_2
   (src/GPU/ShallowWater.cu:299-300).  THY_ERR_NONFINITE if any value is NaN or Inf. */
int thy_upload_state(thy_ctx *ctx, const double *state1, const double *state2);
/* Inverse of thy_upload_state: (h, H) and (qx, qy) as doubles. */
int thy_download_state(thy_ctx *ctx, double *state1, double *state2);

/* eta_max / eta_ini / arrival-time initialisation (src/GPU/ShallowWater.cu:305-309). */
int thy_init_metrics(thy_ctx *ctx);
/* eta_max and arrival update from the current state at time t
   (src/GPU/Metrics.cu:23-42,57-77). */
int thy_update_metrics(thy_ctx *ctx, double t);

/* Okada standard source (src/GPU/Deformacion.cu:118-143; crop :132-142): FP64 displacement,
   crop, H -= U_Z, eta_ini += U_Z.  def_out (optional, host) receives U_Z/H. */
int thy_apply_okada(thy_ctx *ctx, const thy_fault *fault, int crop_flag, double crop_value,
                    double *def_out);

/* Initial dt (src/GPU/Timestep.cu:11-33), non-dimensional. */
int thy_initial_dt(thy_ctx *ctx, double *dt);

/* One step with a given dt (src/GPU/Timestep.cu:70-120); returns the next raw dt
   (global minimum of the per-cell dt). */
int thy_step(thy_ctx *ctx, double dt, double *dt_next);

/* Whole driver loop of shallowWater (src/GPU/ShallowWater.cu:374-453): initial
   dt, fault application and dt truncation at fault times (src/GPU/Timestep.cu:35-68),
   snapshots, gauge series, metrics, steps.  Steps run in captured graph replays with
   device-resident t and dt; the host synchronises once per replay. */
int thy_run(thy_ctx *ctx, const thy_run_config *cfg, thy_run_result *res);

/* Continue the schedule of the last thy_run for up to n_steps more steps (or until
   its end time).  No reference counterpart: the reference loop cannot be resumed;
   this lets a caller time an exact number of steps after a warm-up.  n_steps = 0 takes
   no step but captures and uploads the step sequences the next call replays, so that
   call starts at once.  Any step count replays without idle launches: a sequence is
   kept for each launch slot it can start at (launch n uses control and dt slot n mod 3). */
int thy_run_more(thy_ctx *ctx, int64_t n_steps, thy_run_result *res);

/* Output field conversion to float (src/GPU/netcdf_kernel.cu), [ny*nx] row-major. */
int thy_get_field(thy_ctx *ctx, int field, float *out);

/* Gauge series recorded by the last thy_run: times [n_samples] (seconds) and eta/u/v
   [n_samples * n_gauges] (src/GPU/ShallowWater.cu:69-90).  Pass NULL to skip. */
int thy_get_gauges(thy_ctx *ctx, float *time, float *eta, float *u, float *v,
                   float *eta_min, float *eta_max, float *bathy);

/* Compensated FP64 sum of h*area*cos(phi) (the scheme's mass invariant), for drift checks. */
int thy_mass(thy_ctx *ctx, double *mass);

/* thy_mass without any allocation, for callers that cannot allocate (code under stream
   capture, tight serving loops, harnesses that forbid device or host allocation); no
   counterpart in the reference.  Same kernel and summation as thy_mass, same result.
   The workgroup partial sums go to a file-scope device array of the library (1024
   doubles, 8 KiB per device, held until the process ends) instead of the context's own
   buffer; the grid size does not change the partial count, so no grid is too large.
   Every call on one device shares that array: calls that overlap on one device race. */
int thy_mass_noalloc(thy_ctx *ctx, double *mass);

/* Device bytes held by one context of this grid and parameters (admission). */
size_t thy_footprint(int nx, int ny, const thy_params *params);

/* Plain device-to-device copy kernel of `bytes` bytes run `reps` times; returns the
   time per copy.  No reference counterpart: it calibrates HBM counters and the
   achievable copy bandwidth beside the step kernel. */
int thy_copy_probe(size_t bytes, int reps, double *ms_per_copy);

/* Strip-height autotuner of the row-sweep kernel.  No reference counterpart (the
   reference has one fixed launch shape).  For pairs of strip height and general strip
   height (16/4, 16/8, 24/4, 24/8, 24/12, 32/8, 32/16, 36/18, 40/20, 48/16) it captures 12
   steps from the current state into a graph (each step redoes the same step), times
   `reps` replays (HIP events, median) and keeps the fastest; the state, control
   blocks and dt slots are left as they were.  Call after thy_upload_state, before or
   between runs.  general_rows may be null. */
int thy_autotune(thy_ctx *ctx, int reps, int *strip_rows, int *general_rows, double *ms_per_step);

/* Wall time (ms) of the last kernel launch group, for benchmarking. */
int thy_last_timing(thy_ctx *ctx, double *ms_step_kernel);

/* ---- Batched scenarios --------------------------------------------------------
   No counterpart in the reference, which runs one scenario per process
   (src/GPU/ShallowWater.cu:182-190).  A batch advances many contexts together: a step
   of the batch is three launches over the strips of every running scenario (the
   deep-water strips, the general strips beside them, then the strips the deep-water
   launch handed back), and every scenario keeps its own dt, time, faults, gauges and
   metrics, so its results are those of thy_run on the same context.  Contexts are
   created and uploaded by the caller (thy_create, thy_upload_state), admitted with
   their run configuration, and stay owned by the caller; results are read with
   thy_get_field / thy_get_gauges once the slot is THY_SLOT_FINISHED.  FP32 row-sweep
   contexts only; no snapshots (save_interval < 0).

   The step sequence can run asynchronously: thy_batch_launch returns as soon as it is
   queued, thy_batch_wait collects it.  Between the two the caller may admit
   scenarios (they join the next launch), read and remove finished ones, and create
   or destroy other contexts, while the GPU steps the batch. */
typedef struct thy_batch thy_batch;

/* THY_SLOT_FAILED: the health check found NaN/Inf in the scenario's state; it was
   retired (thy_last_error of its context says after which step) and the batch goes on. */
enum thy_slot_state { THY_SLOT_FREE = 0, THY_SLOT_RUNNING = 1, THY_SLOT_FINISHED = 2, THY_SLOT_FAILED = 3 };

/* capacity scenario slots; hbm_budget bytes of scenario state (0: unlimited);
   steps_per_sync launches per captured step sequence (rounded up to a multiple of 3). */
int thy_batch_create(int capacity, size_t hbm_budget, int steps_per_sync, thy_batch **out);
void thy_batch_destroy(thy_batch *b);
const char *thy_batch_last_error(thy_batch *b);
/* Prepares the run (metrics, initial dt, faults due at t = 0) and admits the context;
   THY_ERR_BUDGET when no slot is free or its footprint would exceed the budget (the
   context is then left as it was).  cfg NULL: the context was prepared by
   thy_batch_prepare, and admission only takes a slot. */
int thy_batch_add(thy_batch *b, thy_ctx *ctx, const thy_run_config *cfg, int *slot);
/* The context-local part of thy_batch_add (run configuration, metrics, initial dt,
   faults due at t = 0), for a thread other than the one that drives the batch: the
   set-up then overlaps the batch's launches instead of holding up its driver.  A later
   upload or thy_run voids it.  No counterpart in the reference (batched serving). */
int thy_batch_prepare(thy_ctx *ctx, const thy_run_config *cfg);
/* Queues one captured sequence of steps_per_sync steps for every running scenario and
   returns; THY_ERR_INVALID while a launch is in flight.  Does nothing when no
   scenario is running. */
int thy_batch_launch(thy_batch *b);
/* *done = 1 when no launch is in flight or it has finished (never waits). */
int thy_batch_ready(thy_batch *b, int *done);
/* Waits for the launch in flight (if any) and does the host work after it: faults
   that came due, finished scenarios.  *n_running: scenarios running afterwards. */
int thy_batch_wait(thy_batch *b, int *n_running);
/* thy_batch_launch, then thy_batch_wait. */
int thy_batch_advance(thy_batch *b, int *n_running);
/* State of a slot (THY_SLOT_*) and its run result; while it runs, steps and t_final
   are those after the last completed launch. */
int thy_batch_status(thy_batch *b, int slot, int *state, thy_run_result *res);
/* Frees a slot (the context is not destroyed; remove a running scenario's slot before
   destroying its context).  Removing a running slot waits for a launch in flight. */
int thy_batch_remove(thy_batch *b, int slot);
int thy_batch_info(thy_batch *b, int *running, size_t *bytes_used, double *gpu_seconds);

/* Releases the device memory, pinned host memory and streams that destroyed contexts
   and batches left in the library's caches (they are otherwise kept for reuse until
   the process ends).  Call with no context or batch alive. */
void thy_release_cached(void);

#ifdef __cplusplus
}
#endif
#endif
