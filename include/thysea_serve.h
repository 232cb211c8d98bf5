/*
 * thysea_serve.h -- request loop for batched earthquake scenarios on one GPU
 * (libthysea_serve.so, C++ over the C-ABI of include/thysea.h).
 *
 * No counterpart in the reference, which runs one scenario per process
 * (src/ShallowWater.cxx:146-151).  Requests arrive over time; each is a grid with its
 * initial state and run configuration, as a caller of thy_create / thy_upload_state /
 * thy_run would pass them.  thy_serve_batched admits them into one batch (thy_batch_*)
 * as they arrive and as slots and the HBM budget free up, keeps their state resident,
 * and advances every running scenario in each captured step sequence; the host work of
 * a request (context creation, upload, run preparation with the faults due at t = 0;
 * read-out of max height, arrival times and gauges, release) runs on worker threads
 * while the GPU steps the batch.  thy_serve_sequential serves the same requests one at
 * a time through thy_run, first come first served: the comparison.
 *
 * Time-to-result of a request = finished_s - arrival_s: from its arrival to its
 * max-height, arrival-time and gauge arrays in the caller's buffers.
 */
#ifndef THYSEA_SERVE_H
#define THYSEA_SERVE_H

#include <stddef.h>
#include <stdint.h>

#include "thysea.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct thy_request {
    double arrival_s;        /* seconds after the loop starts */
    thy_grid grid;           /* as thy_create; arrays stay valid until the call returns */
    thy_params params;       /* FP32 row-sweep kernel for batched runs */
    const double *state1;    /* (h, H) pairs, as thy_upload_state */
    const double *state2;    /* (qx, qy) pairs */
    /* the same state as four arrays [ny*nx], used when state1 is NULL (the loop pairs
       them on a worker thread; requests may share them) */
    const double *h, *H, *qx, *qy;
    thy_run_config run;      /* t_end, faults, gauges, crop; no snapshots, no callbacks */
} thy_request;

typedef struct thy_request_result {
    int status;              /* thy_status of the request (THY_ERR_NONFINITE: health check) */
    double admitted_s;       /* admission (batched) or start (sequential), after the loop start */
    double finished_s;       /* results in the caller's buffers */
    int64_t steps;
    int n_samples;           /* gauge samples recorded */
    /* caller-owned outputs, NULL to skip: [ny*nx] float fields (thy_get_field) and the
       gauge series [gauge_capacity] times and [gauge_capacity * n_gauges] eta/u/v */
    float *max_height, *arrival_times;
    float *g_time, *g_eta, *g_u, *g_v;
    int gauge_capacity;
} thy_request_result;

typedef struct thy_serve_options {
    int capacity;            /* batch slots (0: 512) */
    size_t hbm_budget;       /* bytes of scenario state in the batch (0: unlimited) */
    int steps_per_sync;      /* steps per captured sequence (0: 48) */
    int workers;             /* set-up and read-out threads (0: 8) */
    int burst;               /* admissions while the GPU is idle before the first launch (0: 8) */
} thy_serve_options;

/* Serve n requests through one batch; out[i] receives request i's results.  Returns
   THY_OK when the loop ran (per-request failures are in out[i].status), else the error
   that stopped it.  *gpu_seconds (optional): time of the batch's step launches. */
int thy_serve_batched(const thy_request *req, int n, const thy_serve_options *opt, thy_request_result *out,
                      double *gpu_seconds);

/* The same requests one at a time through thy_run (arrival order). */
int thy_serve_sequential(const thy_request *req, int n, const thy_serve_options *opt, thy_request_result *out);

#ifdef __cplusplus
}
#endif
#endif
