// thy_step.hpp -- fused single-pass time step of the Tsunami-HySEA scheme for gfx950.
//
// One launch = one full step of the reference (src/GPU/Timestep.cu:70-120: mass
// pass, h update, momentum pass, q update, per-cell dt) plus the metric update that
// the reference runs before each step (src/GPU/ShallowWater.cu:423-426) and the gauge
// gather (src/GPU/Volumen_kernel.cu:223-254).  There are no global accumulators:
//   * a TX x TY tile of cells plus a 4-cell halo of (e, D, qx, qy) is staged in LDS;
//   * every mass face of the tile+2 region is evaluated once and accumulated into an
//     LDS accumulator in the reference's order (x even, x odd, y even, y odd faces:
//     src/GPU/Timestep.cu:79-90), giving h* (stored as e*) on tile+2;
//   * every momentum face of the tile is evaluated once from h* and q^n, accumulated
//     the same way, and the cell update, sponge, near-dry filter, velocity cap,
//     semi-implicit Manning friction and per-cell dt are applied in registers;
//   * the per-cell dt is reduced with a wave64 shuffle tree and one atomicMin per
//     workgroup into a device slot that the next launch reads (no host sync).
// Simulated time, dt truncation at fault times (src/GPU/Timestep.cu:35-68) and the
// stop predicate of the driver loop live in a device control block.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "thy_flux.hpp"

namespace thy {

// Control block, triple-buffered by launch index (slot = n mod 3): launch n reads
// ctrl[slot] and the dt slot written by launch n-1, block 0 writes ctrl[slot+1].
struct alignas(16) Ctrl {
    double t;            // non-dimensional time before this step
    double dt;           // finalised dt when dt_valid
    double t_end;
    double next_save;    // snapshot time (non-dim), +inf when disabled
    double next_series;  // gauge sample time, +inf when disabled
    double series_dt;
    double T;            // time scale, for fault times in seconds
    int64_t steps, max_steps;
    int k, nf;           // next unapplied fault, number of faults
    int skip;            // truncation state (saltar_deformacion)
    int dt_valid;
    int cur;             // which state buffer holds the current state
    int nsamples;
    int gcap;
    int snap_done;       // the host took the snapshot due at t: the next launch steps
    int nonfinite;       // a step produced NaN/Inf (the health flag seen by the launch after it)
    int pad;
};

// Arrival tracking keeps one value per cell (src/GPU/Metrics.cu:44-77 keeps two): the
// cell's eta_ini while the wave has not arrived, and once it has, the arrival time t
// stored as t * 2^100 (2^40 for t = 0), exactly (a power-of-two scale).  The arrival
// test reads one array instead of two.  A value at or above 2^39 means "arrived";
// eta_ini values are far below that (the largest initial |eta| is the dry-cell marker
// -1e30, on the other side of zero).
template <typename T> struct Arrival {
    static constexpr float kArrived = 0x1p39f;
    static constexpr float kZero = 0x1p40f;  // arrived at t = 0
    static constexpr float kScale = 0x1p100f;
    __host__ __device__ static T enc(const T t) { return t > T(0) ? t * T(kScale) : T(kZero); }
    __host__ __device__ static bool arrived(const T p) { return p >= T(kArrived); }
    __host__ __device__ static double dec(const T p) { return p == T(kZero) ? 0.0 : (double)p * 0x1p-100; }
};

template <typename T> struct RowGeo {
    T A, invA, ilx, valx;  // area, 1/area, alto*(1/cos(phi)) (x-face length factor), (2*alto/cos)/(2*area)
    T kdx, dxs, dys, kA;   // 0.5 (1 - cvis) valx ilx (2s diffusion factor), ancho and alto for sponge distances, A/4
};

template <typename T> struct FaceGeoY {
    T len, cos0, cos1, A0, A1, valy, ic0, ic1; // ancho, cos/area of rows j-1, j (clamped)
    // face-uniform factors of the deep-water y fluxes, computed on the host in the state
    // type exactly as the general faces compute them: 0.5 (cos0 + cos1), cos1 - cos0,
    // cos1 / cos0 (as cos1 * ic0), the mass flux factor len * cos0, the 2s diffusion
    // factor 0.5 (1 - cvis) valy len; and the positivity limiter's factors cos A / 4 of
    // the two rows
    T gh, gd, g1ic0, k0, k1, lc0, kdy, pad;
};

// The dt minimum of a launch is kept as kDtParts partial minima (float bits; for
// positive floats the unsigned order is the float order), one 64-byte line apart:
// atomics on one line are applied one after another.  The launches of one grid fold
// one minimum per workgroup into partial 0, and the next launch reads that one word.
// The batch launches fold one per wavefront (several thousand per launch for a large
// grid) into partial (wavefront index mod kDtParts), so that the launch does not wait
// for a queue of atomics on one line, and read all partials.  The writer of a launch
// resets every partial of the next launch's slot.
constexpr int kDtParts = 8;
constexpr int kDtStride = 16;  // words: one 64-byte line per partial
constexpr int kDtSlot = kDtParts * kDtStride;
constexpr unsigned int kDtNone = 0x7f7fffffu;  // FLT_MAX: no cell has reported yet

// word of partial `part` (>= 0) of launch slot `slot` (0..2)
__host__ __device__ __forceinline__ int dt_at(const int slot, const int part) {
    return slot * kDtSlot + (part % kDtParts) * kDtStride;
}

template <typename T> struct StepArgs {
    int nx, ny, pitch, periodic;
    int tiles_x, tiles_y, n_tiles, chunk;  // chunk: tiles per XCD chunk (whole tile rows)
    int srows;                             // row-sweep kernel: rows per strip
    int gsrows;                            //   rows per general strip beside a deep-water launch
    int yoff;                              //   split launches: strip row k covers rows k*rows-yoff ..
    int nseg;                              // row-sweep kernel: 56-column segments per row
    const int *order;                      // row-sweep kernel: every block, costliest first (or null)
    int n_order;                           //   number of them
    const int *deep;                       // deep-water strips (strip row * nseg + segment; null: no split)
    int n_deep;
    const int *gen;                        // general strips beside a deep-water launch, costliest first
    int n_gen;
    unsigned int *fb_count;                // 3 slots: strips the deep-water sweep hands back
    int *fb_list;                          // [3][n_deep]
    int sp_sup, sp_inf, sp_izq, sp_der;
    int track_arrival, n_gauges;
    T bsup, binf, bizq, bder;
    T cfl, mf0, vmax, eps_h, hpos, cvis, dif_at, sea_level;
    T Href, Q_over_H;
    T *e[2], *qx[2], *qy[2];
    const T *D;
    T *emax;
    T *arr;                  // arrival tracking: eta_ini until the cell arrives, then the arrival time (Arrival<T>)
    const char *block;       // the arrays above lie in this one allocation (row sweeps: one buffer resource)
    int block_bytes;         //   its size, below 2^31
    const RowGeo<T> *row;
    const FaceGeoY<T> *fy;
    Ctrl *ctrl;
    unsigned int *dtmin;     // 3 launch slots of kDtParts partial minima (dt_at)
    const double *ftime;     // fault times (s), sorted
    const int *gpos;
    float *gbuf;             // [gcap][n_gauges][3]
    float *gtime;            // [gcap]
    double *steplog;         // [logcap][2]: (dt after truncation, t) of each finished iteration
    int logcap;
    unsigned int *health;    // set when a step wrote a non-finite e, qx or qy; stops the run
    int no_publish;          // timing replays: the writer leaves the control blocks as they are
};

__host__ __device__ __forceinline__ void trunc_plain(const Ctrl &c, const double *ft, double t, double &dt, int &skip) {
    if (c.k < c.nf) {
        if ((t + dt) * c.T > ft[c.k]) {
            dt = ft[c.k] / c.T - t;
            skip = 1;
        }
    }
}

__host__ __device__ __forceinline__ void trunc_skipping(const Ctrl &c, const double *ft, double t, double &dt, int &skip) {
    skip = 0;
    if (c.k < c.nf) {
        int i = c.k + 1;
        while (i < c.nf && !(fabs(ft[i] - ft[c.k]) > 2.220446049250313e-16)) i++;
        if (i < c.nf && (t + dt) * c.T > ft[i]) {
            dt = ft[i] / c.T - t;
            skip = 1;
        }
    }
}

// stop predicate of the driver loop, evaluated before each step.  A snapshot is taken
// at most once per iteration (src/GPU/ShallowWater.cu:404-409): after the host took
// the one due at t, the next launch steps even when t is still past the next snapshot
// time (a dt longer than the snapshot interval).
__host__ __device__ __forceinline__ bool must_stop(const Ctrl &c, const double *ft) {
    if (!(c.t < c.t_end)) return true;
    if (c.max_steps > 0 && c.steps >= c.max_steps) return true;
    if (c.k < c.nf && c.t * c.T >= ft[c.k]) return true;
    if (!c.snap_done && c.t >= c.next_save) return true;
    if (c.nonfinite) return true;
    return false;
}

// dt of the launch that reads control block c: the finalised dt when c holds one, else
// the previous step's raw minimum truncated at the next fault time
// (src/GPU/Timestep.cu:35-68, src/GPU/ShallowWater.cu:430-436)
__host__ __device__ __forceinline__ double step_dt(const Ctrl &c, const double *ft, const float raw, int &skip) {
    skip = c.skip;
    if (c.dt_valid) return c.dt;
    double dt = (double)raw;
    if (!skip) trunc_plain(c, ft, c.t, dt, skip);
    else trunc_skipping(c, ft, c.t, dt, skip);
    return dt;
}

template <typename T> struct BlockCtl {
    double t, dt, Tscale;
    int run, sample, slot_sample, cur, gcap;
};

// Start of a step launch (thread 0 of every workgroup): read the control block of this
// launch, finalise dt from the previous launch's minimum (the least of its first
// `parts` dt partials) with the truncation at fault times (src/GPU/Timestep.cu:35-68)
// and the stop predicate; the writer (workgroup 0) also logs the finished iteration,
// resets the dt partials of the next launch and publishes the next control block.
template <typename T>
__device__ __forceinline__ void begin_step(const StepArgs<T> &a, const int slot, const bool writer, BlockCtl<T> &bc,
                                           const int parts = 1) {
    Ctrl c = a.ctrl[slot];
    // the health flag raised by the previous step's cells stops the run here
    if (a.health && *a.health) c.nonfinite = 1;
    int skip;
    unsigned int raw = kDtNone;
    for (int p = 0; p < parts; p++) raw = min(raw, a.dtmin[dt_at((slot + 2) % 3, p)]);
    const double dt = step_dt(c, a.ftime, __uint_as_float(raw), skip);
    const bool stop = must_stop(c, a.ftime);
    const bool sample = !stop && (c.t >= c.next_series);
    bc.t = c.t;
    bc.dt = dt;
    bc.run = stop ? 0 : 1;
    bc.sample = sample ? 1 : 0;
    bc.slot_sample = c.nsamples;
    bc.cur = c.cur;
    bc.gcap = c.gcap;
    bc.Tscale = c.T;
    if (writer) {
        for (int p = 0; p < kDtParts; p++) a.dtmin[dt_at((slot + 1) % 3, p)] = kDtNone;  // the next launch's partials
        if (a.no_publish) return;
        // iteration c.steps has finished: record the dt chosen for the next one and the time
        if (c.steps > 0 && a.steplog) {
            const int64_t li = (c.steps - 1) % a.logcap;
            a.steplog[2 * li] = dt;
            a.steplog[2 * li + 1] = c.t;
        }
        Ctrl o = c;
        o.skip = skip;
        if (stop) {
            o.dt = dt;
            o.dt_valid = 1;
        } else {
            o.t = c.t + dt;
            o.dt_valid = 0;
            o.steps = c.steps + 1;
            o.cur = c.cur ^ 1;
            o.snap_done = 0;
            if (sample) {
                o.next_series = c.next_series + c.series_dt;
                o.nsamples = c.nsamples + 1;
            }
        }
        a.ctrl[(slot + 1) % 3] = o;
    }
}

// gauge gather from the state at time t (src/GPU/Volumen_kernel.cu:223-254), run by
// workgroup 0 of a step launch that takes a sample
template <typename T>
__device__ __forceinline__ void gather_gauges(const StepArgs<T> &a, const BlockCtl<T> &bc, const T *ge,
                                              const T *gqx, const T *gqy, const int tid, const int nthreads) {
    if (!bc.sample || bc.slot_sample >= bc.gcap) return;
    const int nx = a.nx, P = a.pitch;
    for (int g = tid; g < a.n_gauges; g += nthreads) {
        const int p = a.gpos[g];
        float ev = -9999.f, uv = -9999.f, vv = -9999.f;
        if (p >= 0) {
            const int64_t o = (int64_t)(p / nx) * P + (p % nx);
            T e = ge[o], D = a.D[o];
            T h = e + D;
            T f = (h < a.eps_h) ? desing(h, a.eps_h) : Num<T>::rcp(h);
            ev = (float)(e * a.Href);
            uv = (float)(f * gqx[o] * a.Q_over_H);
            vv = (float)(f * gqy[o] * a.Q_over_H);
        }
        float *dst = a.gbuf + ((int64_t)bc.slot_sample * a.n_gauges + g) * 3;
        dst[0] = ev;
        dst[1] = uv;
        dst[2] = vv;
    }
    if (tid == 0) a.gtime[bc.slot_sample] = (float)(bc.t * bc.Tscale);
}

// health check of a stored cell: NaN or Inf in e*, qx or qy
template <typename T>
__device__ __forceinline__ bool bad_cell(const T e, const T qx, const T qy) {
    return !(__builtin_isfinite(e) && __builtin_isfinite(qx) && __builtin_isfinite(qy));
}

// a wavefront that stored a bad cell raises the context's health flag, which stops the
// run at the next launch
template <typename T>
__device__ __forceinline__ void raise_health(const StepArgs<T> &a, const bool bad) {
    if (__any(bad) && (threadIdx.x & 63) == 0 && a.health) atomicOr(a.health, 1u);
}

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}

// friction, filter and velocity cap for h* <= hpos: src/GPU/Volumen_kernel.cu:25-76
template <typename T>
__device__ __forceinline__ void near_dry(T hs, T qxo, T qyo, T &qx, T &qy, T dt, T mf0, T vmax, T eps_h) {
    T f = Num<T>::div(hs, eps_h);
    T f2 = f * f;
    T aux = Num<T>::exp(-dt * Num<T>::rcp(f2 * f2 + Num<T>::eps));
    qx *= aux;
    qy *= aux;
    T fd = desing(hs, eps_h);
    if (vmax > T(0)) {
        T ux = fd * qx, uy = fd * qy;
        T u = Num<T>::sqrt(ux * ux + uy * uy);
        if (u > vmax) {
            T s = Num<T>::div(vmax, u);
            ux *= s;
            uy *= s;
        }
        qx = ux * hs;
        qy = uy * hs;
    }
    T u1x = fd * qx, u1y = fd * qy;
    T u1 = fd * Num<T>::sqrt(qxo * qxo + qyo * qyo);
    if (hs > T(0)) {
        T c1 = Num<T>::div(dt * mf0 * tabs(u1), Num<T>::pow43(hs) + Num<T>::eps);
        T g = Num<T>::div(hs, T(1) + c1);
        qx = g * u1x;
        qy = g * u1y;
    }
}

// sponge of the eta update on cell (i, j) (src/GPU/Volumen_kernel.cu:96-158)
template <typename T>
__device__ __forceinline__ T sponge_eta(const StepArgs<T> &a, T e, const T D, const int i, const int j, const T dx,
                                        const T dy) {
    const int nx = a.nx, ny = a.ny;
    if (i < a.sp_izq || i >= nx - a.sp_der || j < a.sp_sup || j >= ny - a.sp_inf) {
        T dl = (T(a.sp_izq - i) - T(0.5)) * dx; if (dl <= T(0)) dl = T(1e30);
        T dr = (T(i - (nx - 1 - a.sp_der)) - T(0.5)) * dx; if (dr <= T(0)) dr = T(1e30);
        T ds = (T(a.sp_sup - j) - T(0.5)) * dy; if (ds <= T(0)) ds = T(1e30);
        T di = (T(j - (ny - 1 - a.sp_inf)) - T(0.5)) * dy; if (di <= T(0)) di = T(1e30);
        T dist, ls;
        if (dl <= dr && dl <= ds && dl <= di) { dist = dl; ls = dx * T(a.sp_izq); }
        else if (dr <= dl && dr <= ds && dr <= di) { dist = dr; ls = dx * T(a.sp_der); }
        else if (ds <= dl && ds <= dr && ds <= di) { dist = ds; ls = dy * T(a.sp_sup); }
        else { dist = di; ls = dy * T(a.sp_inf); }
        const T r = Num<T>::div(dist, ls);
        const T co = T(0.9) * r * r;
        const T diff = (D + a.sea_level > T(0)) ? (e - a.sea_level) : (e + D);
        e -= co * diff;
        if (!(e + D >= T(0))) e = -D;
    }
    return e;
}

// sponge of the discharge update on cell (i, j) (src/GPU/Volumen_kernel.cu:160-221)
template <typename T>
__device__ __forceinline__ void sponge_q(const StepArgs<T> &a, const int i, const int j, T &qx, T &qy) {
    const int nx = a.nx, ny = a.ny;
    if (i < a.sp_izq) { const T r = (T(a.sp_izq - i) - T(0.5)) / T(a.sp_izq); qx -= T(0.9) * r * r * qx; }
    if (i >= nx - a.sp_der) { const T r = (T(i - (nx - 1 - a.sp_der)) - T(0.5)) / T(a.sp_der); qx -= T(0.9) * r * r * qx; }
    if (j < a.sp_sup) { const T r = (T(a.sp_sup - j) - T(0.5)) / T(a.sp_sup); qy -= T(0.9) * r * r * qy; }
    if (j >= ny - a.sp_inf) { const T r = (T(j - (ny - 1 - a.sp_inf)) - T(0.5)) / T(a.sp_inf); qy -= T(0.9) * r * r * qy; }
}

// max-height and arrival-time update from the state before the step
// (src/GPU/Metrics.cu:23-42,57-77); idempotent within a step
template <typename T>
__device__ __forceinline__ void metrics_cell(const StepArgs<T> &a, const int64_t go, const T eo, const T ho, const T tnow) {
    const T em = a.emax[go];
    if (eo > em && ho > a.eps_h) a.emax[go] = eo;
    if (a.track_arrival) {
        const T p = a.arr[go];
        if (!Arrival<T>::arrived(p) && tabs(eo - p) > a.dif_at) a.arr[go] = Arrival<T>::enc(tnow);
    }
}

// ---------------------------------------------------------------------------
// Deep-water fast path.  When every cell of a tile and its 4-cell halo is deeper
// than hpos and the tile is clear of the borders, sponges and the first-order band,
// every face takes the reference's 2s branch (src/GPU/Arista_kernel.cu:545,653,
// 1105,1214).  The 2s mass flux of a face depends only on the two adjacent cells, so
// each cell evaluates its four mass faces itself; contraction is disabled and the
// result is pinned in a register so that both cells of a face see the same rounded
// value and the update stays conservative.
// ---------------------------------------------------------------------------

// x face: F = q0*alto/cos, positivity (src/GPU/Arista_kernel.cu:79-95,118-130)
template <typename T>
__device__ __forceinline__ T deep_mass_x(T q0, T h0, T h1, T il, T kA, T dt) {
#pragma clang fp contract(off)
    T F = q0 * il;
    T lim = (F > T(0)) ? h0 * kA : h1 * kA;
    T den = habs(F) + Num<T>::eps;
    if (dt * den > lim) {
        T dt0 = lim * Num<T>::rcp(den);
        if (dt > dt0) F *= dt0 * Num<T>::rcp(dt + Num<T>::eps);
    }
    asm volatile("" : "+v"(F));
    return F;
}

// y face: F = ancho*q0*cos0, positivity with cos (src/GPU/Arista_kernel.cu:97-114,674-686)
template <typename T>
__device__ __forceinline__ T deep_mass_y(T q0, T h0, T h1, T len, T c0, T k0, T k1, T dt) {
#pragma clang fp contract(off)
    T F = q0 * (len * c0);
    T lim = (F > T(0)) ? h0 * k0 : h1 * k1;
    T den = habs(F) + Num<T>::eps;
    if (dt * den > lim) {
        T dt0 = lim * Num<T>::rcp(den);
        if (dt > dt0) F *= dt0 * Num<T>::rcp(dt + Num<T>::eps);
    }
    asm volatile("" : "+v"(F));
    return F;
}

template <typename T, int TX, int TY>
struct StepLds {
    static constexpr int IX = TX + 8, IY = TY + 8, SX = TX + 4, SY = TY + 4;
    static constexpr int NIN = IX * IY, NS = SX * SY, NT = TX * TY;
    static constexpr int NFX = (TX + 1) * TY, NFY = TX * (TY + 1);
    static constexpr int R0 = (4 * NIN > 4 * (NFX + NFY)) ? 4 * NIN : 4 * (NFX + NFY);
    static constexpr int R1 = (7 * NS > 2 * NS + 3 * NT) ? 7 * NS : 2 * NS + 3 * NT;
    static constexpr int TOTAL = R0 + R1;
};

template <typename T, int TX, int TY>
__global__ void __launch_bounds__(256) k_step(const StepArgs<T> a, const int slot) {
    constexpr int NT = 256;
    using L = StepLds<T, TX, TY>;
    constexpr int IX = L::IX, IY = L::IY, SX = L::SX, SY = L::SY;
    // one LDS buffer, two layouts:
    //   region 0: input (e, D, qx, qy on tile+4); the fast path overwrites it with the
    //             momentum face fluxes once the cell arrays hold what it needs
    //   region 1: fast path: e*, D, sqrt(h*), u, v, qx, qy on tile+2
    //             general path: e*, mass accumulator, momentum accumulators
    __shared__ T lds[L::TOTAL];
    T *s_e = lds, *s_D = lds + L::NIN, *s_qx = lds + 2 * L::NIN, *s_qy = lds + 3 * L::NIN;
    T *r1 = lds + L::R0;
    T *s_es = r1;
    __shared__ BlockCtl<T> bc;
    __shared__ float s_red[NT / 64];

    const int tid = threadIdx.x;
    // XCD-aware tile order.  Workgroups are dispatched to the 8 XCDs round-robin, so
    // bid & 7 is the XCD and bid >> 3 the workgroup's place in that XCD's sequence.
    // The grid is cut into chunks of whole tile rows, dealt to the XCDs in turn: each
    // XCD walks its chunks in row order (its L2 holds the halos shared along a row),
    // and coastal rows, whose tiles take the slower general path, are spread over all
    // XCDs instead of landing on one.
    const int bid = blockIdx.x;
    const int loc = bid >> 3, ch = loc / a.chunk;
    const int tile = ((ch << 3) + (bid & 7)) * a.chunk + (loc - ch * a.chunk);

    if (tid == 0) begin_step(a, slot, bid == 0, bc);
    __syncthreads();
    if (!bc.run || tile >= a.n_tiles) return;

    const int cur = bc.cur;
    const T *__restrict__ ge = a.e[cur];
    const T *__restrict__ gqx = a.qx[cur];
    const T *__restrict__ gqy = a.qy[cur];
    T *__restrict__ oe = a.e[cur ^ 1];
    T *__restrict__ oqx = a.qx[cur ^ 1];
    T *__restrict__ oqy = a.qy[cur ^ 1];
    const T dt = (T)bc.dt;
    const int nx = a.nx, ny = a.ny, P = a.pitch;
    const bool per = a.periodic != 0;

    if (bid == 0) gather_gauges(a, bc, ge, gqx, gqy, tid, NT);

    const int tyi = tile / a.tiles_x, txi = tile - tyi * a.tiles_x;
    const int x0 = txi * TX, y0 = tyi * TY;
    // clear of borders (first-order band, clamped stencils), sponges and grid wrap
    const bool interior = x0 >= 8 && x0 + TX + 8 <= nx && y0 >= 8 && y0 + TY + 8 <= ny;

    // ---- stage 0: input region (e, D, qx, qy) on tile+4 ----
    if (interior && (sizeof(T) == 4)) {
        constexpr int V = IX / 4;  // float4 per row
        for (int idx = tid; idx < V * IY; idx += NT) {
            const int lj = idx / V, lv = idx - lj * V;
            const int64_t o = (int64_t)(y0 - 4 + lj) * P + (x0 - 4) + 4 * lv;
            const int l = lj * IX + 4 * lv;
            *(float4 *)&s_e[l] = *(const float4 *)&ge[o];
            *(float4 *)&s_D[l] = *(const float4 *)&a.D[o];
            *(float4 *)&s_qx[l] = *(const float4 *)&gqx[o];
            *(float4 *)&s_qy[l] = *(const float4 *)&gqy[o];
        }
    } else {
        for (int idx = tid; idx < IX * IY; idx += NT) {
            const int lj = idx / IX, li = idx - lj * IX;
            int gx = x0 - 4 + li, gy = y0 - 4 + lj;
            if (per) gx = ((gx % nx) + nx) % nx;
            else gx = gx < 0 ? 0 : (gx >= nx ? nx - 1 : gx);
            gy = gy < 0 ? 0 : (gy >= ny ? ny - 1 : gy);
            const int64_t o = (int64_t)gy * P + gx;
            s_e[idx] = ge[o];
            s_D[idx] = a.D[o];
            s_qx[idx] = gqx[o];
            s_qy[idx] = gqy[o];
        }
    }
    __syncthreads();

    // deep vote over the input region
    bool mine = true;
    if (interior) {
        for (int idx = tid; idx < IX * IY; idx += NT) mine = mine && (s_e[idx] + s_D[idx] > a.hpos);
    }
    const bool fast = __syncthreads_and(interior && mine) != 0;

    float dtmin = 3.0e38f;
    bool bad = false;
    const T tnow = (T)bc.t;

    if (fast) {
        // ================= fast path =================
        T *f_D = r1 + L::NS, *f_s = r1 + 2 * L::NS, *f_u = r1 + 3 * L::NS, *f_v = r1 + 4 * L::NS;
        T *f_qx = r1 + 5 * L::NS, *f_qy = r1 + 6 * L::NS;
        // h*, e* on tile+2 with the per-cell quantities of the momentum faces
        bool deep2 = true;
        for (int idx = tid; idx < L::NS; idx += NT) {
            const int sj = idx / SX, si = idx - sj * SX;
            const int gy = y0 - 2 + sj;
            const RowGeo<T> rg = a.row[gy];
            const FaceGeoY<T> fb = a.fy[gy], ft = a.fy[gy + 1];
            const int io = (sj + 2) * IX + si + 2;
            const T e = s_e[io], D = s_D[io], qx = s_qx[io], qy = s_qy[io];
            const T h = e + D;
            const T hl = s_e[io - 1] + s_D[io - 1], hr = s_e[io + 1] + s_D[io + 1];
            const T hb = s_e[io - IX] + s_D[io - IX], ht = s_e[io + IX] + s_D[io + IX];
            const T Fl = deep_mass_x(s_qx[io - 1], hl, h, rg.ilx, rg.kA, dt);
            const T Fr = deep_mass_x(qx, h, hr, rg.ilx, rg.kA, dt);
            const T Fb = deep_mass_y(s_qy[io - IX], hb, h, fb.len, fb.cos0, fb.k0, fb.k1, dt);
            const T Ft = deep_mass_y(qy, h, ht, ft.len, ft.cos0, ft.k0, ft.k1, dt);
            const T acc = (Fl - Fr) + (Fb * fb.ic1 - Ft * ft.ic0);
            const T val = dt * rg.invA;
            T es = Num<T>::mad(val, acc, e);
            if (!(es + D > T(0))) es = -D;
            const T hs = es + D;
            deep2 = deep2 && (hs > a.hpos);
            const T ih = Num<T>::rcp(hs);
            s_es[idx] = es;
            f_D[idx] = D;
            f_s[idx] = Num<T>::sqrt(hs);
            f_u[idx] = qx * ih;
            f_v[idx] = qy * ih;
            f_qx[idx] = qx;
            f_qy[idx] = qy;
            // metrics from the state before the step (src/GPU/Metrics.cu:23-42,57-77); the
            // update is idempotent, so a fall-back to the general path may repeat it
            if (si >= 2 && si < TX + 2 && sj >= 2 && sj < TY + 2) {
                metrics_cell(a, (int64_t)gy * P + x0 - 2 + si, e, h, tnow);
            }
        }
        const bool fast2 = __syncthreads_and(deep2) != 0;
        if (fast2) {
            // momentum faces, 2s branch (src/GPU/Arista_kernel.cu:221-280, 780-843);
            // per face: mn, pn (normal momentum to the low/high cell), ft (tangential,
            // low -= ft, high += ft), w (wave-speed term)
            T *fx = lds, *fyv = lds + 4 * L::NFX;
            for (int idx = tid; idx < L::NFX; idx += NT) {
                const int tj = idx / (TX + 1), fi = idx - tj * (TX + 1);
                const int gy = y0 + tj;
                const RowGeo<T> rg = a.row[gy];
                const int c0 = (tj + 2) * SX + fi + 1, c1 = c0 + 1;
                const T h0 = s_es[c0] + f_D[c0], h1 = s_es[c1] + f_D[c1];
                const T s0 = f_s[c0], s1 = f_s[c1], u0 = f_u[c0], u1 = f_u[c1];
                const T hij = T(0.5) * (h0 + h1);
                const T uij = Num<T>::mad(s0, u0, s1 * u1) * Num<T>::rcp(s0 + s1 + Num<T>::eps);
                const T lmax = habs(uij) + Num<T>::sqrt(hij);
                const T deta = deta_of(s_es[c0], s_es[c1], f_D[c0], f_D[c1]);
                const T il = rg.ilx;
                const bool up = uij < T(0);
                const T uu = up ? u1 : u0;
                const T fn = uu * (up ? f_qx[c1] : f_qx[c0]) * il;
                const T ftg = uu * (up ? f_qy[c1] : f_qy[c0]) * il;
                const T DES = rg.kdx * dt * (lmax * lmax) * (f_qx[c1] - f_qx[c0]);
                fx[4 * idx + 0] = Num<T>::mad(hij * deta, il, fn) - DES;
                fx[4 * idx + 1] = -fn + DES;
                fx[4 * idx + 2] = ftg;
                fx[4 * idx + 3] = il * lmax;
            }
            for (int idx = tid; idx < L::NFY; idx += NT) {
                const int fj = idx / TX, ti = idx - fj * TX;
                const int g = y0 + fj;
                const FaceGeoY<T> fg = a.fy[g];
                const int c0 = (fj + 1) * SX + ti + 2, c1 = c0 + SX;
                const T h0 = s_es[c0] + f_D[c0], h1 = s_es[c1] + f_D[c1];
                const T s0 = f_s[c0], s1 = f_s[c1], v0 = f_v[c0], v1 = f_v[c1];
                const T hij = T(0.5) * (h0 + h1);
                const T uij = Num<T>::mad(s0, v0, s1 * v1) * Num<T>::rcp(s0 + s1 + Num<T>::eps);
                const T lmax = habs(uij) + Num<T>::sqrt(hij);
                const T deta = Num<T>::mad(-(uij * uij), fg.cos1 - fg.cos0,
                                           T(0.5) * (fg.cos0 + fg.cos1) * deta_of(s_es[c0], s_es[c1], f_D[c0], f_D[c1]));
                const bool up = uij < T(0);
                const T uu = up ? v1 : v0;
                const T fn = uu * (up ? f_qy[c1] : f_qy[c0]) * fg.len;
                T mn, pn;
                if (up) { pn = -fn; mn = fn * (fg.cos1 * fg.ic0); }
                else { mn = fn; pn = -fn * fg.cos0 * fg.ic1; }
                mn = Num<T>::mad(hij * deta * fg.len, fg.ic0, mn);
                const T DES = fg.kdy * dt * (lmax * lmax) * Num<T>::mad(f_qy[c1], fg.cos1, -(f_qy[c0] * fg.cos0));
                fyv[4 * idx + 0] = Num<T>::mad(-DES, fg.ic0, mn);
                fyv[4 * idx + 1] = Num<T>::mad(DES, fg.ic1, pn);
                fyv[4 * idx + 2] = uu * (up ? f_qx[c1] : f_qx[c0]) * fg.len;
                fyv[4 * idx + 3] = fg.len * lmax;
            }
            __syncthreads();
            // cell update (src/GPU/Volumen_kernel.cu:160-221; no sponge or friction here)
            for (int idx = tid; idx < TX * TY; idx += NT) {
                const int tj = idx / TX, ti = idx - tj * TX;
                const int gx = x0 + ti, gy = y0 + tj;
                const RowGeo<T> rg = a.row[gy];
                const int so = (tj + 2) * SX + ti + 2;
                const int xl = tj * (TX + 1) + ti, xr = xl + 1;
                const int yb = tj * TX + ti, yt = yb + TX;
                const T ax = -fx[4 * xl + 1] - fx[4 * xr + 0] + fyv[4 * yb + 2] - fyv[4 * yt + 2];
                const T ay = fx[4 * xl + 2] - fx[4 * xr + 2] - fyv[4 * yb + 1] - fyv[4 * yt + 0];
                const T aw = (fx[4 * xl + 3] + fx[4 * xr + 3]) + (fyv[4 * yb + 3] + fyv[4 * yt + 3]);
                const T dtc = (aw < Num<T>::eps) ? Num<T>::big : Num<T>::div(T(2) * a.cfl * rg.A, aw);
                dtmin = fminf(dtmin, (float)dtc);
                const T val = dt * rg.invA;
                const T qx = Num<T>::mad(val, ax, f_qx[so]);
                const T qy = Num<T>::mad(val, ay, f_qy[so]);
                const T es = s_es[so];
                const int64_t go = (int64_t)gy * P + gx;
                bad = bad || bad_cell(es, qx, qy);
                oe[go] = es;
                oqx[go] = qx;
                oqy[go] = qy;
            }
            raise_health(a, bad);
            dtmin = wave_min(dtmin);
            if ((tid & 63) == 0) s_red[tid >> 6] = dtmin;
            __syncthreads();
            if (tid == 0) {
                float m = s_red[0];
#pragma unroll
                for (int w = 1; w < NT / 64; w++) m = fminf(m, s_red[w]);
                atomicMin(&a.dtmin[dt_at(slot, 0)], __float_as_uint(m));
            }
            return;
        }
        // h* fell below hpos somewhere in the halo: continue on the general path with e*
        // already in s_es (the input region is intact)
    }

    // ================= general path =================
    T *s_am = r1 + L::NS;
    T *s_ax = r1 + 2 * L::NS, *s_ay = s_ax + L::NT, *s_aw = s_ay + L::NT;
    Prm<T> px, py;
    px.eps_h = py.eps_h = a.eps_h;
    px.hpos = py.hpos = a.hpos;
    px.cvis = py.cvis = a.cvis;
    px.bl = a.bizq; px.bh = a.bder;
    py.bl = a.bsup; py.bh = a.binf;

    auto in_cell = [&](int li, int lj) {
        const int o = lj * IX + li;
        SC<T> c;
        c.e = s_e[o];
        c.D = s_D[o];
        c.h = c.e + c.D;
        c.qn = s_qx[o];
        c.qt = s_qy[o];
        c.f = desing(c.h, a.eps_h);
        return c;
    };

    if (!fast) {
        for (int idx = tid; idx < SY * SX; idx += NT) s_am[idx] = T(0);
        __syncthreads();
        // ---- stage 1: mass fluxes on tile+2, x faces (even, odd) then y faces ----
        for (int par = 0; par < 2; par++) {
            constexpr int NFX = (SX + 2) / 2;  // faces per row and colour (upper bound)
            for (int idx = tid; idx < NFX * SY; idx += NT) {
                const int sj = idx / NFX;
                const int fi = 2 * (idx - sj * NFX) + par;
                if (fi > SX) continue;
                const int g = x0 - 2 + fi, gy = y0 - 2 + sj;
                if (gy < 0 || gy >= ny) continue;
                if (!per && (g < 0 || g > nx)) continue;
                const int gw = per ? ((g % nx) + nx) % nx : g;
                const int border = per ? 0 : (g == 0 ? 1 : (g == nx ? 2 : 0));
                const bool ac = (gw < 2) || (gw > nx - 2);
                const RowGeo<T> rg = a.row[gy];
                const int lj = sj + 2;
                SC<T> m = in_cell(fi, lj), c0 = in_cell(fi + 1, lj), c1 = in_cell(fi + 2, lj), q = in_cell(fi + 3, lj);
                const T F = mass_flux<T, false>(m, c0, c1, q, border, ac, rg.ilx, T(1), T(1), rg.kA, rg.kA, rg.valx, dt, px);
                const bool lo = fi >= 1 && (per || g - 1 >= 0);
                const bool hi = fi <= SX - 1 && (per || g <= nx - 1);
                if (lo) s_am[sj * SX + fi - 1] -= F;
                if (hi) s_am[sj * SX + fi] += (border == 1) ? -F : F;
            }
            __syncthreads();
        }
        for (int par = 0; par < 2; par++) {
            constexpr int NFY = (SY + 2) / 2;
            for (int idx = tid; idx < NFY * SX; idx += NT) {
                const int k = idx / SX;
                const int si = idx - k * SX;
                const int fj = 2 * k + par;
                if (fj > SY) continue;
                const int g = y0 - 2 + fj;
                if (g < 0 || g > ny) continue;
                const int gx = x0 - 2 + si;
                if (!per && (gx < 0 || gx >= nx)) continue;
                const int border = g == 0 ? 1 : (g == ny ? 2 : 0);
                const bool ac = (g < 2) || (g > ny - 2);
                const FaceGeoY<T> fg = a.fy[g];
                const int li = si + 2;
                SC<T> m = in_cell(li, fj), c0 = in_cell(li, fj + 1), c1 = in_cell(li, fj + 2), q = in_cell(li, fj + 3);
                m.qn = s_qy[fj * IX + li]; c0.qn = s_qy[(fj + 1) * IX + li];
                c1.qn = s_qy[(fj + 2) * IX + li]; q.qn = s_qy[(fj + 3) * IX + li];
                const T F = mass_flux<T, true>(m, c0, c1, q, border, ac, fg.len, fg.cos0, fg.cos1, fg.k0, fg.k1, fg.valy, dt, py);
                const bool lo = fj >= 1 && g - 1 >= 0;
                const bool hi = fj <= SY - 1 && g <= ny - 1;
                if (lo) s_am[(fj - 1) * SX + si] = Num<T>::mad(-F, fg.ic0, s_am[(fj - 1) * SX + si]);
                if (hi) s_am[fj * SX + si] = Num<T>::mad((border == 1) ? -F : F, fg.ic1, s_am[fj * SX + si]);
            }
            __syncthreads();
        }

        // ---- h* (as e*) on tile+2, with the sponge of src/GPU/Volumen_kernel.cu:96-158 ----
        const bool any_sponge = (a.sp_izq | a.sp_der | a.sp_sup | a.sp_inf) != 0;
        for (int idx = tid; idx < SY * SX; idx += NT) {
            const int sj = idx / SX, si = idx - sj * SX;
            const int gx = x0 - 2 + si, gy = y0 - 2 + sj;
            const int io = (sj + 2) * IX + si + 2;
            T e = s_e[io];
            if (gy >= 0 && gy < ny && (per || (gx >= 0 && gx < nx))) {
                const RowGeo<T> rg = a.row[gy];
                const T D = s_D[io];
                e = Num<T>::mad(dt * rg.invA, s_am[idx], e);
                if (!(e + D > T(0))) e = -D;
                if (any_sponge) e = sponge_eta(a, e, D, per ? ((gx % nx) + nx) % nx : gx, gy, rg.dxs, rg.dys);
            }
            s_es[idx] = e;
        }
    }
    for (int idx = tid; idx < TY * TX; idx += NT) {
        s_ax[idx] = T(0);
        s_ay[idx] = T(0);
        s_aw[idx] = T(0);
    }
    __syncthreads();

    // clamp of an e*-region coordinate onto the grid (periodic x needs none)
    auto cxs = [&](int si) {
        if (per) return si;
        int g = x0 - 2 + si;
        g = g < 0 ? 0 : (g >= nx ? nx - 1 : g);
        return g - (x0 - 2);
    };
    auto cys = [&](int sj) {
        int g = y0 - 2 + sj;
        g = g < 0 ? 0 : (g >= ny ? ny - 1 : g);
        return g - (y0 - 2);
    };
    auto st_cell = [&](int si, int sj) {
        SC<T> c;
        const int io = (sj + 2) * IX + si + 2;
        c.e = s_es[sj * SX + si];
        c.D = s_D[io];
        c.h = c.e + c.D;
        c.qn = s_qx[io];
        c.qt = s_qy[io];
        c.f = desing(c.h, a.eps_h);
        return c;
    };

    // ---- stage 2: momentum fluxes on the tile ----
    for (int par = 0; par < 2; par++) {
        constexpr int NFX = (TX + 2) / 2;
        for (int idx = tid; idx < NFX * TY; idx += NT) {
            const int tj = idx / NFX;
            const int fi = 2 * (idx - tj * NFX) + par;
            if (fi > TX) continue;
            const int g = x0 + fi, gy = y0 + tj;
            if (gy >= ny) continue;
            if (g > nx) continue;
            const int gw = per ? (g % nx) : g;
            const int border = per ? 0 : (g == 0 ? 1 : (g == nx ? 2 : 0));
            const bool ac = (gw < 2) || (gw > nx - 2);
            const RowGeo<T> rg = a.row[gy];
            const int sj = tj + 2;
            SC<T> m = st_cell(cxs(fi), sj), c0 = st_cell(cxs(fi + 1), sj);
            SC<T> c1 = st_cell(cxs(fi + 2), sj), q = st_cell(cxs(fi + 3), sj);
            const MF<T> f = mom_flux<T, false>(m, c0, c1, q, border, ac, rg.ilx, T(1), T(1), T(1), T(1), rg.valx, rg.kdx, dt, px);
            const bool lo = fi >= 1;
            const bool hi = fi <= TX - 1 && (per || g <= nx - 1);
            if (lo) {
                const int o = tj * TX + fi - 1;
                s_ax[o] -= f.mn;
                s_ay[o] -= f.mt;
                s_aw[o] += f.w;
            }
            if (hi) {
                const int o = tj * TX + fi;
                if (border == 1) { s_ax[o] -= f.mn; s_ay[o] -= f.mt; }
                else { s_ax[o] -= f.pn; s_ay[o] -= f.pt; }
                s_aw[o] += f.w;
            }
        }
        __syncthreads();
    }
    for (int par = 0; par < 2; par++) {
        constexpr int NFY = (TY + 2) / 2;
        for (int idx = tid; idx < NFY * TX; idx += NT) {
            const int k = idx / TX;
            const int ti = idx - k * TX;
            const int fj = 2 * k + par;
            if (fj > TY) continue;
            const int g = y0 + fj, gx = x0 + ti;
            if (g > ny || gx >= nx) continue;
            const int border = g == 0 ? 1 : (g == ny ? 2 : 0);
            const bool ac = (g < 2) || (g > ny - 2);
            const FaceGeoY<T> fg = a.fy[g];
            const int si = ti + 2;
            SC<T> m = st_cell(si, cys(fj)), c0 = st_cell(si, cys(fj + 1));
            SC<T> c1 = st_cell(si, cys(fj + 2)), q = st_cell(si, cys(fj + 3));
            // y faces: normal = qy, tangential = qx
            T t0;
            t0 = m.qn; m.qn = m.qt; m.qt = t0;
            t0 = c0.qn; c0.qn = c0.qt; c0.qt = t0;
            t0 = c1.qn; c1.qn = c1.qt; c1.qt = t0;
            t0 = q.qn; q.qn = q.qt; q.qt = t0;
            const MF<T> f = mom_flux<T, true>(m, c0, c1, q, border, ac, fg.len, fg.cos0, fg.cos1, fg.ic0, fg.ic1,
                                              fg.valy, fg.kdy, dt, py);
            const bool lo = fj >= 1;
            const bool hi = fj <= TY - 1 && g <= ny - 1;
            if (lo) {
                const int o = (fj - 1) * TX + ti;
                s_ay[o] -= f.mn;
                s_ax[o] -= f.mt;
                s_aw[o] += f.w;
            }
            if (hi) {
                const int o = fj * TX + ti;
                if (border == 1) { s_ay[o] -= f.mn; s_ax[o] -= f.mt; }
                else { s_ay[o] -= f.pn; s_ax[o] -= f.pt; }
                s_aw[o] += f.w;
            }
        }
        __syncthreads();
    }

    // ---- cell update: src/GPU/Volumen_kernel.cu:160-221, metrics src/GPU/Metrics.cu ----
    const bool any_sponge2 = (a.sp_izq | a.sp_der | a.sp_sup | a.sp_inf) != 0;
    for (int idx = tid; idx < TY * TX; idx += NT) {
        const int tj = idx / TX, ti = idx - tj * TX;
        const int gx = x0 + ti, gy = y0 + tj;
        if (gx >= nx || gy >= ny) continue;
        const RowGeo<T> rg = a.row[gy];
        const int io = (tj + 4) * IX + ti + 4;
        const int so = (tj + 2) * SX + ti + 2;
        const T D = s_D[io];
        const T eo = s_e[io];
        const T qxo = s_qx[io], qyo = s_qy[io];
        const int64_t go = (int64_t)gy * P + gx;
        metrics_cell(a, go, eo, eo + D, tnow);  // from the state before the step
        const T aw = s_aw[idx];
        const T dtc = (aw < Num<T>::eps) ? Num<T>::big : Num<T>::div(T(2) * a.cfl * rg.A, aw);
        dtmin = fminf(dtmin, (float)dtc);
        const T val = dt * rg.invA;
        T qx = Num<T>::mad(val, s_ax[idx], qxo);
        T qy = Num<T>::mad(val, s_ay[idx], qyo);
        if (any_sponge2) sponge_q(a, gx, gy, qx, qy);
        const T es = s_es[so];
        const T hs = es + D;
        if (hs <= a.hpos) near_dry(hs, qxo, qyo, qx, qy, dt, a.mf0, a.vmax, a.eps_h);
        bad = bad || bad_cell(es, qx, qy);
        oe[go] = es;
        oqx[go] = qx;
        oqy[go] = qy;
    }
    raise_health(a, bad);
    dtmin = wave_min(dtmin);
    if ((tid & 63) == 0) s_red[tid >> 6] = dtmin;
    __syncthreads();
    if (tid == 0) {
        float m = s_red[0];
#pragma unroll
        for (int w = 1; w < NT / 64; w++) m = fminf(m, s_red[w]);
        atomicMin(&a.dtmin[dt_at(slot, 0)], __float_as_uint(m));
    }
}

}  // namespace thy
