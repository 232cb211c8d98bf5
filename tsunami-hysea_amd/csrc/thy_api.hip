// thy_api.hip -- C-ABI of libthysea.so (declared in include/thysea.h).
//
// Host side of the MI355X time-stepping library: device memory layout, launch
// dispatch, graph-captured step batches and the driver loop with the semantics of
// the reference's shallowWater (src/GPU/ShallowWater.cu:182-500).
//
// Device layout (structure of arrays, row pitch a multiple of 64 elements):
//   e[2], qx[2], qy[2]   ping-pong state (eta relative to sea level, discharges)
//   D                    still-water depth (changes only at Okada time)
//   emax, arr            metrics (arr, when arrival times are tracked: eta_ini until a
//                        cell arrives, then its arrival time; see Arrival<T>)
//   row[ny], fy[ny+1]    per-row / per-face-row geometry constants
//   ctrl[3], dtmin       device control blocks and dt minima (3 launch slots of
//                        kDtParts partials, dt_at)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/thysea.h"
#include "thy_aux.hpp"
#include "thy_pool.hpp"
#include "thy_sweep.hpp"

using namespace thy;

// the device array behind thy_mass_noalloc's partial sums (see kScratchPartials)
static __device__ __align__(256) double scratch[1024];

struct thy_ctx {
    int nx = 0, ny = 0, pitch = 0;
    int device = -1;    // the device the context lives on
    thy_params prm{};
    int prec = 0;       // 0 float, 1 double
    size_t esz = 4;
    std::vector<double> lon, lat, area, cosphi, ancho, alto;
    void *e[2] = {nullptr, nullptr}, *qx[2] = {nullptr, nullptr}, *qy[2] = {nullptr, nullptr};
    void *D = nullptr, *emax = nullptr, *arr = nullptr;
    void *state = nullptr;   // one allocation holding e, qx, qy (x2), D, emax and arr
    size_t state_bytes = 0;
    void *row = nullptr, *fy = nullptr;
    double *area_d = nullptr;
    Ctrl *ctrl = nullptr;
    unsigned int *dtmin = nullptr;
    unsigned int *scratch_u = nullptr;  // [0]: initial dt, [1]: okada max
    double *ftime = nullptr;
    int nf_cap = 0;
    int *gpos = nullptr;
    float *gbuf = nullptr, *gtime = nullptr;
    int ng = 0, gcap = 0;
    double *okada_def = nullptr;
    double *mass_partial = nullptr;
    float *outbuf = nullptr;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;          // general launch beside the deep-water launch
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // two replays in flight: start / end / control-copied events, pinned copies of the
    // control block, the last launch's dt partials and the health flag after each replay
    hipEvent_t rev[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    Ctrl *h_snap = nullptr;
    unsigned int *h_aux = nullptr;     // [2]: health flag
    unsigned int *h_dt = nullptr;      // [2][kDtSlot]: dt partials of one launch slot
    int slot = 0;
    int cur = 0;                 // host mirror of ctrl.cur after a sync
    int tx = 64, ty = 16;
    int chunk_rows = 1;          // tile (strip) rows per XCD chunk of the step kernel's order
    int kernel = 0;              // 0 row sweep (deep-water + general split), 1 2D tiles, 2 row sweep general only
    int srows = 24;              // row sweep: rows per strip
    int gsrows = 32;             //   rows per general strip beside the deep-water launch
    int yoff = 0;                //   row offset of the split launches' strip rows
    int grows_req = 0;           //   requested general strip height (0: the default rule)
    int tuning = 0;              // launches recorded for timing replays publish no control block
    int ncu = 256;               // compute units of the device
    int *order = nullptr;        // row sweep: workgroup -> strip block, costliest first
    int order_n = 0;
    std::vector<int> order_h, cls_h;  // host copies: every block costliest first, and the classes
    // host lists for batched runs, strips of srows rows: the deep-water ones, and the
    // general ones with their cost classes; the general strips also cut into strips of
    // gsrows rows (costliest first) with their classes
    std::vector<int> bdeep_h, bgen_h, bgcls_h, bsub_h, bsubcls_h;
    int n_order = 0;             // blocks in `order` (the general launch's share)
    int *deep = nullptr;         // deep-water strips (n_deep of them)
    int *gen = nullptr;          // general strips beside them (n_gen)
    int n_deep = 0, n_gen = 0, list_cap = 0;
    unsigned int *fb_count = nullptr;  // 3 slots
    int *fb_list = nullptr;            // [3][list_cap]
    unsigned int *health = nullptr;    // non-finite flag raised by the step kernels
    // captured sequences of gk steps, one per launch slot a replay starts at
    hipGraphExec_t gexec[3] = {nullptr, nullptr, nullptr};
    int gk = 0;
    double last_ms = 0.0;
    std::vector<float> g_time, g_eta, g_u, g_v;
    int nsamples = 0;
    bool gauges_stale = false;   // the device holds samples not yet copied to g_*
    // host copy of the control block in launch slot `slot`, when the host knows it
    // (written by the host, or collected after a replay); saves a synchronous read
    Ctrl ctrl_h{};
    bool ctrl_h_valid = false;
    // schedule of the current run (kept so that thy_run_more can continue it)
    std::vector<thy_fault> faults;
    std::vector<double> ftimes;
    int crop_flag = 0;
    double crop_value = 0.0, save_interval = -1.0, next_save = INFINITY;
    int K = 24, nsnap = 0;
    thy_snapshot_fn on_snapshot = nullptr;
    thy_okada_fn on_okada = nullptr;
    thy_step_fn on_step = nullptr;
    void *user = nullptr;
    bool run_active = false;
    bool batch_ready = false;    // prepared by thy_batch_prepare, not yet admitted
    thy_run_result prep_res{};   // its run result so far (initial dt)
    double *steplog = nullptr;
    int logcap = 4096;
    int64_t logged = 0;  // iterations already delivered to on_step
    std::string err;
};

namespace {

constexpr int kMassBlocks = 1024;

// Scratch of thy_mass_noalloc: the kMassBlocks workgroup partial sums of the mass kernel
// (8 KiB, whatever the grid size: the mass kernel always runs kMassBlocks workgroups).
// Each device holds its own copy until the process ends.  Its device address is taken
// once per device, with that device current, when thy_create first runs there; host code
// uses only that address (the array's name is a host-side stand-in, not a device address).
constexpr int kScratchPartials = 1024;
static_assert(kMassBlocks <= kScratchPartials, "the mass partials must fit the scratch");
static_assert(sizeof(scratch) == kScratchPartials * sizeof(double), "scratch size");
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mutex;
double *g_scratch_ptr[kScratchDevices] = {};

int fail(thy_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

// captured step sequences hold the dispatch tables, grid sizes and launch slots of the
// time they were recorded
void drop_graphs(thy_ctx *c) {
    for (hipGraphExec_t &g : c->gexec)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    c->gk = 0;
}

// A context's copies and fills go through its own stream and are complete on return.
// The null stream would order them against every blocking stream, one that another
// thread is capturing into a graph included, and so invalidate that capture.
hipError_t ctx_copy(thy_ctx *c, void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, c->stream);
    return e == hipSuccess ? hipStreamSynchronize(c->stream) : e;
}

hipError_t ctx_fill(thy_ctx *c, void *dst, size_t bytes) {
    hipError_t e = hipMemsetAsync(dst, 0, bytes, c->stream);
    return e == hipSuccess ? hipStreamSynchronize(c->stream) : e;
}

// Entry points run with the calling thread's stream-capture mode relaxed.  Their
// allocations and synchronous copies touch only their own context's or batch's
// streams, none of which another thread can be capturing; under the default (global)
// mode the runtime would invalidate any capture in progress on another thread, such as
// a batch's step sequence being recorded while a worker thread sets up the next context.
struct RelaxedCapture {
    hipStreamCaptureMode m = hipStreamCaptureModeRelaxed;
    RelaxedCapture() { (void)hipThreadExchangeStreamCaptureMode(&m); }
    ~RelaxedCapture() { (void)hipThreadExchangeStreamCaptureMode(&m); }
    RelaxedCapture(const RelaxedCapture &) = delete;
    RelaxedCapture &operator=(const RelaxedCapture &) = delete;
};

#define HIPCHK(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(ctx, e_ == hipErrorOutOfMemory ? THY_ERR_DEVICE_MEMORY : THY_ERR_HIP, \
                        std::string(#call) + ": " + hipGetErrorString(e_));            \
    } while (0)

struct TileShape { int tx, ty; };

// tile shapes compiled into the library (the autotuner chooses among these)
template <typename T> bool tile_supported(int tx, int ty);
template <> bool tile_supported<float>(int tx, int ty) {
    return (tx == 64 && ty == 16) || (tx == 64 && ty == 8) || (tx == 32 && ty == 16) || (tx == 128 && ty == 8) ||
           (tx == 64 && ty == 12) || (tx == 32 && ty == 8) || (tx == 32 && ty == 24);
}
template <> bool tile_supported<double>(int tx, int ty) { return (tx == 32 && ty == 8) || (tx == 64 && ty == 4); }

template <typename T>
StepArgs<T> make_args(thy_ctx *c) {
    StepArgs<T> a;
    memset(&a, 0, sizeof(a));
    const thy_params &p = c->prm;
    a.nx = c->nx; a.ny = c->ny; a.pitch = c->pitch;
    a.periodic = (std::fabs(p.borde_izq - 2.0) < 2.220446049250313e-16 && std::fabs(p.borde_der - 2.0) < 2.220446049250313e-16) ? 1 : 0;
    if (c->kernel == 0 || c->kernel == 2) {
        const int nseg = (c->nx + SW_OUT - 1) / SW_OUT;
        a.tiles_x = (nseg + SW_WAVES - 1) / SW_WAVES;
        a.tiles_y = (c->ny + c->srows - 1) / c->srows;
        a.srows = c->srows;
        a.gsrows = c->gsrows;
        a.yoff = c->yoff;
        a.order = c->order;
        a.n_order = c->n_order;
        a.nseg = nseg;
        a.deep = c->n_deep > 0 ? c->deep : nullptr;
        a.n_deep = c->n_deep;
        a.gen = c->gen;
        a.n_gen = c->n_gen;
        a.fb_count = c->fb_count;
        a.fb_list = c->fb_list;
    } else {
        a.tiles_x = (c->nx + c->tx - 1) / c->tx;
        a.tiles_y = (c->ny + c->ty - 1) / c->ty;
    }
    a.n_tiles = a.tiles_x * a.tiles_y;
    a.chunk = a.tiles_x * c->chunk_rows;
    a.sp_sup = p.sponge_sup; a.sp_inf = p.sponge_inf; a.sp_izq = p.sponge_izq; a.sp_der = p.sponge_der;
    a.track_arrival = p.track_arrival;
    a.n_gauges = c->ng;
    a.bsup = (T)p.borde_sup; a.binf = (T)p.borde_inf; a.bizq = (T)p.borde_izq; a.bder = (T)p.borde_der;
    a.cfl = (T)p.cfl; a.mf0 = (T)p.mf0; a.vmax = (T)p.vmax; a.eps_h = (T)p.eps_h; a.hpos = (T)p.hpos;
    a.cvis = (T)p.cvis; a.dif_at = (T)p.dif_at; a.sea_level = (T)p.sea_level;
    a.Href = (T)p.H; a.Q_over_H = (T)(p.Q / p.H);
    for (int b = 0; b < 2; b++) {
        a.e[b] = (T *)c->e[b]; a.qx[b] = (T *)c->qx[b]; a.qy[b] = (T *)c->qy[b];
    }
    a.D = (const T *)c->D;
    a.emax = (T *)c->emax;
    a.arr = (T *)c->arr;
    a.block = (const char *)c->state;
    a.block_bytes = (int)std::min(c->state_bytes, (size_t)INT32_MAX);  // kernel 0 only when it fits
    a.row = (const RowGeo<T> *)c->row;
    a.fy = (const FaceGeoY<T> *)c->fy;
    a.ctrl = c->ctrl;
    a.dtmin = c->dtmin;
    a.ftime = c->ftime;
    a.gpos = c->gpos;
    a.gbuf = c->gbuf;
    a.gtime = c->gtime;
    a.steplog = c->steplog;
    a.logcap = c->logcap;
    a.health = c->health;
    a.no_publish = c->tuning;
    return a;
}

template <typename T>
hipError_t launch_step_t(thy_ctx *c, int slot) {
    StepArgs<T> a = make_args<T>(c);
    const int n_chunks = (a.tiles_y + c->chunk_rows - 1) / c->chunk_rows;
    dim3 grid(((n_chunks + 7) / 8) * 8 * a.chunk), block(256);
    if (c->kernel == 0 || c->kernel == 2) {
        {
            if (a.deep) {
                // deep-water blocks on the main stream and the general blocks on the side
                // stream run side by side; after the join, the hand-backs of this step
                hipError_t e;
                if ((e = hipEventRecord(c->ev_fork, c->stream)) != hipSuccess) return e;
                if ((e = hipStreamWaitEvent(c->side, c->ev_fork, 0)) != hipSuccess) return e;
                const int g_deep = (a.n_deep + SW_WAVES - 1) / SW_WAVES;
                if (a.n_gen > 0) {
                    hipLaunchKernelGGL((k_sweep<T>), dim3((a.n_gen + SW_WAVES - 1) / SW_WAVES), dim3(64 * SW_WAVES), 0,
                                       c->side, a, slot, 1);
                    if ((e = hipGetLastError()) != hipSuccess) return e;
                }
                hipLaunchKernelGGL((k_sweep_deep<T>), dim3(g_deep), dim3(64 * SW_WAVES), 0, c->stream, a, slot);
                if ((e = hipGetLastError()) != hipSuccess) return e;
                if ((e = hipEventRecord(c->ev_join, c->side)) != hipSuccess) return e;
                if ((e = hipStreamWaitEvent(c->stream, c->ev_join, 0)) != hipSuccess) return e;
                hipLaunchKernelGGL((k_sweep<T>), dim3(g_deep), dim3(64 * SW_WAVES), 0, c->stream, a, slot, 2);
                return hipGetLastError();
            }
        }
        if (a.order) grid = dim3(std::max(1, a.n_order));
        hipLaunchKernelGGL((k_sweep<T>), grid, dim3(64 * SW_WAVES), 0, c->stream, a, slot, 0);
        return hipGetLastError();
    }
#define THY_TILE(TX_, TY_)                                                                            \
    if (c->tx == TX_ && c->ty == TY_) {                                                              \
        hipLaunchKernelGGL((k_step<T, TX_, TY_>), grid, block, 0, c->stream, a, slot);               \
        return hipGetLastError();                                                                    \
    }
    if constexpr (sizeof(T) == 4) {
        THY_TILE(64, 16)
        THY_TILE(64, 8)
        THY_TILE(32, 16)
        THY_TILE(128, 8)
        THY_TILE(64, 12)
        THY_TILE(32, 8)
        THY_TILE(32, 24)
    } else {
        THY_TILE(32, 8)
        THY_TILE(64, 4)
    }
#undef THY_TILE
    return hipErrorInvalidValue;
}

hipError_t launch_step(thy_ctx *c) {
    c->ctrl_h_valid = false;  // the launch publishes a control block the host has not seen
    hipError_t e = c->prec ? launch_step_t<double>(c, c->slot) : launch_step_t<float>(c, c->slot);
    c->slot = (c->slot + 1) % 3;
    return e;
}

int grid_1d(int64_t n) {
    int64_t b = (n + 255) / 256;
    return (int)std::min<int64_t>(std::max<int64_t>(b, 1), 16384);
}

template <typename T>
void fill_geometry(thy_ctx *c, std::vector<RowGeo<T>> &rows, std::vector<FaceGeoY<T>> &fys) {
    const int ny = c->ny;
    // the 2s numerical-diffusion coefficient as the kernels compute it, in T
    const T cd2s = T(0.5) * (T(1) - (T)c->prm.cvis);
    rows.resize(ny);
    fys.resize(ny + 1);
    for (int j = 0; j < ny; j++) {
        const double A = c->area[j], cs = c->cosphi[j], al = c->alto[j];
        RowGeo<T> r;
        r.A = (T)A;
        r.invA = (T)(1.0 / A);
        // alto * (1/cos) rounded in T, as the device would multiply the two
        r.ilx = (T)al * (T)(1.0 / cs);
        r.valx = (T)((2.0 * (al / cs)) / (A + A));
        r.kdx = cd2s * r.valx * r.ilx;
        r.dxs = (T)c->ancho[j];
        r.dys = (T)al;
        r.kA = T(0.25) * r.A;
        rows[j] = r;
    }
    for (int g = 0; g <= ny; g++) {
        const int r0 = std::min(std::max(g - 1, 0), ny - 1), r1 = std::min(std::max(g, 0), ny - 1);
        FaceGeoY<T> f;
        const double len = c->ancho[g];
        f.len = (T)len;
        f.cos0 = (T)c->cosphi[r0];
        f.cos1 = (T)c->cosphi[r1];
        f.A0 = (T)c->area[r0];
        f.A1 = (T)c->area[r1];
        f.valy = (T)((2.0 * len) / (c->area[r0] + c->area[r1]));
        f.ic0 = (T)(1.0 / c->cosphi[r0]);
        f.ic1 = (T)(1.0 / c->cosphi[r1]);
        // in T arithmetic, each a single rounded operation as on the device
        f.gh = T(0.5) * (T)(f.cos0 + f.cos1);
        f.gd = (T)(f.cos1 - f.cos0);
        f.g1ic0 = (T)(f.cos1 * f.ic0);
        f.k0 = (T)(T(0.25) * f.cos0) * f.A0;
        f.k1 = (T)(T(0.25) * f.cos1) * f.A1;
        f.lc0 = (T)(f.len * f.cos0);
        f.kdy = cd2s * f.valy * f.len;
        f.pad = T(0);
        fys[g] = f;
    }
}

int pick_tile(thy_ctx *c) {
    int tx = c->prm.tile_x, ty = c->prm.tile_y;
    if (tx == 0 || ty == 0) {
        if (c->prec) { tx = 32; ty = 8; }
        else { tx = 32; ty = 16; }
    }
    bool ok = c->prec ? tile_supported<double>(tx, ty) : tile_supported<float>(tx, ty);
    if (!ok) return fail(c, THY_ERR_INVALID, "unsupported tile shape " + std::to_string(tx) + "x" + std::to_string(ty));
    c->tx = tx;
    c->ty = ty;
    if (c->prm.kernel < 0 || c->prm.kernel > 2)
        return fail(c, THY_ERR_INVALID, "unknown step kernel " + std::to_string(c->prm.kernel));
    c->kernel = c->prm.kernel;
    if (c->prm.strip_rows < 0 || c->prm.strip_rows > 65536)
        return fail(c, THY_ERR_INVALID, "strip_rows out of range: " + std::to_string(c->prm.strip_rows));
    if (c->prm.strip_rows > 0) c->srows = c->prm.strip_rows;
    if (c->prm.general_rows < 0 || c->prm.general_rows > 65536)
        return fail(c, THY_ERR_INVALID, "general_rows out of range: " + std::to_string(c->prm.general_rows));
    c->grows_req = c->prm.general_rows;
    // THYSEA_CHUNK_ROWS overrides the number of tile rows dealt to one XCD at a time
    if (const char *v = std::getenv("THYSEA_CHUNK_ROWS")) {
        const int r = std::atoi(v);
        if (r >= 1 && r <= 4096) c->chunk_rows = r;
    }
    return THY_OK;
}

// Dispatch order of the row-sweep workgroups.  A strip row whose stencil touches
// shallow water or land runs the general faces, about 2.5x the work of a deep row, so
// coastal strips are the longest-running workgroups; dispatching them first lets the
// deep ones fill in behind them instead of leaving a tail of coastal strips at the end
// of the launch.  Workgroups of equal cost class keep the chunked XCD order of the
// kernel.  The estimate uses the still-water depth: a row counts as coastal for a
// 64-column segment when any cell of it is shallower than 3 hpos (or dry).
// Rows per general strip for strips of S rows: the requested height when it divides S.
// Otherwise, when the general launch is bound by throughput (8-row strips would number
// more than 8 per CU), the largest divisor of S between 8 and min(20, S / 2), whose
// halos are a smaller share; when it is bound by the time of one strip (smaller grids),
// the smallest divisor of S from 4 up, the shortest walk.  S itself if none.
int general_rows_of(int S, int req, size_t n_general, int ncu) {
    if (req > 0 && S % req == 0) return req;
    const bool throughput = n_general * (size_t)((S + 7) / 8) > (size_t)8 * std::max(1, ncu);
    if (throughput) {
        for (int g = std::min(20, S / 2); g >= 8; g--)
            if (S % g == 0) return g;
    } else {
        for (int g = 4; g < S; g++)
            if (S % g == 0) return g;
    }
    return S;
}

// Row offset of the split launches' strip rows (strip row k covers rows k*S-yoff ..
// k*S-yoff+S-1, cut to the grid).  A strip row is deep-water capable only clear of the
// top and bottom border bands and sponges; the offset that makes the rows that are not
// thinnest, so that they cost the least general work (a general row about five times a
// deep-water row, halos included), is taken; ties go to the smaller offset.
int strip_row_offset(int ny, int S, int sps, int spn) {
    int best = 0;
    long best_cost = -1;
    for (int off = 0; off < S; off++) {
        long cost = 0;
        for (int y = -off; y < ny; y += S) {
            const int y0 = std::max(y, 0), y1 = std::min(y + S, ny);
            const bool ok = y0 - 4 >= 2 && y1 + 3 <= ny - 3 && y0 - 2 >= sps && y1 + 1 < ny - spn;
            cost += (long)(y1 - y0 + 8) * (ok ? 1 : 5);
        }
        if (best_cost < 0 || cost < best_cost) {
            best_cost = cost;
            best = off;
        }
    }
    return best;
}

template <typename T>
int build_order(thy_ctx *ctx, const T *D) {
    thy_ctx *c = ctx;
    if (c->kernel == 1) return THY_OK;
    const int nx = c->nx, ny = c->ny, P = c->pitch, S = c->srows;
    const int nseg = (nx + SW_OUT - 1) / SW_OUT;
    const int tiles_x = (nseg + SW_WAVES - 1) / SW_WAVES, tiles_y = (ny + S - 1) / S;
    const int n_tiles = tiles_x * tiles_y;
    const double thr = 3.0 * c->prm.hpos;
    // coastal[j * nseg + s]: row j has a shallow or dry cell within segment s's 64 columns
    std::vector<unsigned char> coastal((size_t)ny * nseg, 0);
    for (int j = 0; j < ny; j++) {
        for (int sgi = 0; sgi < nseg; sgi++) {
            const int xa = std::max(0, sgi * SW_OUT - SW_HALO), xb = std::min(nx, sgi * SW_OUT + SW_OUT + SW_HALO);
            bool any = false;
            for (int i = xa; i < xb && !any; i++) any = (double)D[(size_t)j * P + i] <= thr;
            coastal[(size_t)j * nseg + sgi] = any ? 1 : 0;
        }
    }
    // per strip (segment sgi of strip row ty): cost class, and whether the deep-water
    // launch may take it: rows y0-4 .. y1+3 and its 64 columns clear of the grid edges
    // (no clamping, no first-order band), of the sponges, and of coast
    // the deep-water path addresses each array through a buffer resource with a 32-bit
    // byte range
    const bool split = c->kernel == 0;  // (kernel 0: the state block fits 32-bit offsets)
    const int spi = c->prm.sponge_izq, spd = c->prm.sponge_der, sps = c->prm.sponge_sup, spn = c->prm.sponge_inf;
    // strips of the split launches: rows shifted up by yoff (the block order of the
    // launch without a split keeps unshifted tiles)
    const int yoff = split ? strip_row_offset(ny, S, sps, spn) : 0;
    c->yoff = yoff;
    const int strip_rows = (ny + yoff + S - 1) / S;
    const int n_strips = strip_rows * nseg;
    std::vector<int> scls(n_strips, 0);
    std::vector<unsigned char> sdeep(n_strips, 0);
    for (int ty = 0; ty < strip_rows; ty++) {
        const int y0 = std::max(ty * S - yoff, 0), y1 = std::min(ty * S - yoff + S, ny);
        const int rows = (y1 - y0) + 8;
        const bool yok = y0 - 4 >= 2 && y1 + 3 <= ny - 3 && y0 - 2 >= sps && y1 + 1 < ny - spn;
        for (int sgi = 0; sgi < nseg; sgi++) {
            const int x0 = sgi * SW_OUT;
            int nc = 0;
            for (int j = std::max(0, y0 - 4); j < std::min(ny, y1 + 4); j++) nc += coastal[(size_t)j * nseg + sgi];
            const bool xok = x0 - SW_HALO >= 2 && x0 + SW_OUT + SW_HALO - 1 <= nx - 3 && x0 - SW_HALO >= spi &&
                             x0 + SW_OUT + SW_HALO - 1 < nx - spd;
            scls[ty * nseg + sgi] = (int)(1.5 * nc / rows * 4.0 + 0.5);
            sdeep[ty * nseg + sgi] = (split && yok && xok && nc == 0) ? 1 : 0;
        }
    }
    std::vector<int> cls(n_tiles, 0);
    for (int t = 0; t < n_tiles; t++) {
        const int ty = t / tiles_x, tx = t - ty * tiles_x;
        const int sr = (ty * S + yoff) / S;  // the shifted strip row holding the tile's first row
        for (int w = 0; w < SW_WAVES; w++) {
            const int sgi = tx * SW_WAVES + w;
            if (sgi < nseg) cls[t] = std::max(cls[t], scls[sr * nseg + sgi]);
        }
    }
    // chunked XCD order as the base sequence (the kernel's order without a table)
    std::vector<int> base;
    base.reserve(n_tiles);
    const int chunk = tiles_x * c->chunk_rows;
    const int n_chunks = (tiles_y + c->chunk_rows - 1) / c->chunk_rows;
    const int grid = ((n_chunks + 7) / 8) * 8 * chunk;
    for (int bid = 0; bid < grid; bid++) {
        const int loc = bid >> 3, ch = loc / chunk;
        const int t = ((ch << 3) + (bid & 7)) * chunk + (loc - ch * chunk);
        if (t < n_tiles) base.push_back(t);
    }
    // General strips in the base order, costliest first, cut into strips of G rows
    // (items in units of G rows): short enough that the few slow coastal and border
    // strips do not set the launch's critical path, long enough that their halo rows
    // stay a modest share.
    size_t n_general = 0;
    for (int item = 0; item < n_strips; item++) n_general += sdeep[item] ? 0 : 1;
    const int G = general_rows_of(S, c->grows_req, n_general, c->ncu);
    c->gsrows = G;
    std::vector<int> deep_rm, gen, gcls;
    c->bgen_h.clear();
    c->bgcls_h.clear();
    // the general strips in the chunked XCD order over the shifted strip rows
    std::vector<int> sbase;
    {
        const int st = tiles_x * strip_rows;
        const int s_chunks = (strip_rows + c->chunk_rows - 1) / c->chunk_rows;
        const int s_grid = ((s_chunks + 7) / 8) * 8 * chunk;
        for (int bid = 0; bid < s_grid; bid++) {
            const int loc = bid >> 3, ch = loc / chunk;
            const int t = ((ch << 3) + (bid & 7)) * chunk + (loc - ch * chunk);
            if (t < st) sbase.push_back(t);
        }
    }
    for (int t : sbase) {
        const int ty = t / tiles_x, tx = t - ty * tiles_x;
        for (int w = 0; w < SW_WAVES; w++) {
            const int sgi = tx * SW_WAVES + w;
            if (sgi >= nseg) continue;
            const int item = ty * nseg + sgi;
            if (sdeep[item]) continue;
            c->bgen_h.push_back(item);
            c->bgcls_h.push_back(scls[item]);
            for (int sub = 0; sub < S / G; sub++) {
                const int y = ty * S + sub * G - yoff;  // the sub-strip's first row, before the cut at 0
                if (y + G <= 0 || y >= ny) continue;
                gen.push_back((ty * (S / G) + sub) * nseg + sgi);
                gcls.push_back(scls[item]);
            }
        }
    }
    {
        std::vector<int> idx(gen.size());
        for (size_t i = 0; i < idx.size(); i++) idx[i] = (int)i;
        std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return gcls[a] > gcls[b]; });
        std::vector<int> g2(gen.size()), c2(gen.size());
        for (size_t i = 0; i < idx.size(); i++) {
            g2[i] = gen[idx[i]];
            c2[i] = gcls[idx[i]];
        }
        gen.swap(g2);
        c->bsub_h = gen;
        c->bsubcls_h.swap(c2);
    }
    // Deep-water strips row-major, cut into 8 runs of (nearly) equal length, one per
    // XCD: the workgroup at list slot b runs on XCD b mod 8, so each XCD sweeps one
    // contiguous band and strips side by side share their halo columns in that XCD's
    // L2.  Runs are padded to whole workgroups with -1 (idle wavefronts).
    for (int item = 0; item < n_strips; item++)
        if (sdeep[item]) deep_rm.push_back(item);
    std::vector<int> deep;
    if (!deep_rm.empty()) {
        const int nd = (int)deep_rm.size(), q = nd / 8, rem = nd % 8;
        const int wmax = (q + (rem > 0 ? 1 : 0) + SW_WAVES - 1) / SW_WAVES;
        deep.assign((size_t)wmax * 8 * SW_WAVES, -1);
        for (int x = 0; x < 8; x++) {
            const int size = q + (x < rem ? 1 : 0), start = x * q + std::min(x, rem);
            for (int i = 0; i < size; i++)
                deep[((size_t)(i / SW_WAVES) * 8 + x) * SW_WAVES + i % SW_WAVES] = deep_rm[start + i];
        }
    }
    std::stable_sort(base.begin(), base.end(), [&](int a, int b) { return cls[a] > cls[b]; });
    c->order_h = base;
    c->cls_h = cls;
    c->bdeep_h = deep_rm;
    drop_graphs(c);  // the captured step sequences hold the old tables and grid sizes
    const int cap = std::max({n_strips, strip_rows * (S / G) * nseg, (int)deep.size()});
    if (c->list_cap < cap) {
        void *ptrs[] = {c->order, c->deep, c->gen, c->fb_list};
        for (void *p : ptrs)
            dev_free(p);
        c->order = c->deep = c->gen = c->fb_list = nullptr;
        HIPCHK(dev_alloc(&c->order, (size_t)cap * sizeof(int)));
        HIPCHK(dev_alloc(&c->deep, (size_t)cap * sizeof(int)));
        HIPCHK(dev_alloc(&c->gen, (size_t)cap * sizeof(int)));
        HIPCHK(dev_alloc(&c->fb_list, (size_t)3 * cap * sizeof(int)));
        c->list_cap = cap;
    }
    if (!c->fb_count) HIPCHK(dev_alloc(&c->fb_count, 3 * sizeof(unsigned int)));
    HIPCHK(ctx_fill(ctx, c->fb_count, 3 * sizeof(unsigned int)));
    c->n_order = n_tiles;
    c->n_deep = (int)deep.size();
    c->n_gen = (int)gen.size();
    HIPCHK(ctx_copy(ctx, c->order, base.data(), base.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!deep.empty()) HIPCHK(ctx_copy(ctx, c->deep, deep.data(), deep.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!gen.empty()) HIPCHK(ctx_copy(ctx, c->gen, gen.data(), gen.size() * sizeof(int), hipMemcpyHostToDevice));
    return THY_OK;
}

// rebuild the dispatch order from the depth on the device (after a strip-height change)
template <typename T>
int rebuild_order_t(thy_ctx *ctx) {
    std::vector<T> D((size_t)ctx->pitch * ctx->ny);
    HIPCHK(ctx_copy(ctx, D.data(), ctx->D, D.size() * sizeof(T), hipMemcpyDeviceToHost));
    return build_order<T>(ctx, D.data());
}

template <typename T>
int upload_t(thy_ctx *ctx, const double *s1, const double *s2) {
    const int nx = ctx->nx, ny = ctx->ny, P = ctx->pitch;
    const double hmin = ctx->prm.hmin;
    const size_t n = (size_t)P * ny, b = n * sizeof(T);
    ctx->batch_ready = false;  // a new state: a run prepared for it is void
    // e, D, qx, qy staged in pinned host memory (a cached block), copied asynchronously
    T *st = nullptr;
    HIPCHK(Pool::get().host_alloc((void **)&st, 4 * b));
    struct Staging {
        T *p;
        hipStream_t s;
        ~Staging() {
            (void)hipStreamSynchronize(s);  // the copies read it until they finish
            Pool::get().host_free(p);
        }
    } staging{st, ctx->stream};
    T *e = st, *D = st + n, *qx = st + 2 * n, *qy = st + 3 * n;
    for (int j = 0; j < ny; j++) {
        const size_t r = (size_t)j * nx, ro = (size_t)j * P;
        for (int i = 0; i < nx; i++) {
            const size_t c = r + i, o = ro + i;
            const double h = s1[2 * c], Hs = s1[2 * c + 1];
            if (!std::isfinite(h) || !std::isfinite(Hs) || !std::isfinite(s2[2 * c]) || !std::isfinite(s2[2 * c + 1]))
                return fail(ctx, THY_ERR_NONFINITE,
                            "non-finite input state at cell (" + std::to_string(i) + ", " + std::to_string(j) + ")");
            const double Dd = Hs + hmin;             // unshifted depth
            D[o] = (T)Dd;
            e[o] = (h > 0.0) ? (T)(h - Dd) : -(T)Dd;  // dry: e = -D exactly, so h = 0
            qx[o] = (T)s2[2 * c];
            qy[o] = (T)s2[2 * c + 1];
        }
        for (int i = nx; i < P; i++) e[ro + i] = D[ro + i] = qx[ro + i] = qy[ro + i] = T(0);
    }
    hipStream_t q = ctx->stream;
    HIPCHK(hipMemcpyAsync(ctx->e[0], e, b, hipMemcpyHostToDevice, q));
    HIPCHK(hipMemcpyAsync(ctx->D, D, b, hipMemcpyHostToDevice, q));
    HIPCHK(hipMemcpyAsync(ctx->qx[0], qx, b, hipMemcpyHostToDevice, q));
    HIPCHK(hipMemcpyAsync(ctx->qy[0], qy, b, hipMemcpyHostToDevice, q));
    HIPCHK(hipMemcpyAsync(ctx->e[1], ctx->e[0], b, hipMemcpyDeviceToDevice, q));
    HIPCHK(hipMemcpyAsync(ctx->qx[1], ctx->qx[0], b, hipMemcpyDeviceToDevice, q));
    HIPCHK(hipMemcpyAsync(ctx->qy[1], ctx->qy[0], b, hipMemcpyDeviceToDevice, q));
    ctx->cur = 0;
    HIPCHK(hipMemsetAsync(ctx->health, 0, sizeof(unsigned int), q));  // a new state is healthy until stepped
    const int r = build_order<T>(ctx, D);  // on the host while the copies run
    HIPCHK(hipStreamSynchronize(q));
    return r;
}

template <typename T>
int download_t(thy_ctx *ctx, double *s1, double *s2) {
    const int nx = ctx->nx, ny = ctx->ny, P = ctx->pitch;
    const size_t b = (size_t)P * ny * sizeof(T);
    std::vector<T> e((size_t)P * ny), D((size_t)P * ny), qx((size_t)P * ny), qy((size_t)P * ny);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx_copy(ctx, e.data(), ctx->e[ctx->cur], b, hipMemcpyDeviceToHost));
    HIPCHK(ctx_copy(ctx, D.data(), ctx->D, b, hipMemcpyDeviceToHost));
    HIPCHK(ctx_copy(ctx, qx.data(), ctx->qx[ctx->cur], b, hipMemcpyDeviceToHost));
    HIPCHK(ctx_copy(ctx, qy.data(), ctx->qy[ctx->cur], b, hipMemcpyDeviceToHost));
    for (int j = 0; j < ny; j++) {
        for (int i = 0; i < nx; i++) {
            const size_t c = (size_t)j * nx + i, o = (size_t)j * P + i;
            if (s1) {
                s1[2 * c] = (double)e[o] + (double)D[o];
                s1[2 * c + 1] = (double)D[o] - ctx->prm.hmin;
            }
            if (s2) {
                s2[2 * c] = (double)qx[o];
                s2[2 * c + 1] = (double)qy[o];
            }
        }
    }
    return THY_OK;
}

int read_ctrl(thy_ctx *ctx, Ctrl *c) {
    if (ctx->ctrl_h_valid) {
        *c = ctx->ctrl_h;
        return THY_OK;
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx_copy(ctx, c, &ctx->ctrl[ctx->slot], sizeof(Ctrl), hipMemcpyDeviceToHost));
    ctx->ctrl_h = *c;
    ctx->ctrl_h_valid = true;
    return THY_OK;
}

int write_ctrl(thy_ctx *ctx, const Ctrl &c) {
    ctx->ctrl_h_valid = false;
    HIPCHK(ctx_copy(ctx, &ctx->ctrl[ctx->slot], &c, sizeof(Ctrl), hipMemcpyHostToDevice));
    ctx->ctrl_h = c;
    ctx->ctrl_h_valid = true;
    return THY_OK;
}

// minimum over the dt partials of one launch slot, as copied to the host (kDtSlot words)
unsigned int dt_parts_min(const unsigned int *w) {
    unsigned int m = kDtNone;
    for (int p = 0; p < kDtParts; p++) m = std::min(m, w[dt_at(0, p)]);
    return m;
}

// the raw dt minimum of launch slot s
int read_dt_slot(thy_ctx *ctx, const int s, unsigned int &raw) {
    std::vector<unsigned int> w(kDtSlot);
    HIPCHK(ctx_copy(ctx, w.data(), ctx->dtmin + dt_at(s, 0), kDtSlot * sizeof(unsigned int), hipMemcpyDeviceToHost));
    raw = dt_parts_min(w.data());
    return THY_OK;
}

// every partial of launch slot s (all three slots for s < 0) back to "no cell reported"
int clear_dt_slots(thy_ctx *ctx, const int s) {
    const std::vector<unsigned int> v((s < 0 ? 3 : 1) * kDtSlot, kDtNone);
    HIPCHK(ctx_copy(ctx, ctx->dtmin + (s < 0 ? 0 : dt_at(s, 0)), v.data(), v.size() * sizeof(unsigned int),
                    hipMemcpyHostToDevice));
    return THY_OK;
}

int reset_dt_slots(thy_ctx *ctx) {
    int r;
    if ((r = clear_dt_slots(ctx, -1)) != THY_OK) return r;
    if (ctx->fb_count) HIPCHK(ctx_fill(ctx, ctx->fb_count, 3 * sizeof(unsigned int)));
    return THY_OK;
}

Ctrl idle_ctrl(thy_ctx *ctx) {
    Ctrl c;
    memset(&c, 0, sizeof(c));
    c.t = 0.0;
    c.t_end = 1e300;
    c.next_save = INFINITY;
    c.next_series = INFINITY;
    c.series_dt = 0.0;
    c.T = ctx->prm.T;
    c.cur = ctx->cur;
    c.gcap = ctx->gcap;
    return c;
}

template <typename T>
int apply_okada_t(thy_ctx *ctx, const thy_fault *f, int crop_flag, double crop_value, double *def_out) {
    const int64_t n = (int64_t)ctx->nx * ctx->ny;
    OkadaFault of;
    of.lon_c = f->lon_c; of.lat_c = f->lat_c; of.depth = f->depth_c; of.L = f->length; of.W = f->width;
    of.strike = f->strike; of.dip = f->dip; of.rake = f->rake; of.slip = f->slip; of.H = ctx->prm.H;
    of.lon0 = ctx->lon[0];
    of.dlon = (ctx->lon[ctx->nx - 1] - ctx->lon[0]) / (ctx->nx - 1);
    of.lat0 = ctx->lat[0];
    of.dlat = (ctx->lat[ctx->ny - 1] - ctx->lat[0]) / (ctx->ny - 1);
    HIPCHK(hipMemsetAsync(&ctx->scratch_u[1], 0, sizeof(unsigned int), ctx->stream));
    hipLaunchKernelGGL(k_okada, dim3(grid_1d(n)), dim3(256), 0, ctx->stream, of, ctx->nx, ctx->ny, ctx->okada_def,
                       &ctx->scratch_u[1]);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL((k_okada_apply<T>), dim3(grid_1d(n)), dim3(256), 0, ctx->stream, ctx->okada_def, ctx->nx,
                       ctx->ny, ctx->pitch, crop_flag, crop_value, &ctx->scratch_u[1], (T *)ctx->e[ctx->cur],
                       (T *)nullptr, (T *)ctx->D, (T *)ctx->arr);
    HIPCHK(hipGetLastError());
    if (def_out) {
        HIPCHK(hipMemcpyAsync(def_out, ctx->okada_def, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return THY_OK;
}

template <typename T>
int init_metrics_t(thy_ctx *ctx) {
    const int64_t n = (int64_t)ctx->nx * ctx->ny;
    hipLaunchKernelGGL((k_init_metrics<T>), dim3(grid_1d(n)), dim3(256), 0, ctx->stream, (const T *)ctx->e[ctx->cur],
                       (const T *)ctx->D, ctx->nx, ctx->ny, ctx->pitch, (T)ctx->prm.eps_h, (T *)ctx->emax,
                       (T *)ctx->arr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return THY_OK;
}

template <typename T>
int initial_dt_t(thy_ctx *ctx, double *dt) {
    const int64_t n = (int64_t)ctx->nx * ctx->ny;
    StepArgs<T> a = make_args<T>(ctx);
    unsigned int big = 0x7f7fffffu;
    HIPCHK(ctx_copy(ctx, &ctx->scratch_u[0], &big, sizeof(big), hipMemcpyHostToDevice));
    hipLaunchKernelGGL((k_initial_dt<T>), dim3(grid_1d(n)), dim3(256), 0, ctx->stream, a, ctx->cur, &ctx->scratch_u[0]);
    HIPCHK(hipGetLastError());
    unsigned int r;
    HIPCHK(ctx_copy(ctx, &r, &ctx->scratch_u[0], sizeof(r), hipMemcpyDeviceToHost));
    float f;
    memcpy(&f, &r, 4);
    *dt = (double)f;
    return THY_OK;
}

template <typename T>
int get_field_t(thy_ctx *ctx, int field, float *out) {
    const int64_t n = (int64_t)ctx->nx * ctx->ny;
    if (field == THY_ARRIVAL_TIMES && !ctx->arr) {
        for (int64_t i = 0; i < n; i++) out[i] = -1.0f;
        return THY_OK;
    }
    hipLaunchKernelGGL((k_output<T>), dim3(grid_1d(n)), dim3(256), 0, ctx->stream, field, (const T *)ctx->e[ctx->cur],
                       (const T *)ctx->qx[ctx->cur], (const T *)ctx->qy[ctx->cur], (const T *)ctx->D,
                       (const T *)ctx->emax, (const T *)ctx->arr, ctx->nx, ctx->ny, ctx->pitch, (T)ctx->prm.eps_h,
                       ctx->prm.H, ctx->prm.Q, ctx->prm.T, ctx->outbuf);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, ctx->outbuf, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return THY_OK;
}

template <typename T>
int update_metrics_t(thy_ctx *ctx, double t) {
    const int64_t n = (int64_t)ctx->nx * ctx->ny;
    hipLaunchKernelGGL((k_update_metrics<T>), dim3(grid_1d(n)), dim3(256), 0, ctx->stream,
                       (const T *)ctx->e[ctx->cur], (const T *)ctx->D, ctx->nx, ctx->ny, ctx->pitch,
                       (T)ctx->prm.eps_h, (T)ctx->prm.dif_at, (T)t, (T *)ctx->emax, (T *)ctx->arr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return THY_OK;
}

// Compensated FP64 sum of h * area * cos(phi): kMassBlocks workgroup partials in the
// device buffer `partial`, copied to the host array `host` (kMassBlocks entries).
template <typename T>
int mass_t(thy_ctx *ctx, double *partial, double *host, double *mass) {
    hipLaunchKernelGGL((k_mass<T>), dim3(kMassBlocks), dim3(256), 0, ctx->stream, (const T *)ctx->e[ctx->cur],
                       (const T *)ctx->D, (const RowGeo<T> *)ctx->row, ctx->area_d, ctx->nx, ctx->ny, ctx->pitch,
                       partial);
    HIPCHK(hipGetLastError());
    HIPCHK(ctx_copy(ctx, host, partial, kMassBlocks * sizeof(double), hipMemcpyDeviceToHost));
    double s = 0.0, comp = 0.0;
    for (int k = 0; k < kMassBlocks; k++) {
        const double x = host[k];
        double t = s + x;
        if (std::fabs(s) >= std::fabs(x)) comp += (s - t) + x;
        else comp += (x - t) + s;
        s = t;
    }
    *mass = s + comp;
    return THY_OK;
}

// Capture K consecutive step launches starting at launch slot s0 (slots s0, s0+1, ...
// mod 3) into one graph, and upload it.
int capture_steps(thy_ctx *ctx, int K, int s0) {
    hipGraph_t g;
    const int slot_saved = ctx->slot;
    ctx->slot = s0;
    HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < K; k++) {
        hipError_t e = launch_step(ctx);
        if (e != hipSuccess) {
            (void)hipStreamEndCapture(ctx->stream, &g);
            ctx->slot = slot_saved;
            return fail(ctx, THY_ERR_HIP, std::string("capture: ") + hipGetErrorString(e));
        }
    }
    ctx->slot = slot_saved;
    HIPCHK(hipStreamEndCapture(ctx->stream, &g));
    hipError_t e = hipGraphInstantiate(&ctx->gexec[s0], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(e);
    HIPCHK(hipGraphUpload(ctx->gexec[s0], ctx->stream));
    return THY_OK;
}

// The step sequences replays of K steps need from the current launch slot on: with K a
// multiple of 3 every replay starts at the same slot, otherwise the replays cycle
// through all three (so any step count can be replayed without idle launches).
int build_graphs(thy_ctx *ctx, int K) {
    if (ctx->gk != K) drop_graphs(ctx);
    ctx->gk = K;
    int r;
    for (int m = 0; m < 3; m++) {
        const int s0 = (ctx->slot + m * K) % 3;
        if (!ctx->gexec[s0] && (r = capture_steps(ctx, K, s0)) != THY_OK) return r;
        if (K % 3 == 0) break;
    }
    return THY_OK;
}

// The stop predicate holds for the control block of the next launch, as seen on the
// host after a replay: finalise its dt on the host (truncation at fault times, as the
// launch would have done) and write it back, instead of replaying a sequence of launches
// that would only find the run stopped.  raw: the previous step's dt minimum.
int finalise_on_host(thy_ctx *ctx, Ctrl &c, const unsigned int raw) {
    if (c.dt_valid) return THY_OK;
    float rf;
    memcpy(&rf, &raw, 4);
    int skip;
    const double dt = step_dt(c, ctx->ftimes.data(), rf, skip);
    if (c.steps > 0 && ctx->on_step) {  // the log entry the launch would have written
        const double ent[2] = {dt, c.t};
        HIPCHK(ctx_copy(ctx, ctx->steplog + 2 * ((c.steps - 1) % ctx->logcap), ent, sizeof(ent),
                        hipMemcpyHostToDevice));
    }
    c.dt = dt;
    c.skip = skip;
    c.dt_valid = 1;
    return write_ctrl(ctx, c);
}

// Advance until the stop predicate holds; graph replays of K steps.  After each replay
// the control block, the dt partials of its last launch and the health flag are copied
// to pinned memory behind it.  While the last replay's time, steps and mean dt say that the run goes on
// for two more full replays (before t_end, max_steps, the next fault and the next
// snapshot), the next replay is queued before the one in flight is waited for, so that
// the device does not idle between replays.  When the host sees the stop predicate hold
// after a replay it finalises the control block itself; a queued replay that finds the
// run stopped idles through its launches (the stop predicate holds in every one).
int advance(thy_ctx *ctx, int K, Ctrl *out, double *ms) {
    int r;
    if ((r = build_graphs(ctx, K)) != THY_OK) return r;
    Ctrl last;
    if ((r = read_ctrl(ctx, &last)) != THY_OK) return r;
    if (must_stop(last, ctx->ftimes.data())) {  // stopped before any step
        unsigned int raw = 0;
        if (!last.dt_valid && (r = read_dt_slot(ctx, (ctx->slot + 2) % 3, raw)) != THY_OK) return r;
        if ((r = finalise_on_host(ctx, last, raw)) != THY_OK) return r;
        *out = last;
        if (ms) *ms = 0.0;
        return THY_OK;
    }
    // (a health flag raised before this call stops the first launch on the device)
    double total = 0.0, dt_mean = 0.0;
    int head = 0, queued = 0;
    auto enqueue = [&]() -> hipError_t {
        const int i = (head + queued) % 2;
        hipError_t e;
        if ((e = hipEventRecord(ctx->rev[i][0], ctx->stream)) != hipSuccess) return e;
        if ((e = hipGraphLaunch(ctx->gexec[ctx->slot], ctx->stream)) != hipSuccess) return e;
        if ((e = hipEventRecord(ctx->rev[i][1], ctx->stream)) != hipSuccess) return e;
        ctx->slot = (ctx->slot + K) % 3;
        if ((e = hipMemcpyAsync(&ctx->h_snap[i], &ctx->ctrl[ctx->slot], sizeof(Ctrl), hipMemcpyDeviceToHost,
                                ctx->stream)) != hipSuccess)
            return e;
        if ((e = hipMemcpyAsync(ctx->h_dt + (size_t)i * kDtSlot, ctx->dtmin + dt_at((ctx->slot + 2) % 3, 0),
                                kDtSlot * sizeof(unsigned int), hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess)
            return e;
        if ((e = hipMemcpyAsync(&ctx->h_aux[i], ctx->health, 4, hipMemcpyDeviceToHost, ctx->stream)) !=
            hipSuccess)
            return e;
        if ((e = hipEventRecord(ctx->rev[i][2], ctx->stream)) != hipSuccess) return e;
        queued++;
        return hipSuccess;
    };
    auto goes_on = [&](const Ctrl &c) {
        if (!(dt_mean > 0.0)) return false;
        const double t2 = c.t + 2.5 * K * dt_mean;  // two replays, with a margin for dt to grow
        if (!(t2 < c.t_end) || t2 >= c.next_save) return false;
        if (c.max_steps > 0 && c.steps + 2 * (int64_t)K > c.max_steps) return false;
        if (c.k < c.nf && t2 * c.T >= ctx->ftimes[c.k]) return false;
        return true;
    };
    unsigned int raw = 0;
    auto collect = [&](Ctrl &c) -> hipError_t {
        hipError_t e = hipEventSynchronize(ctx->rev[head][2]);
        if (e != hipSuccess) return e;
        c = ctx->h_snap[head];
        raw = dt_parts_min(ctx->h_dt + (size_t)head * kDtSlot);
        if (ctx->h_aux[head]) c.nonfinite = 1;
        float f = 0.f;
        (void)hipEventElapsedTime(&f, ctx->rev[head][0], ctx->rev[head][1]);
        total += f;
        head ^= 1;
        queued--;
        return hipSuccess;
    };
    HIPCHK(enqueue());
    for (;;) {
        if (queued == 1 && goes_on(last)) HIPCHK(enqueue());
        Ctrl c;
        HIPCHK(collect(c));
        ctx->cur = c.cur;
        if (c.steps > last.steps) dt_mean = (c.t - last.t) / (double)(c.steps - last.steps);
        last = c;
        // After an executed step dt_valid is 0; it stays 1 only when a launch of this
        // replay found the stop predicate true (every later launch then idled).
        const bool host_stop = !c.dt_valid && queued == 0 && must_stop(c, ctx->ftimes.data());
        if (c.dt_valid || host_stop) {
            while (queued > 0) HIPCHK(collect(c));  // a queued replay idled: same control block
            ctx->cur = c.cur;
            if (host_stop) {
                if ((r = finalise_on_host(ctx, c, raw)) != THY_OK) return r;
            } else {
                ctx->ctrl_h = c;  // the control block the stopping launch left in this slot
                ctx->ctrl_h_valid = true;
            }
            *out = c;
            if (ms) *ms = total;
            return THY_OK;
        }
        if (queued == 0) HIPCHK(enqueue());
    }
}

// plain copy, 4 x 16 B per thread per block-chunk, chunks coalesced across the block
__global__ void __launch_bounds__(256) k_copy_probe(const float4 *__restrict__ a, float4 *__restrict__ b, int64_t n) {
    const int64_t chunk = 4 * 256;
    for (int64_t base = (int64_t)blockIdx.x * chunk; base < n; base += (int64_t)gridDim.x * chunk) {
        float4 v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int64_t i = base + k * 256 + threadIdx.x;
            if (i < n) v[k] = a[i];
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int64_t i = base + k * 256 + threadIdx.x;
            if (i < n) b[i] = v[k];
        }
    }
}

}  // namespace

extern "C" {

int thy_copy_probe(size_t bytes, int reps, double *ms_per_copy) {
    RelaxedCapture relaxed;
    thy_ctx *ctx = nullptr;
    const int64_t n = (int64_t)(bytes / sizeof(float4));
    if (n <= 0 || reps <= 0) return THY_ERR_INVALID;
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0, e1;
    HIPCHK(hipMalloc(&a, n * sizeof(float4)));
    HIPCHK(hipMalloc(&b, n * sizeof(float4)));
    HIPCHK(hipMemset(a, 0, n * sizeof(float4)));
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const int grid = (int)std::min<int64_t>((n + 1023) / 1024, 1 << 20);
    hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, 0, (const float4 *)a, (float4 *)b, n);
    HIPCHK(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; r++)
        hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, 0, (const float4 *)a, (float4 *)b, n);
    HIPCHK(hipEventRecord(e1, 0));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    if (ms_per_copy) *ms_per_copy = ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return THY_OK;
}

int thy_probe(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return 1;
    for (int d = 0; d < n; d++) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, d) == hipSuccess && std::strncmp(p.gcnArchName, "gfx950", 6) == 0) return 0;
    }
    return 2;
}

// arrays of the state block: e, qx, qy (two buffers each), D, emax and, with arrival
// tracking, arr
static int state_arrays(const thy_params *p) { return 8 + ((p && p->track_arrival) ? 1 : 0); }

size_t thy_footprint(int nx, int ny, const thy_params *p) {
    const size_t esz = (p && p->precision) ? 8 : 4;
    const size_t pitch = ((size_t)nx + 63) / 64 * 64;
    const size_t n = pitch * ny;
    size_t bytes = (size_t)state_arrays(p) * n * esz;
    bytes += (size_t)nx * ny * (8 + 4);                 // okada scratch (double) + output staging (float)
    bytes += (size_t)(ny + 1) * 64 * 2;                 // geometry
    return bytes;
}

int thy_create(const thy_grid *g, const thy_params *p, thy_ctx **out) {
    RelaxedCapture relaxed;
    if (!g || !p || !out || g->nx < 2 || g->ny < 2) return THY_ERR_INVALID;
    thy_ctx *ctx = new (std::nothrow) thy_ctx();
    if (!ctx) return THY_ERR_HOST_MEMORY;
    *out = ctx;
    ctx->nx = g->nx;
    ctx->ny = g->ny;
    ctx->prm = *p;
    ctx->prec = p->precision ? 1 : 0;
    ctx->esz = ctx->prec ? 8 : 4;
    ctx->pitch = (g->nx + 63) / 64 * 64;
    ctx->lon.assign(g->lon, g->lon + g->nx);
    ctx->lat.assign(g->lat, g->lat + g->ny);
    ctx->area.assign(g->area, g->area + g->ny);
    ctx->cosphi.assign(g->cosphi, g->cosphi + g->ny);
    ctx->ancho.assign(g->ancho, g->ancho + g->ny + 1);
    ctx->alto.assign(g->alto, g->alto + g->ny);
    int r = pick_tile(ctx);
    if (r != THY_OK) return r;
    // the row sweeps address every state array through one buffer resource over the
    // state block, with 32-bit byte offsets; a grid whose block is larger runs the tile
    // kernel
    ctx->state_bytes = (size_t)state_arrays(p) * ctx->pitch * ctx->ny * ctx->esz;
    if (ctx->kernel != 1 && ctx->state_bytes > (size_t)INT32_MAX) ctx->kernel = 1;
    HIPCHK(hipSetDevice(0));
    HIPCHK(Pool::get().get_streams(ctx->stream, ctx->side));
    {
        int dev = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&ctx->ncu, hipDeviceAttributeMultiprocessorCount, dev));
        ctx->device = dev;
        // the device address of this device's copy of the allocation-free scratch, once
        // per device, with the device current
        if (dev < kScratchDevices) {
            std::lock_guard<std::mutex> lock(g_scratch_mutex);
            if (!g_scratch_ptr[dev]) {
                void *p = nullptr;
                HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
                g_scratch_ptr[dev] = (double *)p;
            }
        }
    }
    HIPCHK(hipEventCreate(&ctx->ev0));
    HIPCHK(hipEventCreate(&ctx->ev1));
    HIPCHK(hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
    for (int i = 0; i < 2; i++) {
        HIPCHK(hipEventCreate(&ctx->rev[i][0]));
        HIPCHK(hipEventCreate(&ctx->rev[i][1]));
        HIPCHK(hipEventCreateWithFlags(&ctx->rev[i][2], hipEventDisableTiming));
    }
    HIPCHK(Pool::get().host_alloc((void **)&ctx->h_snap, 2 * sizeof(Ctrl)));
    HIPCHK(Pool::get().host_alloc((void **)&ctx->h_aux, 4 * sizeof(unsigned int)));
    HIPCHK(Pool::get().host_alloc((void **)&ctx->h_dt, 2 * kDtSlot * sizeof(unsigned int)));
    const size_t n = (size_t)ctx->pitch * ctx->ny, b = n * ctx->esz;
    HIPCHK(dev_alloc(&ctx->state, ctx->state_bytes));
    {
        char *blk = (char *)ctx->state;  // arrays of b bytes (whole 256-B rows) in a fixed order
        for (int k = 0; k < 2; k++) {
            ctx->e[k] = blk + (0 + k) * b;
            ctx->qx[k] = blk + (2 + k) * b;
            ctx->qy[k] = blk + (4 + k) * b;
        }
        ctx->D = blk + 6 * b;
        ctx->emax = blk + 7 * b;
        if (p->track_arrival) ctx->arr = blk + 8 * b;
    }
    HIPCHK(dev_alloc((void **)&ctx->okada_def, (size_t)ctx->nx * ctx->ny * sizeof(double)));
    HIPCHK(dev_alloc((void **)&ctx->outbuf, (size_t)ctx->nx * ctx->ny * sizeof(float)));
    HIPCHK(dev_alloc((void **)&ctx->ctrl, 3 * sizeof(Ctrl)));
    HIPCHK(dev_alloc((void **)&ctx->dtmin, 3 * kDtSlot * sizeof(unsigned int)));
    HIPCHK(dev_alloc((void **)&ctx->scratch_u, 4 * sizeof(unsigned int)));
    HIPCHK(dev_alloc((void **)&ctx->health, sizeof(unsigned int)));
    HIPCHK(ctx_fill(ctx, ctx->health, sizeof(unsigned int)));
    HIPCHK(dev_alloc((void **)&ctx->mass_partial, kMassBlocks * sizeof(double)));
    HIPCHK(dev_alloc((void **)&ctx->steplog, (size_t)ctx->logcap * 2 * sizeof(double)));
    HIPCHK(dev_alloc((void **)&ctx->area_d, ctx->ny * sizeof(double)));
    {
        // mass weight per row: area*cos(phi), the invariant of the discrete scheme
        std::vector<double> w(ctx->ny);
        for (int j = 0; j < ctx->ny; j++) w[j] = g->area[j] * g->cosphi[j];
        HIPCHK(ctx_copy(ctx, ctx->area_d, w.data(), ctx->ny * sizeof(double), hipMemcpyHostToDevice));
    }
    if (ctx->prec) {
        std::vector<RowGeo<double>> rows;
        std::vector<FaceGeoY<double>> fys;
        fill_geometry(ctx, rows, fys);
        HIPCHK(dev_alloc(&ctx->row, rows.size() * sizeof(rows[0])));
        HIPCHK(dev_alloc(&ctx->fy, fys.size() * sizeof(fys[0])));
        HIPCHK(ctx_copy(ctx, ctx->row, rows.data(), rows.size() * sizeof(rows[0]), hipMemcpyHostToDevice));
        HIPCHK(ctx_copy(ctx, ctx->fy, fys.data(), fys.size() * sizeof(fys[0]), hipMemcpyHostToDevice));
    } else {
        std::vector<RowGeo<float>> rows;
        std::vector<FaceGeoY<float>> fys;
        fill_geometry(ctx, rows, fys);
        HIPCHK(dev_alloc(&ctx->row, rows.size() * sizeof(rows[0])));
        HIPCHK(dev_alloc(&ctx->fy, fys.size() * sizeof(fys[0])));
        HIPCHK(ctx_copy(ctx, ctx->row, rows.data(), rows.size() * sizeof(rows[0]), hipMemcpyHostToDevice));
        HIPCHK(ctx_copy(ctx, ctx->fy, fys.data(), fys.size() * sizeof(fys[0]), hipMemcpyHostToDevice));
    }
    // zero state so that padding columns hold finite values
    for (int k = 0; k < 2; k++) {
        HIPCHK(ctx_fill(ctx, ctx->e[k], b));
        HIPCHK(ctx_fill(ctx, ctx->qx[k], b));
        HIPCHK(ctx_fill(ctx, ctx->qy[k], b));
    }
    HIPCHK(ctx_fill(ctx, ctx->D, b));
    HIPCHK(ctx_fill(ctx, ctx->emax, b));
    Ctrl c = idle_ctrl(ctx);
    for (int s = 0; s < 3; s++) HIPCHK(ctx_copy(ctx, &ctx->ctrl[s], &c, sizeof(c), hipMemcpyHostToDevice));
    if ((r = reset_dt_slots(ctx)) != THY_OK) return r;
    return THY_OK;
}

void thy_destroy(thy_ctx *ctx) {
    RelaxedCapture relaxed;
    if (!ctx) return;
    // no queued work may still use what goes back to the caches
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    drop_graphs(ctx);
    void *ptrs[] = {ctx->state, ctx->row, ctx->fy, ctx->area_d, ctx->ctrl, ctx->dtmin, ctx->scratch_u,
                    ctx->ftime, ctx->gpos, ctx->gbuf, ctx->gtime, ctx->okada_def, ctx->mass_partial, ctx->outbuf,
                    ctx->steplog, ctx->order, ctx->deep, ctx->gen, ctx->fb_count, ctx->fb_list, ctx->health};
    for (void *q : ptrs)
        dev_free(q);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 3; j++)
            if (ctx->rev[i][j]) (void)hipEventDestroy(ctx->rev[i][j]);
    Pool::get().host_free(ctx->h_snap);
    Pool::get().host_free(ctx->h_aux);
    Pool::get().host_free(ctx->h_dt);
    if (ctx->stream) Pool::get().put_streams(ctx->stream, ctx->side);
    delete ctx;
}

const char *thy_last_error(const thy_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int thy_upload_state(thy_ctx *ctx, const double *s1, const double *s2) {
    RelaxedCapture relaxed;
    if (!ctx || !s1 || !s2) return THY_ERR_INVALID;
    return ctx->prec ? upload_t<double>(ctx, s1, s2) : upload_t<float>(ctx, s1, s2);
}

int thy_download_state(thy_ctx *ctx, double *s1, double *s2) {
    RelaxedCapture relaxed;
    if (!ctx) return THY_ERR_INVALID;
    return ctx->prec ? download_t<double>(ctx, s1, s2) : download_t<float>(ctx, s1, s2);
}

int thy_init_metrics(thy_ctx *ctx) {
    RelaxedCapture relaxed;
    if (!ctx) return THY_ERR_INVALID;
    return ctx->prec ? init_metrics_t<double>(ctx) : init_metrics_t<float>(ctx);
}

int thy_update_metrics(thy_ctx *ctx, double t) {
    RelaxedCapture relaxed;
    if (!ctx) return THY_ERR_INVALID;
    return ctx->prec ? update_metrics_t<double>(ctx, t) : update_metrics_t<float>(ctx, t);
}

int thy_apply_okada(thy_ctx *ctx, const thy_fault *f, int crop_flag, double crop_value, double *def_out) {
    RelaxedCapture relaxed;
    if (!ctx || !f || crop_flag < 0 || crop_flag > 2) return THY_ERR_INVALID;
    return ctx->prec ? apply_okada_t<double>(ctx, f, crop_flag, crop_value, def_out)
                     : apply_okada_t<float>(ctx, f, crop_flag, crop_value, def_out);
}

int thy_initial_dt(thy_ctx *ctx, double *dt) {
    RelaxedCapture relaxed;
    if (!ctx || !dt) return THY_ERR_INVALID;
    return ctx->prec ? initial_dt_t<double>(ctx, dt) : initial_dt_t<float>(ctx, dt);
}

int thy_autotune(thy_ctx *ctx, int reps, int *strip_rows, int *general_rows, double *ms_best) {
    RelaxedCapture relaxed;
    if (!ctx) return THY_ERR_INVALID;
    if (ctx->kernel == 1) return fail(ctx, THY_ERR_INVALID, "autotune covers the row-sweep kernel");
    if (ctx->order_h.empty()) return fail(ctx, THY_ERR_INVALID, "autotune needs an uploaded state");
    if (reps <= 0) reps = 5;
    int r;
    // the run's control blocks and dt slots are restored afterwards
    Ctrl saved[3];
    std::vector<unsigned int> saved_dt(3 * kDtSlot);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx_copy(ctx, saved, ctx->ctrl, sizeof(saved), hipMemcpyDeviceToHost));
    HIPCHK(ctx_copy(ctx, saved_dt.data(), ctx->dtmin, saved_dt.size() * sizeof(unsigned int), hipMemcpyDeviceToHost));
    const int saved_slot = ctx->slot;
    double dt;
    if ((r = thy_initial_dt(ctx, &dt)) != THY_OK) return r;
    Ctrl c = idle_ctrl(ctx);
    c.t = saved[saved_slot].t;
    c.dt = dt;
    c.dt_valid = 1;
    c.cur = ctx->cur;
    // (strip rows, general strip rows)
    const int cand[][2] = {{16, 4}, {16, 8}, {24, 4}, {24, 8}, {24, 12}, {32, 8}, {32, 16}, {36, 18}, {40, 20}, {48, 16}};
    int best = ctx->srows, best_g = ctx->grows_req;
    double best_ms = 1e30;
    for (const auto &sg : cand) {
        const int S = sg[0];
        if (S > ctx->ny && S != cand[0][0]) continue;
        ctx->srows = S;
        ctx->grows_req = sg[1];
        r = ctx->prec ? rebuild_order_t<double>(ctx) : rebuild_order_t<float>(ctx);
        if (r != THY_OK) return r;
        // kSeq steps captured into one graph and replayed, as a run replays its steps.
        // The launches cycle through the three launch slots as in a run (so each step's
        // hand-back list starts empty), but publish no control block: every slot holds
        // the same control block, each step reads the same state and redoes the same
        // step, and the run's state stays.
        constexpr int kSeq = 32;  // long enough that the replay launch is a small share
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        ctx->tuning = 1;
        ctx->slot = 0;
        HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        for (int q = 0; q < kSeq; q++) {
            hipError_t e = launch_step(ctx);
            if (e != hipSuccess) {
                ctx->tuning = 0;
                (void)hipStreamEndCapture(ctx->stream, &g);
                if (g) (void)hipGraphDestroy(g);
                return fail(ctx, THY_ERR_HIP, std::string("autotune capture: ") + hipGetErrorString(e));
            }
        }
        ctx->tuning = 0;
        HIPCHK(hipStreamEndCapture(ctx->stream, &g));
        hipError_t ei = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHK(ei);
        std::vector<float> t;
        const Ctrl c3[3] = {c, c, c};
        for (int k = 0; k <= reps; k++) {  // the first replay is a warm-up
            hipError_t e = hipSuccess;
            if ((e = ctx_copy(ctx, ctx->ctrl, c3, sizeof(c3), hipMemcpyHostToDevice)) == hipSuccess && ctx->fb_count)
                e = ctx_fill(ctx, ctx->fb_count, 3 * sizeof(unsigned int));
            if (e == hipSuccess) e = hipEventRecord(ctx->ev0, ctx->stream);
            if (e == hipSuccess) e = hipGraphLaunch(ge, ctx->stream);
            if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
            if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
            float ms = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
            if (e != hipSuccess) {
                (void)hipGraphExecDestroy(ge);
                return fail(ctx, THY_ERR_HIP, std::string("autotune replay: ") + hipGetErrorString(e));
            }
            if (k > 0) t.push_back(ms / kSeq);
        }
        (void)hipGraphExecDestroy(ge);
        std::sort(t.begin(), t.end());
        const double med = t[t.size() / 2];
        if (med < best_ms) {
            best_ms = med;
            best = S;
            best_g = sg[1];
        }
    }
    ctx->srows = best;
    ctx->grows_req = best_g;
    r = ctx->prec ? rebuild_order_t<double>(ctx) : rebuild_order_t<float>(ctx);
    if (r != THY_OK) return r;
    ctx->slot = saved_slot;
    HIPCHK(ctx_copy(ctx, ctx->ctrl, saved, sizeof(saved), hipMemcpyHostToDevice));
    HIPCHK(ctx_copy(ctx, ctx->dtmin, saved_dt.data(), saved_dt.size() * sizeof(unsigned int), hipMemcpyHostToDevice));
    // the hand-back counts as a run leaves them between steps (the next launch resets its own)
    if (ctx->fb_count) HIPCHK(ctx_fill(ctx, ctx->fb_count, 3 * sizeof(unsigned int)));
    if (strip_rows) *strip_rows = best;
    if (general_rows) *general_rows = ctx->gsrows;
    if (ms_best) *ms_best = best_ms;
    return THY_OK;
}

int thy_step(thy_ctx *ctx, double dt, double *dt_next) {
    RelaxedCapture relaxed;
    if (!ctx) return THY_ERR_INVALID;
    int r;
    Ctrl c = idle_ctrl(ctx);
    c.dt = dt;
    c.dt_valid = 1;
    c.cur = ctx->cur;
    if ((r = write_ctrl(ctx, c)) != THY_OK) return r;
    const int s = ctx->slot;
    if ((r = clear_dt_slots(ctx, s)) != THY_OK) return r;
    if (ctx->fb_count) HIPCHK(ctx_fill(ctx, &ctx->fb_count[s], sizeof(unsigned int)));
    hipError_t e = launch_step(ctx);
    if (e != hipSuccess) return fail(ctx, THY_ERR_HIP, std::string("step launch: ") + hipGetErrorString(e));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    Ctrl o;
    HIPCHK(ctx_copy(ctx, &o, &ctx->ctrl[ctx->slot], sizeof(Ctrl), hipMemcpyDeviceToHost));
    ctx->cur = o.cur;
    unsigned int bad = 0;
    HIPCHK(ctx_copy(ctx, &bad, ctx->health, sizeof(bad), hipMemcpyDeviceToHost));
    if (bad) return fail(ctx, THY_ERR_NONFINITE, "the step produced a non-finite eta, qx or qy");
    if (dt_next) {
        unsigned int v;
        if ((r = read_dt_slot(ctx, s, v)) != THY_OK) return r;
        float f;
        memcpy(&f, &v, 4);
        *dt_next = (double)f;
    }
    return THY_OK;
}

int thy_get_field(thy_ctx *ctx, int field, float *out) {
    RelaxedCapture relaxed;
    if (!ctx || !out || field < 0 || field > 5) return THY_ERR_INVALID;
    return ctx->prec ? get_field_t<double>(ctx, field, out) : get_field_t<float>(ctx, field, out);
}

int thy_mass(thy_ctx *ctx, double *mass) {
    RelaxedCapture relaxed;
    if (!ctx || !mass) return THY_ERR_INVALID;
    std::vector<double> host(kMassBlocks);
    return ctx->prec ? mass_t<double>(ctx, ctx->mass_partial, host.data(), mass)
                     : mass_t<float>(ctx, ctx->mass_partial, host.data(), mass);
}

int thy_mass_noalloc(thy_ctx *ctx, double *mass) {
    RelaxedCapture relaxed;
    if (!ctx || !mass) return THY_ERR_INVALID;
    if (ctx->device < 0 || ctx->device >= kScratchDevices || !g_scratch_ptr[ctx->device])
        return fail(ctx, THY_ERR_INVALID, "the allocation-free scratch of this device was not set up");
    double host[kScratchPartials];  // on the stack: no allocation on the host either
    return ctx->prec ? mass_t<double>(ctx, g_scratch_ptr[ctx->device], host, mass)
                     : mass_t<float>(ctx, g_scratch_ptr[ctx->device], host, mass);
}

int thy_last_timing(thy_ctx *ctx, double *ms) {
    if (!ctx || !ms) return THY_ERR_INVALID;
    *ms = ctx->last_ms;
    return THY_OK;
}

static int collect_gauges(thy_ctx *ctx);

int thy_get_gauges(thy_ctx *ctx, float *time, float *eta, float *u, float *v, float *eta_min, float *eta_max,
                   float *bathy) {
    RelaxedCapture relaxed;
    if (!ctx) return THY_ERR_INVALID;
    if (ctx->gauges_stale) {
        const int r = collect_gauges(ctx);
        if (r != THY_OK) return r;
        ctx->gauges_stale = false;
    }
    const int ns = ctx->nsamples, ng = ctx->ng;
    if (time) std::copy(ctx->g_time.begin(), ctx->g_time.begin() + ns, time);
    if (eta) std::copy(ctx->g_eta.begin(), ctx->g_eta.begin() + (size_t)ns * ng, eta);
    if (u) std::copy(ctx->g_u.begin(), ctx->g_u.begin() + (size_t)ns * ng, u);
    if (v) std::copy(ctx->g_v.begin(), ctx->g_v.begin() + (size_t)ns * ng, v);
    std::vector<int> pos(ng);
    if (ng) HIPCHK(ctx_copy(ctx, pos.data(), ctx->gpos, ng * sizeof(int), hipMemcpyDeviceToHost));
    if (eta_min || eta_max) {
        for (int g = 0; g < ng; g++) {
            float mn = 1e30f, mx = -1e30f;
            if (pos[g] < 0) mn = mx = -9999.0f;
            else
                for (int s = 0; s < ns; s++) {
                    float x = ctx->g_eta[(size_t)s * ng + g];
                    mn = std::min(x, mn);
                    mx = std::max(x, mx);
                }
            if (eta_min) eta_min[g] = mn;
            if (eta_max) eta_max[g] = mx;
        }
    }
    if (bathy && ng) {
        std::vector<float> field((size_t)ctx->nx * ctx->ny);
        int r = thy_get_field(ctx, THY_DEFORMED_BATHY, field.data());
        if (r != THY_OK) return r;
        for (int g = 0; g < ng; g++) bathy[g] = pos[g] < 0 ? -9999.0f : field[pos[g]];
    }
    return THY_OK;
}

// deliver finished iterations [ctx->logged, upto) to the on_step callback
static int drain_log(thy_ctx *ctx, int64_t upto) {
    if (!ctx->on_step || upto <= ctx->logged) {
        if (upto > ctx->logged) ctx->logged = upto;
        return THY_OK;
    }
    std::vector<double> buf((size_t)ctx->logcap * 2);
    HIPCHK(ctx_copy(ctx, buf.data(), ctx->steplog, buf.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int64_t it = std::max(ctx->logged, upto - ctx->logcap); it < upto; it++) {
        const int64_t li = it % ctx->logcap;
        ctx->on_step(ctx->user, it + 1, buf[2 * li], buf[2 * li + 1]);
    }
    ctx->logged = upto;
    return THY_OK;
}


static int run_loop(thy_ctx *ctx, thy_run_result *res) {
    int r;
    double total_ms = 0.0;
    for (;;) {
        Ctrl s;
        double ms = 0.0;
        if ((r = advance(ctx, ctx->K, &s, &ms)) != THY_OK) return r;
        total_ms += ms;
        // after a stop every finished iteration has its log entry; otherwise the last
        // one is completed by the next launch
        if ((r = drain_log(ctx, s.dt_valid ? s.steps : s.steps - 1)) != THY_OK) return r;
        if (s.nonfinite)
            return fail(ctx, THY_ERR_NONFINITE,
                        "non-finite eta, qx or qy after step " + std::to_string((long long)s.steps));
        if (!(s.t < s.t_end) || (s.max_steps > 0 && s.steps >= s.max_steps)) {
            res->steps = s.steps;
            res->t_final = s.t;
            ctx->nsamples = std::min(s.nsamples, ctx->gcap);
            res->n_samples = ctx->nsamples;
            break;
        }
        bool changed = false;
        // src/GPU/ShallowWater.cu:125-147: apply every fault whose time has come
        while (s.k < s.nf && s.t * s.T >= ctx->ftimes[s.k]) {
            if (ctx->on_okada) ctx->on_okada(ctx->user, s.k);
            if ((r = thy_apply_okada(ctx, &ctx->faults[s.k], ctx->crop_flag, ctx->crop_value, nullptr)) != THY_OK)
                return r;
            s.k++;
            changed = true;
        }
        // src/GPU/ShallowWater.cu:404-409: snapshot, at most one per iteration
        if (!s.snap_done && s.t >= ctx->next_save) {
            ctx->next_save += ctx->save_interval;
            s.next_save = ctx->next_save;
            s.snap_done = 1;
            if (ctx->on_snapshot) ctx->on_snapshot(ctx->user, ctx->nsnap, s.t, ctx);
            ctx->nsnap++;
            changed = true;
        }
        if (!changed) return fail(ctx, THY_ERR_INVALID, "driver stalled");
        if ((r = write_ctrl(ctx, s)) != THY_OK) return r;
    }
    res->n_snapshots = ctx->nsnap;
    res->seconds = total_ms * 1e-3;
    ctx->last_ms = total_ms;
    ctx->gauges_stale = true;  // copied to the host when thy_get_gauges asks for them
    return THY_OK;
}

// gauge samples of the finished run into host memory (read by thy_get_gauges)
static int collect_gauges(thy_ctx *ctx) {
    const int ns = ctx->nsamples, ng = ctx->ng;
    ctx->g_time.assign(ns, 0.f);
    ctx->g_eta.assign((size_t)ns * ng, 0.f);
    ctx->g_u.assign((size_t)ns * ng, 0.f);
    ctx->g_v.assign((size_t)ns * ng, 0.f);
    if (ns > 0 && ng > 0) {
        std::vector<float> buf((size_t)ns * ng * 3);
        HIPCHK(ctx_copy(ctx, buf.data(), ctx->gbuf, buf.size() * sizeof(float), hipMemcpyDeviceToHost));
        HIPCHK(ctx_copy(ctx, ctx->g_time.data(), ctx->gtime, ns * sizeof(float), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < (size_t)ns * ng; i++) {
            ctx->g_eta[i] = buf[3 * i];
            ctx->g_u[i] = buf[3 * i + 1];
            ctx->g_v[i] = buf[3 * i + 2];
        }
    }
    return THY_OK;
}

// Everything thy_run does before its first step launch: schedule, gauge buffers,
// metrics, initial dt and truncation, control block in launch slot 0.
static int prepare_run(thy_ctx *ctx, const thy_run_config *cfg, thy_run_result *res) {
    if (cfg->n_faults < 0 || (cfg->n_faults > 0 && !cfg->faults) || (cfg->n_gauges > 0 && !cfg->gauge_pos))
        return fail(ctx, THY_ERR_INVALID, "bad run configuration");
    int r;
    memset(res, 0, sizeof(*res));
    ctx->batch_ready = false;  // a run prepared before is replaced
    ctx->K = cfg->steps_per_sync > 0 ? cfg->steps_per_sync : 24;
    ctx->faults.assign(cfg->faults, cfg->faults + cfg->n_faults);
    ctx->ftimes.assign(std::max(1, cfg->n_faults), 0.0);
    for (int k = 0; k < cfg->n_faults; k++) ctx->ftimes[k] = cfg->faults[k].time;
    ctx->crop_flag = cfg->crop_flag;
    ctx->crop_value = cfg->crop_value;
    ctx->save_interval = cfg->save_interval;
    ctx->on_snapshot = cfg->on_snapshot;
    ctx->on_okada = cfg->on_okada;
    ctx->on_step = cfg->on_step;
    ctx->user = cfg->user;
    ctx->nsnap = 0;
    ctx->logged = 0;
    const std::vector<double> &ft = ctx->ftimes;
    if (ctx->ftime) { dev_free(ctx->ftime); ctx->ftime = nullptr; }
    HIPCHK(dev_alloc((void **)&ctx->ftime, ft.size() * sizeof(double)));
    HIPCHK(ctx_copy(ctx, ctx->ftime, ft.data(), ft.size() * sizeof(double), hipMemcpyHostToDevice));
    // gauges
    if (ctx->gpos) { dev_free(ctx->gpos); ctx->gpos = nullptr; }
    if (ctx->gbuf) { dev_free(ctx->gbuf); ctx->gbuf = nullptr; }
    if (ctx->gtime) { dev_free(ctx->gtime); ctx->gtime = nullptr; }
    ctx->ng = cfg->n_gauges;
    ctx->gcap = 0;
    if (cfg->n_gauges > 0 && cfg->series_interval >= 0.0) {
        double nsd = cfg->series_interval > 0.0 ? cfg->t_end / cfg->series_interval + 4.0 : 1e6;
        if (cfg->max_steps > 0 && cfg->series_interval <= 0.0) nsd = (double)cfg->max_steps + 2.0;
        ctx->gcap = (int)std::min(nsd, 1e6);
    }
    HIPCHK(dev_alloc((void **)&ctx->gpos, std::max(1, ctx->ng) * sizeof(int)));
    if (ctx->ng) HIPCHK(ctx_copy(ctx, ctx->gpos, cfg->gauge_pos, ctx->ng * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(dev_alloc((void **)&ctx->gbuf, std::max<size_t>(1, (size_t)ctx->gcap * ctx->ng * 3) * sizeof(float)));
    HIPCHK(dev_alloc((void **)&ctx->gtime, std::max(1, ctx->gcap) * sizeof(float)));
    drop_graphs(ctx);

    HIPCHK(ctx_fill(ctx, ctx->health, sizeof(unsigned int)));
    // src/GPU/ShallowWater.cu:305-309: metrics from the pre-deformation state
    if ((r = thy_init_metrics(ctx)) != THY_OK) return r;
    // src/GPU/ShallowWater.cu:378-391: initial dt (pre-deformation) and truncation
    double dt0;
    if ((r = thy_initial_dt(ctx, &dt0)) != THY_OK) return r;
    res->dt_initial = dt0;
    Ctrl c = idle_ctrl(ctx);
    c.t = 0.0;
    c.t_end = cfg->t_end;
    c.max_steps = cfg->max_steps;
    c.nf = cfg->n_faults;
    c.k = 0;
    c.T = ctx->prm.T;
    c.next_save = cfg->save_interval >= 0.0 ? 0.0 : INFINITY;
    c.next_series = (cfg->series_interval >= 0.0 && ctx->ng > 0) ? 0.0 : INFINITY;
    c.series_dt = cfg->series_interval;
    c.gcap = ctx->gcap;
    ctx->next_save = c.next_save;
    double dt = dt0;
    int skip = 0;
    if (cfg->n_faults > 0) {
        // the skipping variant when the first fault is at t = 0 (src/GPU/ShallowWater.cu:382-391)
        if (ft[0] < 2.220446049250313e-16) trunc_skipping(c, ft.data(), 0.0, dt, skip);
        else trunc_plain(c, ft.data(), 0.0, dt, skip);
    }
    c.dt = dt;
    c.dt_valid = 1;
    c.skip = skip;
    c.cur = ctx->cur;
    if ((r = reset_dt_slots(ctx)) != THY_OK) return r;
    ctx->slot = 0;  // the next launch reads slot 0 (the dt slots were all reset)
    if ((r = write_ctrl(ctx, c)) != THY_OK) return r;
    ctx->run_active = true;
    return THY_OK;
}

int thy_run(thy_ctx *ctx, const thy_run_config *cfg, thy_run_result *res) {
    RelaxedCapture relaxed;
    if (!ctx || !cfg || !res) return THY_ERR_INVALID;
    int r;
    if ((r = prepare_run(ctx, cfg, res)) != THY_OK) return r;
    double dti = res->dt_initial;
    r = run_loop(ctx, res);
    res->dt_initial = dti;
    return r;
}

int thy_run_more(thy_ctx *ctx, int64_t n_steps, thy_run_result *res) {
    RelaxedCapture relaxed;
    if (!ctx || !res || n_steps < 0) return THY_ERR_INVALID;
    if (!ctx->run_active) return fail(ctx, THY_ERR_INVALID, "thy_run_more without a previous thy_run");
    int r;
    memset(res, 0, sizeof(*res));
    Ctrl c;
    if ((r = read_ctrl(ctx, &c)) != THY_OK) return r;
    c.max_steps = c.steps + n_steps;
    if ((r = write_ctrl(ctx, c)) != THY_OK) return r;
    return run_loop(ctx, res);
}

}  // extern "C"

// ===========================================================================
// Batched scenarios: many contexts advanced together, three launches per step (the
// deep-water strips of every scenario, the general strips beside them on a second
// stream, then the hand-backs), captured once for K steps.
// ===========================================================================
struct thy_batch {
    int capacity = 0;
    size_t budget = 0, used = 0;
    int K = 24;
    hipStream_t stream = nullptr, side = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_fork = nullptr, ev_join = nullptr;
    hipEvent_t ev_done = nullptr;   // end of a launch's work (admissions queue copies after it)
    std::vector<thy_ctx *> ctx;     // per slot; nullptr when free
    std::vector<size_t> bytes;      // admitted footprint per slot
    std::vector<int> srows;         // strip height in the slot's argument block
    std::vector<int> state;         // THY_SLOT_*
    std::vector<thy_run_result> res;
    StepArgs<float> *d_args = nullptr;
    // item lists (scenario slot, strip << 1 | writer) and their device counts; the
    // captured grids cover deep_cap and gen_cap items
    int2 *d_deep = nullptr, *d_gen = nullptr, *d_fb = nullptr;
    int *d_counts = nullptr;        // [0] deep items, [1] general items
    unsigned int *d_fb_count = nullptr;
    int deep_cap = 0, gen_cap = 0;
    int gen_sub = 0;                // general items are gsrows-row sub-strips
    std::vector<int2> h_deep, h_gen;
    int h_counts[2] = {0, 0};
    int fb_grid = 1;
    Ctrl *d_ctrl = nullptr;
    Ctrl *h_ctrl = nullptr;              // pinned: control blocks gathered after a launch
    StepArgs<float> *h_args = nullptr;   // pinned: argument blocks of admitted scenarios
    hipGraphExec_t gexec = nullptr;
    bool dirty = true;
    bool in_flight = false;              // a launch is queued and not yet collected
    std::vector<char> launched;          // per slot: running when the launch was queued
    int launched_hi = 0;
    double gpu_ms = 0.0;
    std::string err;
};

namespace {

int bfail(thy_batch *b, int code, const std::string &msg) {
    b->err = msg;
    return code;
}

#define BHIPCHK(x)                                                                                        \
    do {                                                                                                 \
        hipError_t e_ = (x);                                                                             \
        if (e_ != hipSuccess)                                                                            \
            return bfail(b, e_ == hipErrorOutOfMemory ? THY_ERR_DEVICE_MEMORY : THY_ERR_HIP,             \
                         std::string(#x) + ": " + hipGetErrorString(e_));                                \
    } while (0)

// capacity of an item list: a power of two, at least 256 items
int list_cap(size_t n) {
    int c = 256;
    while ((size_t)c < n) c <<= 1;
    return c;
}

// Item lists of the running scenarios: every deep-water strip, scenario by scenario in
// its own order (neighbouring strips stay together); every general strip, costliest
// cost class first.  A scenario's writer is its first deep-water strip, or its first
// general strip when it has none.  With few general strips in all (a launch bound by
// the time one wavefront takes for one strip rather than by the number of strips) the
// general strips are cut into sub-strips of each scenario's gsrows rows, as in a run
// alone.
int batch_items(thy_batch *b) {
    struct G { int cls; int2 it; };
    std::vector<G> gv;
    b->h_deep.clear();
    size_t n_full = 0;
    for (int i = 0; i < b->capacity; i++)
        if (b->state[i] == THY_SLOT_RUNNING) n_full += b->ctx[i]->bgen_h.size();
    const int sub = n_full < (size_t)8 * b->fb_grid ? 1 : 0;
    for (int i = 0; i < b->capacity; i++) {
        if (b->state[i] != THY_SLOT_RUNNING) continue;
        const thy_ctx *c = b->ctx[i];
        // the strip lists must describe the strips of the argument block
        if (c->srows != b->srows[i]) return bfail(b, THY_ERR_INVALID, "strip height changed after admission");
        const bool nd = c->bdeep_h.empty();
        for (size_t p = 0; p < c->bdeep_h.size(); p++)
            b->h_deep.push_back(make_int2(i, (c->bdeep_h[p] << 1) | (p == 0 ? 1 : 0)));
        const std::vector<int> &gl = sub ? c->bsub_h : c->bgen_h, &gc = sub ? c->bsubcls_h : c->bgcls_h;
        for (size_t p = 0; p < gl.size(); p++)
            gv.push_back({gc[p], make_int2(i, (gl[p] << 1) | (nd && p == 0 ? 1 : 0))});
    }
    // the deep-water items as in a run alone (build_order): row-major, scenario by
    // scenario, cut into 8 runs of (nearly) equal length, one per XCD, padded to whole
    // workgroups with idle items (scenario -1), so that each XCD sweeps one band and
    // strips side by side share their halo columns in its L2
    if (!b->h_deep.empty()) {
        const std::vector<int2> rm = b->h_deep;
        const int nd = (int)rm.size(), q = nd / 8, rem = nd % 8;
        const int wmax = (q + (rem > 0 ? 1 : 0) + SW_WAVES - 1) / SW_WAVES;
        b->h_deep.assign((size_t)wmax * 8 * SW_WAVES, make_int2(-1, 0));
        for (int x = 0; x < 8; x++) {
            const int size = q + (x < rem ? 1 : 0), start = x * q + std::min(x, rem);
            for (int i = 0; i < size; i++)
                b->h_deep[((size_t)(i / SW_WAVES) * 8 + x) * SW_WAVES + i % SW_WAVES] = rm[start + i];
        }
    }
    std::stable_sort(gv.begin(), gv.end(), [](const G &x, const G &y) { return x.cls > y.cls; });
    b->h_gen.resize(gv.size());
    for (size_t k = 0; k < gv.size(); k++) b->h_gen[k] = gv[k].it;
    const int dc = list_cap(b->h_deep.size()), gc = list_cap(b->h_gen.size());
    if (sub != b->gen_sub && b->gexec) {  // captured with the other strip height
        (void)hipGraphExecDestroy(b->gexec);
        b->gexec = nullptr;
    }
    b->gen_sub = sub;
    if (dc != b->deep_cap || gc != b->gen_cap) {
        // the captured launches hold the old lists and grid sizes
        if (b->gexec) {
            (void)hipGraphExecDestroy(b->gexec);
            b->gexec = nullptr;
        }
        void *ptrs[] = {b->d_deep, b->d_gen, b->d_fb};
        for (void *q : ptrs)
            dev_free(q);
        b->d_deep = b->d_gen = b->d_fb = nullptr;
        b->deep_cap = b->gen_cap = 0;
        BHIPCHK(dev_alloc(&b->d_deep, (size_t)dc * sizeof(int2)));
        BHIPCHK(dev_alloc(&b->d_gen, (size_t)gc * sizeof(int2)));
        BHIPCHK(dev_alloc(&b->d_fb, (size_t)3 * dc * sizeof(int2)));
        b->deep_cap = dc;
        b->gen_cap = gc;
    }
    b->h_counts[0] = (int)b->h_deep.size();
    b->h_counts[1] = (int)b->h_gen.size();
    // ordered before the next launch on the batch stream; the host lists live in b
    if (!b->h_deep.empty())
        BHIPCHK(hipMemcpyAsync(b->d_deep, b->h_deep.data(), b->h_deep.size() * sizeof(int2), hipMemcpyHostToDevice,
                               b->stream));
    if (!b->h_gen.empty())
        BHIPCHK(hipMemcpyAsync(b->d_gen, b->h_gen.data(), b->h_gen.size() * sizeof(int2), hipMemcpyHostToDevice,
                               b->stream));
    BHIPCHK(hipMemcpyAsync(b->d_counts, b->h_counts, sizeof(b->h_counts), hipMemcpyHostToDevice, b->stream));
    BHIPCHK(hipStreamSynchronize(b->stream));
    b->dirty = false;
    return THY_OK;
}

// K steps (launch slots 0, 1, 2, ...), captured once: per step the deep-water launch on
// the batch stream, the general launch beside it on the side stream, and after both
// the hand-back launch
int batch_graph(thy_batch *b) {
    if (b->gexec) return THY_OK;
    BatchArgs q;
    q.args = b->d_args;
    q.deep = b->d_deep;
    q.n_deep = b->d_counts;
    q.gen = b->d_gen;
    q.n_gen = b->d_counts + 1;
    q.fb_list = b->d_fb;
    q.fb_count = b->d_fb_count;
    q.fb_cap = b->deep_cap;
    q.gen_sub = b->gen_sub;
    const dim3 blk(64 * SW_WAVES);
    hipGraph_t g;
    BHIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < b->K; k++) {
        const int slot = k % 3;
        BHIPCHK(hipEventRecord(b->ev_fork, b->stream));
        BHIPCHK(hipStreamWaitEvent(b->side, b->ev_fork, 0));
        hipLaunchKernelGGL(k_batch_general<false>, dim3(b->gen_cap / SW_WAVES), blk, 0, b->side, q, slot);
        BHIPCHK(hipEventRecord(b->ev_join, b->side));
        hipLaunchKernelGGL(k_batch_deep, dim3(b->deep_cap / SW_WAVES), blk, 0, b->stream, q, slot);
        BHIPCHK(hipStreamWaitEvent(b->stream, b->ev_join, 0));
        hipLaunchKernelGGL(k_batch_general<true>, dim3(b->fb_grid), blk, 0, b->stream, q, slot);
    }
    BHIPCHK(hipStreamEndCapture(b->stream, &g));
    BHIPCHK(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    return THY_OK;
}

// apply every fault of a scenario whose time has come; Ctrl c is its control block
// faults that came due at the control block's time: Okada, then the control block
// rewritten into launch slot 0; errors are the context's
int ctx_faults(thy_ctx *ctx, Ctrl &cb, bool &changed) {
    changed = false;
    while (cb.k < cb.nf && cb.t * cb.T >= ctx->ftimes[cb.k]) {
        const int r = thy_apply_okada(ctx, &ctx->faults[cb.k], ctx->crop_flag, ctx->crop_value, nullptr);
        if (r != THY_OK) return fail(ctx, r, std::string("okada: ") + thy_last_error(ctx));
        cb.k++;
        changed = true;
    }
    if (changed) {
        HIPCHK(hipStreamSynchronize(ctx->stream));
        ctx->slot = 0;
        return write_ctrl(ctx, cb);
    }
    return THY_OK;
}

int batch_faults(thy_batch *b, thy_ctx *c, Ctrl &cb, bool &changed) {
    const int r = ctx_faults(c, cb, changed);
    return r == THY_OK ? r : bfail(b, r, thy_last_error(c));
}

// the context-local part of admission: run configuration, metrics, initial dt, faults
// due at t = 0 (src/GPU/ShallowWater.cu:305-391)
int prepare_for_batch(thy_ctx *ctx, const thy_run_config *cfg, thy_run_result &res) {
    if (ctx->prec != 0 || ctx->kernel == 1)
        return fail(ctx, THY_ERR_INVALID, "batched scenarios run the FP32 row-sweep kernel only");
    if (cfg->save_interval >= 0.0) return fail(ctx, THY_ERR_INVALID, "batched scenarios take no snapshots");
    if (ctx->order_h.empty()) return fail(ctx, THY_ERR_INVALID, "scenario state not uploaded");
    int r;
    if ((r = prepare_run(ctx, cfg, &res)) != THY_OK) return r;
    Ctrl cb;
    if ((r = read_ctrl(ctx, &cb)) != THY_OK) return r;
    bool changed;
    return ctx_faults(ctx, cb, changed);
}

}  // namespace

extern "C" {

int thy_batch_create(int capacity, size_t hbm_budget, int steps_per_sync, thy_batch **out) {
    RelaxedCapture relaxed;
    if (!out || capacity < 1 || capacity > 65536) return THY_ERR_INVALID;
    thy_batch *b = new (std::nothrow) thy_batch();
    if (!b) return THY_ERR_HOST_MEMORY;
    *out = b;
    b->capacity = capacity;
    b->budget = hbm_budget;
    b->K = steps_per_sync > 0 ? ((steps_per_sync + 2) / 3) * 3 : 24;
    b->ctx.assign(capacity, nullptr);
    b->bytes.assign(capacity, 0);
    b->srows.assign(capacity, 0);
    b->state.assign(capacity, THY_SLOT_FREE);
    b->res.assign(capacity, thy_run_result{});
    b->launched.assign(capacity, 0);
    // cached streams: creating and destroying streams with every batch would shift how
    // the runtime spreads streams over the hardware queues, and the two launches of a
    // step only overlap when the pair sits on two different queues
    BHIPCHK(Pool::get().get_streams(b->stream, b->side));
    BHIPCHK(hipEventCreate(&b->ev0));
    BHIPCHK(hipEventCreate(&b->ev1));
    BHIPCHK(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
    BHIPCHK(hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming));
    BHIPCHK(hipEventCreateWithFlags(&b->ev_done, hipEventDisableTiming));
    BHIPCHK(dev_alloc(&b->d_args, (size_t)capacity * sizeof(StepArgs<float>)));
    BHIPCHK(hipMemsetAsync(b->d_args, 0, (size_t)capacity * sizeof(StepArgs<float>), b->stream));
    BHIPCHK(dev_alloc(&b->d_counts, 2 * sizeof(int)));
    BHIPCHK(hipMemsetAsync(b->d_counts, 0, 2 * sizeof(int), b->stream));
    BHIPCHK(dev_alloc(&b->d_fb_count, 3 * sizeof(unsigned int)));
    BHIPCHK(hipMemsetAsync(b->d_fb_count, 0, 3 * sizeof(unsigned int), b->stream));
    BHIPCHK(dev_alloc(&b->d_ctrl, (size_t)capacity * sizeof(Ctrl)));
    BHIPCHK(Pool::get().host_alloc((void **)&b->h_ctrl, (size_t)capacity * sizeof(Ctrl)));
    BHIPCHK(Pool::get().host_alloc((void **)&b->h_args, (size_t)capacity * sizeof(StepArgs<float>)));
    int dev = 0, ncu = 0;
    BHIPCHK(hipGetDevice(&dev));
    BHIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    b->fb_grid = std::max(1, ncu);
    return THY_OK;
}

void thy_batch_destroy(thy_batch *b) {
    RelaxedCapture relaxed;
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->side) (void)hipStreamSynchronize(b->side);
    Pool::get().host_free(b->h_ctrl);
    Pool::get().host_free(b->h_args);
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    void *ptrs[] = {b->d_args, b->d_deep, b->d_gen, b->d_fb, b->d_counts, b->d_fb_count, b->d_ctrl};
    for (void *p : ptrs)
        dev_free(p);
    hipEvent_t evs[] = {b->ev0, b->ev1, b->ev_fork, b->ev_join, b->ev_done};
    for (hipEvent_t e : evs)
        if (e) (void)hipEventDestroy(e);
    if (b->stream) Pool::get().put_streams(b->stream, b->side);
    delete b;
}

const char *thy_batch_last_error(thy_batch *b) { return b ? b->err.c_str() : "null batch"; }

int thy_batch_prepare(thy_ctx *ctx, const thy_run_config *cfg) {
    RelaxedCapture relaxed;
    if (!ctx || !cfg) return THY_ERR_INVALID;
    ctx->batch_ready = false;
    const int r = prepare_for_batch(ctx, cfg, ctx->prep_res);
    if (r == THY_OK) ctx->batch_ready = true;
    return r;
}

int thy_batch_add(thy_batch *b, thy_ctx *ctx, const thy_run_config *cfg, int *slot_out) {
    RelaxedCapture relaxed;
    if (!b || !ctx || !slot_out) return THY_ERR_INVALID;
    if (!cfg && !ctx->batch_ready) return bfail(b, THY_ERR_INVALID, "context not prepared (thy_batch_prepare)");
    for (int i = 0; i < b->capacity; i++)
        if (b->ctx[i] == ctx && b->state[i] != THY_SLOT_FREE) return bfail(b, THY_ERR_INVALID, "context already admitted");
    int slot = -1;
    for (int i = 0; i < b->capacity && slot < 0; i++)
        if (b->state[i] == THY_SLOT_FREE) slot = i;
    if (slot < 0) return bfail(b, THY_ERR_BUDGET, "no free scenario slot");
    const size_t need = thy_footprint(ctx->nx, ctx->ny, &ctx->prm);
    if (b->budget > 0 && b->used + need > b->budget) return bfail(b, THY_ERR_BUDGET, "HBM budget exceeded");
    thy_run_result res;
    if (cfg) {
        const int r = prepare_for_batch(ctx, cfg, res);
        if (r != THY_OK) return bfail(b, r, thy_last_error(ctx));
    } else {
        res = ctx->prep_res;
    }
    ctx->batch_ready = false;
    ctx->ctrl_h_valid = false;  // the batch's launches advance its control blocks
    const StepArgs<float> a = make_args<float>(ctx);
    // ordered after a launch in flight, which does not read this free slot
    b->h_args[slot] = a;
    BHIPCHK(hipMemcpyAsync(b->d_args + slot, b->h_args + slot, sizeof(a), hipMemcpyHostToDevice, b->stream));
    b->ctx[slot] = ctx;
    b->srows[slot] = a.srows;
    b->bytes[slot] = need;
    b->used += need;
    b->state[slot] = THY_SLOT_RUNNING;
    b->launched[slot] = 0;  // not in a launch in flight: it joins the next one
    b->res[slot] = res;
    b->dirty = true;
    *slot_out = slot;
    return THY_OK;
}

int thy_batch_launch(thy_batch *b) {
    RelaxedCapture relaxed;
    if (!b) return THY_ERR_INVALID;
    if (b->in_flight) return bfail(b, THY_ERR_INVALID, "a launch is in flight");
    int r;
    if (b->dirty && (r = batch_items(b)) != THY_OK) return r;
    int hi = 0;
    for (int i = 0; i < b->capacity; i++) {
        b->launched[i] = b->state[i] == THY_SLOT_RUNNING;
        if (b->launched[i]) hi = i + 1;
    }
    if (hi == 0) return THY_OK;
    if ((r = batch_graph(b)) != THY_OK) return r;
    BHIPCHK(hipEventRecord(b->ev0, b->stream));
    BHIPCHK(hipGraphLaunch(b->gexec, b->stream));
    BHIPCHK(hipEventRecord(b->ev1, b->stream));
    // K is a multiple of 3: the next launch reads slot 0 again
    hipLaunchKernelGGL(k_gather_ctrl, dim3((hi + 63) / 64), dim3(64), 0, b->stream, b->d_args, hi, 0, b->d_ctrl);
    BHIPCHK(hipGetLastError());
    BHIPCHK(hipMemcpyAsync(b->h_ctrl, b->d_ctrl, (size_t)hi * sizeof(Ctrl), hipMemcpyDeviceToHost, b->stream));
    BHIPCHK(hipEventRecord(b->ev_done, b->stream));
    b->launched_hi = hi;
    b->in_flight = true;
    return THY_OK;
}

int thy_batch_ready(thy_batch *b, int *done) {
    RelaxedCapture relaxed;
    if (!b || !done) return THY_ERR_INVALID;
    if (!b->in_flight) {
        *done = 1;
        return THY_OK;
    }
    // the launch's own work: copies that admissions queued after it do not count
    const hipError_t e = hipEventQuery(b->ev_done);
    if (e != hipSuccess && e != hipErrorNotReady) BHIPCHK(e);
    *done = e == hipSuccess ? 1 : 0;
    return THY_OK;
}

int thy_batch_wait(thy_batch *b, int *n_running) {
    RelaxedCapture relaxed;
    if (!b) return THY_ERR_INVALID;
    int r;
    if (b->in_flight) {
        b->in_flight = false;
        BHIPCHK(hipStreamSynchronize(b->stream));
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, b->ev0, b->ev1);
        b->gpu_ms += ms;
        for (int i = 0; i < b->launched_hi; i++) {
            if (!b->launched[i] || b->state[i] != THY_SLOT_RUNNING) continue;
            thy_ctx *c = b->ctx[i];
            Ctrl cb = b->h_ctrl[i];
            c->cur = cb.cur;
            b->res[i].steps = cb.steps;  // progress, final once it has finished
            b->res[i].t_final = cb.t;
            if (!cb.dt_valid) continue;  // stepped through the whole sequence
            if (cb.nonfinite) {
                // the health check stopped it: retired as failed, the batch goes on
                c->err = "non-finite eta, qx or qy after step " + std::to_string((long long)cb.steps);
                b->state[i] = THY_SLOT_FAILED;
                b->dirty = true;
                BHIPCHK(hipMemsetAsync(b->d_args + i, 0, sizeof(StepArgs<float>), b->stream));
                continue;
            }
            // stopped: finished, or a fault is due (the only other stop a batched run has)
            if (!(cb.t < cb.t_end) || (cb.max_steps > 0 && cb.steps >= cb.max_steps)) {
                c->nsamples = std::min(cb.nsamples, c->gcap);
                c->gauges_stale = true;  // read by thy_get_gauges, on the caller's thread
                thy_run_result &res = b->res[i];
                res.steps = cb.steps;
                res.t_final = cb.t;
                res.n_samples = c->nsamples;
                b->state[i] = THY_SLOT_FINISHED;
                b->dirty = true;
                // no later launch or gather reads a finished scenario
                BHIPCHK(hipMemsetAsync(b->d_args + i, 0, sizeof(StepArgs<float>), b->stream));
                continue;
            }
            bool changed;
            if ((r = batch_faults(b, c, cb, changed)) != THY_OK) return r;
            if (!changed) return bfail(b, THY_ERR_INVALID, "batched scenario stalled");
        }
        BHIPCHK(hipStreamSynchronize(b->stream));
    }
    if (n_running) {
        int running = 0;
        for (int i = 0; i < b->capacity; i++) running += b->state[i] == THY_SLOT_RUNNING;
        *n_running = running;
    }
    return THY_OK;
}

int thy_batch_advance(thy_batch *b, int *n_running) {
    RelaxedCapture relaxed;
    int r = thy_batch_launch(b);
    if (r != THY_OK) return r;
    return thy_batch_wait(b, n_running);
}

int thy_batch_status(thy_batch *b, int slot, int *state, thy_run_result *res) {
    if (!b || slot < 0 || slot >= b->capacity) return THY_ERR_INVALID;
    if (state) *state = b->state[slot];
    if (res) *res = b->res[slot];
    return THY_OK;
}

int thy_batch_remove(thy_batch *b, int slot) {
    RelaxedCapture relaxed;
    if (!b || slot < 0 || slot >= b->capacity || b->state[slot] == THY_SLOT_FREE) return THY_ERR_INVALID;
    if (b->state[slot] == THY_SLOT_RUNNING) {
        // a launch in flight may be stepping it: wait, and take it out of every later one
        BHIPCHK(hipStreamSynchronize(b->stream));
        BHIPCHK(hipMemsetAsync(b->d_args + slot, 0, sizeof(StepArgs<float>), b->stream));
    }
    // a finished slot's argument block was cleared when it finished; a launch in flight
    // (already waited for) must not be collected for this slot, nor for a scenario
    // admitted to it later
    b->launched[slot] = 0;
    b->ctx[slot]->run_active = false;
    b->used -= b->bytes[slot];
    b->bytes[slot] = 0;
    b->ctx[slot] = nullptr;
    b->state[slot] = THY_SLOT_FREE;
    b->dirty = true;
    return THY_OK;
}

void thy_release_cached(void) {
    RelaxedCapture relaxed;
    Pool::get().trim();
}

int thy_batch_info(thy_batch *b, int *running, size_t *bytes_used, double *gpu_seconds) {
    if (!b) return THY_ERR_INVALID;
    int n = 0;
    for (int i = 0; i < b->capacity; i++) n += b->state[i] == THY_SLOT_RUNNING;
    if (running) *running = n;
    if (bytes_used) *bytes_used = b->used;
    if (gpu_seconds) *gpu_seconds = b->gpu_ms * 1e-3;
    return THY_OK;
}

}  // extern "C"
