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
    int pad;
};

template <typename T> struct RowGeo {
    T A, invA, gx, valx;  // area, 1/area, 1/cos(phi), (2*alto/cos)/(2*area)
    T lenx, dxs, dys, pad; // alto (x-face length), ancho and alto for sponge distances
};

template <typename T> struct FaceGeoY {
    T len, cos0, cos1, A0, A1, valy, ic0, ic1; // ancho, cos/area of rows j-1, j (clamped)
};

template <typename T> struct StepArgs {
    int nx, ny, pitch, periodic;
    int tiles_x, tiles_y, n_tiles, q_per_xcd;
    int sp_sup, sp_inf, sp_izq, sp_der;
    int track_arrival, n_gauges;
    T bsup, binf, bizq, bder;
    T cfl, mf0, vmax, eps_h, hpos, cvis, dif_at, sea_level;
    T Href, Q_over_H;
    T *e[2], *qx[2], *qy[2];
    const T *D;
    T *emax, *tarr;
    const T *eini;
    const RowGeo<T> *row;
    const FaceGeoY<T> *fy;
    Ctrl *ctrl;
    unsigned int *dtmin;     // 3 slots, float bits
    const double *ftime;     // fault times (s), sorted
    const int *gpos;
    float *gbuf;             // [gcap][n_gauges][3]
    float *gtime;            // [gcap]
};

__device__ __forceinline__ void trunc_plain(const Ctrl &c, const double *ft, double t, double &dt, int &skip) {
    if (c.k < c.nf) {
        if ((t + dt) * c.T > ft[c.k]) {
            dt = ft[c.k] / c.T - t;
            skip = 1;
        }
    }
}

__device__ __forceinline__ void trunc_skipping(const Ctrl &c, const double *ft, double t, double &dt, int &skip) {
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

// stop predicate of the driver loop, evaluated before each step
__device__ __forceinline__ bool must_stop(const Ctrl &c, const double *ft) {
    if (!(c.t < c.t_end)) return true;
    if (c.max_steps > 0 && c.steps >= c.max_steps) return true;
    if (c.k < c.nf && c.t * c.T >= ft[c.k]) return true;
    if (c.t >= c.next_save) return true;
    return false;
}

template <typename T> struct BlockCtl {
    double t, dt, Tscale;
    int run, sample, slot_sample, cur, gcap;
};

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

template <typename T, int TX, int TY>
__global__ void __launch_bounds__(256) k_step(const StepArgs<T> a, const int slot) {
    constexpr int NT = 256;
    constexpr int IX = TX + 8, IY = TY + 8, SX = TX + 4, SY = TY + 4;
    __shared__ T s_e[IY * IX], s_D[IY * IX], s_qx[IY * IX], s_qy[IY * IX];
    __shared__ T s_es[SY * SX];
    __shared__ T s_am[SY * SX];
    __shared__ T s_ax[TY * TX], s_ay[TY * TX], s_aw[TY * TX];
    __shared__ BlockCtl<T> bc;
    __shared__ float s_red[NT / 64];

    const int tid = threadIdx.x;
    // XCD-aware tile order: consecutive workgroups land on different XCDs, so give
    // each XCD a contiguous band of tiles (its own L2 then holds the shared halos).
    const int bid = blockIdx.x;
    const int tile = (bid & 7) * a.q_per_xcd + (bid >> 3);

    if (tid == 0) {
        const Ctrl c = a.ctrl[slot];
        double dt;
        int skip = c.skip;
        if (c.dt_valid) {
            dt = c.dt;
        } else {
            dt = (double)__uint_as_float(a.dtmin[(slot + 2) % 3]);
            if (!skip) trunc_plain(c, a.ftime, c.t, dt, skip);
            else trunc_skipping(c, a.ftime, c.t, dt, skip);
        }
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
        if (bid == 0) {
            a.dtmin[(slot + 1) % 3] = 0x7f7fffffu;  // FLT_MAX: reset the slot of the next launch
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
                if (sample) {
                    o.next_series = c.next_series + c.series_dt;
                    o.nsamples = c.nsamples + 1;
                }
            }
            a.ctrl[(slot + 1) % 3] = o;
        }
    }
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

    // gauge gather from the state at time t (src/GPU/Volumen_kernel.cu:223-254)
    if (bid == 0 && bc.sample && bc.slot_sample < bc.gcap) {
        for (int g = tid; g < a.n_gauges; g += NT) {
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

    const int tyi = tile / a.tiles_x, txi = tile - tyi * a.tiles_x;
    const int x0 = txi * TX, y0 = tyi * TY;

    // ---- stage 0: input region with clamp (or periodic wrap in x) ----
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
    for (int idx = tid; idx < SY * SX; idx += NT) s_am[idx] = T(0);
    __syncthreads();

    Prm<T> px, py;
    px.eps_h = py.eps_h = a.eps_h;
    px.hpos = py.hpos = a.hpos;
    px.cvis = py.cvis = a.cvis;
    px.cd2s = py.cd2s = T(0.5) * (T(1) - a.cvis);
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
        return c;
    };

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
            const T F = mass_flux<T, false>(m, c0, c1, q, border, ac, rg.lenx, rg.gx, rg.gx, rg.A, rg.A, rg.valx, dt, px);
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
            const T F = mass_flux<T, true>(m, c0, c1, q, border, ac, fg.len, fg.cos0, fg.cos1, fg.A0, fg.A1, fg.valy, dt, py);
            const bool lo = fj >= 1 && g - 1 >= 0;
            const bool hi = fj <= SY - 1 && g <= ny - 1;
            if (lo) s_am[(fj - 1) * SX + si] -= F * fg.ic0;
            if (hi) s_am[fj * SX + si] += ((border == 1) ? -F : F) * fg.ic1;
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
            e = e + dt * rg.invA * s_am[idx];
            if (!(e + D > T(0))) e = -D;
            if (any_sponge) {
                const int i = per ? ((gx % nx) + nx) % nx : gx, j = gy;
                if (i < a.sp_izq || i >= nx - a.sp_der || j < a.sp_sup || j >= ny - a.sp_inf) {
                    const T dx = rg.dxs, dy = rg.dys;
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
            }
        }
        s_es[idx] = e;
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
            if (!per && g > nx) continue;
            if (per && g > nx) continue;
            const int gw = per ? (g % nx) : g;
            const int border = per ? 0 : (g == 0 ? 1 : (g == nx ? 2 : 0));
            const bool ac = (gw < 2) || (gw > nx - 2);
            const RowGeo<T> rg = a.row[gy];
            const int sj = tj + 2;
            SC<T> m = st_cell(cxs(fi), sj), c0 = st_cell(cxs(fi + 1), sj);
            SC<T> c1 = st_cell(cxs(fi + 2), sj), q = st_cell(cxs(fi + 3), sj);
            const MF<T> f = mom_flux<T, false>(m, c0, c1, q, border, ac, rg.lenx, rg.gx, rg.gx, rg.valx, dt, px);
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
            const MF<T> f = mom_flux<T, true>(m, c0, c1, q, border, ac, fg.len, fg.cos0, fg.cos1, fg.valy, dt, py);
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
    float dtmin = 3.0e38f;
    const T tnow = (T)bc.t;
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
        // metrics from the state before the step
        {
            const T ho = eo + D;
            const T em = a.emax[go];
            if (eo > em && ho > a.eps_h) a.emax[go] = eo;
            if (a.track_arrival) {
                const T ta = a.tarr[go];
                if (ta < T(0) && tabs(eo - a.eini[go]) > a.dif_at) a.tarr[go] = tnow;
            }
        }
        const T aw = s_aw[idx];
        const T dtc = (aw < Num<T>::eps) ? Num<T>::big : Num<T>::div(T(2) * a.cfl * rg.A, aw);
        dtmin = fminf(dtmin, (float)dtc);
        const T val = dt * rg.invA;
        T qx = qxo + val * s_ax[idx];
        T qy = qyo + val * s_ay[idx];
        if (any_sponge) {
            const int i = gx, j = gy;
            if (i < a.sp_izq) { const T r = (T(a.sp_izq - i) - T(0.5)) / T(a.sp_izq); qx -= T(0.9) * r * r * qx; }
            if (i >= nx - a.sp_der) { const T r = (T(i - (nx - 1 - a.sp_der)) - T(0.5)) / T(a.sp_der); qx -= T(0.9) * r * r * qx; }
            if (j < a.sp_sup) { const T r = (T(a.sp_sup - j) - T(0.5)) / T(a.sp_sup); qy -= T(0.9) * r * r * qy; }
            if (j >= ny - a.sp_inf) { const T r = (T(j - (ny - 1 - a.sp_inf)) - T(0.5)) / T(a.sp_inf); qy -= T(0.9) * r * r * qy; }
        }
        const T es = s_es[so];
        const T hs = es + D;
        if (hs <= a.hpos) near_dry(hs, qxo, qyo, qx, qy, dt, a.mf0, a.vmax, a.eps_h);
        oe[go] = es;
        oqx[go] = qx;
        oqy[go] = qy;
    }
    dtmin = wave_min(dtmin);
    if ((tid & 63) == 0) s_red[tid >> 6] = dtmin;
    __syncthreads();
    if (tid == 0) {
        float m = s_red[0];
#pragma unroll
        for (int w = 1; w < NT / 64; w++) m = fminf(m, s_red[w]);
        atomicMin(&a.dtmin[slot], __float_as_uint(m));
    }
}

}  // namespace thy
