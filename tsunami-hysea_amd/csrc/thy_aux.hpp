// thy_aux.hpp -- kernels around the step: initial dt, Okada (FP64), metric init,
// output conversion, gauges, mass.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "thy_step.hpp"

namespace thy {

// ---------------------------------------------------------------------------
// Initial dt (src/GPU/Timestep.cu:11-33): per cell the sum of the four faces'
// wave speeds (src/GPU/Arista_kernel.cu:1237-1428), dt = 2*CFL*A/sum with CFL
// passed as float (src/GPU/Volumen_kernel.cu:78-94), min over cells.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_initial_dt(const StepArgs<T> a, int cur, unsigned int *out) {
    const int nx = a.nx, ny = a.ny, P = a.pitch;
    const bool per = a.periodic != 0;
    const T *e = a.e[cur], *qx = a.qx[cur], *qy = a.qy[cur];
    float m = 3.0e38f;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < (int64_t)nx * ny;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(c / nx), i = (int)(c - (int64_t)j * nx);
        const int64_t o = (int64_t)j * P + i;
        const T h = e[o] + a.D[o];
        T sum = T(0);
        const RowGeo<T> rg = a.row[j];
        // x faces i (left) and i+1 (right)
        for (int s = 0; s < 2; s++) {
            const int g = i + s;               // face index
            if (per && g == nx) { /* face nx of cell nx-1 == face 0 */ }
            int il = g - 1, ir = g;
            bool lb = false, rb = false;
            if (il < 0) { if (per) il += nx; else { il = 0; lb = true; } }
            if (ir >= nx) { if (per) ir -= nx; else { ir = nx - 1; rb = true; } }
            const int64_t ol = (int64_t)j * P + il, orr = (int64_t)j * P + ir;
            T h0 = e[ol] + a.D[ol], q0 = qx[ol];
            T h1 = e[orr] + a.D[orr], q1 = qx[orr];
            T nrm = T(1);
            if (lb) { q0 *= a.bizq; nrm = T(-1); T t = h0; h0 = h1; h1 = t; t = q0; q0 = q1; q1 = t; }
            if (rb) q1 *= a.bder;
            sum += rg.ilx * init_wave(h0, q0 * nrm, h1, q1 * nrm, a.eps_h);
        }
        for (int s = 0; s < 2; s++) {
            const int g = j + s;
            int jl = g - 1, jr = g;
            bool lb = false, rb = false;
            if (jl < 0) { jl = 0; lb = true; }
            if (jr >= ny) { jr = ny - 1; rb = true; }
            const int64_t ol = (int64_t)jl * P + i, orr = (int64_t)jr * P + i;
            T h0 = e[ol] + a.D[ol], q0 = qy[ol];
            T h1 = e[orr] + a.D[orr], q1 = qy[orr];
            T nrm = T(1);
            if (lb) { q0 *= a.bsup; nrm = T(-1); T t = h0; h0 = h1; h1 = t; t = q0; q0 = q1; q1 = t; }
            if (rb) q1 *= a.binf;
            sum += a.fy[g].len * init_wave(h0, q0 * nrm, h1, q1 * nrm, a.eps_h);
        }
        (void)h;
        const T cflf = (T)(float)a.cfl;
        const T dtc = (sum < Num<T>::eps) ? Num<T>::big : Num<T>::div(T(2) * cflf * rg.A, sum);
        m = fminf(m, (float)dtc);
    }
    m = wave_min(m);
    __shared__ float red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float r = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) r = fminf(r, red[w]);
        atomicMin(out, __float_as_uint(r));
    }
}

// ---------------------------------------------------------------------------
// Okada (1992) rectangular source evaluated on the free surface (z = 0).
// At z = 0 the DC3D real-source part-A terms and the image-source part-A terms
// cancel identically (src/GPU/DC3D.cu:716-835 with d = depth for both), and the
// part-C terms enter multiplied by z, so the vertical displacement is the
// rotated part-B contribution  U_Z = sum_corners (+-) (B1 sin(dip) + B2 cos(dip)).
// Corner guards, singular cases and log branches follow DC3D (:661-714, :72-157).
// Projection: Gauss-Krueger about the fault centroid (src/GPU/prtoxy.cu:33-54).
// ---------------------------------------------------------------------------
struct OkadaFault {
    double lon_c, lat_c, depth, L, W, strike, dip, rake, slip, H;
    double lon0, dlon, lat0, dlat;
};

__device__ inline double okada_uz(double x, double y, double depth, double dipdeg, double al1, double al2,
                                  double aw1, double aw2, double d1, double d2) {
    const double alpha = 2.0 / 3.0;
    const double alp3 = (1.0 - alpha) / alpha;
    const double pi2 = 6.283185307179586;
    double sd = sin(dipdeg * (pi2 / 360.0)), cd = cos(dipdeg * (pi2 / 360.0));
    if (fabs(cd) < 1e-6) {
        cd = 0.0;
        if (sd > 0.0) sd = 1.0;
        if (sd < 0.0) sd = -1.0;
    }
    const double sdcd = sd * cd, cdcd = cd * cd;
    double xi[2] = {x - al1, x - al2};
    if (fabs(xi[0]) < 1e-6) xi[0] = 0.0;
    if (fabs(xi[1]) < 1e-6) xi[1] = 0.0;
    const double p = y * cd + depth * sd;
    double q = y * sd - depth * cd;
    double et[2] = {p - aw1, p - aw2};
    if (fabs(q) < 1e-6) q = 0.0;
    if (fabs(et[0]) < 1e-6) et[0] = 0.0;
    if (fabs(et[1]) < 1e-6) et[1] = 0.0;
    if (q == 0.0 && ((xi[0] * xi[1] <= 0.0 && et[0] * et[1] == 0.0) || (et[0] * et[1] <= 0.0 && xi[0] * xi[1] == 0.0)))
        return 0.0;  // on a fault edge
    int kxi[2] = {0, 0}, ket[2] = {0, 0};
    const double r12 = sqrt(xi[0] * xi[0] + et[1] * et[1] + q * q);
    const double r21 = sqrt(xi[1] * xi[1] + et[0] * et[0] + q * q);
    const double r22 = sqrt(xi[1] * xi[1] + et[1] * et[1] + q * q);
    if (xi[0] < 0.0 && r21 + xi[1] < 1e-6) kxi[0] = 1;
    if (xi[0] < 0.0 && r22 + xi[1] < 1e-6) kxi[1] = 1;
    if (et[0] < 0.0 && r12 + et[1] < 1e-6) ket[0] = 1;
    if (et[0] < 0.0 && r22 + et[1] < 1e-6) ket[1] = 1;
    double uz = 0.0;
    for (int k = 0; k < 2; k++) {
        for (int j = 0; j < 2; j++) {
            const double X = xi[j], E = et[k];
            const double xi2 = X * X, q2 = q * q;
            const double r = sqrt(xi2 + E * E + q2);
            if (r == 0.0) continue;
            const double yy = E * cd + q * sd;
            const double dd = E * sd - q * cd;
            const double tt = (q == 0.0) ? 0.0 : atan(X * E / (q * r));
            double x11, y11, ale;
            if (kxi[k] == 1) x11 = 0.0;
            else x11 = 1.0 / (r * (r + X));
            if (ket[j] == 1) { ale = -log(r - E); y11 = 0.0; }
            else { ale = log(r + E); y11 = 1.0 / (r * (r + E)); }
            const double rd = r + dd;
            double ai3, ai4;
            if (cd != 0.0) {
                if (X == 0.0) ai4 = 0.0;
                else {
                    const double xx = sqrt(xi2 + q2);
                    ai4 = 1.0 / cdcd * (X / rd * sdcd + 2.0 * atan((E * (xx + q * cd) + xx * (r + xx) * sd) / (X * (r + xx) * cd)));
                }
                ai3 = (yy * cd / rd - ale + sd * log(rd)) / cdcd;
            } else {
                const double rd2 = rd * rd;
                ai3 = (E / rd + yy * q / rd2 - ale) / 2.0;
                ai4 = X * yy / rd2 / 2.0;
            }
            const double ai2 = log(rd) + ai3 * sd;
            const double qx = q * x11, qy = q * y11;
            double b1 = 0.0, b2 = 0.0;
            if (d1 != 0.0) {
                const double c = d1 / pi2;
                b1 += c * (-q / r + alp3 * yy / rd * sd);
                b2 += c * (q * qy - alp3 * ai2 * sd);
            }
            if (d2 != 0.0) {
                const double c = d2 / pi2;
                b1 += c * (-E * qx - tt - alp3 * X / rd * sdcd);
                b2 += c * (q * qx + alp3 * ai4 * sdcd);
            }
            const double v = b1 * sd + b2 * cd;
            uz += (j + k != 1) ? v : -v;
        }
    }
    return uz;
}

__device__ inline void gauss_krueger(double lat, double lon, double lat0, double lon0, double &x, double &y) {
    const double D2R = 0.017453292371619689;
    const double rlat = lat * D2R;
    const double sl = sin(rlat), cl = cos(rlat);
    const double v2 = cl * 0.0067397251 * cl + 1.0;
    const double al = lon - lon0;
    const double ph1 = lat + v2 * 0.5 * al * al * sl * cl * D2R;
    const double rph1 = ph1 * D2R;
    const double rph2 = (ph1 + lat0) * 0.5 * D2R;
    const double s1 = sin(rph1), s2 = sin(rph2);
    const double dd = sqrt(1.0 - s2 * 0.0066946053 * s2);
    const double r = 6335.4607362597517 / (dd * (dd * dd));
    const double an = 6378.16 / sqrt(1.0 - s1 * 0.0066946053 * s1);
    y = (ph1 - lat0) / (57.29578 / r);
    x = al * cl / (57.29578 / an) * (al * al * cos(rlat * 2.0) / 19696.838434850401 + 1.0);
}

__global__ void __launch_bounds__(256) k_okada(OkadaFault f, int nx, int ny, double *def, unsigned int *absmax) {
    const double RAD = 3.14159265358979323846 / 180.0;
    float m = 0.f;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < (int64_t)nx * ny;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(c / nx), i = (int)(c - (int64_t)j * nx);
        const double lonp = f.lon0 + i * f.dlon, latp = f.lat0 + j * f.dlat;
        double xp, yp;
        gauss_krueger(latp, lonp, f.lat_c, f.lon_c, xp, yp);
        const double ss = sin(RAD * f.strike), cs = cos(RAD * f.strike);
        const double xo = xp * ss + yp * cs;
        const double yo = -xp * cs + yp * ss;
        const double d1 = f.slip * cos(RAD * f.rake), d2 = f.slip * sin(RAD * f.rake);
        const double uz = okada_uz(xo, yo, f.depth, f.dip, -0.5 * f.L, 0.5 * f.L, -0.5 * f.W, 0.5 * f.W, d1, d2) / f.H;
        def[c] = uz;
        m = fmaxf(m, (float)fabs(uz));
    }
    // float max of |U_Z| for relative cropping (src/GPU/Deformacion.cu:13-20,132-136)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(absmax, __float_as_uint(m));
}

// crop + add: src/GPU/Deformacion.cu:22-53,132-142
template <typename T>
__global__ void __launch_bounds__(256) k_okada_apply(const double *def, int nx, int ny, int P, int crop_flag,
                                                     double crop_value, const unsigned int *absmax, T *e0, T *e1,
                                                     T *D, T *arr) {
    double thr = -1.0;
    if (crop_flag == 1) thr = crop_value * (double)__uint_as_float(*absmax);
    else if (crop_flag == 2) thr = crop_value;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < (int64_t)nx * ny;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(c / nx), i = (int)(c - (int64_t)j * nx);
        double u = def[c];
        if (crop_flag != 0 && fabs(u) < thr) u = 0.0;
        const int64_t o = (int64_t)j * P + i;
        D[o] = (T)((double)D[o] - u);
        e0[o] = (T)((double)e0[o] + u);
        if (e1) e1[o] = (T)((double)e1[o] + u);
        // eta_ini += U_Z (a cell that has arrived keeps its arrival time)
        if (arr && !Arrival<T>::arrived(arr[o])) arr[o] = (T)((double)arr[o] + u);
    }
}

// eta_max / eta_ini / arrival init (src/GPU/Metrics.cu:6-21,44-55 as called at
// src/GPU/ShallowWater.cu:305-309)
template <typename T>
__global__ void __launch_bounds__(256) k_init_metrics(const T *e, const T *D, int nx, int ny, int P, T eps_h,
                                                      T *emax, T *arr) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < (int64_t)nx * ny;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(c / nx), i = (int)(c - (int64_t)j * nx);
        const int64_t o = (int64_t)j * P + i;
        const T h = e[o] + D[o];
        const T v = (h < eps_h) ? T(-1e30) : e[o];
        emax[o] = v;
        if (arr) arr[o] = v;  // eta_ini, by the same kernel as eta_max (src/GPU/ShallowWater.cu:305-308)
    }
}

// metric update outside the step kernel (src/GPU/Metrics.cu:23-42,57-77)
template <typename T>
__global__ void __launch_bounds__(256) k_update_metrics(const T *e, const T *D, int nx, int ny, int P, T eps_h,
                                                        T dif_at, T t, T *emax, T *arr) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < (int64_t)nx * ny;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(c / nx), i = (int)(c - (int64_t)j * nx);
        const int64_t o = (int64_t)j * P + i;
        const T h = e[o] + D[o];
        if (e[o] > emax[o] && h > eps_h) emax[o] = e[o];
        if (arr && !Arrival<T>::arrived(arr[o]) && tabs(e[o] - arr[o]) > dif_at) arr[o] = Arrival<T>::enc(t);
    }
}

// output fields (src/GPU/netcdf_kernel.cu:6-126) into a dense [ny][nx] float array
template <typename T>
__global__ void __launch_bounds__(256) k_output(int which, const T *e, const T *qx, const T *qy, const T *D,
                                                const T *emax, const T *arr, int nx, int ny, int P, T eps_h,
                                                double Href, double Q, double Tscale, float *out) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < (int64_t)nx * ny;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(c / nx), i = (int)(c - (int64_t)j * nx);
        const int64_t o = (int64_t)j * P + i;
        double v;
        if (which == 0) {
            const T h = e[o] + D[o];
            v = (h < eps_h) ? -9999.0 : (double)e[o] * Href;
        } else if (which == 1 || which == 2) {
            const double h = (double)e[o] + (double)D[o];
            const double eh = (double)eps_h;
            const double f = (h < eh) ? 1.4142135623730951 * h / sqrt(h * h * h * h + eh * eh * eh * eh) * (Q / Href)
                                      : Q / (h * Href);
            v = (double)(which == 1 ? qx[o] : qy[o]) * f;
        } else if (which == 3) {
            v = (double)D[o] * Href;
        } else if (which == 4) {
            v = ((double)emax[o] < -1e20) ? -9999.0 : (double)emax[o] * Href;
        } else {
            v = (arr && Arrival<T>::arrived(arr[o])) ? Arrival<T>::dec(arr[o]) * Tscale : -1.0;
        }
        out[c] = (float)v;
    }
}

// per-block FP64 partial sums of h*area for the mass check
template <typename T>
__global__ void __launch_bounds__(256) k_mass(const T *e, const T *D, const RowGeo<T> *row, const double *area,
                                              int nx, int ny, int P, double *partial) {
    double s = 0.0;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < (int64_t)nx * ny;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(c / nx), i = (int)(c - (int64_t)j * nx);
        const int64_t o = (int64_t)j * P + i;
        s += ((double)e[o] + (double)D[o]) * area[j];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) r += red[w];
        partial[blockIdx.x] = r;
    }
}

}  // namespace thy
