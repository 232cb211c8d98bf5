/*
 * thy_oracle.c -- FP64 CPU restatement of the Tsunami-HySEA v1.3.0 time-stepping
 * path.  TEST INFRASTRUCTURE ONLY: the parity checker and the CPU baseline (see
 * thy_oracle.h).  Nothing in the product links this file.
 *
 * Citations are to the reference tree (src/...).  The arithmetic follows the
 * reference expression by expression, including the order of operations where it
 * changes rounding, so that the FP64 restatement tracks the FP64 reference to
 * within round-off.  Accumulation order per launch follows the reference's
 * even/odd edge colouring (src/GPU/Timestep.cu:79-110).
 */
#include "thy_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EPS DBL_EPSILON                      /* src/Constantes.hxx:11 */
#define SQRT2 1.41421356237309504880         /* M_SQRT2 */

struct orc_sim {
    orc_params p;
    int64_t n;
    int periodic;
    double *h, *H, *qx, *qy;       /* state1 = (h, H), state2 = (qx, qy) */
    double *a1x, *a1y, *a2x, *a2y; /* acumulador_1 (x, y), acumulador_2 (x, y) */
    double *dtc;                   /* per-cell dt (d_deltaTVolumenes) */
    double *eta_max, *t_arr, *eta_ini;
    double *area, *cosphi, *ancho, *alto;
};

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

static double *dalloc(int64_t n) { return (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }

orc_sim *orc_create(const orc_params *p, const double *area, const double *cosphi,
                    const double *ancho, const double *alto)
{
    orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
    int64_t n = (int64_t)p->nx * p->ny;
    s->p = *p;
    s->n = n;
    /* src/GPU/ShallowWater.cu:219 */
    s->periodic = (fabs(p->borde_izq - 2.0) < EPS) && (fabs(p->borde_der - 2.0) < EPS);
    s->h = dalloc(n); s->H = dalloc(n); s->qx = dalloc(n); s->qy = dalloc(n);
    s->a1x = dalloc(n); s->a1y = dalloc(n); s->a2x = dalloc(n); s->a2y = dalloc(n);
    s->dtc = dalloc(n);
    s->eta_max = dalloc(n); s->t_arr = dalloc(n); s->eta_ini = dalloc(n);
    s->area = dalloc(p->ny); s->cosphi = dalloc(p->ny);
    s->ancho = dalloc(p->ny + 1); s->alto = dalloc(p->ny);
    memcpy(s->area, area, sizeof(double) * p->ny);
    memcpy(s->cosphi, cosphi, sizeof(double) * p->ny);
    memcpy(s->ancho, ancho, sizeof(double) * (p->ny + 1));
    memcpy(s->alto, alto, sizeof(double) * p->ny);
    return s;
}

void orc_destroy(orc_sim *s)
{
    if (!s) return;
    free(s->h); free(s->H); free(s->qx); free(s->qy);
    free(s->a1x); free(s->a1y); free(s->a2x); free(s->a2y);
    free(s->dtc); free(s->eta_max); free(s->t_arr); free(s->eta_ini);
    free(s->area); free(s->cosphi); free(s->ancho); free(s->alto);
    free(s);
}

void orc_set_state(orc_sim *s, const double *h, const double *H, const double *qx, const double *qy)
{
    size_t b = sizeof(double) * (size_t)s->n;
    if (h) memcpy(s->h, h, b);
    if (H) memcpy(s->H, H, b);
    if (qx) memcpy(s->qx, qx, b);
    if (qy) memcpy(s->qy, qy, b);
}

void orc_get_state(const orc_sim *s, double *h, double *H, double *qx, double *qy)
{
    size_t b = sizeof(double) * (size_t)s->n;
    if (h) memcpy(h, s->h, b);
    if (H) memcpy(H, s->H, b);
    if (qx) memcpy(qx, s->qx, b);
    if (qy) memcpy(qy, s->qy, b);
}

void orc_get_metrics(const orc_sim *s, double *eta_max, double *t_arr, double *eta_ini)
{
    size_t b = sizeof(double) * (size_t)s->n;
    if (eta_max) memcpy(eta_max, s->eta_max, b);
    if (t_arr) memcpy(t_arr, s->t_arr, b);
    if (eta_ini) memcpy(eta_ini, s->eta_ini, b);
}

void orc_get_cell_dt(const orc_sim *s, double *dt_cell)
{
    memcpy(dt_cell, s->dtc, sizeof(double) * (size_t)s->n);
}

/* ------------------------------------------------------------------------- */
/* scalar helpers                                                            */
/* ------------------------------------------------------------------------- */

/* column index with clamp or periodic wrap: src/GPU/Arista_kernel.cu:6-21 */
static inline int col_index(int c, int periodic, int nx)
{
    if (c >= 0 && c < nx) return c;
    if (c < 0) return periodic ? nx + c : 0;
    return periodic ? c - nx : nx - 1;
}

static inline int row_clamp(int r, int ny) { return r < 0 ? 0 : (r >= ny ? ny - 1 : r); }

static inline double sgn(double x) { return (x > 0.0 ? 1.0 : 0.0) - (x < 0.0 ? 1.0 : 0.0); }

/* smooth monotone limiter: src/GPU/Arista_kernel.cu:45-68 */
double orc_limiter(double p0m, double p0, double p1, double p1p, double S)
{
    double c0 = p1 - p0;
    double c1 = (S < 0.0) ? (p1p - p1) : ((S > 0.0) ? (p0 - p0m) : 0.5 * (p1p - p0m));
    double sg = (c0 * c1 >= 0.0) ? 1.0 : 0.0;
    c0 = fabs(c0);
    c1 = fabs(c1);
    if (c0 < EPS) return 1.0;
    double aux = fabs((c1 - c0) / (c1 + c0));
    c1 = 0.5 * (c1 + c0) * (1.0 - aux * aux * aux);
    double vs = (c0 > c1) ? c1 / (c0 + EPS) : c0 / (c1 + EPS);
    return sg * vs * (1.0 + vs) / (1.0 + vs * vs);
}

/* desingularisation factor sqrt(2) h / sqrt(h^4 + max(h,eps)^4):
   src/GPU/Arista_kernel.cu:163-168 */
static inline double desing_factor(double h, double eps_h)
{
    double m = h > eps_h ? h : eps_h;
    return SQRT2 * h / sqrt(h * h * h * h + m * m * m * m);
}

/* desingularised velocity in the form sqrt(2) h q / sqrt(...):
   src/GPU/Arista_kernel.cu:414,418,1257-1259 */
double orc_desing(double h, double q, double eps_h)
{
    double m = h > eps_h ? h : eps_h;
    return SQRT2 * h * q / sqrt(h * h * h * h + m * m * m * m);
}

/* src/GPU/Arista_kernel.cu:70-77 */
static inline void wet_dry(double h0, double h1, double e0, double e1, double H0, double H1,
                           double eps_h, double *q0n, double *q1n)
{
    if ((h0 < eps_h) && (-e1 > H0)) *q1n = 0.0;
    if ((h1 < eps_h) && (-e0 > H1)) *q0n = 0.0;
}

/* src/GPU/Arista_kernel.cu:79-95 */
static inline void positivity(double h0, double h1, double dt, double area0, double area1, double *F)
{
    double dt0;
    if (*F > 0.0) dt0 = 0.25 * h0 * area0 / ((*F) + EPS);
    else          dt0 = 0.25 * h1 * area1 / (-(*F) + EPS);
    if (dt > dt0) *F *= dt0 / (dt + EPS);
}

/* src/GPU/Arista_kernel.cu:97-114 */
static inline void positivity_y(double h0, double h1, double cos0, double cos1, double dt,
                                double area0, double area1, double *F)
{
    double dt0;
    if (*F > 0.0) dt0 = 0.25 * h0 * cos0 * area0 / ((*F) + EPS);
    else          dt0 = 0.25 * h1 * cos1 * area1 / (-(*F) + EPS);
    if (dt > dt0) *F *= dt0 / (dt + EPS);
}

/* blend between the 2s and WAF fluxes: src/GPU/Arista_kernel.cu:205-208 */
static inline double blend_alpha(double hij, double hpos, int alfa_cero)
{
    if (alfa_cero) return 0.0;
    if (hij < hpos) return 1.0 / (1.0 + exp(-15.0 * (1.0 - 0.5 * hpos / (hij + EPS))));
    return 1.0;
}

/* WAF limiter coefficients, mass form (division by (l2-l1)):
   src/GPU/Arista_kernel.cu:183-199 */
static inline double waf_a0_mass(double l1, double l2, double lmax, double val,
                                 double e0m, double e0, double e1, double e1p, int alfa_cero)
{
    double xiL = 0.0, xiR = 0.0;
    if (!alfa_cero) {
        xiL = orc_limiter(e0m, e0, e1, e1p, l1);
        xiR = orc_limiter(e0m, e0, e1, e1p, l2);
    }
    if (fabs(l1 - l2) >= EPS) {
        double alr = sgn(l1) * (1.0 - xiL) + val * l1 * xiL;
        double blr = sgn(l2) * (1.0 - xiR) + val * l2 * xiR;
        return (alr - blr) * (l1 * l2) / (l2 - l1);
    }
    return lmax;
}

/* WAF coefficients a0, a1, momentum form (multiplication by 1/(l2-l1)):
   src/GPU/Arista_kernel.cu:374-393 */
static inline void waf_a0_a1(double l1, double l2, double lmax, double val,
                             double e0m, double e0, double e1, double e1p, int alfa_cero,
                             double *a0, double *a1)
{
    double xiL = 0.0, xiR = 0.0;
    if (!alfa_cero) {
        xiL = orc_limiter(e0m, e0, e1, e1p, l1);
        xiR = orc_limiter(e0m, e0, e1, e1p, l2);
    }
    if (fabs(l1 - l2) >= EPS) {
        double alr = sgn(l1) * (1.0 - xiL) + val * l1 * xiL;
        double blr = sgn(l2) * (1.0 - xiR) + val * l2 * xiR;
        double aux = 1.0 / (l2 - l1);
        *a0 = (alr - blr) * (l1 * l2) * aux;
        *a1 = (blr * l2 - alr * l1) * aux;
    } else {
        *a0 = lmax;
        *a1 = 0.0;
    }
}

/* ------------------------------------------------------------------------- */
/* mass (step 1) face fluxes                                                 */
/* ------------------------------------------------------------------------- */

/* 2s+WAF mass flux across an x-face: src/GPU/Arista_kernel.cu:132-219.
   Returns 0 when both sides are dry (no accumulation), else 1 with *F set. */
static int mass_x_waf(double h0m, double h0, double h1, double h1p, double q0x, double q1x,
                      double H0m, double H0, double H1, double H1p, double normal, double len,
                      double area0, double area1, double dt, double eps_h, double cosphi,
                      double hpos, int alfa_cero, double *F)
{
    if (!((h0 > EPS) || (h1 > EPS))) return 0;
    double il = len / cosphi;
    double q0n = q0x * normal, q1n = q1x * normal;
    double hij = 0.5 * (h0 + h1);
    double e0m = h0m - H0m, e0 = h0 - H0, e1 = h1 - H1, e1p = h1p - H1p;
    wet_dry(h0, h1, e0, e1, H0, H1, eps_h, &q0n, &q1n);
    double u0n = desing_factor(h0, eps_h) * q0n;
    double u1n = desing_factor(h1, eps_h) * q1n;
    q0n = u0n * h0;
    q1n = u1n * h1;
    double F2s = il * q0n * normal;
    double s0 = sqrt(h0), s1 = sqrt(h1);
    double uij = (s0 * u0n + s1 * u1n) / (s0 + s1 + EPS);
    double val = (2.0 * il * dt) / (area0 + area1);
    double c0 = sqrt(hij);
    double l1 = fmin(uij - c0, u0n - s0);
    double l2 = fmax(uij + c0, u1n + s1);
    double lmax = fmax(fabs(l2), fabs(l1));
    double a0 = waf_a0_mass(l1, l2, lmax, val, e0m, e0, e1, e1p, alfa_cero);
    double Hm = fmin(H0, H1);
    double deta = fmax(e1 + Hm, 0.0) - fmax(e0 + Hm, 0.0);
    double Fw = 0.5 * (q1n + q0n - a0 * deta) * il;
    double alfa = blend_alpha(hij, hpos, alfa_cero);
    double f = alfa * F2s + (1.0 - alfa) * Fw;
    positivity(h0, h1, dt, area0, area1, &f);
    *F = f;
    return 1;
}

/* 2s+WAF mass flux across a y-face: src/GPU/Arista_kernel.cu:688-778 */
static int mass_y_waf(double h0m, double h0, double h1, double h1p, double q0y, double q1y,
                      double H0m, double H0, double H1, double H1p, double normal, double len,
                      double area0, double area1, double dt, double eps_h, double cos0, double cos1,
                      double hpos, int alfa_cero, double *F)
{
    if (!((h0 > EPS) || (h1 > EPS))) return 0;
    double q0n = q0y * normal, q1n = q1y * normal;
    double hij = 0.5 * (h0 + h1);
    double c0 = sqrt(hij);
    double e0m = h0m - H0m, e0 = h0 - H0, e1 = h1 - H1, e1p = h1p - H1p;
    wet_dry(h0, h1, e0, e1, H0, H1, eps_h, &q0n, &q1n);
    double u0n = desing_factor(h0, eps_h) * q0n;
    double u1n = desing_factor(h1, eps_h) * q1n;
    q0n = u0n * h0;
    q1n = u1n * h1;
    double F2s = len * q0n * cos0 * normal;
    double s0 = sqrt(h0), s1 = sqrt(h1);
    double uij = (s0 * u0n + s1 * u1n) / (s0 + s1 + EPS);
    double val = (2.0 * len * dt) / (area0 + area1);
    double l1 = fmin(uij - c0, u0n - s0);
    double l2 = fmax(uij + c0, u1n + s1);
    double lmax = fmax(fabs(l2), fabs(l1));
    double a0 = waf_a0_mass(l1, l2, lmax, val, e0m, e0, e1, e1p, alfa_cero);
    double fl1 = q1n * cos1, fl0 = q0n * cos0;
    double Hm = fmin(H0, H1);
    double deta = 0.5 * (cos0 + cos1) * (fmax(e1 + Hm, 0.0) - fmax(e0 + Hm, 0.0))
                  - uij * uij * (cos1 - cos0);
    double Fw = 0.5 * (fl0 + fl1 - a0 * deta) * len;
    double alfa = blend_alpha(hij, hpos, alfa_cero);
    double f = alfa * F2s + (1.0 - alfa) * Fw;
    positivity_y(h0, h1, cos0, cos1, dt, area0, area1, &f);
    *F = f;
    return 1;
}

/* ------------------------------------------------------------------------- */
/* momentum (step 2) face fluxes                                             */
/* ------------------------------------------------------------------------- */

typedef struct { double h, qx, qy; } cell3;
typedef struct { double mx, my, px, py, wave; } mom_flux; /* acc2[vol0] -= m, acc2[vol1] -= p */

/* deep-water 2s momentum flux, x-face: src/GPU/Arista_kernel.cu:221-280 */
static void mom_x_2s(const cell3 *W0, const cell3 *W1, double H0, double H1, double cosphi,
                     double len, double dt, double area0, double area1, double cvis, mom_flux *o)
{
    double il = len / cosphi;
    double hij = 0.5 * (W0->h + W1->h);
    double Hm = fmin(H0, H1);
    double u0n = W0->qx / W0->h;
    double u1n = W1->qx / W1->h;
    double s0 = sqrt(W0->h), s1 = sqrt(W1->h);
    double uij = (s0 * u0n + s1 * u1n) / (s0 + s1 + EPS);
    double c0 = sqrt(hij);
    double lmax = fabs(uij) + c0;
    double deta = fmax(W1->h - H1 + Hm, 0.0) - fmax(W0->h - H0 + Hm, 0.0);
    double mx, my, px, py;
    if (uij < 0.0) {
        mx = u1n * W1->qx * il; px = -mx; mx += hij * deta * il;
        my = u1n * W1->qy * il; py = -my;
    } else {
        mx = u0n * W0->qx * il; px = -mx; mx += hij * deta * il;
        my = u0n * W0->qy * il; py = -my;
    }
    double cd = 0.5 * (1.0 - cvis);
    double val = (2.0 * il * dt) / (area0 + area1);
    double DES = cd * val * lmax * lmax * (W1->qx - W0->qx) * il;
    mx -= DES;
    px += DES;
    o->mx = mx; o->my = my; o->px = px; o->py = py;
    o->wave = il * lmax;
}

/* 2s+WAF momentum flux, x-face: src/GPU/Arista_kernel.cu:282-464 */
static int mom_x_waf(const cell3 *W0m, const cell3 *W0, const cell3 *W1, const cell3 *W1p,
                     double H0m, double H0, double H1, double H1p, double normal, double len,
                     double area0, double area1, double dt, double eps_h, double cosphi,
                     double cvis, double hpos, int alfa_cero, mom_flux *o)
{
    if (!((W0->h > EPS) || (W1->h > EPS))) return 0;
    double il = len / cosphi;
    double h0 = W0->h, h1 = W1->h;
    double q0n = W0->qx * normal, q1n = W1->qx * normal;
    double hij = 0.5 * (h0 + h1);
    double h1p = W1p->h, h0m = W0m->h;
    double e0m = h0m - H0m, e0 = h0 - H0, e1 = h1 - H1, e1p = h1p - H1p;
    wet_dry(h0, h1, e0, e1, H0, H1, eps_h, &q0n, &q1n);
    double f0 = desing_factor(h0, eps_h), f1 = desing_factor(h1, eps_h);
    double u0n = f0 * q0n, u1n = f1 * q1n;
    q0n = u0n * h0;
    q1n = u1n * h1;
    double s0 = sqrt(h0), s1 = sqrt(h1);
    double uij = (s0 * u0n + s1 * u1n) / (s0 + s1 + EPS);
    double val = (2.0 * il * dt) / (area0 + area1);
    double c0 = sqrt(hij);
    double l1 = fmin(uij - c0, u0n - s0);
    double l2 = fmax(uij + c0, u1n + s1);
    double lmax = fmax(fabs(l2), fabs(l1));
    double fl1x = q1n, fl1y = u1n * q1n, fl0x = q0n, fl0y = u0n * q0n;
    double dq = fl1x - fl0x;
    double Hm = fmin(H0, H1);
    double deta = fmax(e1 + Hm, 0.0) - fmax(e0 + Hm, 0.0);
    /* 2s part */
    double smy, spy, smz, spz;
    if (uij < 0.0) {
        smy = u1n * q1n * il * normal; spy = -smy; smy += hij * deta * il;
        smz = u1n * W1->qy * il; spz = -smz;
    } else {
        smy = u0n * q0n * il * normal; spy = -smy; smy += hij * deta * il;
        smz = u0n * W0->qy * il; spz = -smz;
    }
    double cd = 0.5 * (1.0 - cvis / (1.0 + exp(-20.0 * (1.0 - hpos / hij))));
    double des = cd * val * lmax * lmax * dq * il * normal;
    smy -= des;
    spy += des;
    /* WAF part */
    double a0, a1;
    waf_a0_a1(l1, l2, lmax, val, e0m, e0, e1, e1p, alfa_cero, &a0, &a1);
    double mxw = 0.5 * dq, pxw = 0.5 * dq;
    double myw = 0.5 * (fl1y - fl0y + hij * deta), pyw = myw;
    double DESx = 0.5 * a0 * deta;
    double DESy = 0.5 * a0 * dq + a1 * myw;
    mxw += fl0x - DESx;
    myw += fl0y - DESy;
    pxw += DESx - fl1x;
    pyw += DESy - fl1y;
    mxw *= il; myw *= il; pxw *= il; pyw *= il;
    /* transverse momentum: limited upwinding of the tangential velocity */
    double q0mt = W0m->qy * normal, q0t = W0->qy * normal, q1t = W1->qy * normal, q1pt = W1p->qy * normal;
    double p0m = orc_desing(h0m, q0mt, eps_h);
    double p0 = f0 * q0t;
    double p1 = f1 * q1t;
    double p1p = orc_desing(h1p, q1pt, eps_h);
    u0n = p0;
    u1n = p1;
    double sgm = -1.0 * (mxw < 0.0 ? 1.0 : 0.0) + 1.0 * (mxw > 0.0 ? 1.0 : 0.0);
    double us = fabs(uij) * sgm;
    double xi = alfa_cero ? 0.0 : orc_limiter(p0m, p0, p1, p1p, us);
    double uijt = 0.5 * (u0n + u1n) - 0.5 * ((1.0 - xi) * sgm + xi * us * val) * (u1n - u0n);
    double aux = mxw * uijt;
    double f3py = pyw * normal, f3pz = -aux * normal;
    double f3my = myw * normal, f3mz = aux * normal;
    double alfa = blend_alpha(hij, hpos, alfa_cero);
    o->mx = alfa * smy + (1.0 - alfa) * f3my;
    o->my = alfa * smz + (1.0 - alfa) * f3mz;
    o->px = alfa * spy + (1.0 - alfa) * f3py;
    o->py = alfa * spz + (1.0 - alfa) * f3pz;
    if (lmax < eps_h) lmax += eps_h;
    o->wave = il * lmax;
    return 1;
}

/* deep-water 2s momentum flux, y-face: src/GPU/Arista_kernel.cu:780-843 */
static void mom_y_2s(const cell3 *W0, const cell3 *W1, double H0, double H1, double len,
                     double area0, double area1, double dt, double cos0, double cos1,
                     double cvis, mom_flux *o)
{
    double icos0 = 1.0 / cos0, icos1 = 1.0 / cos1;
    double hij = 0.5 * (W0->h + W1->h);
    double Hm = fmin(H0, H1);
    double u0n = W0->qy / W0->h;
    double u1n = W1->qy / W1->h;
    double s0 = sqrt(W0->h), s1 = sqrt(W1->h);
    double uij = (s0 * u0n + s1 * u1n) / (s0 + s1 + EPS);
    double c0 = sqrt(hij);
    double lmax = fabs(uij) + c0;
    double deta = 0.5 * (cos0 + cos1) * (fmax(W1->h - H1 + Hm, 0.0) - fmax(W0->h - H0 + Hm, 0.0))
                  - uij * uij * (cos1 - cos0);
    double my, py, mx, px;
    if (uij < 0.0) {
        my = u1n * W1->qy * len; py = -my;
        my *= cos1 * icos0;
        my += hij * deta * len * icos0;
        mx = u1n * W1->qx * len; px = -mx;
    } else {
        my = u0n * W0->qy * len;
        py = -my * cos0 * icos1;
        my += hij * deta * len * icos0;
        mx = u0n * W0->qx * len; px = -mx;
    }
    double val = (2.0 * len * dt) / (area0 + area1);
    double cd = 0.5 * (1.0 - cvis);
    double DES = cd * val * lmax * lmax * (W1->qy * cos1 - W0->qy * cos0) * len;
    my -= DES * icos0;
    py += DES * icos1;
    o->mx = mx; o->my = my; o->px = px; o->py = py;
    o->wave = len * lmax;
}

/* 2s+WAF momentum flux, y-face: src/GPU/Arista_kernel.cu:845-1026 */
static int mom_y_waf(const cell3 *W0m, const cell3 *W0, const cell3 *W1, const cell3 *W1p,
                     double H0m, double H0, double H1, double H1p, double normal, double len,
                     double area0, double area1, double dt, double eps_h, double cos0, double cos1,
                     double cvis, double hpos, int alfa_cero, mom_flux *o)
{
    if (!((W0->h > EPS) || (W1->h > EPS))) return 0;
    double icos0 = 1.0 / cos0, icos1 = 1.0 / cos1;
    double q0n = W0->qy * normal, q1n = W1->qy * normal;
    double h0 = W0->h, h1 = W1->h;
    double hij = 0.5 * (h0 + h1);
    double c0 = sqrt(hij);
    double e0m = W0m->h - H0m, e0 = W0->h - H0, e1 = W1->h - H1, e1p = W1p->h - H1p;
    wet_dry(h0, h1, e0, e1, H0, H1, eps_h, &q0n, &q1n);
    double f0 = desing_factor(h0, eps_h), f1 = desing_factor(h1, eps_h);
    double u0n = f0 * q0n, u1n = f1 * q1n;
    double s0 = sqrt(h0), s1 = sqrt(h1);
    double uij = (s0 * u0n + s1 * u1n) / (s0 + s1 + EPS);
    double val = (2.0 * len * dt) / (area0 + area1);
    double l1 = fmin(uij - c0, u0n - s0);
    double l2 = fmax(uij + c0, u1n + s1);
    double lmax = fmax(fabs(l2), fabs(l1));
    double Hm = fmin(H0, H1);
    double deta = 0.5 * (cos0 + cos1) * (fmax(e1 + Hm, 0.0) - fmax(e0 + Hm, 0.0))
                  - uij * uij * (cos1 - cos0);
    /* 2s part (z: normal = y momentum, y: tangential = x momentum) */
    double smz, spz, smy, spy;
    if (uij < 0.0) {
        smz = u1n * q1n * len * normal; spz = -smz;
        smz *= cos1 * icos0;
        smz += hij * deta * len * icos0;
        smy = u1n * W1->qx * len; spy = -smy;
    } else {
        smz = u0n * q0n * len * normal;
        spz = -smz * cos0 * icos1;
        smz += hij * deta * len * icos0;
        smy = u0n * W0->qx * len; spy = -smy;
    }
    double cd = 0.5 * (1.0 - cvis / (1.0 + exp(-20.0 * (1.0 - hpos / hij))));
    double fl1x = q1n * cos1, fl1y = u1n * fl1x;
    double fl0x = q0n * cos0, fl0y = u0n * fl0x;
    double dq = fl1x - fl0x;
    double des = cd * val * lmax * lmax * dq * len;
    smz -= des * icos0;
    spz += des * icos1;
    /* WAF part */
    double a0, a1;
    waf_a0_a1(l1, l2, lmax, val, e0m, e0, e1, e1p, alfa_cero, &a0, &a1);
    double mxw = 0.5 * dq, pxw = 0.5 * dq;
    double myw = 0.5 * (fl1y - fl0y + hij * deta), pyw = myw;
    double DESx = 0.5 * a0 * deta;
    double DESy = 0.5 * a0 * dq + a1 * myw;
    mxw += fl0x - DESx;
    myw += fl0y - DESy;
    pxw += DESx - fl1x;
    pyw += DESy - fl1y;
    mxw *= len; myw *= len; pxw *= len; pyw *= len;
    /* tangential velocity */
    double h0m = W0m->h, h1p = W1p->h;
    double q0mt = -W0m->qx * normal, q0t = -W0->qx * normal;
    double q1t = -W1->qx * normal, q1pt = -W1p->qx * normal;
    double p0m = orc_desing(h0m, q0mt, eps_h);
    double p0 = f0 * q0t;
    double p1 = f1 * q1t;
    double p1p = orc_desing(h1p, q1pt, eps_h);
    u0n = p0;
    u1n = p1;
    double sgm = -1.0 * (mxw < 0.0 ? 1.0 : 0.0) + 1.0 * (mxw > 0.0 ? 1.0 : 0.0);
    double us = fabs(uij) * sgm;
    double xi = alfa_cero ? 0.0 : orc_limiter(p0m, p0, p1, p1p, us);
    double uijt = 0.5 * (u0n * icos0 + u1n * icos1)
                  - 0.5 * ((1.0 - xi) * sgm + xi * us * val) * (u1n * icos1 - u0n * icos0);
    double aux = mxw * uijt;
    double f3py = aux * normal, f3pz = pyw * normal * icos1;
    double f3my = -aux * normal, f3mz = myw * normal * icos0;
    if (lmax < eps_h) lmax += eps_h;
    double alfa = blend_alpha(hij, hpos, alfa_cero);
    o->mx = alfa * smy + (1.0 - alfa) * f3my;
    o->my = alfa * smz + (1.0 - alfa) * f3mz;
    o->px = alfa * spy + (1.0 - alfa) * f3py;
    o->py = alfa * spz + (1.0 - alfa) * f3pz;
    o->wave = len * lmax;
    return 1;
}

/* ------------------------------------------------------------------------- */
/* edge sweeps (the reference's 8 edge launches per step)                    */
/* ------------------------------------------------------------------------- */

/* x-faces of one row for the mass pass, faces i = first, first+2, ...:
   src/GPU/Arista_kernel.cu:466-563 */
static void sweep_mass_x_row(orc_sim *s, double dt, int j, int first)
{
    const orc_params *p = &s->p;
    const int nx = p->nx, per = s->periodic;
    const int64_t rb = (int64_t)j * nx;
    const double len = s->alto[j], area = s->area[j], cph = s->cosphi[j];
    for (int i = first; i <= nx; i += 2) {
        int64_t v0 = rb + i - 1, v1 = v0 + 1;
        double normal = 1.0;
        int front = 0;
        int c = col_index(i - 1, per, nx);
        double h0 = s->h[rb + c], H0 = s->H[rb + c], q0 = s->qx[rb + c];
        if ((i - 1 < 0) && !per) { q0 *= p->borde_izq; v0 = rb; v1 = -1; normal = -1.0; front = 1; }
        c = col_index(i, per, nx);
        double h1 = s->h[rb + c], H1 = s->H[rb + c], q1 = s->qx[rb + c];
        if ((i >= nx) && !per) { q1 *= p->borde_der; v0 = rb + nx - 1; v1 = -1; }
        c = col_index(i - 2, per, nx);
        double h0m = s->h[rb + c], H0m = s->H[rb + c];
        c = col_index(i + 1, per, nx);
        double h1p = s->h[rb + c], H1p = s->H[rb + c];
        if (per && i == 0) v0 = -1;
        if (per && i == nx) v1 = -1;
        int alfa_cero = (i < 2) || (i > nx - 2);
        double F;
        if (h0 > p->hpos && h1 > p->hpos && h0m > p->hpos && h1p > p->hpos && !alfa_cero) {
            /* src/GPU/Arista_kernel.cu:118-130 */
            F = q0 * len / cph;
            positivity(h0, h1, dt, area, area, &F);
        } else {
            double t;
            if (front) {
                t = H0m; H0m = H1p; H1p = t;
                t = H0; H0 = H1; H1 = t;
                t = h0m; h0m = h1p; h1p = t;
                t = h0; h0 = h1; h1 = t;
                t = q0; q0 = q1; q1 = t;
            }
            if (!mass_x_waf(h0m, h0, h1, h1p, q0, q1, H0m, H0, H1, H1p, normal, len, area, area,
                            dt, p->eps_h, cph, p->hpos, alfa_cero, &F))
                continue;
        }
        if (v0 != -1) s->a1x[v0] -= F;
        if (v1 != -1) s->a1x[v1] += F;
    }
}

/* y-faces of face-row j for the mass pass: src/GPU/Arista_kernel.cu:1028-1125 */
static void sweep_mass_y_row(orc_sim *s, double dt, int j)
{
    const orc_params *p = &s->p;
    const int nx = p->nx, ny = p->ny;
    const double len = s->ancho[j];
    const int r0 = row_clamp(j - 1, ny), r1 = row_clamp(j, ny);
    const int r0m = row_clamp(j - 2, ny), r1p = row_clamp(j + 1, ny);
    for (int i = 0; i < nx; i++) {
        int64_t v0 = (int64_t)(j - 1) * nx + i, v1 = v0 + nx;
        double normal = 1.0;
        int front = 0;
        int64_t c = (int64_t)r0 * nx + i;
        double h0 = s->h[c], H0 = s->H[c], q0 = s->qy[c];
        double area0 = s->area[r0], cos0 = s->cosphi[r0];
        if (j - 1 < 0) { q0 *= p->borde_sup; v0 = i; v1 = -1; front = 1; normal = -1.0; }
        c = (int64_t)r1 * nx + i;
        double h1 = s->h[c], H1 = s->H[c], q1 = s->qy[c];
        double area1 = s->area[r1], cos1 = s->cosphi[r1];
        if (j >= ny) { q1 *= p->borde_inf; v0 = (int64_t)(j - 1) * nx + i; v1 = -1; }
        c = (int64_t)r0m * nx + i;
        double h0m = s->h[c], H0m = s->H[c];
        c = (int64_t)r1p * nx + i;
        double h1p = s->h[c], H1p = s->H[c];
        int alfa_cero = (j < 2) || (j > ny - 2);
        double F;
        if (h0 > p->hpos && h1 > p->hpos && h0m > p->hpos && h1p > p->hpos && !alfa_cero) {
            /* src/GPU/Arista_kernel.cu:674-686 */
            F = len * q0 * cos0;
            positivity_y(h0, h1, cos0, cos1, dt, area0, area1, &F);
        } else {
            double t;
            if (front) {
                t = H0m; H0m = H1p; H1p = t;
                t = H0; H0 = H1; H1 = t;
                t = h0m; h0m = h1p; h1p = t;
                t = h0; h0 = h1; h1 = t;
                t = q0; q0 = q1; q1 = t;
                t = area0; area0 = area1; area1 = t;
                t = cos0; cos0 = cos1; cos1 = t;
            }
            if (!mass_y_waf(h0m, h0, h1, h1p, q0, q1, H0m, H0, H1, H1p, normal, len, area0, area1,
                            dt, p->eps_h, cos0, cos1, p->hpos, alfa_cero, &F))
                continue;
        }
        if (v0 >= 0) s->a1x[v0] -= F / cos0;
        if (v1 != -1) s->a1x[v1] += F / cos1;
    }
}

static inline void load3(const orc_sim *s, int64_t c, cell3 *w, double *H)
{
    w->h = s->h[c]; w->qx = s->qx[c]; w->qy = s->qy[c];
    *H = s->H[c];
}

static inline void acc_mom(orc_sim *s, int64_t v0, int64_t v1, const mom_flux *f)
{
    if (v0 >= 0) { s->a1y[v0] += f->wave; s->a2x[v0] -= f->mx; s->a2y[v0] -= f->my; }
    if (v1 != -1) { s->a1y[v1] += f->wave; s->a2x[v1] -= f->px; s->a2y[v1] -= f->py; }
}

/* x-faces of one row for the momentum pass: src/GPU/Arista_kernel.cu:565-670 */
static void sweep_mom_x_row(orc_sim *s, double dt, int j, int first)
{
    const orc_params *p = &s->p;
    const int nx = p->nx, per = s->periodic;
    const int64_t rb = (int64_t)j * nx;
    const double len = s->alto[j], area = s->area[j], cph = s->cosphi[j];
    for (int i = first; i <= nx; i += 2) {
        int64_t v0 = rb + i - 1, v1 = v0 + 1;
        double normal = 1.0;
        int front = 0;
        cell3 W0, W1, W0m, W1p;
        double H0, H1, H0m, H1p;
        load3(s, rb + col_index(i - 1, per, nx), &W0, &H0);
        if ((i - 1 < 0) && !per) { W0.qx *= p->borde_izq; v0 = rb; v1 = -1; front = 1; normal = -1.0; }
        load3(s, rb + col_index(i, per, nx), &W1, &H1);
        if ((i >= nx) && !per) { W1.qx *= p->borde_der; v0 = rb + nx - 1; v1 = -1; }
        load3(s, rb + col_index(i - 2, per, nx), &W0m, &H0m);
        load3(s, rb + col_index(i + 1, per, nx), &W1p, &H1p);
        if (per && i == 0) v0 = -1;
        if (per && i == nx) v1 = -1;
        int alfa_cero = (i < 2) || (i > nx - 2);
        mom_flux f;
        if (W0.h > p->hpos && W1.h > p->hpos && W0m.h > p->hpos && W1p.h > p->hpos && !alfa_cero) {
            mom_x_2s(&W0, &W1, H0, H1, cph, len, dt, area, area, p->cvis, &f);
        } else {
            if (front) {
                cell3 tw; double t;
                tw = W0; W0 = W1; W1 = tw;
                tw = W0m; W0m = W1p; W1p = tw;
                t = H0; H0 = H1; H1 = t;
                t = H0m; H0m = H1p; H1p = t;
            }
            if (!mom_x_waf(&W0m, &W0, &W1, &W1p, H0m, H0, H1, H1p, normal, len, area, area, dt,
                           p->eps_h, cph, p->cvis, p->hpos, alfa_cero, &f))
                continue;
        }
        acc_mom(s, v0, v1, &f);
    }
}

/* y-faces of face-row j for the momentum pass: src/GPU/Arista_kernel.cu:1127-1233 */
static void sweep_mom_y_row(orc_sim *s, double dt, int j)
{
    const orc_params *p = &s->p;
    const int nx = p->nx, ny = p->ny;
    const double len = s->ancho[j];
    const int r0 = row_clamp(j - 1, ny), r1 = row_clamp(j, ny);
    const int r0m = row_clamp(j - 2, ny), r1p = row_clamp(j + 1, ny);
    for (int i = 0; i < nx; i++) {
        int64_t v0 = (int64_t)(j - 1) * nx + i, v1 = v0 + nx;
        double normal = 1.0;
        int front = 0;
        cell3 W0, W1, W0m, W1p;
        double H0, H1, H0m, H1p;
        load3(s, (int64_t)r0 * nx + i, &W0, &H0);
        double area0 = s->area[r0], cos0 = s->cosphi[r0];
        if (j - 1 < 0) { W0.qy *= p->borde_sup; v0 = i; v1 = -1; front = 1; normal = -1.0; }
        load3(s, (int64_t)r1 * nx + i, &W1, &H1);
        double area1 = s->area[r1], cos1 = s->cosphi[r1];
        if (j >= ny) { W1.qy *= p->borde_inf; v0 = (int64_t)(j - 1) * nx + i; v1 = -1; }
        load3(s, (int64_t)r0m * nx + i, &W0m, &H0m);
        load3(s, (int64_t)r1p * nx + i, &W1p, &H1p);
        int alfa_cero = (j < 2) || (j > ny - 2);
        mom_flux f;
        if (W0.h > p->hpos && W1.h > p->hpos && W0m.h > p->hpos && W1p.h > p->hpos && !alfa_cero) {
            mom_y_2s(&W0, &W1, H0, H1, len, area0, area1, dt, cos0, cos1, p->cvis, &f);
        } else {
            if (front) {
                cell3 tw; double t;
                tw = W0; W0 = W1; W1 = tw;
                tw = W0m; W0m = W1p; W1p = tw;
                t = H0; H0 = H1; H1 = t;
                t = H0m; H0m = H1p; H1p = t;
                t = area0; area0 = area1; area1 = t;
                t = cos0; cos0 = cos1; cos1 = t;
            }
            if (!mom_y_waf(&W0m, &W0, &W1, &W1p, H0m, H0, H1, H1p, normal, len, area0, area1, dt,
                           p->eps_h, cos0, cos1, p->cvis, p->hpos, alfa_cero, &f))
                continue;
        }
        acc_mom(s, v0, v1, &f);
    }
}

/* ------------------------------------------------------------------------- */
/* cell updates                                                              */
/* ------------------------------------------------------------------------- */

/* src/GPU/Volumen_kernel.cu:96-158 (+ spongeLayerPaso1 :6-14) */
static void cell_update_mass(orc_sim *s, double dt)
{
    const orc_params *p = &s->p;
    const int nx = p->nx, ny = p->ny;
    const int any_sponge = (p->sponge_izq > 0) || (p->sponge_der > 0) || (p->sponge_sup > 0) || (p->sponge_inf > 0);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; j++) {
        double area = s->area[j];
        double val = dt / area;
        for (int i = 0; i < nx; i++) {
            int64_t c = (int64_t)j * nx + i;
            double hs = s->h[c] + val * s->a1x[c];
            hs *= (hs > 0.0) ? 1.0 : 0.0;
            double h0s = fmax(s->H[c] + p->hmin + p->sea_level, 0.0);
            if (any_sponge &&
                ((i < p->sponge_izq) || (i >= nx - p->sponge_der) || (j < p->sponge_sup) || (j >= ny - p->sponge_inf))) {
                double dx = s->ancho[j], dy = s->alto[j];
                double dl = (p->sponge_izq - i - 0.5) * dx; if (dl <= 0.0) dl = 1e30;
                double dr = (i - (nx - 1 - p->sponge_der) - 0.5) * dx; if (dr <= 0.0) dr = 1e30;
                double ds = (p->sponge_sup - j - 0.5) * dy; if (ds <= 0.0) ds = 1e30;
                double di = (j - (ny - 1 - p->sponge_inf) - 0.5) * dy; if (di <= 0.0) di = 1e30;
                double dist, ls;
                if ((dl <= dr) && (dl <= ds) && (dl <= di)) { dist = dl; ls = dx * p->sponge_izq; }
                else if ((dr <= dl) && (dr <= ds) && (dr <= di)) { dist = dr; ls = dx * p->sponge_der; }
                else if ((ds <= dl) && (ds <= dr) && (ds <= di)) { dist = ds; ls = dy * p->sponge_sup; }
                else { dist = di; ls = dy * p->sponge_inf; }
                double r = dist / ls;
                double co = 0.9 * r * r;
                hs -= co * (hs - h0s);
                hs *= (hs >= 0.0) ? 1.0 : 0.0;
            }
            s->h[c] = hs;
        }
    }
}

/* near-dry filter + velocity cap: src/GPU/Volumen_kernel.cu:49-76;
   semi-implicit Manning friction: src/GPU/Volumen_kernel.cu:25-47 */
void orc_friction_filter(double hs, double qxo, double qyo, double *qx, double *qy,
                         double dt, double mf0, double vmax, double eps_h)
{
    double factor = hs / eps_h;
    double aux0 = 1.0 / (factor * factor * factor * factor + EPS);
    double aux = exp(-dt * aux0);
    *qx *= aux;
    *qy *= aux;
    if (vmax > 0.0) {
        double h = hs;
        double m = h > eps_h ? h : eps_h;
        double hm = sqrt(h * h * h * h + m * m * m * m);
        double f = SQRT2 * h / hm;
        double ux = f * (*qx), uy = f * (*qy);
        double u = sqrt(ux * ux + uy * uy);
        if (u > vmax) { ux *= vmax / u; uy *= vmax / u; }
        *qx = ux * h;
        *qy = uy * h;
    }
    {
        double h1 = hs;
        double m = h1 > eps_h ? h1 : eps_h;
        double h1m = sqrt(h1 * h1 * h1 * h1 + m * m * m * m);
        double f = SQRT2 * h1 / h1m;
        double u1x = f * (*qx), u1y = f * (*qy);
        double mw = hs > eps_h ? hs : eps_h;
        double u1 = SQRT2 * sqrt(qxo * qxo + qyo * qyo) * hs / sqrt(hs * hs * hs * hs + mw * mw * mw * mw);
        if (h1 > 0.0) {
            double c1 = dt * mf0 * fabs(u1) / (pow(h1, 4.0 / 3.0) + EPS);
            double g = h1 / (1.0 + c1);
            *qx = g * u1x;
            *qy = g * u1y;
        }
    }
}

/* src/GPU/Volumen_kernel.cu:160-221 (+ spongeLayerPaso2 :16-23) */
static void cell_update_momentum(orc_sim *s, double dt)
{
    const orc_params *p = &s->p;
    const int nx = p->nx, ny = p->ny;
    const int any_sponge = (p->sponge_izq > 0) || (p->sponge_der > 0) || (p->sponge_sup > 0) || (p->sponge_inf > 0);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; j++) {
        double area = s->area[j];
        double val = dt / area;
        for (int i = 0; i < nx; i++) {
            int64_t c = (int64_t)j * nx + i;
            s->dtc[c] = (s->a1y[c] < EPS) ? 1e30 : (2.0 * p->cfl * area) / s->a1y[c];
            double qxo = s->qx[c], qyo = s->qy[c];
            double qx = qxo + val * s->a2x[c];
            double qy = qyo + val * s->a2y[c];
            double h0s = fmax(s->H[c] + p->hmin + p->sea_level, 0.0);
            double v0 = 0.0;
            if (any_sponge) {
                double dx = s->ancho[j], dy = s->alto[j], d, r, co;
                if (i < p->sponge_izq) {
                    d = (p->sponge_izq - i - 0.5) * dx; r = d / (dx * p->sponge_izq); co = 0.9 * r * r;
                    qx -= co * (qx - h0s * v0);
                }
                if (i >= nx - p->sponge_der) {
                    d = (i - (nx - 1 - p->sponge_der) - 0.5) * dx; r = d / (dx * p->sponge_der); co = 0.9 * r * r;
                    qx -= co * (qx - h0s * v0);
                }
                if (j < p->sponge_sup) {
                    d = (p->sponge_sup - j - 0.5) * dy; r = d / (dy * p->sponge_sup); co = 0.9 * r * r;
                    qy -= co * (qy - h0s * v0);
                }
                if (j >= ny - p->sponge_inf) {
                    d = (j - (ny - 1 - p->sponge_inf) - 0.5) * dy; r = d / (dy * p->sponge_inf); co = 0.9 * r * r;
                    qy -= co * (qy - h0s * v0);
                }
            }
            if (s->h[c] <= p->hpos)
                orc_friction_filter(s->h[c], qxo, qyo, &qx, &qy, dt, p->mf0, p->vmax, p->eps_h);
            s->qx[c] = qx;
            s->qy[c] = qy;
        }
    }
}

static void zero_acc(orc_sim *s)
{
    size_t b = sizeof(double) * (size_t)s->n;
    memset(s->a1x, 0, b); memset(s->a1y, 0, b);
    memset(s->a2x, 0, b); memset(s->a2y, 0, b);
}

/* ------------------------------------------------------------------------- */
/* reductions                                                                */
/* ------------------------------------------------------------------------- */

/* launch arithmetic of src/GPU/Reduccion_kernel.cu:130-138 */
static void red_plan(int64_t n, int *blocks, int *threads)
{
    int t = (n < 256 * 2) ? (int)(n / 2) : 256;
    int64_t b = (t > 0) ? n / ((int64_t)t * 2) : 0;
    if (b > 128) b = 128;
    *threads = t;
    *blocks = (int)b;
}

static int red_threads_ok(int t)
{
    return t == 512 || t == 256 || t == 128 || t == 64 || t == 32 || t == 16 || t == 8 ||
           t == 4 || t == 2 || t == 1;
}

/* one reduce6_min launch (src/GPU/Reduccion_kernel.cu:11-85), read from a snapshot
   of the input (race-free reading of the in-place call at :151) */
static void red_stage(double *d, int64_t n, int blocks, int threads)
{
    double *out = (double *)malloc(sizeof(double) * (size_t)(blocks > 0 ? blocks : 1));
    int64_t grid = (int64_t)threads * 2 * blocks;
    for (int b = 0; b < blocks; b++) {
        double bmin = INFINITY;
        for (int t = 0; t < threads; t++) {
            int64_t i = (int64_t)b * threads * 2 + t;
            double m;
            if (i + threads >= n) m = d[i];
            else m = fmin(d[i], d[i + threads]);
            i += grid;
            while (i < n) {
                double v = (i + threads >= n) ? d[i] : fmin(d[i], d[i + threads]);
                m = fmin(m, v);
                i += grid;
            }
            bmin = fmin(bmin, m);
        }
        out[b] = bmin;
    }
    for (int b = 0; b < blocks; b++) d[b] = out[b];
    free(out);
}

/* src/GPU/Reduccion_kernel.cu:141-173 */
double orc_reduce_min(const double *v, int64_t n, int quirk)
{
    if (n <= 0) return 1e30;
    if (!quirk) {
        double m = v[0];
        for (int64_t i = 1; i < n; i++) m = fmin(m, v[i]);
        return m;
    }
    if (n == 1) return v[0];
    double *d = (double *)malloc(sizeof(double) * (size_t)n);
    memcpy(d, v, sizeof(double) * (size_t)n);
    int blocks, threads;
    red_plan(n, &blocks, &threads);
    if (red_threads_ok(threads)) red_stage(d, n, blocks, threads);
    int64_t s = blocks;
    while (s > 1) {
        red_plan(s, &blocks, &threads);
        if (red_threads_ok(threads)) red_stage(d, s, blocks, threads);
        s = s / ((int64_t)threads * 2);
    }
    double r = d[0];
    free(d);
    return r;
}

float orc_reduce_min_f(const float *v, int64_t n, int quirk)
{
    double *d = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; i++) d[i] = (double)v[i];
    float r = (float)orc_reduce_min(d, n, quirk);
    free(d);
    return r;
}

/* ------------------------------------------------------------------------- */
/* initial dt: src/GPU/Timestep.cu:11-33                                     */
/* ------------------------------------------------------------------------- */

/* src/GPU/Arista_kernel.cu:1237-1280 */
static double init_wave(const cell3 *W0, const cell3 *W1, double nx_, double ny_, double len,
                        double eps_h, double cos0)
{
    double q0 = W0->qx * nx_ + W0->qy * ny_;
    double q1 = W1->qx * nx_ + W1->qy * ny_;
    double h0 = W0->h, h1 = W1->h;
    double hij = 0.5 * (h0 + h1);
    double u0n = orc_desing(h0, q0, eps_h);
    double u1n = orc_desing(h1, q1, eps_h);
    double s0 = sqrt(h0), s1 = sqrt(h1);
    double uij = (s0 * u0n + s1 * u1n) / (s0 + s1 + EPS);
    double lmax = fabs(uij) + sqrt(hij);
    if (lmax < eps_h) lmax += eps_h;
    return len * lmax / cos0;
}

double orc_initial_dt(orc_sim *s)
{
    const orc_params *p = &s->p;
    const int nx = p->nx, ny = p->ny, per = s->periodic;
    memset(s->a1y, 0, sizeof(double) * (size_t)s->n);
    /* x-faces, src/GPU/Arista_kernel.cu:1282-1357 */
    for (int first = 0; first < 2; first++) {
#pragma omp parallel for schedule(static)
        for (int j = 0; j < ny; j++) {
            int64_t rb = (int64_t)j * nx;
            double len = s->alto[j], cph = s->cosphi[j];
            for (int i = first; i <= nx; i += 2) {
                int64_t v0 = rb + i - 1, v1 = v0 + 1;
                double normal = 1.0, H0, H1;
                int front = 0;
                cell3 W0, W1;
                load3(s, rb + col_index(i - 1, per, nx), &W0, &H0);
                if ((i - 1 < 0) && !per) { W0.qx *= p->borde_izq; v0 = rb; v1 = -1; normal = -1.0; front = 1; }
                load3(s, rb + col_index(i, per, nx), &W1, &H1);
                if ((i >= nx) && !per) { W1.qx *= p->borde_der; v0 = rb + nx - 1; v1 = -1; }
                if (per && i == 0) v0 = -1;
                if (per && i == nx) v1 = -1;
                if (front) { cell3 t = W0; W0 = W1; W1 = t; }
                double w = init_wave(&W0, &W1, normal, 0.0, len, p->eps_h, cph);
                if (v0 >= 0) s->a1y[v0] += w;
                if (v1 != -1) s->a1y[v1] += w;
            }
        }
    }
    /* y-faces, src/GPU/Arista_kernel.cu:1359-1428 */
    for (int first = 0; first < 2; first++) {
#pragma omp parallel for schedule(static)
        for (int j = first; j <= ny; j += 2) {
            double len = s->ancho[j];
            for (int i = 0; i < nx; i++) {
                int64_t v0 = (int64_t)(j - 1) * nx + i, v1 = v0 + nx;
                double normal = 1.0, H0, H1;
                int front = 0;
                cell3 W0, W1;
                load3(s, (int64_t)row_clamp(j - 1, ny) * nx + i, &W0, &H0);
                if (j - 1 < 0) { W0.qy *= p->borde_sup; v0 = i; v1 = -1; normal = -1.0; front = 1; }
                load3(s, (int64_t)row_clamp(j, ny) * nx + i, &W1, &H1);
                if (j >= ny) { W1.qy *= p->borde_inf; v0 = (int64_t)(j - 1) * nx + i; v1 = -1; }
                if (front) { cell3 t = W0; W0 = W1; W1 = t; }
                double w = init_wave(&W0, &W1, 0.0, normal, len, p->eps_h, 1.0);
                if (v0 >= 0) s->a1y[v0] += w;
                if (v1 != -1) s->a1y[v1] += w;
            }
        }
    }
    /* src/GPU/Volumen_kernel.cu:78-94 (CFL passed as float) */
    double cflf = (double)(float)p->cfl;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; j++) {
        double area = s->area[j];
        for (int i = 0; i < nx; i++) {
            int64_t c = (int64_t)j * nx + i;
            s->dtc[c] = (s->a1y[c] < EPS) ? 1e30 : (2.0 * cflf * area) / s->a1y[c];
        }
    }
    memset(s->a1y, 0, sizeof(double) * (size_t)s->n);
    return orc_reduce_min(s->dtc, s->n, p->quirk_reduction);
}

/* ------------------------------------------------------------------------- */
/* one step: src/GPU/Timestep.cu:70-120                                      */
/* ------------------------------------------------------------------------- */

void orc_mass_pass(orc_sim *s, double dt)
{
    const int ny = s->p.ny;
    zero_acc(s);
    for (int first = 0; first < 2; first++) {
#pragma omp parallel for schedule(static)
        for (int j = 0; j < ny; j++) sweep_mass_x_row(s, dt, j, first);
    }
    for (int first = 0; first < 2; first++) {
#pragma omp parallel for schedule(static)
        for (int j = first; j <= ny; j += 2) sweep_mass_y_row(s, dt, j);
    }
    cell_update_mass(s, dt);
}

double orc_momentum_pass(orc_sim *s, double dt)
{
    const int ny = s->p.ny;
    zero_acc(s);
    for (int first = 0; first < 2; first++) {
#pragma omp parallel for schedule(static)
        for (int j = 0; j < ny; j++) sweep_mom_x_row(s, dt, j, first);
    }
    for (int first = 0; first < 2; first++) {
#pragma omp parallel for schedule(static)
        for (int j = first; j <= ny; j += 2) sweep_mom_y_row(s, dt, j);
    }
    cell_update_momentum(s, dt);
    return orc_reduce_min(s->dtc, s->n, s->p.quirk_reduction);
}

double orc_step(orc_sim *s, double dt)
{
    orc_mass_pass(s, dt);
    return orc_momentum_pass(s, dt);
}

/* ------------------------------------------------------------------------- */
/* metrics: src/GPU/Metrics.cu                                               */
/* ------------------------------------------------------------------------- */

void orc_init_metrics(orc_sim *s)
{
    const double eps_h = s->p.eps_h;
    for (int64_t c = 0; c < s->n; c++) {
        double v = (s->h[c] < eps_h) ? -1e30 : s->h[c] - s->H[c];
        s->eta_max[c] = v;   /* src/GPU/ShallowWater.cu:305-306 */
        s->eta_ini[c] = v;   /* src/GPU/ShallowWater.cu:307-308 (same kernel) */
        s->t_arr[c] = -1.0;  /* src/GPU/Metrics.cu:44-55 */
    }
}

void orc_update_metrics(orc_sim *s, double t)
{
    const double eps_h = s->p.eps_h, dif = s->p.dif_at;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < s->n; c++) {
        double v = s->h[c] - s->H[c];
        if ((v > s->eta_max[c]) && (s->h[c] > eps_h)) s->eta_max[c] = v;
        if ((s->t_arr[c] < 0.0) && (fabs(v - s->eta_ini[c]) > dif)) s->t_arr[c] = t;
    }
}

/* ------------------------------------------------------------------------- */
/* Okada: src/GPU/Deformacion.cu, src/GPU/DC3D.cu, src/GPU/prtoxy.cu         */
/* ------------------------------------------------------------------------- */

/* Gauss-Krueger projection, both directions: src/GPU/prtoxy.cu:6-81.
   The geodetic constants are those of the reference (they fix the projection). */
int orc_prtoxy(double *lat, double *lon, double lat0, double lon0, double *x, double *y, int ind)
{
    const double D2R = 0.017453292371619689;
    if (ind == 0) {
        double rlat = *lat * D2R;
        double sl = sin(rlat), cl = cos(rlat);
        double v2 = cl * 0.0067397251 * cl + 1.0;
        double al = *lon - lon0;
        double ph1 = *lat + v2 * 0.5 * al * al * sl * cl * D2R;
        double rph1 = ph1 * D2R;
        double rph2 = (ph1 + lat0) * 0.5 * D2R;
        double s1 = sin(rph1), s2 = sin(rph2);
        double d = sqrt(1.0 - s2 * 0.0066946053 * s2);
        double r = 6335.4607362597517 / (d * (d * d));
        double an = 6378.16 / sqrt(1.0 - s1 * 0.0066946053 * s1);
        double c1 = 57.29578 / r;
        double c2 = 57.29578 / an;
        *y = (ph1 - lat0) / c1;
        *x = al * cl / c2 * (al * al * cos(rlat * 2.0) / 19696.838434850401 + 1.0);
    } else if (ind == 1) {
        double rlato = lat0 * D2R;
        double so = sin(rlato), co = cos(rlato);
        double den = sqrt(1.0 - so * 0.0066946053 * so);
        double r = 6335.4607362597517 / (den * (den * den));
        double an = 6378.16 / den;
        double v2 = co * 0.0067397251 * co + 1.0;
        double c1 = 57.29578 / r;
        double c2 = 57.29578 / an;
        double ph1 = lat0 + c1 * *y;
        double rph1 = ph1 * D2R;
        double tp = tan(rph1), cp = cos(rph1);
        double bl = c2 * *x;
        *lat = ph1 - bl * 0.5 * bl * v2 * tp * D2R;
        *lon = lon0 + bl / cp * (1.0 - bl * bl * (tp * 2.0 * tp + 1.0) / 19696.838434850401);
    }
    return 0;
}

typedef struct {
    double alp1, alp2, alp3, sd, cd, sdsd, cdcd, sdcd;
} okd_medium;

typedef struct {
    double xi2, et2, q2, r, r2, y, d, tt, alx, ale, x11, y11;
} okd_geom;

/* medium and dip constants: src/GPU/DC3D.cu:22-70 (dccon0) */
static void okd_dccon0(double alpha, double dip, okd_medium *m)
{
    const double pi2 = 6.283185307179586;
    m->alp1 = (1.0 - alpha) / 2.0;
    m->alp2 = alpha / 2.0;
    m->alp3 = (1.0 - alpha) / alpha;
    double p18 = pi2 / 360.0;
    m->sd = sin(dip * p18);
    m->cd = cos(dip * p18);
    if (fabs(m->cd) < 1e-6) {
        m->cd = 0.0;
        if (m->sd > 0.0) m->sd = 1.0;
        if (m->sd < 0.0) m->sd = -1.0;
    }
    m->sdsd = m->sd * m->sd;
    m->cdcd = m->cd * m->cd;
    m->sdcd = m->sd * m->cd;
}

/* station geometry for one corner (displacement terms only): src/GPU/DC3D.cu:72-157
   (dccon2).  The strain terms (ey..hz) do not enter the displacement. */
static void okd_dccon2(double *xi, double *et, double *q, double sd, double cd, int kxi, int ket,
                       okd_geom *g)
{
    if (fabs(*xi) < 1e-6) *xi = 0.0;
    if (fabs(*et) < 1e-6) *et = 0.0;
    if (fabs(*q) < 1e-6) *q = 0.0;
    g->xi2 = *xi * *xi;
    g->et2 = *et * *et;
    g->q2 = *q * *q;
    g->r2 = g->xi2 + g->et2 + g->q2;
    g->r = sqrt(g->r2);
    if (g->r == 0.0) return;
    g->y = *et * cd + *q * sd;
    g->d = *et * sd - *q * cd;
    g->tt = (*q == 0.0) ? 0.0 : atan(*xi * *et / (*q * g->r));
    if (kxi == 1) {
        g->alx = -log(g->r - *xi);
        g->x11 = 0.0;
    } else {
        double rxi = g->r + *xi;
        g->alx = log(rxi);
        g->x11 = 1.0 / (g->r * rxi);
    }
    if (ket == 1) {
        g->ale = -log(g->r - *et);
        g->y11 = 0.0;
    } else {
        double ret = g->r + *et;
        g->ale = log(ret);
        g->y11 = 1.0 / (g->r * ret);
    }
}

/* part A displacement (infinite medium): src/GPU/DC3D.cu:159-269 (ua), u[0..2] */
static void okd_ua(double xi, double et, double q, double d1, double d2, double d3,
                   const okd_medium *m, const okd_geom *g, double u[3])
{
    const double pi2 = 6.283185307179586;
    double qx = q * g->x11, qy = q * g->y11;
    u[0] = u[1] = u[2] = 0.0;
    if (d1 != 0.0) {
        double c = d1 / pi2;
        u[0] += c * (g->tt / 2.0 + m->alp2 * xi * qy);
        u[1] += c * (m->alp2 * q / g->r);
        u[2] += c * (m->alp1 * g->ale - m->alp2 * q * qy);
    }
    if (d2 != 0.0) {
        double c = d2 / pi2;
        u[0] += c * (m->alp2 * q / g->r);
        u[1] += c * (g->tt / 2.0 + m->alp2 * et * qx);
        u[2] += c * (m->alp1 * g->alx - m->alp2 * q * qx);
    }
    if (d3 != 0.0) {
        double c = d3 / pi2;
        u[0] += c * (-m->alp1 * g->ale - m->alp2 * q * qy);
        u[1] += c * (-m->alp1 * g->alx - m->alp2 * q * qx);
        u[2] += c * (g->tt / 2.0 - m->alp2 * (et * qx + xi * qy));
    }
}

/* part B displacement (surface deformation term): src/GPU/DC3D.cu:271-413 (ub), u[0..2] */
static void okd_ub(double xi, double et, double q, double d1, double d2, double d3,
                   const okd_medium *m, const okd_geom *g, double u[3])
{
    const double pi2 = 6.283185307179586;
    double rd = g->r + g->d;
    double ai3, ai4;
    if (m->cd != 0.0) {
        if (xi == 0.0) ai4 = 0.0;
        else {
            double x = sqrt(g->xi2 + g->q2);
            ai4 = 1.0 / m->cdcd * (xi / rd * m->sdcd
                  + 2.0 * atan((et * (x + q * m->cd) + x * (g->r + x) * m->sd) / (xi * (g->r + x) * m->cd)));
        }
        ai3 = (g->y * m->cd / rd - g->ale + m->sd * log(rd)) / m->cdcd;
    } else {
        double rd2 = rd * rd;
        ai3 = (et / rd + g->y * q / rd2 - g->ale) / 2.0;
        ai4 = xi * g->y / rd2 / 2.0;
    }
    double ai1 = -xi / rd * m->cd - ai4 * m->sd;
    double ai2 = log(rd) + ai3 * m->sd;
    double qx = q * g->x11, qy = q * g->y11;
    u[0] = u[1] = u[2] = 0.0;
    if (d1 != 0.0) {
        double c = d1 / pi2;
        u[0] += c * (-xi * qy - g->tt - m->alp3 * ai1 * m->sd);
        u[1] += c * (-q / g->r + m->alp3 * g->y / rd * m->sd);
        u[2] += c * (q * qy - m->alp3 * ai2 * m->sd);
    }
    if (d2 != 0.0) {
        double c = d2 / pi2;
        u[0] += c * (-q / g->r + m->alp3 * ai3 * m->sdcd);
        u[1] += c * (-et * qx - g->tt - m->alp3 * xi / rd * m->sdcd);
        u[2] += c * (q * qx + m->alp3 * ai4 * m->sdcd);
    }
    if (d3 != 0.0) {
        double c = d3 / pi2;
        u[0] += c * (q * qy - m->alp3 * ai3 * m->sdsd);
        u[1] += c * (q * qx + m->alp3 * xi / rd * m->sdsd);
        u[2] += c * (et * qx + xi * qy - g->tt - m->alp3 * ai4 * m->sdsd);
    }
}

/* corner setup shared by the real and image passes: src/GPU/DC3D.cu:672-714 */
static int okd_corners(double y, double dd, const okd_medium *m, double aw1, double aw2,
                       const double xi[2], double et[2], double *q, int kxi[2], int ket[2])
{
    double p = y * m->cd + dd * m->sd;
    *q = y * m->sd - dd * m->cd;
    et[0] = p - aw1;
    et[1] = p - aw2;
    if (fabs(*q) < 1e-6) *q = 0.0;
    if (fabs(et[0]) < 1e-6) et[0] = 0.0;
    if (fabs(et[1]) < 1e-6) et[1] = 0.0;
    /* on a fault edge: singular */
    if ((*q == 0.0) && (((xi[0] * xi[1] <= 0.0) && (et[0] * et[1] == 0.0)) ||
                        ((et[0] * et[1] <= 0.0) && (xi[0] * xi[1] == 0.0))))
        return 1;
    /* on the negative extension of a fault edge */
    kxi[0] = kxi[1] = ket[0] = ket[1] = 0;
    double r12 = sqrt(xi[0] * xi[0] + et[1] * et[1] + *q * *q);
    double r21 = sqrt(xi[1] * xi[1] + et[0] * et[0] + *q * *q);
    double r22 = sqrt(xi[1] * xi[1] + et[1] * et[1] + *q * *q);
    if (xi[0] < 0.0 && r21 + xi[1] < 1e-6) kxi[0] = 1;
    if (xi[0] < 0.0 && r22 + xi[1] < 1e-6) kxi[1] = 1;
    if (et[0] < 0.0 && r12 + et[1] < 1e-6) ket[0] = 1;
    if (et[0] < 0.0 && r22 + et[1] < 1e-6) ket[1] = 1;
    return 0;
}

/* DC3D restricted to the free surface (z = 0) and to displacements:
   src/GPU/DC3D.cu:583-867.  At z = 0 the part-C terms enter only multiplied by z
   (displacements) or into the strain components, so they are omitted. */
void orc_dc3d_surface(double alpha, double x, double y, double depth, double dip,
                      double al1, double al2, double aw1, double aw2,
                      double disl1, double disl2, double disl3,
                      double *ux, double *uy, double *uz, int *iret)
{
    okd_medium m;
    okd_geom g;
    double u[3] = {0.0, 0.0, 0.0};
    double xi[2], et[2], q;
    int kxi[2], ket[2];
    memset(&g, 0, sizeof(g));
    *iret = 0;
    okd_dccon0(alpha, dip, &m);
    xi[0] = x - al1;
    xi[1] = x - al2;
    if (fabs(xi[0]) < 1e-6) xi[0] = 0.0;
    if (fabs(xi[1]) < 1e-6) xi[1] = 0.0;
    /* real source, d = depth + z */
    if (okd_corners(y, depth, &m, aw1, aw2, xi, et, &q, kxi, ket)) goto singular;
    for (int k = 0; k < 2; k++) {
        for (int j = 0; j < 2; j++) {
            double a[3], du[3];
            double xij = xi[j], etk = et[k], qq = q;
            okd_dccon2(&xij, &etk, &qq, m.sd, m.cd, kxi[k], ket[j], &g);
            okd_ua(xij, etk, qq, disl1, disl2, disl3, &m, &g, a);
            du[0] = -a[0];
            du[1] = -a[1] * m.cd + a[2] * m.sd;
            du[2] = -a[1] * m.sd - a[2] * m.cd;
            for (int c = 0; c < 3; c++) {
                if (j + k != 1) u[c] += du[c];
                else u[c] -= du[c];
            }
        }
    }
    /* image source, d = depth - z */
    if (okd_corners(y, depth, &m, aw1, aw2, xi, et, &q, kxi, ket)) goto singular;
    for (int k = 0; k < 2; k++) {
        for (int j = 0; j < 2; j++) {
            double a[3], b[3], du[3];
            double xij = xi[j], etk = et[k], qq = q;
            okd_dccon2(&xij, &etk, &qq, m.sd, m.cd, kxi[k], ket[j], &g);
            okd_ua(xij, etk, qq, disl1, disl2, disl3, &m, &g, a);
            okd_ub(xij, etk, qq, disl1, disl2, disl3, &m, &g, b);
            du[0] = a[0] + b[0];
            du[1] = (a[1] + b[1]) * m.cd - (a[2] + b[2]) * m.sd;
            du[2] = (a[1] + b[1]) * m.sd + (a[2] + b[2]) * m.cd;
            for (int c = 0; c < 3; c++) {
                if (j + k != 1) u[c] += du[c];
                else u[c] -= du[c];
            }
        }
    }
    *ux = u[0]; *uy = u[1]; *uz = u[2];
    return;
singular:
    *iret = 1;
    *ux = *uy = *uz = 0.0;
}

/* src/GPU/Deformacion.cu:118-143 */
void orc_okada(orc_sim *s, double lon0, double dlon, double lat0, double dlat,
               double lon_c, double lat_c, double depth_c, double length, double width,
               double strike, double dip, double rake, double slip,
               int crop_flag, double crop_value, double H, double *def_out)
{
    const int nx = s->p.nx, ny = s->p.ny;
    const int64_t n = s->n;
    double *def = (double *)malloc(sizeof(double) * (size_t)n);
    const double RAD = M_PI / 180.0;
#pragma omp parallel for schedule(dynamic, 4)
    for (int j = 0; j < ny; j++) {
        for (int i = 0; i < nx; i++) {
            double lonp = lon0 + i * dlon;
            double latp = lat0 + j * dlat;
            double sr = sin(RAD * rake), cr = cos(RAD * rake);
            double ss = sin(RAD * strike), cs = cos(RAD * strike);
            double d2 = slip * sr, d1 = slip * cr;
            double xp, yp, ux, uy, uz;
            int iret;
            orc_prtoxy(&latp, &lonp, lat_c, lon_c, &xp, &yp, 0);
            double xo = xp * ss + yp * cs;
            double yo = -xp * cs + yp * ss;
            orc_dc3d_surface(2.0 / 3.0, xo, yo, depth_c, dip, -0.5 * length, 0.5 * length,
                             -0.5 * width, 0.5 * width, d1, d2, 0.0, &ux, &uy, &uz, &iret);
            def[(int64_t)j * nx + i] = uz / H;
        }
    }
    if (def_out) memcpy(def_out, def, sizeof(double) * (size_t)n);
    if (crop_flag == 1) {
        /* relative crop: float reduction of -|U_Z| (src/GPU/Deformacion.cu:132-136, kernel :13-20) */
        float *neg = (float *)malloc(sizeof(float) * (size_t)n);
        for (int64_t c = 0; c < n; c++) neg[c] = (float)(-fabs(def[c]));
        float dmax = -orc_reduce_min_f(neg, n, s->p.quirk_reduction);
        free(neg);
        double thr = crop_value * (double)dmax;
        for (int64_t c = 0; c < n; c++) if (fabs(def[c]) < thr) def[c] = 0.0;
    } else if (crop_flag == 2) {
        for (int64_t c = 0; c < n; c++) if (fabs(def[c]) < crop_value) def[c] = 0.0;
    }
    /* src/GPU/Deformacion.cu:22-53 */
    for (int64_t c = 0; c < n; c++) {
        s->H[c] -= def[c];
        s->eta_ini[c] += def[c];
    }
    free(def);
}

/* ------------------------------------------------------------------------- */
/* outputs                                                                   */
/* ------------------------------------------------------------------------- */

void orc_output_field(const orc_sim *s, int which, double H, double Q, double T, float *out)
{
    const double eps_h = s->p.eps_h, hmin = s->p.hmin;
    for (int64_t c = 0; c < s->n; c++) {
        double h = s->h[c];
        double v;
        switch (which) {
        case 0: /* src/GPU/netcdf_kernel.cu:6-21 */
            v = (h < eps_h) ? -9999.0 : (h - s->H[c] - hmin) * H;
            break;
        case 1: case 2: { /* :23-71 */
            double f = (h < eps_h) ? SQRT2 * h / sqrt(h * h * h * h + eps_h * eps_h * eps_h * eps_h) * (Q / H)
                                   : Q / (h * H);
            v = (which == 1 ? s->qx[c] : s->qy[c]) * f;
            break;
        }
        case 3: /* :73-88 */
            v = (s->H[c] + hmin) * H;
            break;
        case 4: /* :94-109 */
            v = (s->eta_max[c] < -1e20) ? -9999.0 : (s->eta_max[c] - hmin) * H;
            break;
        default: /* :111-126 */
            v = (s->t_arr[c] < 0.0) ? -1.0 : s->t_arr[c] * T;
            break;
        }
        out[c] = (float)v;
    }
}

void orc_gauges(const orc_sim *s, const int *pos, int n, double H, double Q,
                float *eta, float *u, float *v)
{
    const double eps_h = s->p.eps_h, hmin = s->p.hmin;
    for (int g = 0; g < n; g++) {
        int c = pos[g];
        if (c < 0) { eta[g] = u[g] = v[g] = -9999.0f; continue; }
        double h = s->h[c];
        double f = (h < eps_h) ? SQRT2 * h / sqrt(h * h * h * h + eps_h * eps_h * eps_h * eps_h) : 1.0 / h;
        double wx = f * s->qx[c], wy = f * s->qy[c];
        eta[g] = (float)((h - s->H[c] - hmin) * H);
        u[g] = (float)(wx * Q / H);
        v[g] = (float)(wy * Q / H);
    }
}

double orc_mass(const orc_sim *s)
{
    /* Neumaier-compensated sum of h*area*cos(phi).  The reference's cell "area" has no
       cos(phi) factor (src/Problema.cxx:444): x fluxes carry 1/cos(phi) and y fluxes are
       divided by the receiving cell's cos(phi) (src/GPU/Arista_kernel.cu:683-685), so
       the invariant of the discrete scheme is sum h*area*cos(phi). */
    double sum = 0.0, comp = 0.0;
    for (int j = 0; j < s->p.ny; j++) {
        for (int i = 0; i < s->p.nx; i++) {
            double x = s->h[(int64_t)j * s->p.nx + i] * s->area[j] * s->cosphi[j];
            double t = sum + x;
            if (fabs(sum) >= fabs(x)) comp += (sum - t) + x;
            else comp += (x - t) + sum;
            sum = t;
        }
    }
    return sum + comp;
}

/* ------------------------------------------------------------------------- */
/* driver: src/GPU/ShallowWater.cu:374-453, truncation src/GPU/Timestep.cu:35-68 */
/* ------------------------------------------------------------------------- */

static void trunc_next(const orc_run_cfg *c, int k, double t, double *dt, int *skip)
{
    if (k < c->n_faults) {
        if ((t + *dt) * c->T > c->faults[k].time) {
            *dt = c->faults[k].time / c->T - t;
            *skip = 1;
        }
    }
}

static void trunc_next_skipping(const orc_run_cfg *c, int k, double t, double *dt, int *skip)
{
    *skip = 0;
    if (k < c->n_faults) {
        int i = k + 1;
        int found = 0;
        while (i < c->n_faults && !found) {
            if (fabs(c->faults[i].time - c->faults[k].time) > EPS) found = 1;
            else i++;
        }
        if (found && (t + *dt) * c->T > c->faults[i].time) {
            *dt = c->faults[i].time / c->T - t;
            *skip = 1;
        }
    }
}

int orc_run(orc_sim *s, const orc_run_cfg *c, orc_run_out *o)
{
    double t = 0.0, next_save = 0.0, next_series = 0.0;
    int k = 0, skip = 0, nsnap = 0, nsamp = 0;
    int64_t steps = 0;
    memset(o, 0, sizeof(*o));
    orc_init_metrics(s);
    for (int g = 0; g < c->n_gauges; g++) {
        if (c->gauge_pos[g] == -1) { c->g_eta_min[g] = -9999.0f; c->g_eta_max[g] = -9999.0f; }
        else { c->g_eta_min[g] = 1e30f; c->g_eta_max[g] = -1e30f; }
    }
    double dt = orc_initial_dt(s);
    o->dt_initial = dt;
    if (c->n_faults > 0) {
        if (c->faults[0].time < EPS) trunc_next_skipping(c, k, t, &dt, &skip);
        else trunc_next(c, k, t, &dt, &skip);
    }
    while (t < c->t_end) {
        if (c->max_steps > 0 && steps >= c->max_steps) break;
        while (k < c->n_faults && t * c->T >= c->faults[k].time) {
            const orc_fault *f = &c->faults[k];
            orc_okada(s, c->lon0, c->dlon, c->lat0, c->dlat, f->lon_c, f->lat_c, f->depth_c,
                      f->length, f->width, f->strike, f->dip, f->rake, f->slip,
                      c->crop_flag, c->crop_value, c->H, NULL);
            k++;
        }
        if (c->save_interval >= 0.0 && t >= next_save) {
            next_save += c->save_interval;
            if (c->on_snapshot) c->on_snapshot(c->user, nsnap, t);
            nsnap++;
        }
        if (c->series_interval >= 0.0 && t >= next_series) {
            next_series += c->series_interval;
            if (nsamp < c->max_samples && c->n_gauges > 0) {
                float *e = c->g_eta + (int64_t)nsamp * c->n_gauges;
                float *u = c->g_u + (int64_t)nsamp * c->n_gauges;
                float *v = c->g_v + (int64_t)nsamp * c->n_gauges;
                orc_gauges(s, c->gauge_pos, c->n_gauges, c->H, c->Q, e, u, v);
                for (int g = 0; g < c->n_gauges; g++) {
                    if (c->gauge_pos[g] == -1) continue;
                    if (e[g] < c->g_eta_min[g]) c->g_eta_min[g] = e[g];
                    if (e[g] > c->g_eta_max[g]) c->g_eta_max[g] = e[g];
                }
                c->g_time[nsamp] = (float)(t * c->T);
            }
            nsamp++;
        }
        orc_update_metrics(s, t);
        double dtn = orc_step(s, dt);
        t += dt;
        dt = dtn;
        if (!skip) trunc_next(c, k, t, &dt, &skip);
        else trunc_next_skipping(c, k, t, &dt, &skip);
        steps++;
    }
    o->steps = steps;
    o->n_samples = nsamp;
    o->n_snapshots = nsnap;
    o->t_final = t;
    o->dt_last = dt;
    return 0;
}

double orc_time_of(const orc_sim *s) { (void)s; return 0.0; }
