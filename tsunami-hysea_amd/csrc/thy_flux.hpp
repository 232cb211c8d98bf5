// thy_flux.hpp -- per-face 2s+WAF flux math of the Tsunami-HySEA scheme for gfx950.
//
// Restates the reference's edge functions (src/GPU/Arista_kernel.cu) for one face,
// templated on the arithmetic type (float for the production path, double for the
// FP64-everywhere precision study).  State is carried as
//     e = eta relative to sea level (h - D),  D = still-water depth (unshifted),
// instead of the reference's (h, H = D - Hmin).  Every place where the reference
// forms eta_k + min(H0,H1) becomes e_k + min(D0,D1) (the Hmin shift cancels), and the
// hydrostatic difference max(a,0) - max(b,0) with both sides wet is evaluated as
// e1 - e0, which is the same number without the cancellation of two O(1) depths.
#pragma once
#include <hip/hip_runtime.h>

namespace thy {

// Guards keep the reference's DBL_EPSILON (src/Constantes.hxx:11) in both precisions.
template <typename T> struct Num;

template <> struct Num<float> {
    static constexpr float eps = 2.220446049250313e-16f;
    static constexpr float big = 1e30f;
    static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    static __device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
    // a * b + c rounded once (v_fma_f32).  The library is compiled without automatic
    // contraction, so every fused product is written out here, at the same place in each
    // copy of a formula, and all kernels round it the same way.
    static __device__ __forceinline__ float mad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
    // a / sqrt(x) with one transcendental (reciprocal square root)
    static __device__ __forceinline__ float div_sqrt(float a, float x) { return a * __builtin_amdgcn_rsqf(x); }
    static __device__ __forceinline__ float exp(float x) {
        return __builtin_amdgcn_exp2f(x * 1.4426950408889634f);
    }
    // h^(4/3) for h > 0
    static __device__ __forceinline__ float pow43(float x) {
        return __builtin_amdgcn_exp2f(1.3333333333333333f * __builtin_amdgcn_logf(x));
    }
};

template <> struct Num<double> {
    static constexpr double eps = 2.220446049250313e-16;
    static constexpr double big = 1e30;
    static __device__ __forceinline__ double rcp(double x) { return 1.0 / x; }
    static __device__ __forceinline__ double div(double a, double b) { return a / b; }
    static __device__ __forceinline__ double sqrt(double x) { return ::sqrt(x); }
    // FP64 keeps the product and the sum rounded separately, as the oracle rounds them
    static __device__ __forceinline__ double mad(double a, double b, double c) { return a * b + c; }
    static __device__ __forceinline__ double div_sqrt(double a, double x) { return a / ::sqrt(x); }
    static __device__ __forceinline__ double exp(double x) { return ::exp(x); }
    static __device__ __forceinline__ double pow43(double x) { return ::pow(x, 4.0 / 3.0); }
};

template <typename T> struct Prm {
    T eps_h, hpos, cvis;
    T bl, bh;                    // border factors of the low / high side of this face direction
};

// One stencil cell: total depth h, eta e, depth D, normal and tangential discharge, and
// the desingularisation factor desing(h, eps_h) (computed once per cell by the caller,
// which can share it between the faces of the cell).
template <typename T> struct SC {
    T h, e, D, qn, qt, f;
};

template <typename T> __device__ __forceinline__ T tmax(T a, T b) { return a > b ? a : b; }
template <typename T> __device__ __forceinline__ T tmin(T a, T b) { return a < b ? a : b; }
template <typename T> __device__ __forceinline__ T tabs(T a) { return a < T(0) ? -a : a; }
// |a| as a source modifier of the instruction that uses it (no compare and select); where
// it is added to a positive number the sum equals tabs(a) + that number
__device__ __forceinline__ float habs(float a) { return __builtin_fabsf(a); }
__device__ __forceinline__ double habs(double a) { return __builtin_fabs(a); }
template <typename T> __device__ __forceinline__ T sgn(T x) {
    return (x > T(0) ? T(1) : T(0)) - (x < T(0) ? T(1) : T(0));
}

// sqrt(2) h / sqrt(h^4 + max(h,eps)^4): src/GPU/Arista_kernel.cu:163-168
template <typename T> __device__ __forceinline__ T desing(T h, T eps_h) {
    T m = tmax(h, eps_h);
    T h2 = h * h, m2 = m * m;
    return Num<T>::div_sqrt(T(1.4142135623730951) * h, Num<T>::mad(h2, h2, m2 * m2));
}

// smooth monotone limiter: src/GPU/Arista_kernel.cu:45-68
template <typename T> __device__ __forceinline__ T limiter(T p0m, T p0, T p1, T p1p, T S) {
    T c0 = p1 - p0;
    T c1 = (S < T(0)) ? (p1p - p1) : ((S > T(0)) ? (p0 - p0m) : T(0.5) * (p1p - p0m));
    T sg = (c0 * c1 >= T(0)) ? T(1) : T(0);
    c0 = tabs(c0);
    c1 = tabs(c1);
    if (c0 < Num<T>::eps) return T(1);
    T aux = tabs(Num<T>::div(c1 - c0, c1 + c0));
    c1 = T(0.5) * (c1 + c0) * Num<T>::mad(-(aux * aux), aux, T(1));
    // with r = p / q the smaller over the larger, sg r (1 + r) / (1 + r^2) =
    // sg p (q + p) / (q^2 + p^2): one division instead of two
    const bool big0 = c0 > c1;
    const T pn = big0 ? c1 : c0, qd = (big0 ? c0 : c1) + Num<T>::eps;
    return Num<T>::div(sg * pn * (qd + pn), Num<T>::mad(qd, qd, pn * pn));
}

// max(e1 + Dm, 0) - max(e0 + Dm, 0) with Dm = min(D0, D1)  (src/GPU/Arista_kernel.cu:201-202)
template <typename T> __device__ __forceinline__ T deta_of(T e0, T e1, T D0, T D1) {
    T Dm = tmin(D0, D1);
    T a = e1 + Dm, b = e0 + Dm;
    return (a > T(0) && b > T(0)) ? (e1 - e0) : (tmax(a, T(0)) - tmax(b, T(0)));
}

// 2s/WAF blend: src/GPU/Arista_kernel.cu:205-208
template <typename T> __device__ __forceinline__ T blend(T hij, T hpos, bool alfa_cero) {
    if (alfa_cero) return T(0);
    if (hij < hpos)
        return Num<T>::rcp(T(1) + Num<T>::exp(T(-15) * (T(1) - Num<T>::div(T(0.5) * hpos, hij + Num<T>::eps))));
    return T(1);
}

// positivity limiter: src/GPU/Arista_kernel.cu:79-114 (g = 1 for x faces, cos for y).
// The reference limits the flux when dt exceeds dt0 = 0.25 h g A / (|F| + eps); the
// test is evaluated as dt (|F| + eps) > h k with k = g A / 4 of the upwind row, taken
// from the row tables (the same condition, no division unless the limiter acts).  Every
// face evaluation (deep-water forms and general faces) goes through this one function,
// so a face evaluated by two wavefronts has one value.
template <typename T> __device__ __forceinline__ T positivity(T F, T h0, T h1, T k0, T k1, T dt) {
    const T lim = (F > T(0)) ? h0 * k0 : h1 * k1;
    const T den = habs(F) + Num<T>::eps;
    if (dt * den > lim) {
        const T dt0 = Num<T>::div(lim, den);
        if (dt > dt0) F *= Num<T>::div(dt0, dt + Num<T>::eps);
    }
    return F;
}

// wet/dry treatment of the normal discharges: src/GPU/Arista_kernel.cu:70-77
template <typename T> __device__ __forceinline__ void wet_dry(const SC<T> &c0, const SC<T> &c1, T eps_h, T &q0, T &q1) {
    if (c0.h < eps_h && -c1.e > c0.D) q1 = T(0);
    if (c1.h < eps_h && -c0.e > c1.D) q0 = T(0);
}

// Apply the ghost-cell border of a face and orient it: border 1 = low-side border
// (ghost on the low side, swapped so the interior cell is "0", normal -1),
// border 2 = high-side border.  src/GPU/Arista_kernel.cu:506-524,550-556
template <typename T> __device__ __forceinline__ T orient(int border, const Prm<T> &p, SC<T> &m, SC<T> &a, SC<T> &b, SC<T> &q) {
    if (border == 1) {
        a.qn *= p.bl;
        SC<T> t = a; a = b; b = t;
        t = m; m = q; q = t;
        return T(-1);
    }
    if (border == 2) b.qn *= p.bh;
    return T(1);
}

// ---------------------------------------------------------------------------
// Mass (step 1) face flux.  Returns F: low cell -= F, high cell += F (x); the y
// caller divides by the cell's cos.  Border faces return the value that the
// single adjacent interior cell receives with the sign convention documented in
// the step kernel.
//   x: src/GPU/Arista_kernel.cu:118-219,466-563   y: :674-778,1028-1125
// g0/g1: 1/cos(phi_row) for x faces (both sides), cos0/cos1 for y faces; k0/k1: the
// positivity factors g A / 4 of the two rows (A / 4 for x faces).
// ---------------------------------------------------------------------------
template <typename T, bool YFACE>
__device__ __forceinline__ T mass_flux(SC<T> m, SC<T> a, SC<T> b, SC<T> q, int border, bool alfa_cero,
                                       T len, T g0, T g1, T k0, T k1, T valf, T dt, const Prm<T> &p) {
    // deep branch (2s): all four depths above hpos and outside the first-order band
    if (!alfa_cero && m.h > p.hpos && a.h > p.hpos && b.h > p.hpos && q.h > p.hpos) {
        T F = a.qn * (len * g0);
        return positivity(F, a.h, b.h, k0, k1, dt);
    }
    T normal = orient(border, p, m, a, b, q);
    if (border == 1) { T t = g0; g0 = g1; g1 = t; t = k0; k0 = k1; k1 = t; }
    if (!(a.h > Num<T>::eps || b.h > Num<T>::eps)) return T(0);
    T q0 = a.qn * normal, q1 = b.qn * normal;
    T hij = T(0.5) * (a.h + b.h);
    wet_dry(a, b, p.eps_h, q0, q1);
    T u0 = a.f * q0;
    T u1 = b.f * q1;
    q0 = u0 * a.h;
    q1 = u1 * b.h;
    T s0 = Num<T>::sqrt(a.h), s1 = Num<T>::sqrt(b.h);
    T uij = Num<T>::div(Num<T>::mad(s0, u0, s1 * u1), s0 + s1 + Num<T>::eps);
    T val = valf * dt;   // 2*len*dt/(A0+A1) with len/cos for x
    T c0 = Num<T>::sqrt(hij);
    T l1 = tmin(uij - c0, u0 - s0);
    T l2 = tmax(uij + c0, u1 + s1);
    T lmax = tmax(tabs(l2), tabs(l1));
    T xiL = T(0), xiR = T(0);
    if (!alfa_cero) {
        xiL = limiter(m.e, a.e, b.e, q.e, l1);
        xiR = limiter(m.e, a.e, b.e, q.e, l2);
    }
    T a0;
    if (tabs(l1 - l2) >= Num<T>::eps) {
        T alr = Num<T>::mad(val * l1, xiL, sgn(l1) * (T(1) - xiL));
        T blr = Num<T>::mad(val * l2, xiR, sgn(l2) * (T(1) - xiR));
        a0 = Num<T>::div((alr - blr) * (l1 * l2), l2 - l1);
    } else {
        a0 = lmax;
    }
    T F2s, Fw;
    if (YFACE) {
        F2s = len * q0 * g0 * normal;
        T deta = Num<T>::mad(-(uij * uij), g1 - g0, T(0.5) * (g0 + g1) * deta_of(a.e, b.e, a.D, b.D));
        Fw = T(0.5) * Num<T>::mad(-a0, deta, Num<T>::mad(q0, g0, q1 * g1)) * len;
    } else {
        T il = len * g0;
        F2s = il * q0 * normal;
        T deta = deta_of(a.e, b.e, a.D, b.D);
        Fw = T(0.5) * Num<T>::mad(-a0, deta, q1 + q0) * il;
    }
    T alfa = blend(hij, p.hpos, alfa_cero);
    T F = Num<T>::mad(alfa, F2s, (T(1) - alfa) * Fw);
    return positivity(F, a.h, b.h, k0, k1, dt);
}

// Momentum face result: low cell acc2 -= (mn, mt), high cell acc2 -= (pn, pt),
// both cells acc1.y += w.  (n = normal component, t = tangential.)
template <typename T> struct MF {
    T mn, mt, pn, pt, w;
};

// ---------------------------------------------------------------------------
// Momentum (step 2) face flux.
//   x: src/GPU/Arista_kernel.cu:221-464,565-670   y: :780-1026,1127-1233
// For x faces: g0 = g1 = 1/cos(phi_row) (len*g = len/cos), qn = qx, qt = qy.
// For y faces: g0 = cos0, g1 = cos1 with ic0 = 1/cos0, ic1 = 1/cos1 (the row table's
// values, shared with the deep-water forms), qn = qy, qt = qx.
// ---------------------------------------------------------------------------
template <typename T, bool YFACE>
__device__ __forceinline__ MF<T> mom_flux(SC<T> m, SC<T> a, SC<T> b, SC<T> q, int border, bool alfa_cero,
                                          T len, T g0, T g1, T ic0, T ic1, T valf, T kdes, T dt, const Prm<T> &p) {
    MF<T> o;
    if (!alfa_cero && m.h > p.hpos && a.h > p.hpos && b.h > p.hpos && q.h > p.hpos) {
        // deep-water 2s branch
        T hij = T(0.5) * (a.h + b.h);
        T ia = Num<T>::rcp(a.h), ib = Num<T>::rcp(b.h);
        T u0 = a.qn * ia, u1 = b.qn * ib;
        T s0 = Num<T>::sqrt(a.h), s1 = Num<T>::sqrt(b.h);
        T uij = Num<T>::div(Num<T>::mad(s0, u0, s1 * u1), s0 + s1 + Num<T>::eps);
        T lmax = habs(uij) + Num<T>::sqrt(hij);
        T deta = deta_of(a.e, b.e, a.D, b.D);
        bool up = uij < T(0);
        T uu = up ? u1 : u0;
        T qnn = up ? b.qn : a.qn;
        T qtt = up ? b.qt : a.qt;
        if (YFACE) {
            deta = Num<T>::mad(-(uij * uij), g1 - g0, T(0.5) * (g0 + g1) * deta);
            T fn = uu * qnn * len;
            T mn, pn;
            if (up) { pn = -fn; mn = fn * (g1 * ic0); }
            else { mn = fn; pn = -fn * g0 * ic1; }
            mn = Num<T>::mad(hij * deta * len, ic0, mn);
            T DES = kdes * dt * (lmax * lmax) * Num<T>::mad(b.qn, g1, -(a.qn * g0));
            o.mn = Num<T>::mad(-DES, ic0, mn);
            o.pn = Num<T>::mad(DES, ic1, pn);
            T ft = uu * qtt * len;
            o.mt = ft;
            o.pt = -ft;
            o.w = len * lmax;
        } else {
            T il = len * g0;
            T fn = uu * qnn * il;
            T DES = kdes * dt * (lmax * lmax) * (b.qn - a.qn);
            o.mn = Num<T>::mad(hij * deta, il, fn) - DES;
            o.pn = -fn + DES;
            T ft = uu * qtt * il;
            o.mt = ft;
            o.pt = -ft;
            o.w = il * lmax;
        }
        return o;
    }
    T normal = orient(border, p, m, a, b, q);
    if (border == 1) {
        T t = g0; g0 = g1; g1 = t;
        t = ic0; ic0 = ic1; ic1 = t;
    }
    if (!(a.h > Num<T>::eps || b.h > Num<T>::eps)) {
        o.mn = o.mt = o.pn = o.pt = o.w = T(0);
        return o;
    }
    if (!YFACE) ic0 = ic1 = T(1);
    T L = YFACE ? len : len * g0;   // face length factor (len/cos for x)
    T h0 = a.h, h1 = b.h;
    T q0 = a.qn * normal, q1 = b.qn * normal;
    T hij = T(0.5) * (h0 + h1);
    wet_dry(a, b, p.eps_h, q0, q1);
    T f0 = a.f, f1 = b.f;
    T u0 = f0 * q0, u1 = f1 * q1;
    if (!YFACE) { q0 = u0 * h0; q1 = u1 * h1; }
    T s0 = Num<T>::sqrt(h0), s1 = Num<T>::sqrt(h1);
    T uij = Num<T>::div(Num<T>::mad(s0, u0, s1 * u1), s0 + s1 + Num<T>::eps);
    T val = valf * dt;
    T c0 = Num<T>::sqrt(hij);
    T l1 = tmin(uij - c0, u0 - s0);
    T l2 = tmax(uij + c0, u1 + s1);
    T lmax = tmax(tabs(l2), tabs(l1));
    T deta = deta_of(a.e, b.e, a.D, b.D);
    T fl0x, fl0y, fl1x, fl1y;
    if (YFACE) {
        deta = Num<T>::mad(-(uij * uij), g1 - g0, T(0.5) * (g0 + g1) * deta);
        fl1x = q1 * g1; fl1y = u1 * fl1x;
        fl0x = q0 * g0; fl0y = u0 * fl0x;
    } else {
        fl1x = q1; fl1y = u1 * q1;
        fl0x = q0; fl0y = u0 * q0;
    }
    T dq = fl1x - fl0x;
    // 2s part (s: normal momentum, z: tangential)
    bool up = uij < T(0);
    T sm_n, sp_n, sm_t, sp_t;
    {
        T uu = up ? u1 : u0;
        T qq = up ? q1 : q0;
        T qt = up ? b.qt : a.qt;
        T fn = uu * qq * L * normal;
        if (YFACE) {
            if (up) { sp_n = -fn; sm_n = fn * (g1 * ic0); }
            else { sm_n = fn; sp_n = -fn * g0 * ic1; }
            sm_n += hij * deta * L * ic0;
        } else {
            sp_n = -fn;
            sm_n = fn + hij * deta * L;
        }
        T ft = uu * qt * L;
        sm_t = ft;
        sp_t = -ft;
        T cd = T(0.5) * (T(1) - Num<T>::div(p.cvis, T(1) + Num<T>::exp(T(-20) * (T(1) - Num<T>::div(p.hpos, hij)))));
        T des = cd * val * lmax * lmax * dq * L;
        if (YFACE) { sm_n -= des * ic0; sp_n += des * ic1; }
        else { des *= normal; sm_n -= des; sp_n += des; }
    }
    // WAF part
    T xiL = T(0), xiR = T(0);
    if (!alfa_cero) {
        xiL = limiter(m.e, a.e, b.e, q.e, l1);
        xiR = limiter(m.e, a.e, b.e, q.e, l2);
    }
    T a0, a1;
    if (tabs(l1 - l2) >= Num<T>::eps) {
        T alr = Num<T>::mad(val * l1, xiL, sgn(l1) * (T(1) - xiL));
        T blr = Num<T>::mad(val * l2, xiR, sgn(l2) * (T(1) - xiR));
        T aux = Num<T>::rcp(l2 - l1);
        a0 = (alr - blr) * (l1 * l2) * aux;
        a1 = Num<T>::mad(blr, l2, -(alr * l1)) * aux;
    } else {
        a0 = lmax;
        a1 = T(0);
    }
    T wy = T(0.5) * Num<T>::mad(hij, deta, fl1y - fl0y);
    T DESx = T(0.5) * a0 * deta;
    T DESy = Num<T>::mad(a1, wy, T(0.5) * a0 * dq);
    T mxw = Num<T>::mad(T(0.5), dq, fl0x - DESx) * L;
    T myw = (wy + (fl0y - DESy)) * L;
    T pyw = (wy + (DESy - fl1y)) * L;
    // limited upwinding of the tangential velocity
    T tsign = YFACE ? T(-1) : T(1);
    T q0mt = tsign * m.qt * normal, q0t = tsign * a.qt * normal;
    T q1t = tsign * b.qt * normal, q1pt = tsign * q.qt * normal;
    T p0m = m.f * q0mt;
    T p0 = f0 * q0t;
    T p1 = f1 * q1t;
    T p1p = q.f * q1pt;
    T sgm = sgn(mxw);
    T us = tabs(uij) * sgm;
    T xi = alfa_cero ? T(0) : limiter(p0m, p0, p1, p1p, us);
    T uijt;
    const T wt = T(-0.5) * Num<T>::mad(xi * us, val, (T(1) - xi) * sgm);
    if (YFACE) uijt = Num<T>::mad(wt, Num<T>::mad(p1, ic1, -(p0 * ic0)), T(0.5) * Num<T>::mad(p0, ic0, p1 * ic1));
    else uijt = Num<T>::mad(wt, p1 - p0, T(0.5) * (p0 + p1));
    T aux = mxw * uijt;
    T f3m_n, f3m_t, f3p_n, f3p_t;
    if (YFACE) {
        f3p_t = aux * normal;  f3p_n = pyw * normal * ic1;
        f3m_t = -aux * normal; f3m_n = myw * normal * ic0;
    } else {
        f3p_n = pyw * normal;  f3p_t = -aux * normal;
        f3m_n = myw * normal;  f3m_t = aux * normal;
    }
    T alfa = blend(hij, p.hpos, alfa_cero);
    T oma = T(1) - alfa;
    o.mn = Num<T>::mad(alfa, sm_n, oma * f3m_n);
    o.mt = Num<T>::mad(alfa, sm_t, oma * f3m_t);
    o.pn = Num<T>::mad(alfa, sp_n, oma * f3p_n);
    o.pt = Num<T>::mad(alfa, sp_t, oma * f3p_t);
    if (lmax < p.eps_h) lmax += p.eps_h;
    o.w = L * lmax;
    return o;
}

// Initial-dt wave speed across a face: src/GPU/Arista_kernel.cu:1237-1280
template <typename T>
__device__ __forceinline__ T init_wave(T h0, T q0, T h1, T q1, T eps_h) {
    T hij = T(0.5) * (h0 + h1);
    T u0 = desing(h0, eps_h) * q0;
    T u1 = desing(h1, eps_h) * q1;
    T s0 = Num<T>::sqrt(h0), s1 = Num<T>::sqrt(h1);
    T uij = Num<T>::div(Num<T>::mad(s0, u0, s1 * u1), s0 + s1 + Num<T>::eps);
    T lmax = tabs(uij) + Num<T>::sqrt(hij);
    if (lmax < eps_h) lmax += eps_h;
    return lmax;
}

}  // namespace thy
