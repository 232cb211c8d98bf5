// thy_sweep.hpp -- row-sweep step kernel of the Tsunami-HySEA scheme for gfx950.
//
// Same step as k_step (src/GPU/Timestep.cu:70-120 plus the metric update and the
// gauge gather), organised for the CDNA execution model instead of 2D LDS tiles:
//   * one wavefront owns 64 adjacent columns (56 output columns plus a 4-column halo
//     on each side, the reach of the 2s+WAF stencil through the two passes) and walks
//     a strip of rows from top to bottom;
//   * each row flows through a register pipeline: load row r, mass x-faces of row r,
//     mass y-face between rows r-2 and r-1, h* of row r-2, momentum x-faces of row
//     r-2, momentum y-face between rows r-4 and r-3, cell update of row r-4;
//   * x-neighbours come from adjacent lanes (cross-lane moves), y-neighbours from the
//     previous rows kept in registers, so every face is evaluated exactly once and no
//     LDS, barrier or halo re-load is needed inside a strip;
//   * the row geometry is uniform across the wavefront (scalar loads) and each row is
//     fetched one iteration before it is used;
//   * each stage votes across the wavefront: when every lane's stencil is deeper than
//     hpos and clear of the first-order band, the stage uses the deep-water 2s forms
//     below; otherwise it uses the general per-face functions of k_step
//     (thy_flux.hpp), with the same border, first-order band and clamped-stencil
//     conventions.
#pragma once
#include "thy_step.hpp"

namespace thy {

constexpr int SW_HALO = 4;
constexpr int SW_OUT = 64 - 2 * SW_HALO;  // output columns per wavefront
constexpr int SW_WAVES = 4;               // wavefronts per workgroup (side by side in x)

// Cross-lane moves: lane l receives the value of lane l-1 / l-2 / l+1 (a lane without
// a source keeps its own value; lanes 0, 1 and 63 lie outside every output cell's
// stencil).  These go through ds_bpermute on the LDS pipe, which leaves the VALU (the
// unit this kernel saturates) to the flux arithmetic; DPP row moves were measured
// slower here for that reason.
template <typename T> __device__ __forceinline__ T from_left(T v) { return __shfl_up(v, 1, 64); }
template <typename T> __device__ __forceinline__ T from_left2(T v) { return __shfl_up(v, 2, 64); }
template <typename T> __device__ __forceinline__ T from_right(T v) { return __shfl_down(v, 1, 64); }

// The same moves as VALU data-parallel-primitive moves across the whole wavefront
// (wave_shr:1 / wave_shl:1; the lane without a source keeps its own value): a few cycles
// of latency instead of an LDS round trip.  The general strips, whose wavefronts are few
// and walk long dependency chains, use these; the deep-water strips keep the LDS moves,
// which leave the VALU to their flux arithmetic.  (dpp_left2 differs from from_left2 only
// in lane 1, which lies outside every output cell's stencil.)
__device__ __forceinline__ int dpp_shr1(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xf, 0xf, false); }
__device__ __forceinline__ int dpp_shl1(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x130, 0xf, 0xf, false); }
__device__ __forceinline__ float dpp_left(float v) { return __int_as_float(dpp_shr1(__float_as_int(v))); }
__device__ __forceinline__ float dpp_right(float v) { return __int_as_float(dpp_shl1(__float_as_int(v))); }
__device__ __forceinline__ double dpp_left(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_shr1((int)(b & 0xffffffffll)), hi = dpp_shr1((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double dpp_right(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_shl1((int)(b & 0xffffffffll)), hi = dpp_shl1((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <typename T> __device__ __forceinline__ T dpp_left2(T v) { return dpp_left(dpp_left(v)); }

// read-only table entry at a wavefront-uniform address (scalar load)
template <typename S> __device__ __forceinline__ S uniform_load(const S *p) {
    static_assert(sizeof(S) % 4 == 0, "table entries are whole dwords");
    S out;
    __builtin_memcpy(&out, (const __attribute__((address_space(4))) S *)p, sizeof(S));
    return out;
}

// ---------------------------------------------------------------------------
// Deep-water forms of the face fluxes: the 2s branch that mass_flux / mom_flux take
// when all four stencil depths exceed hpos outside the first-order band
// (src/GPU/Arista_kernel.cu:545,653,1105,1214), written with the per-cell square
// roots and velocities computed once per cell (the same arithmetic as the per-lane 2s
// branch of the general faces, so either path gives a face the same value).
// ---------------------------------------------------------------------------
// x face between cells a (left) and b: src/GPU/Arista_kernel.cu:79-95,118-130 (il =
// alto / cos of the row)
template <typename T>
__device__ __forceinline__ T fast_mass_x(T qa, T ha, T hb, T il, T kA, T dt) {
    const T F = qa * il;
    return positivity(F, ha, hb, kA, kA, dt);
}

// y face between rows a (g-1) and b (g): src/GPU/Arista_kernel.cu:97-114,674-686
template <typename T>
__device__ __forceinline__ T fast_mass_y(T qa, T ha, T hb, const FaceGeoY<T> &f, T dt) {
    const T F = qa * f.lc0;
    return positivity(F, ha, hb, f.k0, f.k1, dt);
}

// one cell of a deep momentum face: total depth, eta, depth, normal and tangential
// discharge, sqrt(h), normal velocity
template <typename T> struct DC {
    T h, e, D, qn, qt, s, un;
};

// x face, 2s momentum flux: src/GPU/Arista_kernel.cu:221-280
template <typename T>
__device__ __forceinline__ MF<T> fast_mom_x(const DC<T> &a, const DC<T> &b, T il, T kd, T dt) {
    MF<T> o;
    const T hij = T(0.5) * (a.h + b.h);
    const T uij = Num<T>::div(Num<T>::mad(a.s, a.un, b.s * b.un), a.s + b.s + Num<T>::eps);
    const T lmax = habs(uij) + Num<T>::sqrt(hij);
    const T deta = deta_of(a.e, b.e, a.D, b.D);
    const bool up = uij < T(0);
    const T uu = up ? b.un : a.un;
    const T qnn = up ? b.qn : a.qn;
    const T qtt = up ? b.qt : a.qt;
    const T fn = uu * qnn * il;
    const T DES = kd * dt * (lmax * lmax) * (b.qn - a.qn);
    o.mn = Num<T>::mad(hij * deta, il, fn) - DES;
    o.pn = -fn + DES;
    const T ft = uu * qtt * il;
    o.mt = ft;
    o.pt = -ft;
    o.w = il * lmax;
    return o;
}

// y face, 2s momentum flux: src/GPU/Arista_kernel.cu:780-843
template <typename T>
__device__ __forceinline__ MF<T> fast_mom_y(const DC<T> &a, const DC<T> &b, const FaceGeoY<T> &f, T dt) {
    MF<T> o;
    const T g0 = f.cos0, g1 = f.cos1, ic0 = f.ic0, ic1 = f.ic1, len = f.len;
    const T hij = T(0.5) * (a.h + b.h);
    const T uij = Num<T>::div(Num<T>::mad(a.s, a.un, b.s * b.un), a.s + b.s + Num<T>::eps);
    const T lmax = habs(uij) + Num<T>::sqrt(hij);
    const T deta = Num<T>::mad(-(uij * uij), f.gd, f.gh * deta_of(a.e, b.e, a.D, b.D));
    const bool up = uij < T(0);
    const T uu = up ? b.un : a.un;
    const T qnn = up ? b.qn : a.qn;
    const T qtt = up ? b.qt : a.qt;
    const T fn = uu * qnn * len;
    T mn, pn;
    if (up) { pn = -fn; mn = fn * f.g1ic0; }
    else { mn = fn; pn = -fn * g0 * ic1; }
    mn = Num<T>::mad(hij * deta * len, ic0, mn);
    const T DES = f.kdy * dt * (lmax * lmax) * Num<T>::mad(b.qn, g1, -(a.qn * g0));
    o.mn = Num<T>::mad(-DES, ic0, mn);
    o.pn = Num<T>::mad(DES, ic1, pn);
    const T ft = uu * qtt * len;
    o.mt = ft;
    o.pt = -ft;
    o.w = len * lmax;
    return o;
}

template <typename T> struct RowCell {
    T e, D, qx, qy;
};

// The row sweeps address the state arrays through one buffer resource over the state
// block that holds them all: the lane's column is the vector offset, the array's offset
// in the block plus the row's the scalar offset, all 32-bit (the block is under 2^31
// bytes; thy_create keeps larger grids off these kernels).  One resource instead of one
// per array leaves the scalar registers to the rest of the loop.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t whole_array(const void *p, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, bytes, 0x00020000);
}

// byte offsets in the state block of the arrays one launch reads and writes; without
// arrival tracking the arrival load reads emax instead (its stores never apply)
template <typename T> struct StateOffsets {
    int e, D, qx, qy, oe, oqx, oqy, emax, arr;
    template <typename A>
    __device__ __forceinline__ StateOffsets(const A &a, const T *ie, const T *iD, const T *iqx, const T *iqy,
                                            const T *we, const T *wqx, const T *wqy)
        : e(at(a, ie)), D(at(a, iD)), qx(at(a, iqx)), qy(at(a, iqy)), oe(at(a, we)), oqx(at(a, wqx)),
          oqy(at(a, wqy)), emax(at(a, a.emax)), arr(at(a, a.track_arrival ? a.arr : a.emax)) {}
    template <typename A> static __device__ __forceinline__ int at(const A &a, const T *p) {
        return (int)((const char *)p - a.block);
    }
};

template <typename T> __device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, int voff, int soff);
template <> __device__ __forceinline__ float buf_load<float>(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
template <> __device__ __forceinline__ double buf_load<double>(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

// Vector offset of a buffer store that must not happen: at or beyond the resource's
// record count, so the hardware drops it.  A store whose lanes may all be masked is
// issued this way instead of under a branch, so that every iteration of a row loop
// issues the same number of memory instructions and the compiler can wait for a row's
// loads without waiting for the stores issued after them (one counter covers both).
constexpr int kDropOffset = (int)0x80000000;

template <typename T>
__device__ __forceinline__ void buf_store(T v, __amdgpu_buffer_rsrc_t r, int voff, int soff, bool on);
template <>
__device__ __forceinline__ void buf_store<float>(float v, __amdgpu_buffer_rsrc_t r, int voff, int soff, bool on) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, on ? voff : kDropOffset, soff, 0);
}
template <>
__device__ __forceinline__ void buf_store<double>(double v, __amdgpu_buffer_rsrc_t r, int voff, int soff, bool on) {
    using V2 = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v), r, on ? voff : kDropOffset, soff, 0);
}

// a stencil cell with its desingularisation factor f (shared by the x faces of a lane
// through lane moves) ...
template <typename T> __device__ __forceinline__ SC<T> sc_of(T e, T D, T qn, T qt, T f) {
    SC<T> c;
    c.e = e;
    c.D = D;
    c.h = e + D;
    c.qn = qn;
    c.qt = qt;
    c.f = f;
    return c;
}
// ... or computed for this cell alone (the compiler drops it where no face uses it)
template <typename T> __device__ __forceinline__ SC<T> sc_of(T e, T D, T qn, T qt, const Prm<T> &p) {
    return sc_of(e, D, qn, qt, desing(e + D, p.eps_h));
}

// One wavefront's share of a step: the strip of a.srows rows whose 56 output columns
// start at x0 and whose first row is y0, with bc filled by begin_step for this launch.
// Returns this lane's minimum of the per-cell dt of the next step (3e38 when it
// updated no cell, or when x0 lies past the grid).
// EDGE false: the caller guarantees a strip 8 cells or more inside the borders and their
// sponges, not periodic (no clamping, border faces, first-order band or sponge).
template <typename T, bool EDGE = true>
__device__ __forceinline__ float sweep_strip(const StepArgs<T> &a, const BlockCtl<T> &bc, const int x0,
                                             const int y0, const int rows) {
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    // (selects rather than indexing, so that a copy of the arguments stays in registers)
    const bool c1 = bc.cur != 0;
    const T *__restrict__ ge = c1 ? a.e[1] : a.e[0];
    const T *__restrict__ gqx = c1 ? a.qx[1] : a.qx[0];
    const T *__restrict__ gqy = c1 ? a.qy[1] : a.qy[0];
    const T *__restrict__ gD = a.D;
    T *__restrict__ oe = c1 ? a.e[0] : a.e[1];
    T *__restrict__ oqx = c1 ? a.qx[0] : a.qx[1];
    T *__restrict__ oqy = c1 ? a.qy[0] : a.qy[1];
    const T dt = (T)bc.dt;
    const T tnow = (T)bc.t;
    const int nx = a.nx, ny = a.ny, P = a.pitch;
    const bool per = EDGE && a.periodic != 0;
    const int y1 = min(y0 + rows, ny);
    float dtmin = 3.0e38f;
    bool bad = false;

    if (x0 < nx) {
        const int gx = x0 - SW_HALO + lane;
        // column loaded by this lane: wrapped (periodic) or clamped onto the grid
        const int cx = !EDGE ? gx : (per ? ((gx % nx) + nx) % nx : (gx < 0 ? 0 : (gx >= nx ? nx - 1 : gx)));
        // lanes outside the grid take the e* of the clamped column (clamped stencils)
        const bool xedge = EDGE && !per && (x0 - SW_HALO < 0 || x0 + SW_OUT + SW_HALO > nx);
        const int src = cx - (x0 - SW_HALO);
        // this lane's x face is the left face of its cell: g = gx
        const int xb = (!EDGE || per) ? 0 : (gx == 0 ? 1 : (gx == nx ? 2 : 0));
        const int gw = per ? cx : gx;
        const bool xac = EDGE && ((gw < 2) || (gw > nx - 2));
        const bool out_col = lane >= SW_HALO && lane < SW_HALO + SW_OUT && (!EDGE || gx < nx);
        const bool any_sponge = EDGE && (a.sp_izq | a.sp_der | a.sp_sup | a.sp_inf) != 0;
        const int isp = per ? cx : gx;

        Prm<T> px, py;
        px.eps_h = py.eps_h = a.eps_h;
        px.hpos = py.hpos = a.hpos;
        px.cvis = py.cvis = a.cvis;
        px.bl = a.bizq; px.bh = a.bder;
        py.bl = a.bsup; py.bh = a.binf;

        auto clamp_row = [&](int r) { return !EDGE ? r : (r < 0 ? 0 : (r >= ny ? ny - 1 : r)); };
        auto clamp_face = [&](int g) { return !EDGE ? g : (g < 0 ? 0 : (g > ny ? ny : g)); };
        // buffer loads: the row (clamped onto the grid) in the scalar offset, the lane's
        // column in the vector offset; never predicated, so the compiler can keep the
        // prefetches in flight across iterations
        const __amdgpu_buffer_rsrc_t rs = whole_array(a.block, a.block_bytes);
        const StateOffsets<T> o(a, ge, gD, gqx, gqy, oe, oqx, oqy);
        // the loads carry each array's offset in the vector offset (one register per array,
        // set once) and only the row in the scalar offset: this loop is short of scalar
        // registers, not of vector ones
        const int vcx = cx * (int)sizeof(T);
        const int ve = vcx + o.e, vD = vcx + o.D, vqx = vcx + o.qx, vqy = vcx + o.qy;
        auto load_row = [&](int r) {
            const int so = clamp_row(r) * P * (int)sizeof(T);
            RowCell<T> c;
            c.e = buf_load<T>(rs, ve, so);
            c.D = buf_load<T>(rs, vD, so);
            c.qx = buf_load<T>(rs, vqx, so);
            c.qy = buf_load<T>(rs, vqy, so);
            return c;
        };
        const int vgx = (EDGE ? (gx < 0 ? 0 : (gx >= nx ? nx - 1 : gx)) : gx) * (int)sizeof(T);
        const int vm = vgx + o.emax, va = vgx + o.arr;
        auto load_metrics = [&](int jn, T &m, T &t) {
            const int so = (jn < y0 ? y0 : (jn >= y1 ? y1 - 1 : jn)) * P * (int)sizeof(T);
            m = buf_load<T>(rs, vm, so);
            t = buf_load<T>(rs, va, so);
        };

        // the segment touches the first-order band of the x borders: no x face of it
        // takes the deep form
        const bool seg_ac = EDGE && __any(xac) != 0;
        const T hpos = a.hpos;

        // pipeline registers: input rows r-5..r, e* rows r-5..r-2 (with sqrt(h*) and
        // qy/h* of rows r-4..r-2 for the deep y faces), mass accumulators rows
        // r-2..r-1, momentum accumulators rows r-4..r-3, and per-row wavefront votes
        // "every lane deeper than hpos" for the inputs (rows r-3..r) and h* (r-5..r-2)
        // and "every lane dry" (h <= eps) for the inputs (rows r-2..r) and h* (r-4..r-2)
        RowCell<T> i5, i4, i3, i2, i1, i0;
        i5 = i4 = i3 = i2 = i1 = i0 = RowCell<T>{T(0), T(0), T(0), T(0)};
        T s5 = T(0), s4 = T(0), s3 = T(0);
        T sq4 = T(0), sq3 = T(0), v4 = T(0), v3 = T(0);
        T m2 = T(0), m1 = T(0);
        T X4 = T(0), Y4 = T(0), W4 = T(0), X3 = T(0), Y3 = T(0), W3 = T(0);
        // the wavefront votes of the rows in the pipeline, one bit each in one word (one
        // scalar register instead of one per vote): bit k of each byte is the vote of the
        // k-th row back; byte 0 "deeper than hpos" (inputs, rows r..r-3), byte 1 "dry"
        // (inputs, rows r..r-2), byte 2 "deeper than hpos" (h*, rows r-2..r-5), byte 3 "dry"
        // (h*, rows r-2..r-4)
        unsigned int votes = 0u;
        // A face both of whose cells are dry (h <= eps) gets a zero flux from the face
        // functions (src/GPU/Arista_kernel.cu:140,298); where every face of a stage is such
        // a face (land), the stage uses those zeros without evaluating the faces.
        constexpr T kZero = T(0);
        const MF<T> kDryMF{kZero, kZero, kZero, kZero, kZero};
        // One pipeline iteration for row r.  buf holds row r on entry and receives row
        // r+1; (pm, pt) hold the metrics inputs of row r-4 on entry and receive those of
        // row r-3: every load has one iteration of work behind it.  (Keeping loads two
        // iterations ahead needs a two-way unrolled loop, whose register use costs more
        // occupancy than the extra latency cover gains.)
        auto iteration = [&](const int r, RowCell<T> &buf, T &pm, T &pt) {
            i0 = buf;
            buf = load_row(r + 1);
            const int ju = r - 4;  // row updated in this iteration
            const T cmax = pm, carr = pt;
            load_metrics(ju + 1, pm, pt);
            const T h0 = i0.e + i0.D;
            const bool din0 = __all(h0 > hpos) != 0;
            const bool dry0 = __all(!(h0 > Num<T>::eps)) != 0;
            votes |= (din0 ? 0x1u : 0u) | (dry0 ? 0x100u : 0u);

            // Stages run only for the rows and faces that some output row of the strip
            // depends on (h* rows y0-2 .. y1+1, momentum faces of rows y0 .. y1-1 and
            // y faces y0 .. y1); the halo iterations skip the rest.
            // ---- mass x-faces of row r ----
            T m0 = T(0);
            if (r >= y0 - 2 && r <= y1 + 1) {
                const RowGeo<T> rg = uniform_load(a.row + clamp_row(r));
                if (din0 && !seg_ac) {
                    const T F = fast_mass_x(dpp_left(i0.qx), dpp_left(h0), h0, rg.ilx, rg.kA, dt);
                    m0 = F - dpp_right(F);
                } else if (dry0) {
                    m0 = ((xb == 1) ? -kZero : kZero) - kZero;
                } else {
                    const T eL = dpp_left(i0.e), DL = dpp_left(i0.D), qL = dpp_left(i0.qx);
                    const T eLL = dpp_left2(i0.e), DLL = dpp_left2(i0.D);
                    const T eR = dpp_right(i0.e), DR = dpp_right(i0.D);
                    const T f0 = desing(h0, px.eps_h);  // the lane's cell, moved to its right neighbour
                    const SC<T> M = sc_of(eLL, DLL, T(0), T(0), T(0)), A = sc_of(eL, DL, qL, T(0), dpp_left(f0));
                    const SC<T> B = sc_of(i0.e, i0.D, i0.qx, T(0), f0), Q = sc_of(eR, DR, T(0), T(0), T(0));
                    const T F = mass_flux<T, false>(M, A, B, Q, xb, xac, rg.ilx, T(1), T(1), rg.kA, rg.kA, rg.valx, dt, px);
                    m0 = ((xb == 1) ? -F : F) - dpp_right(F);
                }
            }
            // ---- mass y-face g = r-1 (rows r-3..r) ----
            if (r - 1 >= y0 - 2 && r - 1 <= y1 + 2) {
                const int g = r - 1;
                const FaceGeoY<T> fg = uniform_load(a.fy + clamp_face(g));
                const int yb = !EDGE ? 0 : (g == 0 ? 1 : (g == ny ? 2 : 0));
                const bool yac = EDGE && ((g < 2) || (g > ny - 2));
                T F;
                if ((votes & 0xfu) == 0xfu && !yac) {
                    F = fast_mass_y(i2.qy, i2.e + i2.D, i1.e + i1.D, fg, dt);
                } else if ((votes & 0x600u) == 0x600u) {
                    F = kZero;
                } else {
                    const SC<T> M = sc_of(i3.e, i3.D, T(0), T(0), T(0)), A = sc_of(i2.e, i2.D, i2.qy, T(0), py);
                    const SC<T> B = sc_of(i1.e, i1.D, i1.qy, T(0), py), Q = sc_of(i0.e, i0.D, T(0), T(0), T(0));
                    F = mass_flux<T, true>(M, A, B, Q, yb, yac, fg.len, fg.cos0, fg.cos1, fg.k0, fg.k1, fg.valy, dt, py);
                }
                m2 = Num<T>::mad(-F, fg.ic0, m2);
                m1 = Num<T>::mad((yb == 1) ? -F : F, fg.ic1, m1);
            }
            // ---- h* (as e*) of row r-2, with the sponge ----
            T s2;
            const RowGeo<T> rg2 = uniform_load(a.row + clamp_row(r - 2));
            const bool hrow = r - 2 >= y0 - 2 && r - 2 <= y1 + 1;  // h* of row r-2 is needed
            if (hrow) {
                T es = Num<T>::mad(dt * rg2.invA, m2, i2.e);
                if (!(es + i2.D > T(0))) es = -i2.D;
                if (any_sponge) es = sponge_eta(a, es, i2.D, isp, r - 2, rg2.dxs, rg2.dys);
                if (xedge) es = __shfl(es, src, 64);
                s2 = es;
            } else {
                s2 = i2.e;
            }
            const T hs2 = s2 + i2.D;
            const bool ds2 = hrow && __all(hs2 > hpos) != 0;
            const bool dsd2 = __all(!(hs2 > Num<T>::eps)) != 0;
            votes |= (ds2 ? 0x10000u : 0u) | (dsd2 ? 0x1000000u : 0u);
            T sq2 = T(0), u2 = T(0), v2 = T(0);
            if (ds2) {
                sq2 = Num<T>::sqrt(hs2);
                const T ih = Num<T>::rcp(hs2);
                u2 = i2.qx * ih;
                v2 = i2.qy * ih;
            }
            // ---- momentum x-faces of row r-2 ----
            T X2 = T(0), Y2 = T(0), W2 = T(0);
            if (r - 2 < y0 || r - 2 >= y1) {
                // halo row: no output row uses its x faces
            } else if (ds2 && !seg_ac) {
                DC<T> A, B;
                A.h = dpp_left(hs2); A.e = dpp_left(s2); A.D = dpp_left(i2.D);
                A.qn = dpp_left(i2.qx); A.qt = dpp_left(i2.qy); A.s = dpp_left(sq2); A.un = dpp_left(u2);
                B.h = hs2; B.e = s2; B.D = i2.D; B.qn = i2.qx; B.qt = i2.qy; B.s = sq2; B.un = u2;
                const MF<T> f = fast_mom_x(A, B, rg2.ilx, rg2.kdx, dt);
                X2 = -f.pn - dpp_right(f.mn);
                Y2 = -f.pt - dpp_right(f.mt);
                W2 = f.w + dpp_right(f.w);
            } else if (dsd2) {
                const MF<T> &f = kDryMF;
                X2 = -((xb == 1) ? f.mn : f.pn) - f.mn;
                Y2 = -((xb == 1) ? f.mt : f.pt) - f.mt;
                W2 = f.w + f.w;
            } else {
                const T eL = dpp_left(s2), DL = dpp_left(i2.D), qxL = dpp_left(i2.qx), qyL = dpp_left(i2.qy);
                const T eLL = dpp_left2(s2), DLL = dpp_left2(i2.D), qyLL = dpp_left2(i2.qy);
                const T eR = dpp_right(s2), DR = dpp_right(i2.D), qyR = dpp_right(i2.qy);
                const T fs = desing(hs2, px.eps_h);  // the lane's cell, moved to its neighbours
                const SC<T> M = sc_of(eLL, DLL, T(0), qyLL, dpp_left2(fs)), A = sc_of(eL, DL, qxL, qyL, dpp_left(fs));
                const SC<T> B = sc_of(s2, i2.D, i2.qx, i2.qy, fs), Q = sc_of(eR, DR, T(0), qyR, dpp_right(fs));
                const MF<T> f = mom_flux<T, false>(M, A, B, Q, xb, xac, rg2.ilx, T(1), T(1), T(1), T(1),
                                                   rg2.valx, rg2.kdx, dt, px);
                X2 = -((xb == 1) ? f.mn : f.pn) - dpp_right(f.mn);
                Y2 = -((xb == 1) ? f.mt : f.pt) - dpp_right(f.mt);
                W2 = f.w + dpp_right(f.w);
            }
            // ---- momentum y-face g = r-3 (rows r-5..r-2) ----
            if (r - 3 >= y0 && r - 3 <= y1) {
                const int g = r - 3;
                const FaceGeoY<T> fg = uniform_load(a.fy + clamp_face(g));
                const int yb = !EDGE ? 0 : (g == 0 ? 1 : (g == ny ? 2 : 0));
                const bool yac = EDGE && ((g < 2) || (g > ny - 2));
                MF<T> f;
                if ((votes & 0xf0000u) == 0xf0000u && !yac) {
                    DC<T> A, B;
                    A.h = s4 + i4.D; A.e = s4; A.D = i4.D; A.qn = i4.qy; A.qt = i4.qx; A.s = sq4; A.un = v4;
                    B.h = s3 + i3.D; B.e = s3; B.D = i3.D; B.qn = i3.qy; B.qt = i3.qx; B.s = sq3; B.un = v3;
                    f = fast_mom_y(A, B, fg, dt);
                } else if ((votes & 0x6000000u) == 0x6000000u && (!EDGE || (g - 1 >= 0 && g < ny))) {
                    f = kDryMF;  // the face's cells are rows r-4 and r-3 (no clamping here)
                } else {
                    SC<T> M = sc_of(s5, i5.D, i5.qy, i5.qx, py), A = sc_of(s4, i4.D, i4.qy, i4.qx, py);
                    SC<T> B = sc_of(s3, i3.D, i3.qy, i3.qx, py), Q = sc_of(s2, i2.D, i2.qy, i2.qx, py);
                    // rows outside the grid take the e* of the edge row (clamped stencil)
                    if (EDGE) {
                        if (g - 2 < 0) M = (g - 1 < 0) ? B : A;
                        if (g - 1 < 0) A = B;
                        if (g >= ny) B = A;
                        if (g + 1 >= ny) Q = (g >= ny) ? A : B;
                    }
                    f = mom_flux<T, true>(M, A, B, Q, yb, yac, fg.len, fg.cos0, fg.cos1, fg.ic0, fg.ic1, fg.valy,
                                          fg.kdy, dt, py);
                }
                Y4 -= f.mn;
                X4 -= f.mt;
                W4 += f.w;
                if (yb == 1) { Y3 -= f.mn; X3 -= f.mt; }
                else { Y3 -= f.pn; X3 -= f.pt; }
                W3 += f.w;
            }
            // ---- cell update of row r-4 (src/GPU/Volumen_kernel.cu:160-221) ----
            // The arithmetic runs for output rows only; the five stores are issued in
            // every iteration and dropped where they do not apply.
            {
                const bool row_on = ju >= y0 && ju < y1;
                const bool on = row_on && out_col;
                const int soff = (row_on ? ju : y0) * P * (int)sizeof(T);
                const int voff = on ? gx * (int)sizeof(T) : kDropOffset;
                const T eo = i4.e;
                bool max_on = false, arr_on = false;
                T qx = T(0), qy = T(0);
                if (row_on) {
                    const RowGeo<T> rg = uniform_load(a.row + ju);
                    const T D = i4.D;
                    // metrics from the state before the step (src/GPU/Metrics.cu:23-42,57-77)
                    max_on = on && eo > cmax && eo + D > a.eps_h;
                    arr_on = on && a.track_arrival && !Arrival<T>::arrived(carr) && tabs(eo - carr) > a.dif_at;
                    const T dtc = (W4 < Num<T>::eps) ? Num<T>::big : Num<T>::div(T(2) * a.cfl * rg.A, W4);
                    dtmin = fminf(dtmin, (float)dtc);  // halo lanes are dropped after the loop
                    const T val = dt * rg.invA;
                    qx = Num<T>::mad(val, X4, i4.qx);
                    qy = Num<T>::mad(val, Y4, i4.qy);
                    if (any_sponge) sponge_q(a, gx, ju, qx, qy);
                    const T hs = s4 + D;
                    if (hs <= hpos) near_dry(hs, i4.qx, i4.qy, qx, qy, dt, a.mf0, a.vmax, a.eps_h);
                    bad = bad || (on && bad_cell(s4, qx, qy));
                }
                buf_store<T>(eo, rs, voff, o.emax + soff, max_on);
                buf_store<T>(Arrival<T>::enc(tnow), rs, voff, o.arr + soff, arr_on);
                buf_store<T>(s4, rs, voff, o.oe + soff, on);
                buf_store<T>(qx, rs, voff, o.oqx + soff, on);
                buf_store<T>(qy, rs, voff, o.oqy + soff, on);
            }
            // ---- advance the pipeline ----
            i5 = i4; i4 = i3; i3 = i2; i2 = i1; i1 = i0;
            s5 = s4; s4 = s3; s3 = s2;
            sq4 = sq3; sq3 = sq2; v4 = v3; v3 = v2;
            m2 = m1; m1 = m0;
            X4 = X3; Y4 = Y3; W4 = W3;
            X3 = X2; Y3 = Y2; W3 = W2;
            votes = (votes << 1) & 0x060e060eu;  // the bits of rows that stay in the pipeline
        };

        // rows y0-4 .. y1+3
        const int rbeg = y0 - 4, rend = y1 + 3;
        RowCell<T> buf = load_row(rbeg);
        T pm, pt;
        load_metrics(rbeg - 4, pm, pt);
        // enter the loop with nothing outstanding, so that the wait at its head only
        // accounts for the previous iteration (see kDropOffset)
        __builtin_amdgcn_s_waitcnt(0);
        for (int r = rbeg; r <= rend; r++) iteration(r, buf, pm, pt);
        dtmin = out_col ? dtmin : 3.0e38f;
    }
    raise_health(a, bad);
    return dtmin;
}

// The workgroup's four wavefronts on strip block `tile` (side by side in x).
template <typename T>
__device__ __forceinline__ float sweep_block(const StepArgs<T> &a, const BlockCtl<T> &bc, const int tile,
                                             const bool gauges) {
    if (gauges) {
        const bool c1 = bc.cur != 0;
        gather_gauges(a, bc, c1 ? a.e[1] : a.e[0], c1 ? a.qx[1] : a.qx[0], c1 ? a.qy[1] : a.qy[0], threadIdx.x,
                      64 * SW_WAVES);
    }
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tyi = tile / a.tiles_x, txi = tile - tyi * a.tiles_x;
    return sweep_strip(a, bc, (txi * SW_WAVES + wv) * SW_OUT, tyi * a.srows, a.srows);
}

// wavefront work items of the split launches: strip row * nseg + segment (-1: idle),
// strip rows in units of `rows`, shifted up by a.yoff rows so that the strip rows along
// the top and bottom borders are thin (the first one is cut at row 0).  n: the strip's
// rows inside the grid.
template <typename T>
__device__ __forceinline__ void item_origin(const StepArgs<T> &a, const int item, const int rows, int &x0, int &y0,
                                            int &n) {
    const int strip = item / a.nseg;
    x0 = (item - strip * a.nseg) * SW_OUT;
    const int y = strip * rows - a.yoff;
    y0 = max(y, 0);
    n = min(y + rows, a.ny) - y0;
}

// origin of a deep-water strip (never cut at row 0: it lies clear of the top border band)
template <typename T>
__device__ __forceinline__ void deep_origin(const StepArgs<T> &a, const int item, int &x0, int &y0) {
    const int strip = item / a.nseg;
    x0 = (item - strip * a.nseg) * SW_OUT;
    y0 = strip * a.srows - a.yoff;
}

// workgroup minimum of the lanes' dt, folded into the launch's dt slot of the grid
template <typename T>
__device__ __forceinline__ void block_dt_min(const StepArgs<T> &a, const int slot, float dtmin, float *s_red) {
    const int tid = threadIdx.x;
    dtmin = wave_min(dtmin);
    if ((tid & 63) == 0) s_red[tid >> 6] = dtmin;
    __syncthreads();
    if (tid == 0) {
        float m = s_red[0];
#pragma unroll
        for (int w = 1; w < SW_WAVES; w++) m = fminf(m, s_red[w]);
        atomicMin(&a.dtmin[dt_at(slot, 0)], __float_as_uint(m));
    }
}

// One step of one grid, general faces.  mode 0: one workgroup per strip block, every
// block (the table's order, costliest first, or without a table the chunked XCD
// order).  With a deep-water launch in the step, one wavefront per strip: mode 1 takes
// the general strips (costliest first, cut into strips of a.gsrows rows so that the
// few, slow coastal and border strips do not set the step's critical path), mode 2 the
// strips the deep-water launch of this step handed back (the grid has room for all of
// them; workgroups beyond the hand-backs return before the prologue).
template <typename T>
__global__ void __launch_bounds__(64 * SW_WAVES) k_sweep(const StepArgs<T> a, const int slot, const int mode) {
    __shared__ BlockCtl<T> bc;
    __shared__ float s_red[SW_WAVES];
    const int tid = threadIdx.x;
    const int bid = blockIdx.x;
    // the step's writer is the deep-water launch when there is one, else workgroup 0
    const bool writer = bid == 0 && mode == 0;
    int n_fb = 0;
    if (mode == 2) {  // hand-backs: workgroups beyond them return before the prologue
        n_fb = min((int)__builtin_amdgcn_readfirstlane(a.fb_count[slot]), a.n_deep);
        if (bid * SW_WAVES >= n_fb) return;
    }
    if (tid == 0) begin_step(a, slot, writer, bc);
    __syncthreads();
    if (!bc.run) return;
    float dtmin;
    if (mode == 0) {
        int tile;
        if (!a.order) {
            const int loc = bid >> 3, ch = loc / a.chunk;
            tile = ((ch << 3) + (bid & 7)) * a.chunk + (loc - ch * a.chunk);
            if (tile >= a.n_tiles) return;
        } else {
            if (bid >= a.n_order) return;
            tile = a.order[bid];
        }
        dtmin = sweep_block(a, bc, __builtin_amdgcn_readfirstlane(tile), writer);
    } else {
        const int k = bid * SW_WAVES + __builtin_amdgcn_readfirstlane(tid >> 6);
        int item = -1;
        if (mode == 1) {
            if (k < a.n_gen) item = a.gen[k];
        } else {
            if (k < n_fb) item = a.fb_list[slot * a.n_deep + k];
        }
        item = __builtin_amdgcn_readfirstlane(item);
        const int rows = mode == 1 ? a.gsrows : a.srows;
        dtmin = 3.0e38f;
        if (item >= 0) {
            int x0, y0, n;
            item_origin(a, item, rows, x0, y0, n);
            // strips whose rows and 64 columns lie 8 cells or more inside the borders and
            // their sponges (no clamped rows or columns, no first-order band, no border
            // faces, no sponge) take the strip compiled without the edge cases
            constexpr int m = 8;
            const bool interior = !a.periodic && x0 - SW_HALO - m >= a.sp_izq &&
                                  x0 + SW_OUT + SW_HALO + m <= a.nx - a.sp_der && y0 - m >= a.sp_sup &&
                                  y0 + n + m <= a.ny - a.sp_inf;
            if (n > 0) dtmin = interior ? sweep_strip<T, false>(a, bc, x0, y0, n) : sweep_strip<T, true>(a, bc, x0, y0, n);
        }
    }
    block_dt_min(a, slot, dtmin, s_red);
}

// ---------------------------------------------------------------------------
// Deep-water sweep.  Blocks whose strips and halos lie clear of the borders, the
// first-order bands and the sponges, and whose still-water depth exceeds 3 hpos, run
// this lean form of the pipeline: only the 2s faces, no sponge and no near-dry update,
// no clamped columns or rows, and fewer pipeline registers.  Every wavefront checks at
// run time the conditions under which the general faces take the 2s branch (every
// loaded row deeper than hpos; every h* row that a face's stencil reaches deeper than
// hpos).  A workgroup where a check fails commits no dt and hands its block to the
// general launch of the same step, which recomputes and rewrites all of it; the two
// launches evaluate the deep faces with the same functions, so the split changes no
// result.
// ---------------------------------------------------------------------------


// One wavefront's strip on the deep-water path.  The 2s faces reach one row and one
// column, so the pipeline lags two rows: iteration r loads row r, evaluates the mass
// faces of row r and the y face between rows r-1 and r, h* of row r-1, the momentum
// faces of row r-1 and the y face between rows r-2 and r-1, and updates row r-2.  Per
// cell the accumulation order is that of the general pipeline (x faces, lower y face,
// upper y face), so the values are the same.  The conditions under which the general
// faces take the 2s branch are checked for the whole strip and combined: every row
// y0-4 .. y1+3 and every h* row y0-2 .. y1+1 deeper than hpos on every lane.  Returns
// whether they held; if not, the caller hands the strip back (its writes are then
// rewritten by the general launch and its dt is dropped).
template <typename T>
__device__ __forceinline__ bool deep_strip(const StepArgs<T> &a, const BlockCtl<T> &bc, const int x0, const int y0,
                                           float &dtmin_out) {
    const int lane = threadIdx.x & 63;
    const bool c1 = bc.cur != 0;
    T *__restrict__ oe = c1 ? a.e[0] : a.e[1];
    T *__restrict__ oqx = c1 ? a.qx[0] : a.qx[1];
    T *__restrict__ oqy = c1 ? a.qy[0] : a.qy[1];
    const T dt = (T)bc.dt;
    const T tnow = (T)bc.t;
    const int P = a.pitch;
    const int y1 = min(y0 + a.srows, a.ny);
    float dtmin = 3.0e38f;
    bool valid = true, bad = false;
    if (x0 < a.nx) {
        // the strip classification keeps columns x0-4 .. x0+59 and rows y0-4 .. y1+3
        // inside the grid: no clamping.  Loads are buffer loads: the lane's column is the
        // vector offset, the row the scalar offset.
        const __amdgpu_buffer_rsrc_t rs = whole_array(a.block, a.block_bytes);
        const StateOffsets<T> o(a, c1 ? a.e[1] : a.e[0], a.D, c1 ? a.qx[1] : a.qx[0], c1 ? a.qy[1] : a.qy[0], oe, oqx,
                                oqy);
        const int gx = x0 - SW_HALO + lane;
        const int voff = gx * (int)sizeof(T);
        // byte addresses of the left and right neighbour lanes for ds_bpermute (the same
        // moves as from_left / from_right, computed once per strip; FP64 values move as
        // two dwords through from_left / from_right)
        const int adl = (lane > 0 ? lane - 1 : lane) << 2, adr = (lane < 63 ? lane + 1 : lane) << 2;
        auto left = [&](const T v) -> T {
            if constexpr (sizeof(T) == 4) return __int_as_float(__builtin_amdgcn_ds_bpermute(adl, __float_as_int(v)));
            else return from_left(v);
        };
        auto right = [&](const T v) -> T {
            if constexpr (sizeof(T) == 4) return __int_as_float(__builtin_amdgcn_ds_bpermute(adr, __float_as_int(v)));
            else return from_right(v);
        };
        auto ld = [&](const int off, const int row) { return buf_load<T>(rs, voff, off + row * P * (int)sizeof(T)); };
        auto load_row = [&](const int r) {
            RowCell<T> c;
            c.e = ld(o.e, r);
            c.D = ld(o.D, r);
            c.qx = ld(o.qx, r);
            c.qy = ld(o.qy, r);
            return c;
        };
        auto metrics_row = [&](const int j) { return j < y0 ? y0 : (j >= y1 ? y1 - 1 : j); };
        const bool out_col = lane >= SW_HALO && lane < SW_HALO + SW_OUT;
        const T hpos = a.hpos;
        // rows y0-4 and y1+3 enter only the branch conditions of the outermost mass faces
        valid = __all(ld(o.e, y0 - 4) + ld(o.D, y0 - 4) > hpos) && __all(ld(o.e, y1 + 3) + ld(o.D, y1 + 3) > hpos);
        // pipeline registers: input rows r-2, r-1; h* row r-2 with sqrt(h*), qy/h*;
        // mass accumulator of row r-1; momentum accumulators of row r-2
        RowCell<T> i2 = RowCell<T>{T(0), T(0), T(0), T(0)}, i1 = i2;
        T s2 = T(0), sq2 = T(0), v2 = T(0), m1 = T(0);
        T X2 = T(0), Y2 = T(0), W2 = T(0);
        RowCell<T> buf = load_row(y0 - 3);
        T pm = ld(o.emax, y0), pt = ld(o.arr, y0);
        // One row of the pipeline.  In the steady state (rows y0+2 .. y1) every stage
        // applies, so the stage conditions drop out at compile time; the other rows keep
        // them.  Both forms evaluate the same expressions.
        auto row = [&](const int r, auto steady_tag) {
            constexpr bool kS = decltype(steady_tag)::value;
            const RowCell<T> i0 = buf;
            buf = load_row((kS || r < y1 + 2) ? r + 1 : r);
            const int ju = r - 2;  // row updated in this iteration
            const T cmax = pm, carr = pt;
            {
                const int jn = kS ? ju + 1 : metrics_row(ju + 1);
                pm = ld(o.emax, jn);
                pt = ld(o.arr, jn);
            }
            const T h0 = i0.e + i0.D;
            valid = valid && __all(h0 > hpos);
            // mass x-faces of row r (h* rows y0-2 .. y1+1)
            T m0 = T(0);
            if (kS || (r >= y0 - 2 && r <= y1 + 1)) {
                const RowGeo<T> rg = uniform_load(a.row + r);
                const T F = fast_mass_x(left(i0.qx), left(h0), h0, rg.ilx, rg.kA, dt);
                m0 = F - right(F);
            }
            // mass y-face g = r between rows r-1 and r (faces y0-2 .. y1+2)
            if (kS || r >= y0 - 2) {
                const FaceGeoY<T> fg = uniform_load(a.fy + r);
                const T F = fast_mass_y(i1.qy, i1.e + i1.D, h0, fg, dt);
                m1 = Num<T>::mad(-F, fg.ic0, m1);
                m0 = Num<T>::mad(F, fg.ic1, m0);
            }
            // h* of row r-1 (rows y0-2 .. y1+1)
            T s1 = T(0), hs1 = T(0), sq1 = T(0), u1 = T(0), v1 = T(0);
            const int j1 = r - 1;
            const RowGeo<T> rg1 = uniform_load(a.row + ((kS || j1 >= y0 - 2) ? j1 : y0 - 2));
            if (kS || (j1 >= y0 - 2 && j1 <= y1 + 1)) {
                T es = Num<T>::mad(dt * rg1.invA, m1, i1.e);
                if (!(es + i1.D > T(0))) es = -i1.D;
                s1 = es;
                hs1 = es + i1.D;
                valid = valid && __all(hs1 > hpos);
                sq1 = Num<T>::sqrt(hs1);
                const T ih = Num<T>::rcp(hs1);
                u1 = i1.qx * ih;
                v1 = i1.qy * ih;
            }
            // momentum x-faces of row r-1 (rows y0 .. y1-1)
            T X1 = T(0), Y1 = T(0), W1 = T(0);
            if (kS || (j1 >= y0 && j1 < y1)) {
                DC<T> A, B;
                A.h = left(hs1); A.e = left(s1); A.D = left(i1.D);
                A.qn = left(i1.qx); A.qt = left(i1.qy); A.s = left(sq1); A.un = left(u1);
                B.h = hs1; B.e = s1; B.D = i1.D; B.qn = i1.qx; B.qt = i1.qy; B.s = sq1; B.un = u1;
                const MF<T> f = fast_mom_x(A, B, rg1.ilx, rg1.kdx, dt);
                X1 = -f.pn - right(f.mn);
                Y1 = -f.pt - right(f.mt);
                W1 = f.w + right(f.w);
            }
            // momentum y-face g = r-1 between rows r-2 and r-1 (faces y0 .. y1)
            if (kS || (j1 >= y0 && j1 <= y1)) {
                const FaceGeoY<T> fg = uniform_load(a.fy + j1);
                DC<T> A, B;
                A.h = s2 + i2.D; A.e = s2; A.D = i2.D; A.qn = i2.qy; A.qt = i2.qx; A.s = sq2; A.un = v2;
                B.h = hs1; B.e = s1; B.D = i1.D; B.qn = i1.qy; B.qt = i1.qx; B.s = sq1; B.un = v1;
                const MF<T> f = fast_mom_y(A, B, fg, dt);
                Y2 -= f.mn;
                X2 -= f.mt;
                W2 += f.w;
                Y1 -= f.pn;
                X1 -= f.pt;
                W1 += f.w;
            }
            // cell update of row r-2 (src/GPU/Volumen_kernel.cu:160-221; h* > hpos here).
            // Every iteration issues the five stores; those of halo rows and halo lanes,
            // and metric stores that do not apply, are dropped.
            {
                const bool in_rows = kS || (ju >= y0 && ju < y1);
                const bool on = in_rows && out_col;
                const int jr = in_rows ? ju : y0;  // a row inside the arrays
                const RowGeo<T> rg = uniform_load(a.row + jr);
                const int soff = jr * P * (int)sizeof(T);
                const T D = i2.D, eo = i2.e;
                buf_store<T>(eo, rs, voff, o.emax + soff, on && eo > cmax && eo + D > a.eps_h);
                buf_store<T>(Arrival<T>::enc(tnow), rs, voff, o.arr + soff,
                              on && a.track_arrival && !Arrival<T>::arrived(carr) && tabs(eo - carr) > a.dif_at);
                const T dtc = (W2 < Num<T>::eps) ? Num<T>::big : Num<T>::div(T(2) * a.cfl * rg.A, W2);
                dtmin = in_rows ? fminf(dtmin, (float)dtc) : dtmin;  // halo lanes are dropped after the loop
                const T val = dt * rg.invA;
                const T qx = Num<T>::mad(val, X2, i2.qx), qy = Num<T>::mad(val, Y2, i2.qy);
                bad = bad || (on && bad_cell(s2, qx, qy));
                buf_store<T>(s2, rs, voff, o.oe + soff, on);
                buf_store<T>(qx, rs, voff, o.oqx + soff, on);
                buf_store<T>(qy, rs, voff, o.oqy + soff, on);
            }
            i2 = i1; i1 = i0;
            s2 = s1; sq2 = sq1; v2 = v1;
            m1 = m0;
            X2 = X1; Y2 = Y1; W2 = W1;
        };
        using Edge = std::integral_constant<bool, false>;
        using Steady = std::integral_constant<bool, true>;
        // Enter the loops with no memory instruction outstanding: the wait before a row
        // then only has the previous rows to account for (their stores may stay in flight
        // while the next row loads), instead of a full wait every row.
        __builtin_amdgcn_s_waitcnt(0);
        // rows y0-3 .. y1+2: the halo rows at either end with the stage conditions, the
        // steady state without them
        const int pe = min(y0 + 1, y1 + 2);
        int r = y0 - 3;
        for (; r <= pe; r++) row(r, Edge{});
#pragma unroll 1
        for (; r <= y1; r++) row(r, Steady{});
        for (; r <= y1 + 2; r++) row(r, Edge{});
        dtmin = out_col ? dtmin : 3.0e38f;
    }
    // a strip handed back is redone (and checked) by the general launch
    if (valid) raise_health(a, bad);
    dtmin_out = dtmin;
    return valid;
}

template <typename T>
__global__ void __launch_bounds__(64 * SW_WAVES) k_sweep_deep(const StepArgs<T> a, const int slot) {
    __shared__ BlockCtl<T> bc;
    __shared__ float s_red[SW_WAVES];
    const int tid = threadIdx.x, bid = blockIdx.x;
    if (tid == 0) {
        begin_step(a, slot, bid == 0, bc);  // this launch is the step's writer
        if (bid == 0) a.fb_count[(slot + 1) % 3] = 0u;  // the next launch's hand-back list
    }
    __syncthreads();
    if (!bc.run) return;
    if (bid == 0) {
        const bool c1 = bc.cur != 0;
        gather_gauges(a, bc, c1 ? a.e[1] : a.e[0], c1 ? a.qx[1] : a.qx[0], c1 ? a.qy[1] : a.qy[0], tid,
                      64 * SW_WAVES);
    }
    const int k = bid * SW_WAVES + __builtin_amdgcn_readfirstlane(tid >> 6);
    const int item = __builtin_amdgcn_readfirstlane(k < a.n_deep ? a.deep[k] : -1);
    float dtmin = 3.0e38f;
    if (item >= 0) {
        int x0, y0;
        deep_origin(a, item, x0, y0);
        if (!deep_strip(a, bc, x0, y0, dtmin)) {
            // hand the strip back; its dt comes from the general launch that redoes it
            dtmin = 3.0e38f;
            if ((tid & 63) == 0) {
                const unsigned n = atomicAdd(&a.fb_count[slot], 1u);
                if (n < (unsigned)a.n_deep) a.fb_list[slot * a.n_deep + n] = item;
            }
        }
    }
    block_dt_min(a, slot, dtmin, s_red);
}

}  // namespace thy

namespace thy {

// ---------------------------------------------------------------------------
// Many grids per launch.  Each scenario keeps its own state, control block and dt
// slots; a step of the batch advances every running scenario by one of its own steps
// in three launches, as for one grid: the deep-water strips of all scenarios, beside
// them (on a second stream) the general strips of all scenarios, costliest first, and
// after both the strips that the deep-water launch handed back.  Every wavefront takes
// one item (scenario, strip) and runs the step prologue of its scenario itself; the
// item flagged as its scenario's writer publishes the next control block.  Item
// counts are read on the device, so one captured launch sequence serves a changing
// item list up to the captured grid sizes.
// ---------------------------------------------------------------------------
struct BatchArgs {
    const StepArgs<float> *args;  // per scenario slot (zeroed while the slot is free)
    const int2 *deep;             // (scenario slot, strip << 1 | writer): deep-water strips
    const int *n_deep;
    const int2 *gen;              // (scenario slot, strip << 1 | writer): general strips
    const int *n_gen;
    int2 *fb_list;                // 3 slots of fb_cap hand-backs (scenario slot, strip)
    unsigned int *fb_count;       // 3 slots: launch n appends to slot n mod 3
    int fb_cap;
    int gen_sub;                  // general items are strips of gsrows rows (else srows)
};

// the step prologue of one wavefront's scenario (every lane computes the same values;
// lane 0 of the writer item publishes the next control block and samples the gauges)
__device__ __forceinline__ BlockCtl<float> wave_begin(const StepArgs<float> &a, const int slot, const bool writer) {
    const int lane = threadIdx.x & 63;
    BlockCtl<float> bc;
    begin_step(a, slot, writer && lane == 0, bc, kDtParts);
    if (writer && bc.run) {
        const bool c1 = bc.cur != 0;
        gather_gauges(a, bc, c1 ? a.e[1] : a.e[0], c1 ? a.qx[1] : a.qx[0], c1 ? a.qy[1] : a.qy[0], lane, 64);
    }
    return bc;
}

__device__ __forceinline__ void wave_dt_min(const StepArgs<float> &a, const int slot, const float dtmin) {
    const float m = wave_min(dtmin);
    if ((threadIdx.x & 63) == 0)
        atomicMin(&a.dtmin[dt_at(slot, blockIdx.x * SW_WAVES + (threadIdx.x >> 6))], __float_as_uint(m));
}

__global__ void __launch_bounds__(64 * SW_WAVES) k_batch_deep(const BatchArgs q, const int slot) {
    const int lane = threadIdx.x & 63;
    const int k = blockIdx.x * SW_WAVES + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (k == 0 && lane == 0) q.fb_count[(slot + 1) % 3] = 0u;  // the next launch's hand-back list
    const int n = __builtin_amdgcn_readfirstlane(*q.n_deep);
    if (k >= n) return;
    const int2 w = q.deep[k];
    const int sc = __builtin_amdgcn_readfirstlane(w.x), code = __builtin_amdgcn_readfirstlane(w.y);
    if (sc < 0) return;  // padding of an XCD run
    const StepArgs<float> a = uniform_load(q.args + sc);
    if (!a.ctrl) return;
    const BlockCtl<float> bc = wave_begin(a, slot, (code & 1) != 0);
    if (!bc.run) return;
    int x0, y0;
    deep_origin(a, code >> 1, x0, y0);
    float dtmin;
    if (!deep_strip(a, bc, x0, y0, dtmin)) {
        // hand the strip back; its dt comes from the launch that redoes it
        dtmin = 3.0e38f;
        if (lane == 0) {
            const unsigned m = atomicAdd(&q.fb_count[slot], 1u);
            if (m < (unsigned)q.fb_cap) q.fb_list[slot * q.fb_cap + m] = make_int2(sc, code >> 1);
        }
    }
    wave_dt_min(a, slot, dtmin);
}

// one wavefront's general strip of the batch (item code: strip << 1 | writer; strips
// of gsrows rows when sub, else of srows rows)
__device__ __forceinline__ void batch_general_item(const BatchArgs &q, const int slot, const int sc, const int code,
                                                   const bool sub) {
    const StepArgs<float> a = uniform_load(q.args + sc);
    if (!a.ctrl) return;
    const BlockCtl<float> bc = wave_begin(a, slot, (code & 1) != 0);
    if (!bc.run) return;
    const int rows = sub ? a.gsrows : a.srows;
    int x0, y0, n;
    item_origin(a, code >> 1, rows, x0, y0, n);
    wave_dt_min(a, slot, n > 0 ? sweep_strip(a, bc, x0, y0, n) : 3.0e38f);
}

// HANDBACK false: the general strips, one wavefront each.  HANDBACK true: the strips
// that this step's deep-water launch handed back; a small grid strides over them.
template <bool HANDBACK>
__global__ void __launch_bounds__(64 * SW_WAVES) k_batch_general(const BatchArgs q, const int slot) {
    const int k0 = blockIdx.x * SW_WAVES + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (!HANDBACK) {
        const int n = __builtin_amdgcn_readfirstlane(*q.n_gen);
        if (k0 >= n) return;
        const int2 w = q.gen[k0];
        batch_general_item(q, slot, __builtin_amdgcn_readfirstlane(w.x), __builtin_amdgcn_readfirstlane(w.y),
                           q.gen_sub != 0);
    } else {
        const int n = __builtin_amdgcn_readfirstlane((int)min(q.fb_count[slot], (unsigned)q.fb_cap));
        for (int k = k0; k < n; k += gridDim.x * SW_WAVES) {
            const int2 w = q.fb_list[slot * q.fb_cap + k];
            batch_general_item(q, slot, __builtin_amdgcn_readfirstlane(w.x),
                               __builtin_amdgcn_readfirstlane(w.y) << 1, false);
        }
    }
}

// control blocks of the scenarios at launch slot `slot`, gathered for one copy (a free
// scenario slot has a zeroed argument block and is skipped)
__global__ void k_gather_ctrl(const StepArgs<float> *args, const int n, const int slot, Ctrl *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && args[i].ctrl) out[i] = args[i].ctrl[slot];
}

}  // namespace thy
