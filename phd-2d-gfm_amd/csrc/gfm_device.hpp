// gfm_device.hpp -- device-side types and numerics for the GFM hot path on
// gfx950: instance descriptors, stiffened-gas EOS, the exact Riemann solver,
// the boundary-condition gather map and the level-set helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace gfm {

enum { BC_T = 0, BC_R = 1, BC_P = 2 };
enum { V_ORIGINAL = 0, V_REAL = 1, V_RIEMANN = 2, V_MODIFIED = 3 };

// double scalars per instance
enum { SD_T = 0, SD_DT = 1, SD_MASS1 = 2, SD_MASS2 = 3, SD_N = 8 };
// int scalars per instance
enum {
    SI_NSTEP = 0, SI_ACTIVE = 1, SI_NIFACE = 2, SI_DTCOUNT = 3,
    SI_BBOX = 4,   // 8 ints: fluid1 imin,imax,jmin,jmax ; fluid2 ...
    SI_STATUS = 12, SI_ROUNDS = 13, SI_N = 16
};
enum { ST_UNPHYS = 1, ST_RS = 2, ST_IOVF = 4, ST_IROUNDS = 8 };

// One simulation.  Plane arrays are row-major with row pitch `pitch`
// elements; a 4-component field is 4 consecutive planes of `plane` elements.
struct Inst {
    int Nx, Ny, g, W, H, pitch;
    long long plane;
    double x0, y0, dx, dy;
    int bc[4];                 // L, R, B, T
    double gam[2], pinf[2];
    double cfl, T;
    int max_steps;
    int ntile_dt;
    int icap;                  // interfacial list capacity
    int use_bbox;              // OGFM narrows the per-fluid update boxes
    void* Q[2][2];             // conserved [fluid][buffer]
    void* P[2][2];             // primitives [fluid][buffer]
    void* phi;
    void* phiA;
    void* phiB;
    void* vel;                 // 2 planes
    int* imap;                 // interfacial list index per cell, -1 if none
    int* ilist;                // row-major linear index (i*W+j) per entry
    double* side;              // 8 doubles per interfacial entry
    double* partials;          // 4 doubles per dt tile
    double* sd;                // SD_* scalars
    int* si;                   // SI_* scalars
    double* mass_hist;         // (t, mass1, mass2) per step, hist_cap entries
    int hist_cap;
    void* irec;                // interfacial records, icap entries
    int* idone;                // 1 when an entry's star state is final after the parallel pass
};

// HOUC first-derivative stencil for positive velocity: offsets o0..o0+n-1
struct Houc { double c[10]; int n, o0; };

__device__ __forceinline__ long long cidx(const Inst& I, int i, int j) {
    return (long long)i * I.pitch + j;
}

// ---------------------------------------------------------------------------
// stiffened-gas EOS (closed forms; the reference takes them from PhD-Common
// stiffened_gas_eos.hpp, used at euler_misc.hpp:46,59,77,156)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T eos_p(T g, T pinf, T e, T rho) { return (g - T(1)) * rho * e - g * pinf; }
template <typename T>
__device__ __forceinline__ T eos_e(T g, T pinf, T p, T rho) { return (p + g * pinf) / ((g - T(1)) * rho); }
template <typename T>
__device__ __forceinline__ T eos_c(T g, T pinf, T p, T rho) { return sqrt(g * (p + pinf) / rho); }

template <typename T>
struct V4 { T v[4]; };

// euler_misc.hpp:52-67
template <typename T>
__device__ __forceinline__ V4<T> cons2prims(T g, T pinf, const V4<T>& U) {
    V4<T> W;
    T rho = U.v[0];
    T u = U.v[1] / rho;
    T v = U.v[2] / rho;
    T e = (U.v[3] / rho) - T(0.5) * (u * u + v * v);
    W.v[0] = rho; W.v[1] = u; W.v[2] = v; W.v[3] = eos_p(g, pinf, e, rho);
    return W;
}

// euler_misc.hpp:89-108
template <typename T>
__device__ __forceinline__ V4<T> prims2cons(T g, T pinf, const V4<T>& W) {
    V4<T> U;
    T rho = W.v[0], u = W.v[1], v = W.v[2];
    T e = eos_e(g, pinf, W.v[3], rho);
    U.v[0] = rho; U.v[1] = rho * u; U.v[2] = rho * v;
    U.v[3] = rho * e + T(0.5) * rho * (u * u + v * v);
    return U;
}

// euler_misc.hpp:32-41
template <typename T>
__device__ __forceinline__ bool phys_state(const V4<T>& U) {
    T rho = U.v[0];
    T u = U.v[1] / rho, v = U.v[2] / rho;
    T e = (U.v[3] / rho) - T(0.5) * (u * u + v * v);
    return rho >= T(0) && e >= T(0);
}

// euler_misc.hpp:44-49
template <typename T>
__device__ __forceinline__ bool phys_prims(T g, T pinf, const V4<T>& W) {
    T e = eos_e(g, pinf, W.v[3], W.v[0]);
    return e >= T(0) && W.v[0] >= T(0);
}

// euler_misc.hpp:70-86
template <typename T>
__device__ __forceinline__ V4<T> flux_prims(T g, T pinf, const V4<T>& W) {
    V4<T> F;
    T rho = W.v[0], u = W.v[1], v = W.v[2], p = W.v[3];
    T e = eos_e(g, pinf, p, rho);
    T E = rho * e + T(0.5) * rho * (u * u + v * v);
    F.v[0] = rho * u;
    F.v[1] = rho * u * u + p;
    F.v[2] = rho * u * v;
    F.v[3] = u * (E + p);
    return F;
}

// ---------------------------------------------------------------------------
// Exact Riemann solver for two stiffened gases (Toro ch. 4 with p -> p+pinf
// per side; the reference links it from exact_rs_stiffenedgas, call sites
// pure_RS_exact.cpp:29-60 and mixed_RS_exact.cpp:33-47).  The Newton
// iteration for p* runs in FP64 in every precision mode.
// ---------------------------------------------------------------------------
struct RSSide { double g, pinf, rho, u, p, c, pbar; };

__device__ __forceinline__ void rs_side(RSSide& s, double g, double pinf, double rho, double u, double p) {
    s.g = g; s.pinf = pinf; s.rho = rho; s.u = u; s.p = p;
    s.pbar = p + pinf;
    s.c = sqrt(g * s.pbar / rho);
}

__device__ __forceinline__ double rs_f(const RSSide& s, double p, double& fd) {
    double pbar = p + s.pinf;
    if (p > s.p) {
        double A = 2.0 / ((s.g + 1.0) * s.rho);
        double B = (s.g - 1.0) / (s.g + 1.0) * s.pbar;
        double q = sqrt(A / (pbar + B));
        fd = q * (1.0 - 0.5 * (p - s.p) / (B + pbar));
        return (p - s.p) * q;
    }
    double r = pbar / s.pbar;
    double pw = pow(r, (s.g - 1.0) / (2.0 * s.g));
    fd = pw / (s.rho * s.c * r);
    return 2.0 * s.c / (s.g - 1.0) * (pw - 1.0);
}

// Starting pressure (after Toro 9.5.2): PVRS when the pressures differ by less
// than a factor 2; otherwise the two-rarefaction estimate (exact for two
// rarefactions of one gas) when PVRS is below both, the two-shock estimate
// when not.  Barred pressures p + pinf throughout.
__device__ __forceinline__ double rs_guess(const RSSide& L, const RSSide& R, double pfloor) {
    const double du = R.u - L.u;
    const double ppv = 0.5 * (L.p + R.p) - 0.125 * du * (L.rho + R.rho) * (L.c + R.c);
    const double pmin = fmin(L.p, R.p);
    const double bmin = fmin(L.pbar, R.pbar), bmax = fmax(L.pbar, R.pbar);
    double p;
    if (bmax < 2.0 * bmin) {   // weak waves: PVRS is already close
        p = ppv;
    } else if (ppv < pmin && L.g == R.g && L.pinf == R.pinf) {
        const double z = (L.g - 1.0) / (2.0 * L.g);
        const double num = L.c + R.c - 0.5 * (L.g - 1.0) * du;
        const double den = L.c / pow(L.pbar, z) + R.c / pow(R.pbar, z);
        p = (num > 0.0) ? pow(num / den, 1.0 / z) - L.pinf : pfloor;
    } else {
        const double ps = fmax(pfloor, ppv);
        const double gL = sqrt(2.0 / ((L.g + 1.0) * L.rho) / (ps + L.pinf + (L.g - 1.0) / (L.g + 1.0) * L.pbar));
        const double gR = sqrt(2.0 / ((R.g + 1.0) * R.rho) / (ps + R.pinf + (R.g - 1.0) / (R.g + 1.0) * R.pbar));
        p = (gL * L.p + gR * R.p - du) / (gL + gR);
    }
    return p > pfloor ? p : pfloor;
}

struct Star { double p, u, rl, rr; };

// star state; returns status bits (ST_RS on non-convergence / vacuum / NaN)
// p0: Newton starting pressure, used when it lies above -min(pinf); otherwise
// the adaptive estimate.  tol: relative pressure change that ends the Newton
// iteration: 1e-6 (Toro's TOLPRE; ~1e-12 after the quadratic last step) for
// the face fluxes, 1e-12 for the interfacial star states.
__device__ __forceinline__ unsigned rs_star(const RSSide& L, const RSSide& R, Star& S, double p0 = -1e300,
                                            double tol = 1e-6) {
    if (L.rho == R.rho && L.u == R.u && L.p == R.p && L.g == R.g && L.pinf == R.pinf) {
        S.p = L.p; S.u = L.u; S.rl = L.rho; S.rr = R.rho;
        return 0;
    }
    unsigned st = 0;
    double du = R.u - L.u;
    double plow = -fmin(L.pinf, R.pinf);
    if (2.0 * L.c / (L.g - 1.0) + 2.0 * R.c / (R.g - 1.0) <= du) st |= ST_RS;
    const double pfloor = plow + 1e-6 * (0.5 * (L.p + R.p) - plow);
    double p = p0 > pfloor ? p0 : rs_guess(L, R, pfloor);
    bool conv = false;
    for (int it = 0; it < 60; ++it) {
        double dL, dR;
        double fL = rs_f(L, p, dL);
        double fR = rs_f(R, p, dR);
        double pn = p - (fL + fR + du) / (dL + dR);
        if (pn <= plow) pn = 0.5 * (p + plow);
        double chg = fabs(pn - p);
        double ref = 0.5 * (pn + p) - plow;
        p = pn;
        if (chg <= tol * ref) { conv = true; break; }
    }
    if (!conv || p != p) st |= ST_RS;
    double dL, dR;
    double fL = rs_f(L, p, dL);
    double fR = rs_f(R, p, dR);
    S.p = p;
    S.u = 0.5 * (L.u + R.u) + 0.5 * (fR - fL);
    {
        double ratio = (p + L.pinf) / L.pbar;
        if (p > L.p) { double g6 = (L.g - 1.0) / (L.g + 1.0); S.rl = L.rho * (ratio + g6) / (g6 * ratio + 1.0); }
        else S.rl = L.rho * pow(ratio, 1.0 / L.g);
    }
    {
        double ratio = (p + R.pinf) / R.pbar;
        if (p > R.p) { double g6 = (R.g - 1.0) / (R.g + 1.0); S.rr = R.rho * (ratio + g6) / (g6 * ratio + 1.0); }
        else S.rr = R.rho * pow(ratio, 1.0 / R.g);
    }
    return st;
}

// Toro 4.5 sampling at x/t = S (barred pressures); out = (rho, u_n, p)
__device__ __forceinline__ void rs_sample(const RSSide& L, const RSSide& R, const Star& st, double S,
                                          double& orho, double& ou, double& op) {
    if (S <= st.u) {
        const RSSide& s = L;
        if (st.p > s.p) {
            double SL = s.u - s.c * sqrt((s.g + 1.0) / (2.0 * s.g) * (st.p + s.pinf) / s.pbar + (s.g - 1.0) / (2.0 * s.g));
            if (S <= SL) { orho = s.rho; ou = s.u; op = s.p; }
            else { orho = st.rl; ou = st.u; op = st.p; }
        } else {
            double SHL = s.u - s.c;
            if (S <= SHL) { orho = s.rho; ou = s.u; op = s.p; }
            else {
                double cml = s.c * pow((st.p + s.pinf) / s.pbar, (s.g - 1.0) / (2.0 * s.g));
                double STL = st.u - cml;
                if (S > STL) { orho = st.rl; ou = st.u; op = st.p; }
                else {
                    double cf = 2.0 / (s.g + 1.0) * (s.c + 0.5 * (s.g - 1.0) * (s.u - S));
                    orho = s.rho * pow(cf / s.c, 2.0 / (s.g - 1.0));
                    ou = 2.0 / (s.g + 1.0) * (s.c + 0.5 * (s.g - 1.0) * s.u + S);
                    op = s.pbar * pow(cf / s.c, 2.0 * s.g / (s.g - 1.0)) - s.pinf;
                }
            }
        }
    } else {
        const RSSide& s = R;
        if (st.p > s.p) {
            double SR = s.u + s.c * sqrt((s.g + 1.0) / (2.0 * s.g) * (st.p + s.pinf) / s.pbar + (s.g - 1.0) / (2.0 * s.g));
            if (S >= SR) { orho = s.rho; ou = s.u; op = s.p; }
            else { orho = st.rr; ou = st.u; op = st.p; }
        } else {
            double SHR = s.u + s.c;
            if (S >= SHR) { orho = s.rho; ou = s.u; op = s.p; }
            else {
                double cmr = s.c * pow((st.p + s.pinf) / s.pbar, (s.g - 1.0) / (2.0 * s.g));
                double STR = st.u + cmr;
                if (S <= STR) { orho = st.rr; ou = st.u; op = st.p; }
                else {
                    double cf = 2.0 / (s.g + 1.0) * (s.c - 0.5 * (s.g - 1.0) * (s.u - S));
                    orho = s.rho * pow(cf / s.c, 2.0 / (s.g - 1.0));
                    ou = 2.0 / (s.g + 1.0) * (-s.c + 0.5 * (s.g - 1.0) * s.u + S);
                    op = s.pbar * pow(cf / s.c, 2.0 * s.g / (s.g - 1.0)) - s.pinf;
                }
            }
        }
    }
}

// sampling at x/t = 0 in the storage precision: FP64 for the plain path,
// FP32 in mixed precision (the star state itself comes from the FP64 Newton)
template <typename T>
__device__ __forceinline__ void rs_sample_t(const RSSide& L, const RSSide& R, const Star& st, T& orho, T& ou, T& op) {
    if (sizeof(T) == sizeof(double)) {
        double r, u, p;
        rs_sample(L, R, st, 0.0, r, u, p);
        orho = (T)r; ou = (T)u; op = (T)p;
        return;
    }
    const bool left = 0.0 <= st.u;
    RSSide s;
    s.g = left ? L.g : R.g; s.pinf = left ? L.pinf : R.pinf; s.rho = left ? L.rho : R.rho;
    s.u = left ? L.u : R.u; s.p = left ? L.p : R.p; s.c = left ? L.c : R.c; s.pbar = left ? L.pbar : R.pbar;
    const T sg = left ? T(1) : T(-1);   // mirror the right side onto the left formulas
    const T g = (T)s.g, pinf = (T)s.pinf, rho = (T)s.rho, c = (T)s.c, pbar = (T)s.pbar;
    const T uk = sg * (T)s.u, us = sg * (T)st.u, ps = (T)st.p;
    const T rs = left ? (T)st.rl : (T)st.rr;
    if (st.p > s.p) {
        const T Sk = uk - c * sqrt((g + T(1)) / (T(2) * g) * (ps + pinf) / pbar + (g - T(1)) / (T(2) * g));
        if (T(0) <= Sk) { orho = rho; ou = (T)s.u; op = (T)s.p; }
        else { orho = rs; ou = (T)st.u; op = ps; }
        return;
    }
    const T SH = uk - c;
    if (T(0) <= SH) { orho = rho; ou = (T)s.u; op = (T)s.p; return; }
    const T cm = c * pow((ps + pinf) / pbar, (g - T(1)) / (T(2) * g));
    const T ST = us - cm;
    if (T(0) > ST) { orho = rs; ou = (T)st.u; op = ps; return; }
    const T cf = T(2) / (g + T(1)) * (c + T(0.5) * (g - T(1)) * uk);
    orho = rho * pow(cf / c, T(2) / (g - T(1)));
    ou = sg * (T(2) / (g + T(1)) * (c + T(0.5) * (g - T(1)) * uk));
    op = pbar * pow(cf / c, T(2) * g / (g - T(1))) - pinf;
}

// flux from the sampled solution (pure_RS_exact.cpp:45-71): tangential
// velocity upwinded on u*, y-faces swap the momentum components
template <typename T>
__device__ __forceinline__ T tang(bool px, const V4<T>& W) { return px ? W.v[2] : W.v[1]; }
template <typename T>
__device__ __forceinline__ T norm_u(bool px, const V4<T>& W) { return px ? W.v[1] : W.v[2]; }

template <typename T>
__device__ __forceinline__ V4<T> rs_flux_out(bool px, T g, T pinf, T r, T u, T p, double ustar, const V4<T>& WL,
                                             const V4<T>& WR) {
    V4<T> soln;
    soln.v[0] = r;
    soln.v[1] = u;
    const T tr = tang(px, WR), tl = tang(px, WL);
    soln.v[2] = ustar <= 0.0 ? tr : tl;
    soln.v[3] = p;
    V4<T> F = flux_prims(g, pinf, soln);
    if (!px) { T tmp = F.v[1]; F.v[1] = F.v[2]; F.v[2] = tmp; }
    return F;
}

// identical normal states (rho, u_n, p): the exact solution is the state
// itself, bit for bit what the Newton iteration below returns
template <typename T>
__device__ __forceinline__ bool rs_trivial(bool px, const V4<T>& WL, const V4<T>& WR) {
    return WL.v[0] == WR.v[0] && norm_u(px, WL) == norm_u(px, WR) && WL.v[3] == WR.v[3];
}

template <typename T>
__device__ __forceinline__ V4<T> rs_flux_trivial(bool px, T g, T pinf, const V4<T>& WL, const V4<T>& WR) {
    const T un = norm_u(px, WL);
    return rs_flux_out<T>(px, g, pinf, WL.v[0], un, WL.v[3], (double)un, WL, WR);
}

// the general case: FP64 Newton for the star state, sampling at x/t = 0
template <typename T>
__device__ __forceinline__ V4<T> rs_flux_full(bool px, T g, T pinf, const V4<T>& WL, const V4<T>& WR, unsigned& st) {
    RSSide L, R;
    rs_side(L, (double)g, (double)pinf, (double)WL.v[0], (double)norm_u(px, WL), (double)WL.v[3]);
    rs_side(R, (double)g, (double)pinf, (double)WR.v[0], (double)norm_u(px, WR), (double)WR.v[3]);
    Star S;
    st |= rs_star(L, R, S);
    T r, u, p;
    rs_sample_t<T>(L, R, S, r, u, p);
    return rs_flux_out<T>(px, g, pinf, r, u, p, S.u, WL, WR);
}

// pure_RS_exact::solve_RP (pure_RS_exact.cpp:9-73): conserved L/R -> flux
template <typename T>
__device__ __forceinline__ V4<T> pure_rs_flux(bool px, T g, T pinf, const V4<T>& UL, const V4<T>& UR,
                                              unsigned& st) {
    if (!phys_state(UL) || !phys_state(UR)) st |= ST_UNPHYS;
    const V4<T> WL = cons2prims(g, pinf, UL);
    const V4<T> WR = cons2prims(g, pinf, UR);
    if (rs_trivial(px, WL, WR)) return rs_flux_trivial<T>(px, g, pinf, WL, WR);
    return rs_flux_full<T>(px, g, pinf, WL, WR, st);
}

// ---------------------------------------------------------------------------
// boundary conditions as a gather (euler_bc.hpp:5-103): every ghost cell is
// a copy of one interior cell with optional sign flips.  The right-hand
// reflective flip tests BC_L, as the reference does (euler_bc.hpp:51).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int src_col(const Inst& I, int j, bool& flip) {
    const int g = I.g, Nx = I.Nx;
    flip = false;
    if (I.bc[0] == BC_P && I.bc[1] == BC_P) { int r = (j - g) % Nx; if (r < 0) r += Nx; return r + g; }
    if (j < g) { flip = I.bc[0] == BC_R; return 2 * g - j - 1; }
    if (j >= Nx + g) { flip = I.bc[0] == BC_R; return Nx + g - 1 - (j - Nx - g); }
    return j;
}

__device__ __forceinline__ int src_row(const Inst& I, int i, bool& flip) {
    const int g = I.g, Ny = I.Ny;
    flip = false;
    if (I.bc[2] == BC_P && I.bc[3] == BC_P) { int r = (i - g) % Ny; if (r < 0) r += Ny; return r + g; }
    if (i < g) { flip = I.bc[2] == BC_R; return 2 * g - i - 1; }
    if (i >= Ny + g) { flip = I.bc[3] == BC_R; return Ny + g - 1 - (i - Ny - g); }
    return i;
}

__device__ __forceinline__ bool interior(const Inst& I, int i, int j) {
    return i >= I.g && i < I.g + I.Ny && j >= I.g && j < I.g + I.Nx;
}

}  // namespace gfm
