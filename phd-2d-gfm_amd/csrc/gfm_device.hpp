// gfm_device.hpp -- device-side types and numerics for the GFM hot path on
// gfx950: instance descriptors, stiffened-gas EOS, the exact Riemann solver,
// the boundary-condition gather map and the level-set helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "gfm_fastmath.hpp"

namespace gfm {

enum { BC_T = 0, BC_R = 1, BC_P = 2 };
enum { V_ORIGINAL = 0, V_REAL = 1, V_RIEMANN = 2, V_MODIFIED = 3, V_VOF = 4, V_OGFM_VOF = 5 };

// double scalars per instance
// SD_HOLD_LAST / SD_HOLD_IVL: an instance with SI_HOLD set takes no step while
// t - last > interval (the output test of sim_twofluid.cpp:79), until the
// host moves `last` forward
enum { SD_T = 0, SD_DT = 1, SD_MASS1 = 2, SD_MASS2 = 3, SD_HOLD_LAST = 4, SD_HOLD_IVL = 5, SD_N = 8 };
// int scalars per instance
enum {
    SI_NSTEP = 0, SI_ACTIVE = 1, SI_NIFACE = 2, SI_DTCOUNT = 3,
    SI_BBOX = 4,   // 8 ints: fluid1 imin,imax,jmin,jmax ; fluid2 ...
    SI_STATUS = 12, SI_ROUNDS = 13,
    SI_NLIST = 14,   // entries of the last interfacial list recorded in imap
    SI_HOLD = 15,
    SI_N = 16
};
enum { ST_UNPHYS = 1, ST_RS = 2, ST_IOVF = 4, ST_IROUNDS = 8, ST_ILIST = 16 };
// interfacial entries whose flags the ordered solve holds in LDS (1 byte each)
constexpr int ICAP_LDS = 49152;

// Per-fluid constants of the mixed-precision face Riemann solver, set on the
// host at commit: z = (g-1)/(2g), g6 = (g-1)/(g+1), a2 = 2/(g+1), ig = 1/g,
// k2 = 2/(g-1), bc[k-1] = C(z, k) for k = 1..9; FP32 copies for the
// sampling: sk = (g+1)/(2g), z, 1/g, 2/(g+1), 2/(g-1), 2g/(g-1), g6,
// g pinf, 1/(g-1).
struct RSFluid {
    double g, pinf, z, g6, a2, ig, k2;
    double bc[9];
    float fsk, fz, fig, fa2, fk2, fk3, fg6, fgpinf, figm1;
    float fg, fpinf, fiz;
};

__host__ __device__ inline void rs_fluid(RSFluid& F, double g, double pinf) {
    F.g = g; F.pinf = pinf;
    F.z = (g - 1.0) / (2.0 * g);
    F.g6 = (g - 1.0) / (g + 1.0);
    F.a2 = 2.0 / (g + 1.0);
    F.ig = 1.0 / g;
    F.k2 = 2.0 / (g - 1.0);
    double c = F.z;
    F.bc[0] = c;
    for (int k = 2; k <= 9; ++k) {
        c = c * (F.z - (double)(k - 1)) / (double)k;
        F.bc[k - 1] = c;
    }
    F.fsk = (float)((g + 1.0) / (2.0 * g));
    F.fz = (float)F.z;
    F.fig = (float)F.ig;
    F.fa2 = (float)F.a2;
    F.fk2 = (float)F.k2;
    F.fk3 = (float)(2.0 * g / (g - 1.0));
    F.fg6 = (float)F.g6;
    const float gf = (float)g, pf = (float)pinf;
    F.fgpinf = gf * pf;
    F.figm1 = 1.0f / (gf - 1.0f);
    F.fg = gf;
    F.fpinf = pf;
    F.fiz = (float)(1.0 / F.z);
}

// One simulation.  Plane arrays are row-major with row pitch `pitch`
// elements; a 4-component field is 4 consecutive planes of `plane` elements.
struct Inst {
    int Nx, Ny, g, W, H, pitch;
    long long plane;
    double x0, y0, dx, dy;
    int bc[4];                 // L, R, B, T
    double gam[2], pinf[2];
    RSFluid rsf[2];
    double cfl, T;
    int max_steps;
    int ntile_dt;
    int icap;                  // interfacial list capacity
    int icap_lds;              // lists up to this long keep the ordered solve's flags in LDS
    int use_bbox;              // OGFM narrows the per-fluid update boxes
    int use_slip;              // GFM_base::use_slip_BCs (GFM_modified: the ghost keeps its tangential velocity)
    double band;               // FP32 whole steps: ghost states farther than this (phi) from their
                               // fluid's real region are not updated by the flux kernel (0: all are)
    void* Q[2][2];             // conserved [fluid][buffer]
    void* P[2][2];             // primitives [fluid][buffer]
    void* phi;
    void* phiA;
    void* phiB;
    void* vel;                 // 2 planes
    int* imap;                 // interfacial list index per cell, -1 if none
    int* ilist;                // row-major linear index (i*W+j) per entry
    double* side;              // 8 doubles per interfacial entry (VSIDE per mixed cell for the VOF GFMs)
    double* partials;          // 4 doubles per dt tile
    double* sd;                // SD_* scalars
    int* si;                   // SI_* scalars
    double* mass_hist;         // (t, mass1, mass2) per step, hist_cap entries
    int hist_cap;
    void* irec;                // interfacial records, icap entries
    int* idone;                // 1 when an entry's star state is final after the parallel pass
    int* iseg;                 // interfacial cells per 256-cell row segment, then their offsets
    int nseg;                  // row segments (Ny x ceil(Nx / 256))
    int ntr;                   // passive tracers (0..TR_MAX)
    int tr_keep_y;             // tracers return to their initial y after each step
    double* trk;               // TR_HDR doubles (positions, initial positions), then
                               // hist_cap entries of 1 + 4 ntr doubles: t, (x, y, u, v) per tracer
    double* zv;                // CLSVOF tracker: 3 FP64 planes (z, first sweep, flags); null otherwise
};

constexpr int TR_MAX = 4;
constexpr int TR_HDR = 4 * TR_MAX;

// HOUC first-derivative stencil for positive velocity: offsets o0..o0+n-1
struct Houc { double c[10]; int n, o0; };

__device__ __forceinline__ long long cidx(const Inst& I, int i, int j) {
    return (long long)i * I.pitch + j;
}

// ---------------------------------------------------------------------------
// stiffened-gas EOS (closed forms; the reference takes them from PhD-Common
// stiffened_gas_eos.hpp, used at euler_misc.hpp:46,59,77,156)
// ---------------------------------------------------------------------------
// Conversions round every operation on its own (no FMA contraction), as the
// reference's compiled C++ does: each kernel that converts the same state then
// gets the same bits, whatever the surrounding code lets the compiler fuse.
template <typename T>
__device__ __forceinline__ T eos_p(T g, T pinf, T e, T rho) {
#pragma clang fp contract(off)
    return (g - T(1)) * rho * e - g * pinf;
}
template <typename T>
__device__ __forceinline__ T eos_e(T g, T pinf, T p, T rho) {
#pragma clang fp contract(off)
    return (p + g * pinf) / ((g - T(1)) * rho);
}
template <typename T>
__device__ __forceinline__ T eos_c(T g, T pinf, T p, T rho) { return sqrt(g * (p + pinf) / rho); }

template <typename T>
struct V4 { T v[4]; };

// euler_misc.hpp:52-67
template <typename T>
__device__ __forceinline__ V4<T> cons2prims(T g, T pinf, const V4<T>& U) {
#pragma clang fp contract(off)
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
#pragma clang fp contract(off)
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
// FP32 storage (mixed precision): the conversions and checks above with one
// reciprocal per state in place of a division per component, rho e =
// (p + g pinf) / (g - 1) from a per-fluid reciprocal, and the positivity
// tests in division-free form: for rho > 0, e >= 0 iff p + g pinf >= 0 (W)
// and iff 2 E rho >= |m|^2 (U); both are false for rho <= 0 or NaN.  No FMA
// contraction, so that every kernel converting the same state gets the same
// bits.
// ---------------------------------------------------------------------------
struct FluidF { float g, pinf, gpinf, igm1; };

__device__ __forceinline__ FluidF fluid_f(const RSFluid& R) {
    FluidF F;
    F.g = (float)R.g; F.pinf = (float)R.pinf;
    F.gpinf = R.fgpinf;
    F.igm1 = R.figm1;
    return F;
}

__device__ __forceinline__ V4<float> cons2prims_f(const FluidF& F, const V4<float>& U) {
#pragma clang fp contract(off)
    V4<float> W;
    const float rho = U.v[0], ir = __builtin_amdgcn_rcpf(rho);   // 1 ulp
    const float u = U.v[1] * ir, v = U.v[2] * ir;
    W.v[0] = rho; W.v[1] = u; W.v[2] = v;
    W.v[3] = (F.g - 1.0f) * (U.v[3] - 0.5f * (U.v[1] * u + U.v[2] * v)) - F.gpinf;
    return W;
}

__device__ __forceinline__ V4<float> prims2cons_f(const FluidF& F, const V4<float>& W) {
#pragma clang fp contract(off)
    V4<float> U;
    const float rho = W.v[0], u = W.v[1], v = W.v[2];
    U.v[0] = rho; U.v[1] = rho * u; U.v[2] = rho * v;
    U.v[3] = (W.v[3] + F.gpinf) * F.igm1 + 0.5f * rho * (u * u + v * v);
    return U;
}

__device__ __forceinline__ bool phys_prims_f(const FluidF& F, const V4<float>& W) {
    return W.v[0] > 0.0f && W.v[3] + F.gpinf >= 0.0f;
}

__device__ __forceinline__ bool phys_state_f(const V4<float>& U) {
    return U.v[0] > 0.0f && 2.0f * U.v[3] * U.v[0] >= U.v[1] * U.v[1] + U.v[2] * U.v[2];
}

__device__ __forceinline__ V4<float> flux_prims_f(const FluidF& F, const V4<float>& W) {
#pragma clang fp contract(off)
    V4<float> Fl;
    const float rho = W.v[0], u = W.v[1], v = W.v[2], p = W.v[3];
    const float E = (p + F.gpinf) * F.igm1 + 0.5f * rho * (u * u + v * v);
    Fl.v[0] = rho * u;
    Fl.v[1] = rho * u * u + p;
    Fl.v[2] = rho * u * v;
    Fl.v[3] = u * (E + p);
    return Fl;
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

// x^y: the C library's pow, or pow_fast (FP64, a few ulp) on the
// mixed-precision path
template <bool FAST>
__device__ __forceinline__ double rpow(double x, double y) {
    if constexpr (FAST) return pow_fast(x, y);
    else return pow(x, y);
}

template <bool FAST = false>
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
    double pw = rpow<FAST>(r, (s.g - 1.0) / (2.0 * s.g));
    fd = pw / (s.rho * s.c * r);
    return 2.0 * s.c / (s.g - 1.0) * (pw - 1.0);
}

// Starting pressure (after Toro 9.5.2): PVRS when the pressures differ by less
// than a factor 2; otherwise the two-rarefaction estimate (exact for two
// rarefactions of one gas) when PVRS is below both, the two-shock estimate
// when not.  Barred pressures p + pinf throughout.
template <bool FAST = false>
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
        const double den = L.c / rpow<FAST>(L.pbar, z) + R.c / rpow<FAST>(R.pbar, z);
        p = (num > 0.0) ? rpow<FAST>(num / den, 1.0 / z) - L.pinf : pfloor;
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
template <bool FAST = false>
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
    double p = p0 > pfloor ? p0 : rs_guess<FAST>(L, R, pfloor);
    bool conv = false;
    for (int it = 0; it < 60; ++it) {
        double dL, dR;
        double fL = rs_f<FAST>(L, p, dL);
        double fR = rs_f<FAST>(R, p, dR);
        double pn = p - (fL + fR + du) / (dL + dR);
        if (pn <= plow) pn = 0.5 * (p + plow);
        double chg = fabs(pn - p);
        double ref = 0.5 * (pn + p) - plow;
        p = pn;
        if (chg <= tol * ref) { conv = true; break; }
    }
    if (!conv || p != p) st |= ST_RS;
    double dL, dR;
    double fL = rs_f<FAST>(L, p, dL);
    double fR = rs_f<FAST>(R, p, dR);
    S.p = p;
    S.u = 0.5 * (L.u + R.u) + 0.5 * (fR - fL);
    {
        double ratio = (p + L.pinf) / L.pbar;
        if (p > L.p) { double g6 = (L.g - 1.0) / (L.g + 1.0); S.rl = L.rho * (ratio + g6) / (g6 * ratio + 1.0); }
        else S.rl = L.rho * rpow<FAST>(ratio, 1.0 / L.g);
    }
    {
        double ratio = (p + R.pinf) / R.pbar;
        if (p > R.p) { double g6 = (R.g - 1.0) / (R.g + 1.0); S.rr = R.rho * (ratio + g6) / (g6 * ratio + 1.0); }
        else S.rr = R.rho * rpow<FAST>(ratio, 1.0 / R.g);
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

// ---------------------------------------------------------------------------
// FP64 pieces of the mixed-precision solvers (rs_star_m2).  The star
// pressure and velocity come from rs_star's Newton iteration, in FP64,
// on the same pressure functions (Toro 4.6-4.7).  What differs is the
// evaluation: the fluid's constants are computed once per instance,
// (pbar / pbarK)^z - 1 is summed as its binomial series when
// |pbar / pbarK - 1| < 1/64 (no cancellation) and from pow_fast otherwise,
// and the rarefaction derivative uses z k r^(z-1) / pbarK = (c / g) r^z / pbar.
// ---------------------------------------------------------------------------
// FP32 hardware reciprocal, square root (about 1 ulp) and power
__device__ __forceinline__ float rcp_hw(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float sqrt_hw(float x) { return __builtin_amdgcn_sqrtf(x); }

// x^y = exp2(y log2 x), x >= 0: relative error about (2 + |y log2 x|) ulp
__device__ __forceinline__ float pow_hw(float x, float y) {
    return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}

// one side: rho, u_n, p, pbar = p + pinf, 1 / pbar, 1 / rho, c; the
// shock and rarefaction constants A = a2 / rho, B = g6 pbar, k = k2 c and
// c / g are formed where used
struct RSSideM { double rho, u, p, pbar, ipb, irho, c; };

// FP64 reciprocal and square root for the mixed-precision Newton iteration:
// the hardware estimates refined by two Newton-Raphson steps (a few ulp from
// the correctly rounded results, against one ulp-exact but several times
// longer IEEE sequences), for normal arguments (|x| >= 2^-1022; the pressures,
// densities and sound speeds here are far inside that range).  Zero,
// infinite and NaN arguments give the IEEE results (1/0 = inf, 1/inf = 0,
// sqrt(-0) = -0, sqrt(x < 0) = NaN) by a select, not by an IEEE fallback: the
// compiler if-converts a fallback branch (these are hoisted out of loops),
// which would then run the whole IEEE sequence for every argument.
__device__ __forceinline__ double rcp_nr(double x) {
    const double r0 = __builtin_amdgcn_rcp(x);         // exact for 0, inf, NaN
    double r = fma(fma(-x, r0, 1.0), r0, r0);
    r = fma(fma(-x, r, 1.0), r, r);
    return (fabs(x) < INFINITY && x != 0.0) ? r : r0;
}

__device__ __forceinline__ double sqrt_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);                // ~ 1/sqrt(x)
    y = y * fma(-0.5 * x * y, y, 1.5);
    y = y * fma(-0.5 * x * y, y, 1.5);
    const double s = x * y;                            // ~ sqrt(x)
    const double r = fma(fma(-s, s, x), 0.5 * y, s);   // one Newton step on s
    return (x > 0.0 && x < INFINITY) ? r : (x >= 0.0 ? x : __builtin_nan(""));
}

// 1 / sqrt(x) for normal x > 0: the hardware estimate refined by two
// Newton-Raphson steps (a few ulp); 0, inf, x < 0 and NaN give the
// hardware estimate, which is the IEEE result there, by a select (rcp_nr)
__device__ __forceinline__ double rsq_nr(double x) {
    const double y0 = __builtin_amdgcn_rsq(x);
    double y = y0 * fma(-0.5 * x * y0, y0, 1.5);
    y = y * fma(-0.5 * x * y, y, 1.5);
    return (x > 0.0 && x < INFINITY) ? y : y0;         // y0 exact for 0, inf, x < 0, NaN
}

__device__ __forceinline__ void rs_side_m(RSSideM& s, const RSFluid& F, double rho, double u, double p) {
    s.rho = rho; s.u = u; s.p = p;
    s.pbar = p + F.pinf;
    s.ipb = rcp_nr(s.pbar);
    s.irho = rcp_nr(rho);
    s.c = sqrt_nr(F.g * s.pbar * s.irho);
}

// (1 + eps)^z - 1, r = 1 + eps: the binomial series for |eps| < 1/64, else
// pow_fast
__device__ __forceinline__ double powm1_m(const RSFluid& F, double eps, double r) {
    if (fabs(eps) < 0.015625) {
        double q = F.bc[8];
#pragma unroll
        for (int k = 7; k >= 0; --k) q = fma(q, eps, F.bc[k]);
        return q * eps;
    }
    return pow_fast(r, F.z) - 1.0;
}

__device__ __forceinline__ double rs_f_m(const RSFluid& F, const RSSideM& s, double p, double& fd) {
    const double pbar = p + F.pinf;
    if (p > s.p) {   // shock (Toro 4.6)
        const double t = rcp_nr(pbar + F.g6 * s.pbar);
        const double q = sqrt_nr(F.a2 * s.irho * t);
        fd = q * (1.0 - 0.5 * (p - s.p) * t);
        return (p - s.p) * q;
    }
    // rarefaction (Toro 4.7): f = k ((pbar/pbarK)^z - 1)
    const double w = powm1_m(F, (p - s.p) * s.ipb, pbar * s.ipb);
    fd = s.c * F.ig * (1.0 + w) * rcp_nr(pbar);
    return F.k2 * s.c * w;
}

// rs_star's FP64 Newton iteration for two stiffened gases on the
// mixed-precision evaluation: per-side constants, Newton start p0 when it
// lies above the floor (else rs_guess's choice), tolerance tol, and with
// DENS the star densities (Toro 4.5.1, 4.5.2: shock rho (r + g6) / (g6 r +
// 1), rarefaction rho r^(1/g), r = (p* + pinf) / (p + pinf)).  Only the
// evaluation differs from rs_star.  Used for the interfacial problems (the
// faces use rs_star_face_m below).
template <bool DENS = true>
__device__ __forceinline__ unsigned rs_star_m2(const RSFluid& FL, const RSFluid& FR, const RSSideM& L,
                                               const RSSideM& R, Star& S, double p0, double tol) {
    const bool same = FL.g == FR.g && FL.pinf == FR.pinf;
    if (same && L.rho == R.rho && L.u == R.u && L.p == R.p) {
        S.p = L.p; S.u = L.u; S.rl = L.rho; S.rr = R.rho;
        return 0;
    }
    unsigned st = 0;
    const double du = R.u - L.u;
    const double plow = -fmin(FL.pinf, FR.pinf);
    if (FL.k2 * L.c + FR.k2 * R.c <= du) st |= ST_RS;
    const double pfloor = plow + 1e-6 * (0.5 * (L.p + R.p) - plow);
    double p;
    if (p0 > pfloor) {
        p = p0;
    } else {
        const double ppv = 0.5 * (L.p + R.p) - 0.125 * du * (L.rho + R.rho) * (L.c + R.c);
        const double bmin = fmin(L.pbar, R.pbar), bmax = fmax(L.pbar, R.pbar);
        if (bmax < 2.0 * bmin) {
            p = ppv;
        } else if (ppv < fmin(L.p, R.p) && same) {
            const double num = L.c + R.c - 0.5 * (FL.g - 1.0) * du;
            const double den = L.c / pow_fast(L.pbar, FL.z) + R.c / pow_fast(R.pbar, FL.z);
            p = (num > 0.0) ? pow_fast(num / den, 1.0 / FL.z) - FL.pinf : pfloor;
        } else {
            const double ps = fmax(pfloor, ppv);
            const double gL = sqrt(FL.a2 * L.irho / (ps + FL.pinf + FL.g6 * L.pbar));
            const double gR = sqrt(FR.a2 * R.irho / (ps + FR.pinf + FR.g6 * R.pbar));
            p = (gL * L.p + gR * R.p - du) / (gL + gR);
        }
        p = p > pfloor ? p : pfloor;
    }
    bool conv = false;
    double fL = 0.0, fR = 0.0, dL = 0.0, dR = 0.0, dp = 0.0;
    for (int it = 0; it < 60; ++it) {
        fL = rs_f_m(FL, L, p, dL);
        fR = rs_f_m(FR, R, p, dR);
        double pn = p - (fL + fR + du) * rcp_nr(dL + dR);
        if (pn <= plow) pn = 0.5 * (p + plow);
        const double chg = fabs(pn - p);
        const double ref = 0.5 * (pn + p) - plow;
        dp = pn - p;
        p = pn;
        if (chg <= tol * ref) { conv = true; break; }
    }
    if (!conv || p != p) st |= ST_RS;
    S.p = p;
    S.u = 0.5 * (L.u + R.u) + 0.5 * ((fR + dR * dp) - (fL + dL * dp));
    if (DENS) {
        {
            const double r = (p + FL.pinf) * L.ipb;
            S.rl = p > L.p ? L.rho * (r + FL.g6) * rcp_nr(FL.g6 * r + 1.0) : L.rho * pow_fast(r, FL.ig);
        }
        {
            const double r = (p + FR.pinf) * R.ipb;
            S.rr = p > R.p ? R.rho * (r + FR.g6) * rcp_nr(FR.g6 * r + 1.0) : R.rho * pow_fast(r, FR.ig);
        }
    }
    return st;
}

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

// FP32 storage: the same flux with flux_prims_f
__device__ __forceinline__ V4<float> rs_flux_out_f(bool px, const FluidF& Ff, float r, float u, float p, double ustar,
                                                   const V4<float>& WL, const V4<float>& WR) {
    V4<float> soln;
    soln.v[0] = r;
    soln.v[1] = u;
    const float tr = tang(px, WR), tl = tang(px, WL);
    soln.v[2] = ustar <= 0.0 ? tr : tl;
    soln.v[3] = p;
    V4<float> F = flux_prims_f(Ff, soln);
    if (!px) { float tmp = F.v[1]; F.v[1] = F.v[2]; F.v[2] = tmp; }
    return F;
}

__device__ __forceinline__ V4<float> rs_flux_trivial_f(bool px, const FluidF& Ff, const V4<float>& WL,
                                                       const V4<float>& WR) {
    const float un = norm_u(px, WL);
    return rs_flux_out_f(px, Ff, WL.v[0], un, WL.v[3], (double)un, WL, WR);
}

// the general case, FP64 storage: rs_star, then Toro's sampling at x/t = 0
__device__ __forceinline__ V4<double> rs_flux_full(bool px, double g, double pinf, const V4<double>& WL,
                                                   const V4<double>& WR, unsigned& st) {
    RSSide L, R;
    rs_side(L, g, pinf, WL.v[0], norm_u(px, WL), WL.v[3]);
    rs_side(R, g, pinf, WR.v[0], norm_u(px, WR), WR.v[3]);
    Star S;
    st |= rs_star(L, R, S);
    double r, u, p;
    rs_sample(L, R, S, 0.0, r, u, p);
    return rs_flux_out<double>(px, g, pinf, r, u, p, S.u, WL, WR);
}

// ---------------------------------------------------------------------------
// FP32 pre-iteration of the mixed-precision face solver: rs_star's iteration (Toro
// 4.6-4.7 pressure functions, guess, floor, tolerance 1e-6 on the relative
// pressure change, status bits) evaluated in FP32 with the hardware
// reciprocal, square root, exp2 and log2 (about 1 ulp each).  The
// rarefaction function k ((pbar / pbarK)^z - 1) is formed without
// cancellation (powm1_f).  The iteration also stops once it is at FP32
// resolution: a step of at most 4 ulp of p (matters where p + pinf << pinf),
// or a residual fL + fR + du within 4 ulp of its terms (strong waves, where
// the rounding of the terms moves the Newton step by more than 1e-6 p).
// ---------------------------------------------------------------------------
// (1 + eps)^z - 1 for eps > -1 and |z| < 1, with r = 1 + eps as computed by
// the caller (used only where |eps| > 1/4): log(1 + eps) = 2 atanh(s),
// s = eps / (2 + eps), summed to s^9 for |eps| <= 1/4 (|s| <= 1/7, truncation
// < 1e-9 relative), else ln2 log2(r); then expm1(y), y = z log(1 + eps), by
// its Taylor series to y^7 for |y| <= 1/4 (truncation < 2e-9 relative), else
// exp2(y / ln2) - 1 (|result| > 0.22).  About 3 ulp relative.
__device__ __forceinline__ float powm1_f(float eps, float r, float z) {
    float L;
    if (fabsf(eps) <= 0.25f) {
        const float s = eps * rcp_hw(2.0f + eps);
        const float s2 = s * s;
        float q = 1.0f / 9.0f;
        q = fmaf(q, s2, 1.0f / 7.0f);
        q = fmaf(q, s2, 1.0f / 5.0f);
        q = fmaf(q, s2, 1.0f / 3.0f);
        L = 2.0f * fmaf(s * s2, q, s);
    } else {
        L = 0.693147180559945f * __builtin_amdgcn_logf(r);
    }
    const float y = z * L;
    if (fabsf(y) <= 0.25f) {
        float e = 1.0f / 5040.0f;
        e = fmaf(e, y, 1.0f / 720.0f);
        e = fmaf(e, y, 1.0f / 120.0f);
        e = fmaf(e, y, 1.0f / 24.0f);
        e = fmaf(e, y, 1.0f / 6.0f);
        e = fmaf(e, y, 0.5f);
        e = fmaf(e, y, 1.0f);
        return y * e;
    }
    return __builtin_amdgcn_exp2f(y * 1.44269504088896f) - 1.0f;
}

struct RSSideF { float rho, u, p, pbar, ipb, irho, c; };

__device__ __forceinline__ void rs_side_f(RSSideF& s, const RSFluid& F, float rho, float u, float p) {
    s.rho = rho; s.u = u; s.p = p;
    s.pbar = p + F.fpinf;
    s.ipb = rcp_hw(s.pbar);
    s.irho = rcp_hw(rho);
    s.c = sqrt_hw(F.fg * s.pbar * s.irho);
}

__device__ __forceinline__ float rs_f_f(const RSFluid& F, const RSSideF& s, float p, float& fd) {
    const float pbar = p + F.fpinf;
    const float dp = p - s.p;
    if (p > s.p) {   // shock (Toro 4.6)
        const float t = rcp_hw(pbar + F.fg6 * s.pbar);
        const float q = sqrt_hw(F.fa2 * s.irho * t);
        fd = q * (1.0f - 0.5f * dp * t);
        return dp * q;
    }
    // rarefaction (Toro 4.7): f = k ((pbar / pbarK)^z - 1), f' = (c / g) (pbar / pbarK)^z / pbar
    const float w = powm1_f(dp * s.ipb, pbar * s.ipb, F.fz);
    fd = s.c * F.fig * (1.0f + w) * rcp_hw(pbar);
    return F.fk2 * s.c * w;
}

struct StarF { float p, u; };

__device__ __forceinline__ unsigned rs_star_f(const RSFluid& F, const RSSideF& L, const RSSideF& R, StarF& S,
                                              float tol = 1e-6f) {
    if (L.rho == R.rho && L.u == R.u && L.p == R.p) {
        S.p = L.p; S.u = L.u;
        return 0;
    }
    unsigned st = 0;
    const float du = R.u - L.u;
    const float plow = -F.fpinf;
    if (F.fk2 * (L.c + R.c) <= du) st |= ST_RS;
    const float pfloor = plow + 1e-6f * (0.5f * (L.p + R.p) - plow);
    float p;
    {
        const float ppv = 0.5f * (L.p + R.p) - 0.125f * du * (L.rho + R.rho) * (L.c + R.c);
        const float bmin = fminf(L.pbar, R.pbar), bmax = fmaxf(L.pbar, R.pbar);
        if (bmax < 2.0f * bmin) {
            p = ppv;
        } else if (ppv < fminf(L.p, R.p)) {   // two rarefactions: exact for one gas
            const float num = L.c + R.c - 0.5f * (F.fg - 1.0f) * du;
            const float den = L.c * rcp_hw(pow_hw(L.pbar, F.fz)) + R.c * rcp_hw(pow_hw(R.pbar, F.fz));
            p = (num > 0.0f) ? pow_hw(num * rcp_hw(den), F.fiz) - F.fpinf : pfloor;
        } else {                              // two shocks
            const float ps = fmaxf(pfloor, ppv);
            const float gL = sqrt_hw(F.fa2 * L.irho * rcp_hw(ps + F.fpinf + F.fg6 * L.pbar));
            const float gR = sqrt_hw(F.fa2 * R.irho * rcp_hw(ps + F.fpinf + F.fg6 * R.pbar));
            p = (gL * L.p + gR * R.p - du) * rcp_hw(gL + gR);
        }
        p = p > pfloor ? p : pfloor;
    }
    bool conv = false;
    float fL = 0.0f, fR = 0.0f, dL = 0.0f, dR = 0.0f, dp = 0.0f;
    for (int it = 0; it < 60; ++it) {
        fL = rs_f_f(F, L, p, dL);
        fR = rs_f_f(F, R, p, dR);
        const float res = fL + fR + du;
        float pn = p - res * rcp_hw(dL + dR);
        if (pn <= plow) pn = 0.5f * (p + plow);
        const float chg = fabsf(pn - p);
        const float ref = 0.5f * (pn + p) - plow;
        dp = pn - p;
        p = pn;
        if (chg <= tol * ref || chg <= 4.8e-7f * fabsf(p) ||
            fabsf(res) <= 4.8e-7f * (fabsf(fL) + fabsf(fR) + fabsf(du))) {
            conv = true;
            break;
        }
    }
    if (!conv || p != p) st |= ST_RS;
    // u* from the pressure functions at the converged p, to first order in
    // the last Newton step
    S.p = p;
    S.u = 0.5f * (L.u + R.u) + 0.5f * ((fR + dR * dp) - (fL + dL * dp));
    return st;
}

// The faces' FP64 pressure function of one side (Toro 4.6-4.7, evaluated as
// rs_f_m) with only what its value needs: shock
// f = (p - pK) / sqrt(rho (pbar + g6 pbarK) / a2), rarefaction
// f = k2 c ((pbar / pbarK)^z - 1) with c = g pbarK / sqrt(g pbarK rho).  The
// derivative uses the hardware FP64 reciprocal estimate: f' only scales the
// Newton step, so a relative error e in it leaves about e times the step as
// error after the step; the fixed point is that of the FP64 values and the
// stopping test is on the FP64 step.
__device__ __forceinline__ double rs_f_face(const RSFluid& F, double rho, double pK, double p, double& fd) {
    const double pbK = pK + F.pinf, pbar = p + F.pinf, dp = p - pK;
    if (p > pK) {
        const double y = pbar + F.g6 * pbK;
        const double q = rsq_nr(y * rho * (0.5 * F.g + 0.5));
        fd = q * (1.0 - 0.5 * dp * __builtin_amdgcn_rcp(y));
        return dp * q;
    }
    const double ipb = rcp_nr(pbK);
    const double w = powm1_m(F, dp * ipb, pbar * ipb);
    const double gp = F.g * pbK;
    const double c = gp * rsq_nr(gp * rho);
    fd = c * F.ig * (1.0 + w) * __builtin_amdgcn_rcp(pbar);
    return F.k2 * c * w;
}

// rs_star's FP64 Newton iteration for a face (one fluid) from the FP32
// iteration's p*, p0: rs_star_m2's floor, iteration, tolerance and status,
// on the pressure functions of rs_f_face, with the vacuum test on the FP32
// sound speeds.  A start at or below the floor (the FP32 iteration ended
// there, near vacuum) or NaN is replaced by the floor: the pressure function
// is increasing and concave, so the iteration converges to the same p* from
// any start above -pinf, and from the floor within a few steps.
__device__ __forceinline__ unsigned rs_star_face_m(const RSFluid& F, const RSSideF& Lf, const RSSideF& Rf, float p0,
                                                   Star& S, double tol) {
    const double rl = Lf.rho, ul = Lf.u, pl = Lf.p, rr = Rf.rho, ur = Rf.u, pr = Rf.p;
    const double du = ur - ul;
    const double plow = -F.pinf;
    const double pfloor = plow + 1e-6 * (0.5 * (pl + pr) - plow);
    unsigned st = ((double)(F.fk2 * (Lf.c + Rf.c)) <= du) ? ST_RS : 0u;
    double p = ((double)p0 > pfloor) ? (double)p0 : pfloor;
    bool conv = false;
    double fL = 0.0, fR = 0.0, dL = 0.0, dR = 0.0, dp = 0.0;
    for (int it = 0; it < 60; ++it) {
        fL = rs_f_face(F, rl, pl, p, dL);
        fR = rs_f_face(F, rr, pr, p, dR);
        double pn = p - (fL + fR + du) * __builtin_amdgcn_rcp(dL + dR);
        if (pn <= plow) pn = 0.5 * (p + plow);
        const double chg = fabs(pn - p);
        const double ref = 0.5 * (pn + p) - plow;
        dp = pn - p;
        p = pn;
        if (chg <= tol * ref) { conv = true; break; }
    }
    if (!conv || p != p) st |= ST_RS;
    S.p = p;
    S.u = 0.5 * (ul + ur) + 0.5 * ((fR + dR * dp) - (fL + dL * dp));
    return st;
}

// Toro 4.5 sampling at x/t = 0, the side mirrored onto the left-hand formulas
__device__ __forceinline__ void rs_sample_f(const RSFluid& F, const RSSideF& L, const RSSideF& R, const StarF& st,
                                            float& orho, float& ou, float& op) {
    typedef float T;
    const bool left = T(0) <= st.u;
    const T rho = left ? L.rho : R.rho, su = left ? L.u : R.u, sp = left ? L.p : R.p;
    const T c = left ? L.c : R.c, sipb = left ? L.ipb : R.ipb, spbar = left ? L.pbar : R.pbar;
    const T sg = left ? T(1) : T(-1);
    const T uk = sg * su, us = sg * st.u, ps = st.p;
    const T ratio = (ps + F.fpinf) * sipb;
    if (ps > sp) {
        const T Sk = uk - c * sqrt_hw(F.fsk * ratio + F.fz);
        if (T(0) <= Sk) { orho = rho; ou = su; op = sp; }
        else {
            orho = rho * (ratio + F.fg6) * rcp_hw(F.fg6 * ratio + T(1));
            ou = st.u; op = ps;
        }
        return;
    }
    const T SH = uk - c;
    if (T(0) <= SH) { orho = rho; ou = su; op = sp; return; }
    const T cm = c * pow_hw(ratio, F.fz);
    const T ST = us - cm;
    if (T(0) > ST) { orho = rho * pow_hw(ratio, F.fig); ou = st.u; op = ps; return; }
    const T cf = F.fa2 * (c + T(0.5) * (F.fg - T(1)) * uk);
    const T q = cf * rcp_hw(c);
    orho = rho * pow_hw(q, F.fk2);
    ou = sg * cf;
    op = spbar * pow_hw(q, F.fk3) - F.fpinf;
}

// The mixed-precision face solve: the FP32 iteration above gives the start
// of rs_star's FP64 Newton iteration (rs_star_face_m, tolerance 1e-6),
// which then needs about one step; p* and u* come from the FP64 pressure
// functions; the sampling is FP32.  In two parts (the flux kernel runs each
// over all its faces in turn); rs_flux_full_f is both.
__device__ __forceinline__ float rs_face_start_f(const RSFluid& F, bool px, const V4<float>& WL,
                                                 const V4<float>& WR) {
    RSSideF Lf, Rf;
    rs_side_f(Lf, F, WL.v[0], norm_u(px, WL), WL.v[3]);
    rs_side_f(Rf, F, WR.v[0], norm_u(px, WR), WR.v[3]);
    StarF Sf;
    // a start only (the FP64 iteration decides the status): stopped at a
    // relative step of 1e-4, after which the start is within ~1e-8 and the
    // FP64 iteration's first step is below its 1e-6 tolerance
    (void)rs_star_f(F, Lf, Rf, Sf, 1e-4f);
    return Sf.p;
}

__device__ __forceinline__ V4<float> rs_face_finish_f(const RSFluid& F, const FluidF& Ff, bool px,
                                                      const V4<float>& WL, const V4<float>& WR, float p0,
                                                      unsigned& st) {
    RSSideF Lf, Rf;
    rs_side_f(Lf, F, WL.v[0], norm_u(px, WL), WL.v[3]);
    rs_side_f(Rf, F, WR.v[0], norm_u(px, WR), WR.v[3]);
    Star S;
    st |= rs_star_face_m(F, Lf, Rf, p0, S, 1e-6);
    StarF S2;
    S2.p = (float)S.p;
    S2.u = (float)S.u;
    float r, u, p;
    rs_sample_f(F, Lf, Rf, S2, r, u, p);
    return rs_flux_out_f(px, Ff, r, u, p, S.u, WL, WR);
}

__device__ __forceinline__ V4<float> rs_flux_full_f(const RSFluid& F, const FluidF& Ff, bool px, const V4<float>& WL,
                                                    const V4<float>& WR, unsigned& st) {
    return rs_face_finish_f(F, Ff, px, WL, WR, rs_face_start_f(F, px, WL, WR), st);
}

// pure_RS_exact::solve_RP (pure_RS_exact.cpp:9-73): conserved L/R -> flux
template <typename T>
__device__ __forceinline__ V4<T> pure_rs_flux(bool px, T g, T pinf, const V4<T>& UL, const V4<T>& UR,
                                              unsigned& st) {
    if (!phys_state(UL) || !phys_state(UR)) st |= ST_UNPHYS;
    const V4<T> WL = cons2prims(g, pinf, UL);
    const V4<T> WR = cons2prims(g, pinf, UR);
    if (rs_trivial(px, WL, WR)) return rs_flux_trivial<T>(px, g, pinf, WL, WR);
    return rs_flux_full(px, g, pinf, WL, WR, st);
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
