// gfm_ic.cpp -- initial conditions of the settings-file test cases
// (sim_onefluid::set_sim_ICs, sim_twofluid::set_sim_ICs) and the
// Rankine-Hugoniot shock state used to parameterise them.  Host code; the
// grids are filled in the reference layout (rows of cells of 4 doubles).
#include <cmath>
#include <cstring>

#include "gfm_hip.h"

namespace {

struct Geo {
    int W, H, g;
    double x0, y0, dx, dy;
    explicit Geo(const gfm_sim_info* s)
        : W(s->Nx + 2 * s->numGC), H(s->Ny + 2 * s->numGC), g(s->numGC), x0(s->x0), y0(s->y0), dx(s->dx), dy(s->dy) {}
    // sim_info::cellcentre_coord
    double xc(int j) const { return x0 + (j - g + 0.5) * dx; }
    double yc(int i) const { return y0 + (i - g + 0.5) * dy; }
    size_t at(int i, int j) const { return (size_t)i * W + j; }
};

// euler_misc.hpp:89-108 with the stiffened-gas energy
void to_cons(double g, double pinf, const double* W, double* U) {
    const double rho = W[0], u = W[1], v = W[2], p = W[3];
    const double e = (p + g * pinf) / ((g - 1.0) * rho);
    U[0] = rho;
    U[1] = rho * u;
    U[2] = rho * v;
    U[3] = rho * e + 0.5 * rho * (u * u + v * v);
}

// 10x10 sub-cell samples inside a circle (sim_onefluid.cpp:381-406)
double circle_fraction(const Geo& G, int i, int j, double xc, double yc, double R) {
    const int N = 10;
    int inside = 0;
    const double bx = G.xc(j) - 0.5 * G.dx, by = G.yc(i) - 0.5 * G.dy;
    const double ddx = G.dx / N, ddy = G.dy / N;
    for (int a = 0; a < N; ++a)
        for (int b = 0; b < N; ++b) {
            const double sx = bx + (a + 0.5) * ddx - xc;
            const double sy = by + (b + 0.5) * ddy - yc;
            if (std::sqrt(sx * sx + sy * sy) <= R) inside++;
        }
    return (double)inside / (N * N);
}

}  // namespace

extern "C" {

// sim_onefluid.cpp:196-422.  test_case: 1..5 = TTC1..5_x, 11..15 = TTC1..5_y,
// 100 = circularexplosion.
int gfm_ic_onefluid(const gfm_sim_info* s, const gfm_eos* e, int tc, double* grid) {
    if (!s || !e || !grid) return GFM_E_ARG;
    static const double ttc[5][2][4] = {
        {{1.0, 0.0, 0.0, 1.0}, {0.125, 0.0, 0.0, 0.1}},
        {{1.0, -2.0, 0.0, 0.4}, {1.0, 2.0, 0.0, 0.4}},
        {{1.0, 0.0, 0.0, 1000.0}, {1.0, 0.0, 0.0, 0.01}},
        {{1.0, 0.0, 0.0, 0.1}, {1.0, 0.0, 0.0, 100.0}},
        {{5.99924, 19.5975, 0.0, 460.894}, {5.99242, -6.19633, 0.0, 46.0950}},
    };
    const Geo G(s);
    const double g = e->gamma1, pinf = e->pinf1;
    if (tc == 100) {
        const double L[4] = {1.0, 0.0, 0.0, 1.0}, R[4] = {0.125, 0.0, 0.0, 0.1};
        for (int i = 0; i < G.H; ++i)
            for (int j = 0; j < G.W; ++j) {
                const double f = circle_fraction(G, i, j, 1.0, 1.0, 0.4);
                double W[4];
                for (int k = 0; k < 4; ++k) W[k] = f * L[k] + (1.0 - f) * R[k];
                to_cons(g, pinf, W, grid + G.at(i, j) * 4);
            }
        return GFM_OK;
    }
    const int k = tc % 10;
    const bool ydir = tc > 10;
    if (k < 1 || k > 5) return GFM_E_ARG;
    double WL[4], WR[4], UL[4], UR[4];
    std::memcpy(WL, ttc[k - 1][0], sizeof WL);
    std::memcpy(WR, ttc[k - 1][1], sizeof WR);
    if (ydir) { std::swap(WL[1], WL[2]); std::swap(WR[1], WR[2]); }
    to_cons(g, pinf, WL, UL);
    to_cons(g, pinf, WR, UR);
    // set_planar_IC (sim_onefluid.cpp:426-449): a*x + b*y <= 0.5 -> left state
    for (int i = 0; i < G.H; ++i)
        for (int j = 0; j < G.W; ++j) {
            const double v = ydir ? G.yc(i) : G.xc(j);
            std::memcpy(grid + G.at(i, j) * 4, v <= 0.5 ? UL : UR, sizeof UL);
        }
    return GFM_OK;
}

// shocked bubble (sim_twofluid.cpp:774-859): fluid 1 = bubble, phi < 0 inside
int gfm_ic_bubble(const gfm_sim_info* s, const gfm_eos* e, double xc, double yc, double R, const double* bub,
                  const double* pre, const double* post, double shock_x, double* grid1, double* grid2,
                  double* phi) {
    if (!s || !e || !bub || !pre || !post || !grid1 || !grid2 || !phi) return GFM_E_ARG;
    const Geo G(s);
    double Ub[4], Upre[4], Upost[4];
    to_cons(e->gamma1, e->pinf1, bub, Ub);
    to_cons(e->gamma2, e->pinf2, pre, Upre);
    to_cons(e->gamma2, e->pinf2, post, Upost);
    for (int i = 0; i < G.H; ++i)
        for (int j = 0; j < G.W; ++j) {
            const double x = G.xc(j), y = G.yc(i);
            const size_t q = G.at(i, j);
            std::memcpy(grid2 + q * 4, x < shock_x ? Upre : Upost, sizeof Ub);
            std::memcpy(grid1 + q * 4, Ub, sizeof Ub);
            const double ddx = x - xc, ddy = y - yc;
            phi[q] = std::sqrt(ddx * ddx + ddy * ddy) - R;
        }
    return GFM_OK;
}

// single-mode Richtmyer-Meshkov interface after sim_twofluid.cpp:1245-1347:
// fluid 1 (light, phi < 0) right of x_I(y) = x_if - amp cos(2 pi modes (y-y0)/Ly),
// fluid 2 (heavy) left of it, post-shock light gas for x >= shock_x.  phi is
// the first-order signed distance (x_I - x)/sqrt(1 + x_I'^2).
int gfm_ic_rmi(const gfm_sim_info* s, const gfm_eos* e, double x_if, double amp, int modes, const double* light,
               const double* post, const double* heavy, double shock_x, double* grid1, double* grid2, double* phi) {
    if (!s || !e || !light || !post || !heavy || !grid1 || !grid2 || !phi) return GFM_E_ARG;
    const Geo G(s);
    double Ul[4], Up[4], Uh[4];
    to_cons(e->gamma1, e->pinf1, light, Ul);
    to_cons(e->gamma1, e->pinf1, post, Up);
    to_cons(e->gamma2, e->pinf2, heavy, Uh);
    const double Ly = s->Ny * s->dy;
    const double kk = 2.0 * 3.14159265358979323846 * modes / Ly;
    for (int i = 0; i < G.H; ++i)
        for (int j = 0; j < G.W; ++j) {
            const double x = G.xc(j), y = G.yc(i);
            const size_t q = G.at(i, j);
            const double arg = kk * (y - s->y0);
            const double xi = x_if - amp * std::cos(arg);
            const double dxi = amp * kk * std::sin(arg);
            std::memcpy(grid1 + q * 4, x < shock_x ? Ul : Up, sizeof Ul);
            std::memcpy(grid2 + q * 4, Uh, sizeof Uh);
            phi[q] = (xi - x) / std::sqrt(1.0 + dxi * dxi);
        }
    return GFM_OK;
}

// planar two-material tubes (ST1_x / ST2_y, sim_twofluid.cpp:675-742)
int gfm_ic_planar(const gfm_sim_info* s, const gfm_eos* e, int axis, double pos, const double* p1, const double* p2,
                  double* grid1, double* grid2, double* phi) {
    if (!s || !e || !p1 || !p2 || !grid1 || !grid2 || !phi) return GFM_E_ARG;
    const Geo G(s);
    double U1[4], U2[4];
    to_cons(e->gamma1, e->pinf1, p1, U1);
    to_cons(e->gamma2, e->pinf2, p2, U2);
    for (int i = 0; i < G.H; ++i)
        for (int j = 0; j < G.W; ++j) {
            const size_t q = G.at(i, j);
            std::memcpy(grid1 + q * 4, U1, sizeof U1);
            std::memcpy(grid2 + q * 4, U2, sizeof U2);
            phi[q] = axis == 0 ? G.xc(j) - pos : G.yc(i) - pos;
        }
    return GFM_OK;
}

// Rankine-Hugoniot state behind a shock of Mach number mach moving in
// direction (+1/-1) into `pre` (stiffened gas, barred pressures)
int gfm_shock_state(double g, double pinf, const double* pre, double mach, int direction, double* post) {
    if (!pre || !post) return GFM_E_ARG;
    const double c1 = std::sqrt(g * (pre[3] + pinf) / pre[0]);
    const double m2 = mach * mach;
    const double pr = 1.0 + 2.0 * g / (g + 1.0) * (m2 - 1.0);
    const double rr = (g + 1.0) * m2 / ((g - 1.0) * m2 + 2.0);
    post[0] = pre[0] * rr;
    post[1] = pre[1] + direction * mach * c1 * (1.0 - 1.0 / rr);
    post[2] = pre[2];
    post[3] = (pre[3] + pinf) * pr - pinf;
    return GFM_OK;
}

}  // extern "C"
