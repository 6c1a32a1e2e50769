/*
 * gfm_oracle.h -- FP64 CPU restatement of the PhD-2D-GFM per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product library never links or calls it.
 *
 * Layout follows the reference's grideuler2type: a grid is (Ny+2g) rows of
 * (Nx+2g) cells, each cell four doubles [rho, rho*u, rho*v, E] (conserved) or
 * [rho, u, v, p] (primitive), stored cell-major: grid[((i*W)+j)*4 + k] with
 * W = Nx+2g, i the row (y) index and j the column (x) index.
 * Scalar planes (level set phi) are phi[i*W + j]; the velocity field is
 * vfield[(i*W + j)*2 + {0,1}].
 */
#ifndef GFM_ORACLE_H
#define GFM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_BC_TRANSMISSIVE = 0, OR_BC_REFLECTIVE = 1, OR_BC_PERIODIC = 2 };
enum { OR_GFM_ORIGINAL = 0, OR_GFM_REAL = 1, OR_GFM_RIEMANN = 2, OR_GFM_MODIFIED = 3,
       OR_GFM_VOF = 4, OR_GFM_OGFM_VOF = 5 };
enum { OR_FS_MUSCL = 0, OR_FS_GODUNOV = 1 };

/* status bits (the reference aborts on assert; the oracle records instead) */
enum {
    OR_ST_UNPHYSICAL = 1,   /* a live assert of the reference would fire */
    OR_ST_RS_FAIL = 2,      /* exact RS: no convergence / vacuum / NaN     */
    OR_ST_BADARG = 4
};

typedef struct {
    int Nx, Ny, g;
    double x0, y0, dx, dy;
    int bcL, bcR, bcB, bcT;
} or_grid;

typedef struct {
    double gamma1, pinf1, gamma2, pinf2;
} or_eos;

/* two-fluid run description (settings + methods) */
typedef struct {
    int gfm;            /* OR_GFM_* */
    int fs;             /* OR_FS_*  */
    int ls_order;       /* 1,3,5,7,9 (HOUC order) */
    int reinit_every;   /* 0 = never */
    int reinit_iters;
    double cfl;
    double T;           /* end time (loop runs while t < T and steps < max_steps) */
    int max_steps;
    int use_extension_vfield;   /* GFM_base flags (GFM_base.hpp:20-21); 0 as sim_twofluid.cpp:551 */
    int use_slip_BCs;
    int itm;            /* OR_ITM_*: the interface tracker (settings ITM) */
} or_run;

enum { OR_ITM_LS = 0, OR_ITM_CLSVOF = 1 };

/* A1/A2: EOS + conversions */
double or_eos_pressure(double g, double pinf, double e, double rho);
double or_eos_specific_ie(double g, double pinf, double p, double rho);
double or_eos_soundspeed(double g, double pinf, double p, double rho);
double or_eos_isentropic(double g, double pinf, double rho_old, double p_old, double p_new);
void or_cons_to_prims(double g, double pinf, const double* U, double* W);
void or_prims_to_cons(double g, double pinf, const double* W, double* U);
void or_flux_prims(double g, double pinf, const double* W, double* F);

/* A3/A6: exact stiffened-gas Riemann solver */
int or_exact_rs_star(double gL, double pinfL, double gR, double pinfR,
                     const double* WL, const double* WR,      /* (rho, u_n, p) */
                     double* out4, int* iters);               /* p*, u*, rho*L, rho*R */
void or_exact_rs_sample(double gL, double pinfL, double gR, double pinfR,
                        const double* WL, const double* WR, const double* star4,
                        double S, double* W3);
int or_pure_rs_flux(int paralleltox, double g, double pinf,
                    const double* UL, const double* UR, double* flux);
int or_mixed_rs(double gL, double pinfL, double gR, double pinfR,
                const double* UL, const double* UR, double* out4);

/* A13: boundary conditions */
void or_apply_bcs(const or_grid* G, double* grid);
void or_apply_bcs_scalar(const or_grid* G, double* phi);
void or_apply_bcs_vector(const or_grid* G, double* vfield);

/* A4/A5: flux solvers; bbox = {imin, imax, jmin, jmax} (half-open) */
int or_fs_muscl(const or_grid* G, double g, double pinf, double dt, const int* bbox,
                double* grid, double* future);
int or_fs_godunov(const or_grid* G, double g, double pinf, double dt, const int* bbox,
                  double* grid, double* future);

/* A14 / A16 */
double or_compute_dt_onefluid(const or_grid* G, double g, double pinf, double cfl,
                              const double* grid, double T, double t);
double or_compute_dt_twofluid(const or_grid* G, const or_eos* E, double cfl, const double* phi,
                              const double* g1, const double* g2, double T, double t);
double or_fluid_mass(const or_grid* G, double interiorsign, const double* grid, const double* phi);

/* A15: level set */
void or_ls_normal(const or_grid* G, const double* phi, int i, int j, double* n2);
int or_ls_is_interfacial(const or_grid* G, const double* phi, int i, int j);
void or_bilinear(const or_grid* G, const double* grid, double x, double y, double* out4);
/* diagnostics: RK4 tracers in the velocity field, the volume fraction z */
void or_tracers_step(const or_grid* G, const double* vfield, int n, double* pos, const double* pos0,
                     int keep_y, double t, double dt, double* entry);
void or_volume_fraction(const or_grid* G, const double* phi, double* z);
/* the VOF interface the VOF-based GFMs read, from the level set: z, unit normal, chord midpoint */
void or_vof_interface(const or_grid* G, const double* phi, double* z, double* nrm2, double* loc2);
void or_houc_coeffs(int order, double* c, int* offset0);
void or_ls_advect(const or_grid* G, int order, double* phi, const double* vfield, double dt);
void or_ls_reinit(const or_grid* G, double* phi, int iters);

/* A12 */
int or_extrapolate_vector(const or_grid* G, const double* phi, int N,
                          double* prims1, double* prims2, int mgfm);

/* A7-A11 */
int or_set_ghost_states(const or_grid* G, const or_eos* E, int variant, const double* phi,
                        double* grid1, double* grid2, int* bbox1, int* bbox2, double* vfield);
int or_set_ghost_states_flags(const or_grid* G, const or_eos* E, int variant, const double* phi,
                              double* grid1, double* grid2, int* bbox1, int* bbox2, double* vfield,
                              int use_extension_vfield, int use_slip_BCs);
/* extrapolate_extension_vfield.cpp:4-122 */
void or_extrapolate_extension_vfield(const or_grid* G, const double* phi, int N, double* vfield);

/* A17: step orchestration */
int or_onefluid_step(const or_grid* G, double g, double pinf, int fs, double cfl_eff,
                     double T, double* t, double* grid, double* scratch, double* dt_out);
int or_twofluid_step(const or_grid* G, const or_eos* E, const or_run* R, int step_index,
                     double* t, double* grid1, double* grid2, double* phi,
                     double* vfield, double* scratch1, double* scratch2, double* dt_out);
/* CLSVOF tracker (restated method; see gfm_oracle.c) */
double or_plic_line(double z, double nx, double ny, double dx, double dy);
double or_plic_area(double s, double nx, double ny, double w, double h);
void or_vof_advect(const or_grid* G, int step_index, double* z, const double* phi, const double* vf, double dt);
void or_clsvof_correct(const or_grid* G, double* phi, const double* z);
void or_clsvof_init(const or_grid* G, const double* phi, double* z);
/* or_twofluid_step / or_twofluid_run with the tracker's volume fraction z
 * (read and written when R->itm is OR_ITM_CLSVOF; run: NULL = from phi) */
int or_twofluid_step_z(const or_grid* G, const or_eos* E, const or_run* R, int step_index,
                       double* t, double* grid1, double* grid2, double* phi, double* z,
                       double* vfield, double* scratch1, double* scratch2, double* dt_out);
int or_twofluid_run_z(const or_grid* G, const or_eos* E, const or_run* R,
                      double* grid1, double* grid2, double* phi, double* z, int* steps_done, double* t_out);
int or_twofluid_run(const or_grid* G, const or_eos* E, const or_run* R,
                    double* grid1, double* grid2, double* phi,
                    int* steps_done, double* t_out);
int or_onefluid_run(const or_grid* G, double g, double pinf, int fs, double cfl, double T,
                    int max_steps, double* grid, int* steps_done, double* t_out);

/* initial conditions */
void or_cellcentre(const or_grid* G, int i, int j, double* x, double* y);
int or_ic_onefluid(const or_grid* G, double g, double pinf, int test_case, double* grid);
void or_ic_bubble(const or_grid* G, const or_eos* E,
                  double xc, double yc, double R,
                  const double* bubble_prims, const double* pre_prims, const double* post_prims,
                  double shock_x, double* grid1, double* grid2, double* phi);
void or_ic_rmi(const or_grid* G, const or_eos* E, double x_if, double amp, int modes,
               const double* light_prims, const double* post_prims, const double* heavy_prims,
               double shock_x, double* grid1, double* grid2, double* phi);
void or_ic_planar_twofluid(const or_grid* G, const or_eos* E, int axis, double pos,
                           const double* prims1, const double* prims2,
                           double* grid1, double* grid2, double* phi);
void or_shock_state(double g, double pinf, const double* pre, double mach, int direction,
                    double* post);

#ifdef __cplusplus
}
#endif
#endif
