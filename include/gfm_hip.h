/*
 * gfm_hip.h -- C ABI of the MI355X (gfx950) implementation of the
 * PhD-2D-GFM per-timestep hot path.
 *
 * Plain C types only.  Every entry point returns 0 on success or a negative
 * GFM_E_* code; numerical faults found on the device are reported per
 * instance through gfm_status() (GFM_ST_* bits), mirroring the reference's
 * live asserts and its "Exact RS returned NaN" message.
 *
 * Host grids passed through this ABI use the reference's grideuler2type
 * layout flattened row-major: (Ny+2g) rows x (Nx+2g) cells x 4 doubles,
 * cell (i,j) component k at [((i*(Nx+2g))+j)*4+k]; i is the y (row) index.
 * phi is (Ny+2g) x (Nx+2g) doubles; the velocity field is x2.
 *
 * A context holds one batch of instances that share the methods (flux
 * solver, GFM variant, level-set order and reinitialisation schedule,
 * precision).  Instances may differ in grid size, geometry, EOS, BCs, CFL,
 * end time and step budget; each keeps its own dt, t and step count on the
 * device, so a whole step is launched without host synchronisation and can
 * be captured in one hipGraph.
 */
#ifndef GFM_HIP_H
#define GFM_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* boundary conditions (sim_info::BC_L/BC_R/BC_B/BC_T strings) */
enum { GFM_BC_TRANSMISSIVE = 0, GFM_BC_REFLECTIVE = 1, GFM_BC_PERIODIC = 2 };
/* sim types (GFM_settingsfile::sim_type) */
enum { GFM_SIM_ONEFLUID = 0, GFM_SIM_TWOFLUID = 1 };
/* flux solvers (GFM_settingsfile::FS) */
enum { GFM_FS_MUSCL = 0, GFM_FS_GODUNOV = 1 };
/* ghost fluid methods (GFM_settingsfile::GFM; "OGFM", "realGFM", "riemannGFM", "MGFM", "VOFGFM",
 * "OGFMVOF", sim_twofluid.cpp:522-549).  GFM_VOFGFM and GFM_OGFMVOF (GFM_VOF.cpp, OGFM_VOF.cpp)
 * read the volume fraction, mixed-cell normal and interface location of the level set's PLIC
 * reconstruction (the reference takes them from external VOF trackers) and need numGC >= 6. */
enum { GFM_ORIGINAL = 0, GFM_REAL = 1, GFM_RIEMANN = 2, GFM_MODIFIED = 3, GFM_VOFGFM = 4, GFM_OGFMVOF = 5 };
/* precision */
enum { GFM_FP64 = 0, GFM_MIXED = 1 };
/* interface tracker (settings ITM, sim_twofluid.cpp:555-589): the level set ("LS1".."LS9"), or
 * the CLSVOF tracker ("CLSVOF", :571-574): the level set plus a volume fraction z of fluid 1
 * advected conservatively by split PLIC fluxes (Weymouth and Yue 2010), phi given the PLIC line
 * of z in every cell after each step (the reference's tracker classes are external; this is the
 * published method).  z follows phi whenever the host sets or reinitialises phi (upload,
 * gfm_ic_device, gfm_ls_reinit); gfm_download_z returns it. */
enum { GFM_ITM_LS = 0, GFM_ITM_CLSVOF = 1 };

/* error codes */
enum {
    GFM_OK = 0,
    GFM_E_ARG = -1,
    GFM_E_HIP = -2,
    GFM_E_STATE = -3,
    GFM_E_NOMEM = -4
};

/* device fault bits, per instance */
enum {
    GFM_ST_UNPHYSICAL = 1,      /* an assert of the reference would fire      */
    GFM_ST_RS_FAIL = 2,         /* exact RS: no convergence, vacuum or NaN    */
    GFM_ST_IFACE_OVERFLOW = 4,  /* interfacial list larger than its capacity  */
    GFM_ST_IFACE_ROUNDS = 8,    /* interfacial dependency rounds not drained  */
    GFM_ST_IFACE_LIST = 16      /* interfacial cell counts and list disagree  */
};

/* sim_info (external PhD-Common sim_info.hpp; fields used at
 * sim_twofluid.cpp:343-360 and euler_bc.hpp:5-103) */
typedef struct {
    int Nx, Ny, numGC;
    double x0, y0, dx, dy;
    int BC_L, BC_R, BC_B, BC_T;
} gfm_sim_info;

/* binarySGparams (sim_twofluid.cpp:347-350) */
typedef struct {
    double gamma1, pinf1, gamma2, pinf2;
} gfm_eos;

/* methods shared by a context (sim_twofluid::set_sim_methods,
 * sim_twofluid.cpp:495-594, plus the precision / reinitialisation knobs) */
typedef struct {
    int sim_type;      /* GFM_SIM_*                                  */
    int fs;            /* GFM_FS_*                                   */
    int gfm;           /* GFM_* (two-fluid only)                     */
    int ls_order;      /* 1,3,5,7,9: HOUC order ("LS1".."LS9")       */
    int reinit_every;  /* reinitialise phi every N steps (0 = never) */
    int reinit_iters;  /* pseudo-time iterations per reinit          */
    int precision;     /* GFM_FP64 or GFM_MIXED                      */
    int itm;           /* GFM_ITM_LS or GFM_ITM_CLSVOF               */
} gfm_methods;

typedef struct gfm_ctx gfm_ctx;

/* context lifetime: create, add instances, commit (allocates), use, destroy */
int gfm_ctx_create(const gfm_methods* methods, gfm_ctx** out);
int gfm_ctx_destroy(gfm_ctx* ctx);

/* sim_twofluid::run_sim set-up (sim_twofluid.cpp:27-45): one simulation.
 * cfl and T as in GFM_settingsfile::CFL and sim_info::T; max_steps bounds
 * the step loop in addition to t < T. */
int gfm_instance_add(gfm_ctx* ctx, const gfm_sim_info* info, const gfm_eos* eos,
                     double cfl, double T, int max_steps, int* inst_id);
int gfm_ctx_commit(gfm_ctx* ctx);
int gfm_num_instances(const gfm_ctx* ctx);
/* GFM_base's public flags use_extension_vfield and use_slip_BCs
 * (GFM_base.hpp:20-21), set between gfm_ctx_create and gfm_ctx_commit.  Both
 * default to 0 here, as sim_twofluid::set_sim_methods sets them
 * (sim_twofluid.cpp:551-552); GFM_base's own constructor defaults
 * use_extension_vfield to 1.  use_extension_vfield: the stored velocity field
 * is the interfacial cells' velocity (u* n, or the real fluid's for
 * GFM_original) carried away from the interface by 20 iterations of
 * extrapolate_extension_vfield (extrapolate_extension_vfield.cpp:4-122)
 * instead of the real fluid's velocity.  use_slip_BCs: GFM_modified's ghost
 * fluid keeps its own tangential velocity (GFM_modified.cpp:152-169); no
 * effect on GFM_real and GFM_riemann, refused for GFM_original, whose
 * reference asserts it is off (GFM_original.cpp:21). */
int gfm_set_gfm_flags(gfm_ctx* ctx, int use_extension_vfield, int use_slip_BCs);

/* state transfer (host FP64, reference layout).  grid2 and phi are ignored
 * for one-fluid contexts and may be NULL. */
int gfm_upload_state(gfm_ctx* ctx, int inst, const double* grid1, const double* grid2,
                     const double* phi);
int gfm_download_state(gfm_ctx* ctx, int inst, double* grid1, double* grid2, double* phi);
int gfm_upload_vfield(gfm_ctx* ctx, int inst, const double* vfield);
int gfm_download_vfield(gfm_ctx* ctx, int inst, double* vfield);

/* time bookkeeping (t, step count, last dt) */
int gfm_set_time(gfm_ctx* ctx, int inst, double t, int nstep);
int gfm_get_time(gfm_ctx* ctx, int inst, double* t, int* nstep, double* dt);
int gfm_set_dt(gfm_ctx* ctx, int inst, double dt);

/* step phases over all instances, stream ordered, no host synchronisation.
 * They follow the reference loop body sim_twofluid.cpp:64-88. */
int gfm_compute_dt(gfm_ctx* ctx);             /* sim_twofluid::compute_dt :108-157, + mass sums */
int gfm_set_ghost_states(gfm_ctx* ctx);       /* GFM_base::set_ghost_states, GFM_base.hpp:37-47 */
int gfm_fs_update(gfm_ctx* ctx);              /* FS_base::pure_fluid_update, FS_base.hpp:23-32 (both fluids) */
int gfm_ls_update(gfm_ctx* ctx);              /* ls->update_interface_coupled, sim_twofluid.cpp:72 (+ reinit) */
int gfm_apply_bcs(gfm_ctx* ctx);              /* apply_BCs_euler, euler_bc.hpp:5-103, then t += dt */
int gfm_extrapolate(gfm_ctx* ctx, int N);     /* extrapolate_vector, extrapolate_vector.hpp:8-15,
                                                 on grids holding primitives */
int gfm_ls_reinit(gfm_ctx* ctx, int iters);   /* reinitialise phi now */
int gfm_get_bbox(gfm_ctx* ctx, int inst, int* bbox1, int* bbox2);   /* realcells1/2 */
/* set realcells1/2 ({imin, imax, jmin, jmax}, half-open) for the next
 * gfm_fs_update (FS_base::pure_fluid_update's BBrange argument) */
int gfm_set_bbox(gfm_ctx* ctx, int inst, const int* bbox1, const int* bbox2);

/* whole steps: the loop body above, nsteps times.  gfm_step launches
 * eagerly; gfm_step_graph captures one step as a hipGraph and replays it. */
int gfm_step(gfm_ctx* ctx, int nsteps);
int gfm_step_graph(gfm_ctx* ctx, int nsteps);
int gfm_sync(gfm_ctx* ctx);
/* nsteps eager steps with HIP events around each phase on the context's
 * stream; ms[0..4] receive the summed milliseconds of dt, ghost fill, flux
 * update, level set and BCs (no counterpart in the reference, which only
 * times the whole loop with omp_get_wtime, sim_twofluid.cpp:62,90) */
int gfm_profile_steps(gfm_ctx* ctx, int nsteps, double* ms);

int gfm_status(gfm_ctx* ctx, int inst, unsigned* flags);
/* fluid masses after the last gfm_compute_dt: sim_twofluid::fluid_mass
 * (sim_twofluid.cpp:1803-1827) for fluid 1 (phi<0) and fluid 2 (phi>0) */
int gfm_mass_sums(gfm_ctx* ctx, int inst, double* out2);
/* (t_n, mass1, mass2) recorded by the dt reduction of step n for
 * n = 0 .. n-1 (sim_twofluid::output_mass_error, sim_twofluid.cpp:298-331);
 * capacity min(max_steps + 1, 2^20) per two-fluid instance */
int gfm_mass_history(gfm_ctx* ctx, int inst, int n, double* out3);
/* number of interfacial cells and dependency rounds of the last ghost fill */
int gfm_iface_info(gfm_ctx* ctx, int inst, int* ncells, int* rounds);

/* Output timing (sim_twofluid.cpp:54-83: output at t = 0, then after each step
 * with t - lastoutputtime > T / output_freq).  While interval >= 0, the
 * instance takes no further step once t - last > interval holds, so that a
 * host driving whole graphs of steps still stops at the step the reference
 * writes its output after; the host writes, moves `last` to t and steps on.
 * interval < 0 turns the hold off (the default).  Two-fluid and one-fluid. */
int gfm_set_output_hold(gfm_ctx* ctx, int inst, double last, double interval);
/* The perturbation-amplitude tracers of sim_twofluid.cpp:1831-1887 (RMI_pert):
 * n <= 4 points (xy = x0 y0 x1 y1 ...), each moved after every step by one
 * classical RK4 step of that step's dt in the velocity field its ghost fill
 * stored (time_step_RK4 and the solved velocity field, external: restated as
 * cell-centred bilinear interpolation); keep_y puts each back on its initial
 * y afterwards (:1882-1884).  Called before gfm_ctx_commit to declare them
 * (the history is allocated there), after it to reset their positions.
 * Two-fluid contexts only. */
int gfm_set_tracers(gfm_ctx* ctx, int inst, int n, const double* xy, int keep_y);
/* entries 0 .. n-1 of 1 + 4 ntr doubles: entry k (k >= 1) holds t_k and, per
 * tracer, its position and velocity after step k, before that step's RK4 move
 * (the values output_perturbation_amplitude prints, :1869-1874); entry 0 is
 * zero.  Capacity as gfm_mass_history. */
int gfm_tracer_history(gfm_ctx* ctx, int inst, int n, double* out);
/* z (sim_twofluid.cpp:199, the level set's get_z, external): the fraction of
 * each cell's area on fluid 1's side of the plane through the cell centre with
 * the level set's value and central-difference unit normal there, over the
 * padded grid (the outermost ring 1 or 0 by the sign of phi).  Uses the level
 * set's stage buffer: call between steps.  With GFM_ITM_CLSVOF: the tracker's
 * own volume fraction (the ghost frame by the BCs of a scalar). */
int gfm_download_z(gfm_ctx* ctx, int inst, double* z);
/* No counterpart in the reference, whose trackers keep z to themselves: sets the CLSVOF tracker's
 * volume fraction over the padded grid (after gfm_upload_state, which derives it from phi), so that a
 * caller can restore a checkpoint taken with gfm_download_z.  GFM_ITM_CLSVOF contexts only. */
int gfm_upload_z(gfm_ctx* ctx, int inst, const double* z);

/* Riemann solvers over n faces / interfaces, host pointers, AoS x4.
 * pure_RS_base::solve_RP (pure_RS_base.hpp:13-20): conserved L/R -> flux.
 * mixed_RS_base::solve_mixed_RP (mixed_RS_base.hpp:12-24): conserved L/R
 * -> (p*, u*, rho*_L, rho*_R). */
int gfm_pure_rs_solve(int n, int paralleltox, double gamma, double pinf,
                      const double* UL, const double* UR, double* flux, unsigned* status);
int gfm_mixed_rs_solve(int n, double gammaL, double pinfL, double gammaR, double pinfR,
                       const double* UL, const double* UR, double* out4, unsigned* status);
/* pure_RS_base::solve_RP as the mixed-precision flux update evaluates it:
 * the conserved states rounded to FP32, the exact solver iterated in FP32,
 * the FP32 flux returned widened to double. */
int gfm_pure_rs_solve_f32(int n, int paralleltox, double gamma, double pinf,
                          const double* UL, const double* UR, double* flux, unsigned* status);

/* pure_RS_base::solve_RP over n faces, as gfm_pure_rs_solve, without
 * allocating: no counterpart in the reference.  For callers that cannot
 * allocate device memory (code under hipGraph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The host arrays are staged through a
 * fixed device array of the library (24 MiB per GPU, held until the process
 * ends), so n is at most 262144; larger n returns GFM_E_ARG.  Calls that
 * overlap on one GPU share that array and race; the status word is cleared
 * by every call.  Runs on the null stream of the current device. */
int gfm_pure_rs_solve_noalloc(int n, int paralleltox, double gamma, double pinf,
                              const double* UL, const double* UR, double* flux, unsigned* status);

/* initial conditions of the settings-file test cases (host code, no GPU):
 * sim_onefluid::set_sim_ICs (sim_onefluid.cpp:196-422; test_case 1..5 =
 * TTC1..5_x, 11..15 = TTC1..5_y, 100 = circularexplosion) and
 * sim_twofluid::set_sim_ICs (shocked bubble :774-859, RMI :1245-1347,
 * planar tubes :675-742), plus the Rankine-Hugoniot post-shock state. */
int gfm_ic_onefluid(const gfm_sim_info* info, const gfm_eos* eos, int test_case, double* grid);
int gfm_ic_bubble(const gfm_sim_info* info, const gfm_eos* eos, double xc, double yc, double R,
                  const double* bubble_prims, const double* pre_prims, const double* post_prims,
                  double shock_x, double* grid1, double* grid2, double* phi);
int gfm_ic_rmi(const gfm_sim_info* info, const gfm_eos* eos, double x_if, double amp, int modes,
               const double* light_prims, const double* post_prims, const double* heavy_prims,
               double shock_x, double* grid1, double* grid2, double* phi);
int gfm_ic_planar(const gfm_sim_info* info, const gfm_eos* eos, int axis, double pos,
                  const double* prims1, const double* prims2, double* grid1, double* grid2, double* phi);
int gfm_shock_state(double gamma, double pinf, const double* pre_prims, double mach, int direction,
                    double* post_prims);

/* GPU initial-condition library: the same initial conditions, computed on
 * the device directly into a committed context's state (conserved grids and
 * phi; one-fluid contexts: the current grid), for n listed instances in one
 * launch, and t = 0, step 0.  FP64 arithmetic as the host functions above
 * (rounded to FP32 in a mixed-precision context), so that both give the same
 * values (cos / sin of the RMI interface to within their last bits). */
enum {
    GFM_IC_TTC = 1,             /* test_case 1..5 = TTC1..5_x, 11..15 = TTC1..5_y (one fluid)   */
    GFM_IC_EXPLOSION = 2,       /* circular explosion, 10x10 sub-cell mixing (one fluid)        */
    GFM_IC_BUBBLE = 3,          /* gfm_ic_bubble: xc, yc, R, w1 bubble, w2 pre, w3 post, shock_x  */
    GFM_IC_RMI = 4,             /* gfm_ic_rmi: x_if, amp, modes, w1 light, w2 post, w3 heavy, shock_x */
    GFM_IC_PLANAR = 5           /* gfm_ic_planar: axis, pos, w1 fluid 1, w2 fluid 2            */
};
typedef struct {
    int kind;                   /* GFM_IC_* */
    int test_case;              /* GFM_IC_TTC */
    int axis;                   /* GFM_IC_PLANAR: 0 x, 1 y */
    int modes;                  /* GFM_IC_RMI */
    double xc, yc, R;           /* GFM_IC_BUBBLE */
    double x_if, amp;           /* GFM_IC_RMI */
    double pos;                 /* GFM_IC_PLANAR */
    double shock_x;             /* GFM_IC_BUBBLE, GFM_IC_RMI */
    double w1[4], w2[4], w3[4]; /* primitive states (rho, u, v, p) */
} gfm_ic_desc;
int gfm_ic_device(gfm_ctx* ctx, int n, const int* inst, const gfm_ic_desc* desc);

/* last HIP error string (thread local) */
const char* gfm_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
