// gfm_kernels.hpp -- launch interface between the host context code and the
// gfx950 kernels.
#pragma once
#include <climits>
#include "gfm_device.hpp"

namespace gfm {

// work tables (device pointers) and their lengths for one context
struct LaunchCtx {
    hipStream_t stream;
    const Inst* insts;
    const int4* w_full;  int n_full;    // 64x4 tiles over the padded arrays
    const int4* w_int;   int n_int;     // 64x4 tiles over the interior (.w = tile index in instance)
    const int4* w_fs;    int n_fs;      // flux tiles (FS_SHAPES) x fluids (.w = fluid)
    const int4* w_frame; int n_frame;   // 256-cell chunks of the ghost frame
    const int4* w_inst;  int n_inst;    // one entry per instance
    const int4* w_ich;   int n_ich;     // 256-entry chunks of the interfacial lists
    const int4* w_tb;    int n_tb;      // TB_SHAPES interior tiles (temporally blocked extrapolation)
    const int4* w_rt;    int n_rt;      // 64x32 interior tiles (temporally blocked reinitialisation)
    const int4* w_seg;   int n_seg;     // 256-cell interior row segments (dt, interfacial list)
    const int4* w_vseg;  int n_vseg;    // 256-cell row segments of the interior grown by three cells
    const int4* w_p16;   int n_p16;     // 64x16 tiles of the padded array (CLSVOF contexts only)
                                        // (the VOF GFMs' mixed-cell list; .w = segment index)
    int tb;                             // temporally blocked kernels (no x-periodic instance)
    int mirror_y, periodic_y;           // some instance of the group has mirror / periodic y BCs
    int iface_plain;                    // interfacial solve by plain rounds instead of levels
    int fs_shape;                       // flux tile: FS_SHAPES[fs_shape]
    int fs_persist;                     // > 0: the persistent flux kernel, on fs_persist CUs' worth of
                                        // workgroups (the GPU's CU count); 0: one workgroup per tile
    unsigned* fs_next;                  // its eight queue counters (device, one set per launch context)
    int tb_shape;                       // blocked extrapolation tile: TB_SHAPES[tb_shape]
};

struct Houc;
constexpr int IFREC_BYTES = 384;   // >= sizeof(IfRec), checked in the kernels
constexpr int VOFREC_BYTES = 448;  // >= sizeof(VofRec): a mixed cell's record (VOF GFMs)
constexpr int VSIDE = 16;          // side-buffer doubles per mixed cell (VOF GFMs)
// flux-update tile shapes {x, y} (one cell per thread); a context uses one,
// FS_SHAPES[LaunchCtx::fs_shape], chosen at commit (GFM_FS_TILE)
constexpr int FS_NSHAPE = 4;
constexpr int FS_SHAPES[FS_NSHAPE][2] = {{32, 16}, {64, 8}, {16, 32}, {32, 8}};
// tile shapes {x, y, threads} of the temporally blocked extrapolation
// (TB_SHAPES[LaunchCtx::tb_shape], chosen at commit: GFM_TB_TILE)
constexpr int TB_NSHAPE = 3;
constexpr int TB_SHAPES[TB_NSHAPE][3] = {{64, 16, 512}, {32, 16, 256}, {64, 8, 256}};
constexpr int SEG_PER_WG = 4;   // 256-cell row segments per workgroup (one per wavefront)

template <typename R> void launch_dt(const LaunchCtx& L, int twofluid, int q, int narrow_boxes, int count_iface);
// qorig: the ghost fill's input buffer in a whole two-fluid step (-1: none)
template <typename R> void launch_fs(const LaunchCtx& L, bool muscl, int qin, int qout, int copy_inactive,
                                     int qorig = -1);
template <typename R> void launch_extrap(const LaunchCtx& L, int mode, int bin, int bout);
// band: a whole FP32 step's fill (input Q[.][0]): on tiles farther than
// Inst::band from the interface the first chain writes nothing and the last
// passes the fill's input through; a chain that starts from primitives reads
// the fill's input at cells farther than that
template <typename R> bool launch_extrap_tb(const LaunchCtx& L, int nit, int min, int mout, int bin, int bout,
                                            int band = 0);
template <typename R> void launch_iface(const LaunchCtx& L, int variant, int pbuf, int counted);
template <typename R> void launch_mgfm_final(const LaunchCtx& L, int qin, int qout, bool prims_out = false);
// the ghost fill of GFM_VOF (ogfm false) or OGFM_VOF (true): Q[.][0] -> Q[.][1] and the velocity
// field (OGFM_VOF also narrows the update boxes); the BCs of the result are the caller's
template <typename R> void launch_vof(const LaunchCtx& L, bool ogfm);
// the extension velocity field of a ghost fill (use_extension_vfield): the
// interfacial cells' velocity carried into the rest of the grid by 20
// iterations, written to the velocity field
template <typename R> void launch_vext(const LaunchCtx& L, bool ogfm);
// clsvof: the CLSVOF tracker's z sweeps and coupling around the level set's
// reinitialisation (Inst::zv)
template <typename R> void launch_ls(const LaunchCtx& L, const Houc& H, int every, int iters, bool clsvof);
// z of the CLSVOF tracker from the level set (only >= 0: that instance)
template <typename R> void launch_clsvof_init(const LaunchCtx& L, int only);
template <typename R> void launch_reinit_now(const LaunchCtx& L, int iters);
template <typename R> void launch_bc(const LaunchCtx& L, int q, int nfluid, int advance);
template <typename R> void launch_reset(const LaunchCtx& L, int narrow_boxes);
// z, the fraction of each cell on fluid 1's side of phi's local plane, into phiA
template <typename R> void launch_zfrac(const LaunchCtx& L);
// initial conditions on the device: instance k gets descs[map[k]] (map[k] < 0: untouched), its
// conserved grids Q[.][q] and phi
struct IcDesc {   // the layout of gfm_ic_desc (include/gfm_hip.h)
    int kind, test_case, axis, modes;
    double xc, yc, R, x_if, amp, pos, shock_x;
    double w1[4], w2[4], w3[4];
};
template <typename R> void launch_ic(const LaunchCtx& L, const IcDesc* descs, const int* map, int q, int twofluid);
void launch_pure_rs(hipStream_t s, int n, int px, double g, double pinf, const double* UL, const double* UR,
                    double* F, unsigned* st);
// FP64 two-fluid contexts: Q[.][q] of both fluids and phi rounded to FP32 and back
void launch_round_state_f32(const LaunchCtx& L, int q);
// p[0..n) = v (a kernel, so that it may be captured in a graph)
void launch_fill_u32(hipStream_t s, unsigned* p, int n, unsigned v);
// the mixed-precision path's face solver on the same problem: FP32 states
// (the conserved inputs rounded to FP32), FP32 flux
void launch_pure_rs_f32(hipStream_t s, int n, int px, double g, double pinf, const double* UL, const double* UR,
                        double* F, unsigned* st);
void launch_mixed_rs(hipStream_t s, int n, double gL, double pL, double gR, double pR, const double* UL,
                     const double* UR, double* out, unsigned* st);

}  // namespace gfm
