// gfm_kernels.hip -- gfx950 kernels of the GFM hot path.
//
// Every kernel is batched: blockIdx.x indexes a work table of int4 entries
// whose .x is the instance, so one launch advances every instance of a
// context.  Kernels read the instance's dt / activity from device memory,
// which keeps a whole step free of host round trips and graph-capturable.
// R is the storage type (double, or float in mixed precision).
#include "gfm_kernels.hpp"

#include <algorithm>

namespace gfm {

// every array of an instance is device global memory: pointers cast to the
// global address space so that accesses compile to global (not flat)
// instructions, which do not count against the LDS / scalar wait counter.
// The host pass of this file only parses device code; it keeps the pointer.
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gmem(T* p) {
    return (__attribute__((address_space(1))) T*)p;
}
#else
template <typename T>
__device__ __forceinline__ T* gmem(T* p) { return p; }
#endif

template <typename R>
__device__ __forceinline__ V4<R> ld4(const void* base, const Inst& I, long long c) {
    const auto p = gmem((const R*)base);
    V4<R> a;
    a.v[0] = p[c];
    a.v[1] = p[c + I.plane];
    a.v[2] = p[c + 2 * I.plane];
    a.v[3] = p[c + 3 * I.plane];
    return a;
}

template <typename R>
__device__ __forceinline__ void st4(void* base, const Inst& I, long long c, const V4<R>& a) {
    const auto p = gmem((R*)base);
    p[c] = a.v[0];
    p[c + I.plane] = a.v[1];
    p[c + 2 * I.plane] = a.v[2];
    p[c + 3 * I.plane] = a.v[3];
}

template <typename R>
__device__ __forceinline__ R ld1(const void* base, long long c) { return gmem((const R*)base)[c]; }

__device__ __forceinline__ void flag(const Inst& I, unsigned bits) {
    if (bits) atomicOr((unsigned*)&I.si[SI_STATUS], bits);
}


// update boxes (realcells1/2): the full domain, or empty when GFM_original is
// about to narrow them to the bounding box of each fluid's real cells
// (GFM_original.cpp:47-55)
__device__ __forceinline__ void reset_boxes(const Inst& I, int narrow) {
    for (int f = 0; f < 2; ++f) {
        const auto b = gmem(I.si) + SI_BBOX + 4 * f;
        if (narrow) { b[0] = I.g + I.Ny; b[1] = I.g; b[2] = I.g + I.Nx; b[3] = I.g; }
        else { b[0] = I.g; b[1] = I.g + I.Ny; b[2] = I.g; b[3] = I.g + I.Nx; }
    }
}

// ---------------------------------------------------------------------------
// Row segments: the CFL/mass pass and the interfacial list work on 256-cell
// row segments, one segment per wavefront and four cells per lane (cell
// q*64 + lane of the segment, q = 0..3), so that every lane has its loads for
// four cells in flight at once.
// work: {inst, interior row, segment column, segment index within the instance}
// ---------------------------------------------------------------------------
constexpr int SEGW = 256, SEG_Q = SEGW / 64;

// the interfacial list of the previous build is erased from imap entry by
// entry (imap is -1 everywhere else); wavefront s erases list entries
// [s SEGW, (s+1) SEGW), which covers the list since it has at most one entry
// per interior cell
__device__ __forceinline__ void erase_prev_list(const Inst& I, int s, int lane) {
    const int n = gmem(I.si)[SI_NLIST];
#pragma unroll
    for (int q = 0; q < SEG_Q; ++q) {
        const int k = s * SEGW + q * 64 + lane;
        if (k < n) {
            const int v = I.ilist[k];
            const int i = v / I.W, j = v - (v / I.W) * I.W;
            if (v >= 0 && interior(I, i, j)) I.imap[cidx(I, i, j)] = -1;   // a bad entry was flagged when listed
        }
    }
}

// interfacial cells of a row segment (external Levelset::is_interfacial_cell,
// restated as a 4-neighbour change of side with the side taken as phi <= 0)
// for cells q*64 + lane, as a bit per q, given phi of the segment's cells and of the first ghost column
// after the last interior cell (ph[q], loaded where j <= jend): the rows
// above and below are loaded, the left and right neighbours come from the
// adjacent lanes (the segment is contiguous), plus one load at each end of
// the segment.  row = cidx(I, i, 0); j0 = this lane's column in chunk 0.
template <typename R>
__device__ __forceinline__ unsigned seg_iface_ph(const Inst& I, long long row, int j0, int jend, int lane,
                                                 const R (&ph)[SEG_Q]) {
    R up[SEG_Q], dn[SEG_Q];
#pragma unroll
    for (int k = 0; k < SEG_Q; ++k) {
        const int j = j0 + 64 * k;
        up[k] = dn[k] = R(0);
        if (j < jend) { up[k] = ld1<R>(I.phi, row + I.pitch + j); dn[k] = ld1<R>(I.phi, row - I.pitch + j); }
    }
    R el = R(0), er = R(0);
    if (lane == 0) el = ld1<R>(I.phi, row + j0 - 1);
    // lane 63's last cell is j0 + 64 (SEG_Q - 1)
    if (lane == 63 && j0 + 64 * (SEG_Q - 1) < jend) er = ld1<R>(I.phi, row + j0 + 64 * (SEG_Q - 1) + 1);
    R lft[SEG_Q], rgt[SEG_Q];
    R carry = el;   // lane 0: the left neighbour of chunk k's first cell
#pragma unroll
    for (int k = 0; k < SEG_Q; ++k) {
        const R rot = __shfl(ph[k], (lane + 63) & 63);   // lane l: lane l - 1 (lane 0: lane 63)
        lft[k] = lane == 0 ? carry : rot;
        carry = rot;
    }
    carry = er;     // lane 63: the right neighbour of chunk k's last cell
#pragma unroll
    for (int k = SEG_Q - 1; k >= 0; --k) {
        const R rot = __shfl(ph[k], (lane + 1) & 63);    // lane l: lane l + 1 (lane 63: lane 0)
        rgt[k] = lane == 63 ? carry : rot;
        carry = rot;
    }
    unsigned m = 0;
#pragma unroll
    for (int k = 0; k < SEG_Q; ++k) {
        const bool s0 = ph[k] <= R(0);
        const bool ifc = ((rgt[k] <= R(0)) != s0) || ((lft[k] <= R(0)) != s0) || ((up[k] <= R(0)) != s0) ||
                         ((dn[k] <= R(0)) != s0);
        if (j0 + 64 * k < jend && ifc) m |= 1u << k;
    }
    return m;
}

// ---------------------------------------------------------------------------
// A14 + A16: CFL dt (sim_twofluid.cpp:108-157, sim_onefluid.cpp:68-91) and
// fluid masses (sim_twofluid.cpp:1803-1827).  Segment partials, then one
// block per instance reduces them in segment order (deterministic) and also
// resets the per-step counters used later in the step.  With `count` it also
// counts each segment's interfacial cells for the interfacial list
// (k_iface_count's work, from the same phi) and erases the previous list.
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void k_dt(const Inst* __restrict__ insts, const int4* __restrict__ work,
                                            int nwork, int twofluid, int q, int count) {
    const int sw = blockIdx.x * SEG_PER_WG + (int)(threadIdx.x >> 6);
    if (sw >= nwork) return;
    const int4 w = work[sw];
    const Inst& I = insts[w.x];
    const int lane = threadIdx.x & 63;
    const int i = I.g + w.y, j0 = I.g + w.z * SEGW + lane, jend = I.g + I.Nx;
    if (count) erase_prev_list(I, w.w, lane);
    const long long row = cidx(I, i, 0);
    R ph[SEG_Q];
    V4<R> U[SEG_Q];
    // phi of the segment's cells and of the first ghost column after the
    // last interior cell (a right neighbour)
#pragma unroll
    for (int k = 0; k < SEG_Q; ++k)
        ph[k] = ((twofluid || count) && j0 + 64 * k <= jend) ? ld1<R>(I.phi, row + j0 + 64 * k) : R(0);
    // the interfacial test's loads do not wait for the states
    const unsigned ifm = count ? seg_iface_ph<R>(I, row, j0, jend, lane, ph) : 0u;
#pragma unroll
    for (int k = 0; k < SEG_Q; ++k) {
        const int j = j0 + 64 * k;
        const int f = ph[k] < R(0) ? 0 : 1;   // :128 strict
        U[k] = ld4<R>(I.Q[twofluid ? f : 0][q], I, row + (j < jend ? j : I.g));
    }
    double mu = 0.0, mv = 0.0, m1 = 0.0, m2 = 0.0;
    const double dA = I.dx * I.dy;
    int nc = 0;
    // FP32 storage: the wave speeds from the FP32 primitives (cons2prims_f,
    // c = sqrt(g (p + pinf) / rho) with the hardware square root)
    const FluidF F0 = fluid_f(I.rsf[0]), F1 = fluid_f(I.rsf[1]);
#pragma unroll
    for (int k = 0; k < SEG_Q; ++k) {
        const int j = j0 + 64 * k;
        bool iface = false;
        if (j < jend) {
            const double phi = (double)ph[k];
            const int f = twofluid ? (phi < 0.0 ? 0 : 1) : 0;
            const double U0 = U[k].v[0], U1 = U[k].v[1], U2 = U[k].v[2];
            if constexpr (sizeof(R) == 4) {
                const FluidF& Ff = f ? F1 : F0;
                const V4<float> W = cons2prims_f(Ff, U[k]);
                const float c = sqrt_hw(Ff.g * (W.v[3] + Ff.pinf) * rcp_hw(W.v[0]));
                mu = fmax(mu, (double)(fabsf(W.v[1]) + c));
                mv = fmax(mv, (double)(fabsf(W.v[2]) + c));
            } else {
                const double U3 = U[k].v[3];
                const double rho = U0;
                const double u = U1 / rho, v = U2 / rho;
                const double e = (U3 / rho) - 0.5 * (u * u + v * v);
                const double p = eos_p(I.gam[f], I.pinf[f], e, rho);
                const double cs = eos_c(I.gam[f], I.pinf[f], p, rho);
                mu = fmax(mu, fabs(U1) / rho + cs);
                mv = fmax(mv, fabs(U2) / rho + cs);
            }
            if (twofluid) {
                // the density of the fluid whose mass is summed is the loaded one
                if (-phi > 0.0) m1 += U0 * dA;
                if (phi > 0.0) m2 += U0 * dA;
            }
            iface = (ifm >> k) & 1u;
        }
        if (count) nc += __popcll(__ballot(iface));
    }
    // fixed-order reduction: a butterfly within the wavefront
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        mu = fmax(mu, __shfl_xor(mu, off));
        mv = fmax(mv, __shfl_xor(mv, off));
        m1 += __shfl_xor(m1, off);
        m2 += __shfl_xor(m2, off);
    }
    if (lane == 0) {
        double* pp = I.partials + 4 * (long long)w.w;
        pp[0] = mu; pp[1] = mv; pp[2] = m1; pp[3] = m2;
        if (count) I.iseg[w.w] = nc;
    }
}

// second pass, one workgroup per instance: fixed-order reduction of the
// block partials (deterministic), dt with the CFL ramp and the end-time clip
// (sim_twofluid.cpp:66, :152-156), fluid masses, per-step counter resets
// work: one entry per instance
__global__ __launch_bounds__(256) void k_dt_final(const Inst* __restrict__ insts, const int4* __restrict__ work,
                                                  int twofluid, int narrow_boxes) {
    const Inst& I = insts[work[blockIdx.x].x];
    const int nstep = gmem(I.si)[SI_NSTEP];
    const double t = gmem(I.sd)[SD_T];
    const bool held = gmem(I.si)[SI_HOLD] != 0 && t - gmem(I.sd)[SD_HOLD_LAST] > gmem(I.sd)[SD_HOLD_IVL];
    const bool active = (t < I.T) && (nstep < I.max_steps) && !held;
    __shared__ double red[4][256];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    // eight partials' loads in flight at once; accumulated in the same order
    constexpr int U = 8;
    for (int k0 = threadIdx.x; k0 < I.ntile_dt; k0 += 256 * U) {
        double v[U][4];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 + 256 * u;
            const double* pp = I.partials + 4 * (long long)k;
            for (int q = 0; q < 4; ++q) v[u][q] = k < I.ntile_dt ? pp[q] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (k0 + 256 * u >= I.ntile_dt) break;
            a0 = fmax(a0, v[u][0]);
            a1 = fmax(a1, v[u][1]);
            a2 += v[u][2];
            a3 += v[u][3];
        }
    }
    red[0][threadIdx.x] = a0; red[1][threadIdx.x] = a1; red[2][threadIdx.x] = a2; red[3][threadIdx.x] = a3;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            red[0][threadIdx.x] = fmax(red[0][threadIdx.x], red[0][threadIdx.x + s]);
            red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][threadIdx.x + s]);
            red[2][threadIdx.x] += red[2][threadIdx.x + s];
            red[3][threadIdx.x] += red[3][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double m0 = red[0][0], m1 = red[1][0], m2 = red[2][0], m3 = red[3][0];
    gmem(I.si)[SI_ACTIVE] = active ? 1 : 0;
    gmem(I.si)[SI_NIFACE] = 0;
    reset_boxes(I, narrow_boxes);
    if (active) {
        const double cfl = nstep < 5 ? fmin(I.cfl, 0.2) : I.cfl;
        double dt = cfl * fmin(I.dx / m0, I.dy / m1);
        if (t + dt > I.T) dt = I.T - t;
        gmem(I.sd)[SD_DT] = dt;
    }
    gmem(I.sd)[SD_MASS1] = m2;
    gmem(I.sd)[SD_MASS2] = m3;
    // fluid masses of the state at t_n (output_mass_error, sim_twofluid.cpp:298-331)
    if (twofluid && nstep < I.hist_cap) {
        double* h = I.mass_hist + 3 * (long long)nstep;
        h[0] = t; h[1] = m2; h[2] = m3;
    }
}

// LDS record of four values (a cell's primitives, a face's state or flux),
// aligned so that it moves in one 16-byte (FP32) or two (FP64) LDS accesses
template <typename R>
struct alignas(4 * sizeof(R) > 16 ? 16 : 4 * sizeof(R)) V4A { R v[4]; };

template <typename R>
__device__ __forceinline__ V4<R> tb_get(const V4A<R>* s, int k) {
    const V4A<R> a = s[k];
    V4<R> W;
#pragma unroll
    for (int q = 0; q < 4; ++q) W.v[q] = a.v[q];
    return W;
}

template <typename R>
__device__ __forceinline__ void tb_put(V4A<R>* s, int k, const V4<R>& W) {
    V4A<R> a;
#pragma unroll
    for (int q = 0; q < 4; ++q) a.v[q] = W.v[q];
    s[k] = a;
}

// ---------------------------------------------------------------------------
// A4/A5: one unsplit MUSCL-Hancock (or first-order Godunov) update of one
// fluid on a FS_TILE_X x FS_TILE_Y tile, one thread per cell
// (FS_MUSCL.cpp:148-327, FS_godunov.cpp:5-50).  LDS holds the tile's cell
// primitives and one record per face of the tile (x-faces first, then
// y-faces): the face's two solver inputs, later its flux.
//  A  primitives of the tile + 2-cell halo -> sW
//  B  per cell (tile + 1 halo): minmod slopes, boundary-extrapolated values,
//     cell-level positivity fallback, Hancock half step; each evolved
//     boundary state goes to the record of the face it sits on
//  C1 per face: the face-level fallback (FS_MUSCL.cpp:276-285) and, in
//     FP64, the W->U->W round trip give the solver inputs; identical normal
//     states take their exact flux now, the others keep their inputs in the
//     record and go to a compact list so that C2 runs the Newton iteration
//     on full wavefronts
//  C2 the exact Riemann solves of the listed faces
//  D  flux-difference update in the reference's accumulation order
// work: {inst, tile_y, tile_x, fluid}
// ---------------------------------------------------------------------------
// the tile's cells and the sizes that follow from them: primitives with a
// 2-cell halo (W), cells that evolve a face state (E: tile + 1), faces
template <int TX, int TY>
struct FsTile {
    static constexpr int X = TX, Y = TY, THREADS = TX * TY;
    static constexpr int WX = TX + 4, WY = TY + 4, EX = TX + 2, EY = TY + 2;
    static constexpr int NFX = TY * (TX + 1), NFY = (TY + 1) * TX, NFACE = NFX + NFY;
};

template <typename T>
__device__ __forceinline__ T minmod1(T dL, T dR) {
    return dR > T(0) ? fmax(T(0), fmin(dL, dR)) : fmin(T(0), fmax(dL, dR));
}

// face c of a TX x TY tile: its two cells (lo = left / below) in tile coordinates
template <int TX, int TY>
__device__ __forceinline__ void face_cells(int c, bool& px, int& ra, int& ca, int& rb, int& cb) {
    constexpr int FTX = TX, NFX = FsTile<TX, TY>::NFX;
    if (c < NFX) {
        const int r = c / (FTX + 1), fc = c - r * (FTX + 1);
        px = true; ra = r; ca = fc - 1; rb = r; cb = fc;
    } else {
        const int cc = c - NFX;
        const int r = cc / FTX, fc = cc - r * FTX;
        px = false; ra = r - 1; ca = fc; rb = r; cb = fc;
    }
}

// First-order (Godunov) solver inputs of face c from the stored states
// (pure_RS_exact.cpp:18-19 checks, W of both cells)
template <typename R, typename T>
__device__ __forceinline__ void face_input_godunov(const Inst& I, int i0, int j0, int ra, int ca, int rb, int cb,
                                                   T g, T pinf, const FluidF& Ff, const void* Qi, V4<T>& WL,
                                                   V4<T>& WR, unsigned& st) {
    const V4<R> a = ld4<R>(Qi, I, cidx(I, i0 + ra, j0 + ca));
    const V4<R> b = ld4<R>(Qi, I, cidx(I, i0 + rb, j0 + cb));
    V4<T> UL, UR;
#pragma unroll
    for (int k = 0; k < 4; ++k) { UL.v[k] = (T)a.v[k]; UR.v[k] = (T)b.v[k]; }
    if constexpr (sizeof(T) == 4) {
        if (!phys_state_f(UL) || !phys_state_f(UR)) st |= ST_UNPHYS;
        WL = cons2prims_f(Ff, UL);
        WR = cons2prims_f(Ff, UR);
    } else {
        if (!phys_state(UL) || !phys_state(UR)) st |= ST_UNPHYS;
        WL = cons2prims(g, pinf, UL);
        WR = cons2prims(g, pinf, UR);
    }
}

// minimum waves per SIMD (the second __launch_bounds__ argument on AMD
// GPUs).  FP32 storage: as many as the LDS lets a CU hold, at most 6 (the
// 32x16 tile: 3 workgroups of 512 per CU, 48 KB of LDS each, with at most 80
// VGPRs); FP64 is insensitive to the cap
template <typename R, int TX, int TY>
struct FsWaves {
    typedef FsTile<TX, TY> S;
    static constexpr int lds = 4 * S::WY * S::WX * (int)sizeof(R) + 2 * S::NFACE * (int)sizeof(V4A<R>) +
                               S::NFACE * (int)sizeof(short) + (sizeof(R) == 4 ? S::NFACE * 4 : 4) + 64;
    static constexpr int per_cu = (160 * 1024 / lds) * (S::THREADS / 64) / 4;
    static constexpr int value = sizeof(R) == 4 ? (per_cu < 6 ? per_cu : 6) : 1;
};

// qorig >= 0 (FP32 storage, whole two-fluid steps): Q[.][qorig] holds the
// ghost fill's input.  A cell whose fluid f is real there and not interfacial
// leaves the fill unchanged, and its fill output Q[.][qin] is that input
// through the fill's U->W->U conversion, which in FP32 costs several ulp on
// every step (E - rho |u|^2 / 2 and back): the cell's own state for the update
// D is therefore taken from Q[.][qorig], exactly as stored.  The
// reconstruction keeps reading Q[.][qin].  FP64 takes every state from qin.
// the flux kernel's LDS for one tile
template <typename T, int TX, int TY>
struct FsLds {
    typedef FsTile<TX, TY> S;
    T sW[4][S::WY][S::WX];
    V4A<T> sF[2][S::NFACE];   // per face: inputs lo / hi, then sF[0] = flux
    short sHard[S::NFACE];
    float sP0[sizeof(T) == 4 ? S::NFACE : 1];   // FP32 storage: start of each hard face's FP64 Newton
    T sLam[2];                                  // dt / dx, dt / dy for D
    int nHard;
};

// one work item {inst, tile_y, tile_x, fluid} of the flux update; every
// early return is taken by the whole workgroup
template <typename R, bool MUSCL, int TX, int TY>
__device__ __forceinline__ void fs_tile(const Inst* __restrict__ insts, const int4 w, int qin, int qout,
                                        int copy_inactive, int qorig, FsLds<R, TX, TY>& L) {
    typedef R T;
    typedef FsTile<TX, TY> S;
    constexpr int FTX = S::X, FTY = S::Y, FS_THREADS = S::THREADS, FWX = S::WX, FWY = S::WY;
    constexpr int FEX = S::EX, FEY = S::EY, NFX = S::NFX, NFACE = S::NFACE;
    auto& sW = L.sW;
    auto& sF = L.sF;
    auto& sHard = L.sHard;
    auto& sP0 = L.sP0;
    auto& sLam = L.sLam;
    int& nHard = L.nHard;
    const Inst& I = insts[w.x];
    const int f = w.w;
    const int tid = threadIdx.x;
    if (!gmem(I.si)[SI_ACTIVE]) {
        // a finished instance keeps its state; the one-fluid ping-pong still
        // needs it in the output buffer
        if (copy_inactive) {
            const int ci = I.g + w.y * FTY + (tid / FTX), cj = I.g + w.z * FTX + (tid % FTX);
            if (ci < I.g + I.Ny && cj < I.g + I.Nx) {
                const long long c = cidx(I, ci, cj);
                st4<R>(I.Q[f][qout], I, c, ld4<R>(I.Q[f][qin], I, c));
            }
        }
        return;
    }
    const T g = (T)I.gam[f], pinf = (T)I.pinf[f];
    const FluidF Ff = fluid_f(I.rsf[f]);   // FP32 constants (mixed precision only)
    const double dtd = gmem(I.sd)[SD_DT];
    const T dtodx = (T)(dtd / I.dx), dtody = (T)(dtd / I.dy);
    const int i0 = I.g + w.y * FTY, j0 = I.g + w.z * FTX;
    int bb0 = I.g, bb1 = I.g + I.Ny, bb2 = I.g, bb3 = I.g + I.Nx;
    if (I.use_bbox) {
        const auto b = gmem(I.si) + SI_BBOX + 4 * f;
        bb0 = max(b[0], I.g); bb1 = min(b[1], I.g + I.Ny);
        bb2 = max(b[2], I.g); bb3 = min(b[3], I.g + I.Nx);
    }
    const void* Qi = I.Q[f][qin];
    void* Qo = I.Q[f][qout];
    const bool tile_active = !(i0 + FTY <= bb0 || i0 >= bb1 || j0 + FTX <= bb2 || j0 >= bb3);
    if (tid == 0) { nHard = 0; sLam[0] = dtodx; sLam[1] = dtody; }
    const int ci = i0 + (tid / FTX), cj = j0 + (tid % FTX);
    if (!tile_active) {
        if (ci < I.g + I.Ny && cj < I.g + I.Nx) {
            const long long c = cidx(I, ci, cj);
            st4<R>(Qo, I, c, ld4<R>(Qi, I, c));
        }
        return;
    }
    // the state of this thread's own cell, for D, loaded now
    const bool own_in = ci < I.g + I.Ny && cj < I.g + I.Nx;
    const long long cown = own_in ? cidx(I, ci, cj) : 0;
    V4<R> Uown = {{R(0), R(0), R(0), R(0)}};
    // the own cell's side and interfacial index (FP32 with qorig) are loaded
    // first, the tile's states next, and the own state, whose buffer depends
    // on them, last, so that no load waits behind another's latency
    R ph_own = R(0);
    int im_own = 0;
    const bool pick = sizeof(R) == 4 && qorig >= 0 && own_in;
    if (pick) {
        ph_own = ld1<R>(I.phi, cown);
        im_own = gmem(I.imap)[cown];
    }
    // FP32 whole steps with a ghost band (I.band > 0): a tile on which fluid
    // f is a ghost at more than I.band (phi distance) from its real region
    // is not updated; those ghost states keep their values (no real cell
    // reads them within the step, and their weight in the states the
    // interface does read decays geometrically from step to step)
    if (sizeof(R) == 4 && qorig >= 0 && I.band > 0.0) {
        const R band = (R)I.band;
        const bool near = own_in && (f == 0 ? ph_own < band : ph_own > -band);
        if (!__syncthreads_or(near)) return;
    }
    auto load_own = [&]() {
        if (!own_in) return;
        const bool orig = pick && ((ph_own <= R(0)) == (f == 0)) && im_own < 0;
        Uown = ld4<R>(orig ? I.Q[f][qorig] : Qi, I, cown);
    };
    // the first-order form has no barrier in A or B: the counter and dt/dx,
    // dt/dy set by thread 0 must be visible before any wave reaches C1
    if (!MUSCL) {
        load_own();
        __syncthreads();
    }
    // A: the loads of both passes first, then the conversions
    if (MUSCL) {
        constexpr int NA = (FWY * FWX + FS_THREADS - 1) / FS_THREADS;
        V4<R> UA[NA];
#pragma unroll
        for (int q = 0; q < NA; ++q) {
            const int c = tid + q * FS_THREADS;
            UA[q] = {{R(0), R(0), R(0), R(0)}};
            if (c < FWY * FWX) {
                const int a = c / FWX, b = c % FWX;
                const int i = min(max(i0 - 2 + a, 0), I.H - 1);
                const int j = min(max(j0 - 2 + b, 0), I.W - 1);
                UA[q] = ld4<R>(Qi, I, cidx(I, i, j));
            }
        }
        load_own();
#pragma unroll
        for (int q = 0; q < NA; ++q) {
            const int c = tid + q * FS_THREADS;
            if (c < FWY * FWX) {
                const int a = c / FWX, b = c % FWX;
                V4<T> W;
                if constexpr (sizeof(T) == 4) W = cons2prims_f(Ff, UA[q]);
                else W = cons2prims(g, pinf, UA[q]);
                sW[0][a][b] = W.v[0]; sW[1][a][b] = W.v[1]; sW[2][a][b] = W.v[2]; sW[3][a][b] = W.v[3];
            }
        }
        __syncthreads();
    }
    const T hx = T(0.5) * dtodx, hy = T(0.5) * dtody;
    if (MUSCL) {
        // B: cell (r, c) of the tile (r in -1..FTY, c in -1..FTX) is at
        // (r + 2, c + 2) in sW; its L / R states sit on x-faces (r, c) /
        // (r, c + 1), its B / T states on y-faces (r, c) / (r + 1, c)
        for (int e = tid; e < FEY * FEX; e += FS_THREADS) {
            const int r = e / FEX - 1, cc = e % FEX - 1;
            const bool rin = r >= 0 && r < FTY, cin = cc >= 0 && cc < FTX;
            if (!rin && !cin) continue;   // tile corners: on no face
            const int a = r + 2, b = cc + 2;
            V4<T> Wc, BL, BR, BB, BT;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const T wc = sW[k][a][b];
                const T sx = minmod1(wc - sW[k][a][b - 1], sW[k][a][b + 1] - wc);
                const T sy = minmod1(wc - sW[k][a - 1][b], sW[k][a + 1][b] - wc);
                Wc.v[k] = wc;
                BL.v[k] = wc - T(0.5) * sx; BR.v[k] = wc + T(0.5) * sx;
                BB.v[k] = wc - T(0.5) * sy; BT.v[k] = wc + T(0.5) * sy;
            }
            T rc2, ir;
            const T rho = Wc.v[0], u = Wc.v[1], v = Wc.v[2];
            if constexpr (sizeof(T) == 4) {
                if (!phys_prims_f(Ff, BL) || !phys_prims_f(Ff, BR)) { BL = Wc; BR = Wc; }
                if (!phys_prims_f(Ff, BB) || !phys_prims_f(Ff, BT)) { BB = Wc; BT = Wc; }
                ir = __builtin_amdgcn_rcpf(rho);   // 1 ulp
                rc2 = g * (Wc.v[3] + pinf);   // rho c^2
            } else {
                if (!phys_prims(g, pinf, BL) || !phys_prims(g, pinf, BR)) { BL = Wc; BR = Wc; }
                if (!phys_prims(g, pinf, BB) || !phys_prims(g, pinf, BT)) { BB = Wc; BT = Wc; }
                const T cs = eos_c(g, pinf, Wc.v[3], rho);
                rc2 = rho * cs * cs;
                ir = T(1) / rho;
            }
            // Hancock increment hx*A(W)(BL-BR) + hy*B(W)(BB-BT) (euler_misc.hpp:149-180)
            const T x0 = BL.v[0] - BR.v[0], x1 = BL.v[1] - BR.v[1], x2 = BL.v[2] - BR.v[2], x3 = BL.v[3] - BR.v[3];
            const T y0 = BB.v[0] - BT.v[0], y1 = BB.v[1] - BT.v[1], y2 = BB.v[2] - BT.v[2], y3 = BB.v[3] - BT.v[3];
            T d[4];
            d[0] = hx * (u * x0 + rho * x1) + hy * (v * y0 + rho * y2);
            d[1] = hx * (u * x1 + ir * x3) + hy * (v * y1);
            d[2] = hx * (u * x2) + hy * (v * y2 + ir * y3);
            d[3] = hx * (u * x3 + rc2 * x1) + hy * (v * y3 + rc2 * y2);
            if (rin) {
                V4<T> EL, ER;
#pragma unroll
                for (int k = 0; k < 4; ++k) { EL.v[k] = BL.v[k] + d[k]; ER.v[k] = BR.v[k] + d[k]; }
                if (cc >= 0) tb_put<T>(sF[1], r * (FTX + 1) + cc, EL);        // right input of x-face (r, cc)
                if (cc < FTX) tb_put<T>(sF[0], r * (FTX + 1) + cc + 1, ER);   // left input of x-face (r, cc + 1)
            }
            if (cin) {
                V4<T> EB, ET;
#pragma unroll
                for (int k = 0; k < 4; ++k) { EB.v[k] = BB.v[k] + d[k]; ET.v[k] = BT.v[k] + d[k]; }
                if (r >= 0) tb_put<T>(sF[1], NFX + r * FTX + cc, EB);         // upper input of y-face (r, cc)
                if (r < FTY) tb_put<T>(sF[0], NFX + (r + 1) * FTX + cc, ET);  // lower input of y-face (r + 1, cc)
            }
        }
        __syncthreads();
    }
    // C1
    const int lim_i = min(bb1, i0 + FTY), lim_j = min(bb3, j0 + FTX);
    const int lo_i = max(bb0, i0), lo_j = max(bb2, j0);
    unsigned st = 0;
    for (int c = tid; c < NFACE; c += FS_THREADS) {
        bool px;
        int ra, ca, rb, cb;
        face_cells<TX, TY>(c, px, ra, ca, rb, cb);
        const bool lo_upd = i0 + ra >= lo_i && i0 + ra < lim_i && j0 + ca >= lo_j && j0 + ca < lim_j;
        const bool hi_upd = i0 + rb >= lo_i && i0 + rb < lim_i && j0 + cb >= lo_j && j0 + cb < lim_j;
        if (!lo_upd && !hi_upd) continue;   // no updated cell of the tile needs this face
        V4<T> WL, WR;
        if (MUSCL) {
            const V4<T> SL = tb_get<T>(sF[0], c), SR = tb_get<T>(sF[1], c);
            if constexpr (sizeof(T) == 4) {
                // FP32 storage: the evolved values (or, if unphysical, the
                // cell values) go to the solver as primitives, without the
                // W->U->W round trip, which only adds rounding
                WL = SL; WR = SR;
                if (!phys_prims_f(Ff, WR) || !phys_prims_f(Ff, WL)) {
#pragma unroll
                    for (int k = 0; k < 4; ++k) { WL.v[k] = sW[k][ra + 2][ca + 2]; WR.v[k] = sW[k][rb + 2][cb + 2]; }
                    if (!phys_prims_f(Ff, WR) || !phys_prims_f(Ff, WL)) st |= ST_UNPHYS;
                }
            } else {
                V4<T> PL = SL, PR = SR;
                if (!(phys_prims(g, pinf, PR) && phys_prims(g, pinf, PL))) {
#pragma unroll
                    for (int k = 0; k < 4; ++k) { PL.v[k] = sW[k][ra + 2][ca + 2]; PR.v[k] = sW[k][rb + 2][cb + 2]; }
                }
                const V4<T> UL = prims2cons(g, pinf, PL), UR = prims2cons(g, pinf, PR);
                if (!phys_state(UL) || !phys_state(UR)) st |= ST_UNPHYS;   // pure_RS_exact.cpp:18-19
                WL = cons2prims(g, pinf, UL);
                WR = cons2prims(g, pinf, UR);
            }
        } else {
            face_input_godunov<R, T>(I, i0, j0, ra, ca, rb, cb, g, pinf, Ff, Qi, WL, WR, st);
        }
        if (rs_trivial(px, WL, WR)) {
            V4<T> F;
            if constexpr (sizeof(T) == 4) F = rs_flux_trivial_f(px, Ff, WL, WR);
            else F = rs_flux_trivial<T>(px, g, pinf, WL, WR);
            tb_put<T>(sF[0], c, F);
        } else {
            tb_put<T>(sF[0], c, WL);
            tb_put<T>(sF[1], c, WR);
            sHard[atomicAdd(&nHard, 1)] = (short)c;
        }
    }
    __syncthreads();
    // the own cell's state waits in the primitive tile, unused from here on,
    // while the face solves need the registers (D reads it back, and
    // recomputes its cell index and takes dt / dx, dt / dy from LDS)
    T* const park = &sW[0][0][0];
    if (own_in) {
#pragma unroll
        for (int k = 0; k < 4; ++k) park[k * FWY * FWX + tid] = Uown.v[k];
    }
    // C2
    const int nh = nHard;
    const RSFluid& rf = I.rsf[f];
    if constexpr (sizeof(T) == 4) {
        // FP32 storage, in two passes: the FP32 start of every hard face,
        // then the FP64 Newton step, sampling and flux (rs_flux_full_f split)
        for (int h = tid; h < nh; h += FS_THREADS) {
            const int c = sHard[h];
            const bool px = c < NFX;
            const V4<T> WL = tb_get<T>(sF[0], c), WR = tb_get<T>(sF[1], c);
            sP0[c] = rs_face_start_f(rf, px, WL, WR);
        }
        __syncthreads();
        for (int h = tid; h < nh; h += FS_THREADS) {
            const int c = sHard[h];
            const bool px = c < NFX;
            const V4<T> WL = tb_get<T>(sF[0], c), WR = tb_get<T>(sF[1], c);
            tb_put<T>(sF[0], c, rs_face_finish_f(rf, Ff, px, WL, WR, sP0[c], st));
        }
    } else {
        for (int h = tid; h < nh; h += FS_THREADS) {
            const int c = sHard[h];
            const bool px = c < NFX;
            const V4<T> WL = tb_get<T>(sF[0], c), WR = tb_get<T>(sF[1], c);
            tb_put<T>(sF[0], c, rs_flux_full(px, g, pinf, WL, WR, st));
        }
    }
    flag(I, st);
    __syncthreads();
    // D
    int t = tid;
    asm volatile("" : "+v"(t));   // recomputed below, not kept live through C2
    const int di = i0 + t / FTX, dj = j0 + t % FTX;
    if (di < I.g + I.Ny && dj < I.g + I.Nx) {
        const long long c = cidx(I, di, dj);
        V4<R> U;
#pragma unroll
        for (int k = 0; k < 4; ++k) U.v[k] = park[k * FWY * FWX + t];
        const T dtodx = sLam[0], dtody = sLam[1];
        if (di >= bb0 && di < bb1 && dj >= bb2 && dj < bb3) {
            bool ok;
            if constexpr (sizeof(T) == 4) ok = phys_state_f(U);
            else ok = phys_state(U);
            if (!ok) flag(I, ST_UNPHYS);   // FS_MUSCL.cpp:167
            const int rr = t / FTX, fc = t % FTX;
            const int fxl = rr * (FTX + 1) + fc, fyb = NFX + rr * FTX + fc;
            const V4<T> Fl = tb_get<T>(sF[0], fxl), Fr = tb_get<T>(sF[0], fxl + 1);
            const V4<T> Fb = tb_get<T>(sF[0], fyb), Ft = tb_get<T>(sF[0], fyb + FTX);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                T u = (T)U.v[k];
                u = u + (dtodx * Fl.v[k] + dtody * Fb.v[k]);   // :290
                u = u - dtodx * Fr.v[k];                         // :291 of the next column
                u = u - dtody * Ft.v[k];                         // :292 of the next row
                U.v[k] = (R)u;
            }
        }
        st4<R>(Qo, I, c, U);
    }
}

template <typename R, bool MUSCL, int TX, int TY>
__global__ __launch_bounds__(TX * TY, (FsWaves<R, TX, TY>::value)) void k_fs(const Inst* __restrict__ insts,
                                                                      const int4* __restrict__ work, int qin,
                                                                      int qout, int copy_inactive, int qorig) {
    __shared__ FsLds<R, TX, TY> L;
    fs_tile<R, MUSCL, TX, TY>(insts, work[blockIdx.x], qin, qout, copy_inactive, qorig, L);
}

// Persistent form of k_fs: a grid of as many workgroups as the GPU holds at
// once drains the work table.  The table is in XCD order (entry p belongs to
// XCD p mod 8), so the workgroups of XCD x first take x, x + 8, x + 16, ...
// from queue x (counter next[x], zero at launch), then help the other
// queues in turn.  Thread 0 keeps the queue it is on and claims its next
// item before processing the current one; the claim's counter value is
// looked at only after that item, so the atomic's latency overlaps it.
// Every workgroup stops once all eight queues are empty.
template <typename R, bool MUSCL, int TX, int TY>
__global__ __launch_bounds__(TX * TY, (FsWaves<R, TX, TY>::value)) void k_fs_persist(
        const Inst* __restrict__ insts, const int4* __restrict__ work, int n, unsigned* __restrict__ next, int qin,
        int qout, int copy_inactive, int qorig) {
    __shared__ FsLds<R, TX, TY> L;
    __shared__ int s_item;
    const unsigned x = blockIdx.x & 7u;
    unsigned tried = 0;   // thread 0: queues found empty (x, x + 1, ... mod 8)
    unsigned pend = 0;    // thread 0: counter value of the claim on queue x + tried
    if (threadIdx.x == 0) pend = atomicAdd(&next[x], 1u);
    for (;;) {
        __syncthreads();   // the previous item's LDS and s_item are no longer read
        if (threadIdx.x == 0) {
            int p = -1;
            while (tried < 8u) {
                const long long cand = (long long)((x + tried) & 7u) + 8ll * pend;
                if (cand < n) { p = (int)cand; break; }
                if (++tried < 8u) pend = atomicAdd(&next[(x + tried) & 7u], 1u);
            }
            s_item = p;
            if (p >= 0) pend = atomicAdd(&next[(x + tried) & 7u], 1u);   // the next claim, read next round
        }
        __syncthreads();
        const int p = s_item;
        if (p < 0) return;
        fs_tile<R, MUSCL, TX, TY>(insts, work[p], qin, qout, copy_inactive, qorig, L);
    }
}

// ---------------------------------------------------------------------------
// A12: one Jacobi iteration of extrapolate_vector (extrapolate_vector.cpp:45-130)
// for both fluids, with the boundary conditions of :128-129 folded in as a
// gather: every cell of the padded array computes the new value of its BC
// source cell.  Mode bits:
//   MF_FIRST   read conserved Q[.][qin] and convert (GFM U->W, e.g. GFM_riemann.cpp:33-44)
//   MF_INJECT  interfacial star states from the side buffer replace the real
//              fluid's value on read (the in-place writes of GFM_riemann.cpp:127-140)
//   MF_LAST    write conserved Q[.][qout] and the velocity field (W->U and vfield)
//   MF_OGFM    apply GFM_original's ghost-state rule (GFM_original.cpp:59-128)
//   MF_RAW     read/write Q buffers as primitives (standalone extrapolation)
// work: {inst, tile_y, tile_x, 0}, 64x4 tiles over the padded array
// ---------------------------------------------------------------------------
enum { MF_FIRST = 1, MF_INJECT = 2, MF_LAST = 4, MF_OGFM = 8, MF_RAW = 16, MF_HOLDIF = 32 };

template <typename R>
__device__ __forceinline__ void unit_normal(const Inst& I, int i, int j, double& nx, double& ny) {
#pragma clang fp contract(off)
    nx = ((double)ld1<R>(I.phi, cidx(I, i, j + 1)) - (double)ld1<R>(I.phi, cidx(I, i, j - 1))) / (2.0 * I.dx);
    ny = ((double)ld1<R>(I.phi, cidx(I, i + 1, j)) - (double)ld1<R>(I.phi, cidx(I, i - 1, j))) / (2.0 * I.dy);
    const double m = sqrt(nx * nx + ny * ny);
    if (m < 1e-12) { nx = 1.0; ny = 0.0; }
    else { nx /= m; ny /= m; }
}

// per-instance constants of the extrapolation sweep, computed once per
// thread: the pseudo-time step over dx and dy (ax, ay) and 1/(2dx), 1/(2dy)
// for the FP32 normal
template <typename R>
struct ExK { R ax, ay; float i2dx, i2dy; };

template <typename R>
__device__ __forceinline__ ExK<R> ex_k(const Inst& I) {
#pragma clang fp contract(off)
    ExK<R> k;
    const R pdt = R(0.5 * fmin(I.dx, I.dy));
    k.ax = pdt / (R)I.dx;
    k.ay = pdt / (R)I.dy;
    k.i2dx = (float)(0.5 / I.dx);
    k.i2dy = (float)(0.5 / I.dy);
    return k;
}

// unit normal of phi in the storage precision from phi at the east, west,
// north and south neighbours: FP64 as unit_normal; FP32 (mixed precision)
// from one reciprocal square root
template <typename R>
__device__ __forceinline__ void unit_normal_v(const Inst& I, const ExK<R>& K, R pe, R pw, R pn, R ps, R& nx, R& ny) {
#pragma clang fp contract(off)
    if constexpr (sizeof(R) == 8) {
        double a = ((double)pe - (double)pw) / (2.0 * I.dx);
        double b = ((double)pn - (double)ps) / (2.0 * I.dy);
        const double m = sqrt(a * a + b * b);
        if (m < 1e-12) { a = 1.0; b = 0.0; }
        else { a /= m; b /= m; }
        nx = a; ny = b;
    } else {
        const float gx = (pe - pw) * K.i2dx;
        const float gy = (pn - ps) * K.i2dy;
        const float m2 = gx * gx + gy * gy;
        if (m2 < 1e-24f) { nx = 1.0f; ny = 0.0f; }
        else { const float r = rsqrtf(m2); nx = gx * r; ny = gy * r; }
    }
}

template <typename R>
__device__ __forceinline__ void unit_normal_r(const Inst& I, const ExK<R>& K, int i, int j, R& nx, R& ny) {
    unit_normal_v<R>(I, K, ld1<R>(I.phi, cidx(I, i, j + 1)), ld1<R>(I.phi, cidx(I, i, j - 1)),
                     ld1<R>(I.phi, cidx(I, i + 1, j)), ld1<R>(I.phi, cidx(I, i - 1, j)), nx, ny);
}

template <typename R, int MODE>
__device__ __forceinline__ V4<R> ex_load(const Inst& I, int f, int i, int j, const void* src) {
    const long long c = cidx(I, i, j);
    V4<R> W;
    if (MODE & MF_FIRST) {
        if constexpr (sizeof(R) == 4) W = cons2prims_f(fluid_f(I.rsf[f]), ld4<R>(src, I, c));
        else W = cons2prims<R>((R)I.gam[f], (R)I.pinf[f], ld4<R>(src, I, c));
    } else {
        W = ld4<R>(src, I, c);
    }
    if (MODE & MF_INJECT) {
        const int k = gmem(I.imap)[c];
        if (k >= 0) {
            const int fr = ld1<R>(I.phi, c) <= R(0) ? 0 : 1;
            if (fr == f) {
                const auto s = gmem(I.side) + 8 * (long long)k;
                W.v[0] = (R)s[0]; W.v[1] = (R)s[1]; W.v[2] = (R)s[2]; W.v[3] = (R)s[3];
            }
        }
    }
    return W;
}

// upwind neighbour offsets of extrapolate_vector.cpp:74-121: the fluid-2
// ghost takes information along -n, the fluid-1 ghost along +n
template <typename R>
__device__ __forceinline__ void extrap_dirs(int fg, R nx, R ny, int& dj, int& di) {
    dj = ((fg == 1) == (nx > R(0))) ? 1 : -1;
    di = ((fg == 1) == (ny > R(0))) ? 1 : -1;
}

// one Jacobi update of the ghost fluid's primitives from the cell (c0) and
// its upwind x (xo) and y (yo) neighbours, pseudo-time step h/2:
//   fluid 2 ghost (:74-97):  c0 + ax nx dx+ + ay ny dy+, d+ = +-(xo - c0)
//   fluid 1 ghost (:98-121): c0 - ax nx dx- - ay ny dy-, d- = -+(xo - c0)
// with the neighbours taken upwind as extrap_dirs does; both are exactly
// c0 + wx (xo - c0) + wy (yo - c0), wx = ax |nx|, wy = ay |ny| (IEEE
// negation is exact), which is what is evaluated: in FP64 with the
// reference's separate roundings, in FP32 (mixed precision) as two fused
// multiply-adds
template <typename R>
__device__ __forceinline__ V4<R> extrap_combine(R wx, R wy, const V4<R>& c0, const V4<R>& xo, const V4<R>& yo) {
#pragma clang fp contract(off)
    V4<R> Wg;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if constexpr (sizeof(R) == 4)
            Wg.v[k] = fmaf(wy, yo.v[k] - c0.v[k], fmaf(wx, xo.v[k] - c0.v[k], c0.v[k]));
        else
            Wg.v[k] = (c0.v[k] + wx * (xo.v[k] - c0.v[k])) + wy * (yo.v[k] - c0.v[k]);
    }
    return Wg;
}

template <typename R>
__device__ __forceinline__ void extrap_weights(const ExK<R>& K, R nx, R ny, R& wx, R& wy) {
#pragma clang fp contract(off)
    wx = K.ax * fabs(nx);
    wy = K.ay * fabs(ny);
}

// GFM_original.cpp:65-116: the ghost fluid takes the real fluid's velocity
// and pressure, its density by isentropic extrapolation; false if unphysical
template <typename R>
__device__ __forceinline__ bool ogfm_ghost(const Inst& I, int fg, const V4<R>& Wr, V4<R>& Wg) {
#pragma clang fp contract(off)
    const double gg = I.gam[fg], pg = I.pinf[fg];
    const double rn = (double)Wg.v[0] * pow(((double)Wr.v[3] + pg) / ((double)Wg.v[3] + pg), 1.0 / gg);
    Wg.v[0] = (R)rn; Wg.v[1] = Wr.v[1]; Wg.v[2] = Wr.v[2]; Wg.v[3] = Wr.v[3];
    return rn >= 0.0;
}

// one cell's output in the form the mode asks for, before any boundary flip:
// (a, b) = fluid 1 and 2 (conserved for MF_LAST, else primitives) and the
// velocity of the fluid real in the cell.  A ghost cell's output is its
// source's with the momentum / velocity components negated (negation commutes
// exactly with W->U); every ghost image of a cell is written from the one
// conversion, so no two conversions of equal data can round differently.
template <typename R>
struct ExOut { V4<R> a, b; R vu, vv; };

// W -> U of fluid f in the storage precision's form
template <typename R>
__device__ __forceinline__ V4<R> to_cons(const Inst& I, int f, const V4<R>& W) {
    if constexpr (sizeof(R) == 4) return prims2cons_f(fluid_f(I.rsf[f]), W);
    else return prims2cons<R>((R)I.gam[f], (R)I.pinf[f], W);
}

template <typename R, int MODE>
__device__ __forceinline__ ExOut<R> extrap_out(const Inst& I, int fr, const V4<R>& Wr, const V4<R>& Wg) {
    ExOut<R> o;
    V4<R> A, B;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        A.v[k] = fr == 0 ? Wr.v[k] : Wg.v[k];
        B.v[k] = fr == 0 ? Wg.v[k] : Wr.v[k];
    }
    if (MODE & MF_LAST) {
        o.a = to_cons<R>(I, 0, A);
        o.b = to_cons<R>(I, 1, B);
    } else {
        o.a = A;
        o.b = B;
    }
    o.vu = Wr.v[1];
    o.vv = Wr.v[2];
    return o;
}

template <typename R, int MODE>
__device__ __forceinline__ void extrap_put(const Inst& I, int i, int j, int bout, ExOut<R> o, bool fu, bool fv) {
    if (fu) { o.a.v[1] = -o.a.v[1]; o.b.v[1] = -o.b.v[1]; o.vu = -o.vu; }
    if (fv) { o.a.v[2] = -o.a.v[2]; o.b.v[2] = -o.b.v[2]; o.vv = -o.vv; }
    const long long c = cidx(I, i, j);
    if (MODE & MF_LAST) {
        st4<R>(I.Q[0][bout], I, c, o.a);
        st4<R>(I.Q[1][bout], I, c, o.b);
        const auto vel = gmem((R*)I.vel);
        vel[c] = o.vu;
        vel[c + I.plane] = o.vv;
    } else if (MODE & MF_RAW) {
        st4<R>(I.Q[0][bout], I, c, o.a);
        st4<R>(I.Q[1][bout], I, c, o.b);
    } else {
        st4<R>(I.P[0][bout], I, c, o.a);
        st4<R>(I.P[1][bout], I, c, o.b);
    }
}

template <typename R, int MODE>
__device__ __forceinline__ void extrap_cell(const Inst& I, int i, int j, int bin, int bout, int* sbb) {
    bool fv, fu;
    const int si = src_row(I, i, fv);
    const int sj = src_col(I, j, fu);
    const bool qsrc = (MODE & (MF_FIRST | MF_RAW)) != 0;
    const long long cs = cidx(I, si, sj);
    const R lsval = ld1<R>(I.phi, cs);
    R nx, ny;
    const ExK<R> K = ex_k<R>(I);
    unit_normal_r<R>(I, K, si, sj, nx, ny);
    const int fr = lsval <= R(0) ? 0 : 1, fg = 1 - fr;
    const void* in_r = qsrc ? I.Q[fr][bin] : I.P[fr][bin];
    const void* in_g = qsrc ? I.Q[fg][bin] : I.P[fg][bin];
    R wx, wy;
    extrap_weights<R>(K, nx, ny, wx, wy);
    V4<R> Wr = ex_load<R, MODE>(I, fr, si, sj, in_r);
    int dj, di;
    extrap_dirs<R>(fg, nx, ny, dj, di);
    V4<R> Wg = extrap_combine<R>(wx, wy, ex_load<R, MODE>(I, fg, si, sj, in_g),
                                 ex_load<R, MODE>(I, fg, si, sj + dj, in_g),
                                 ex_load<R, MODE>(I, fg, si + di, sj, in_g));
    // MF_HOLDIF: interfacial cells keep their ghost state (extrapolate_vector_mgfm,
    // extrapolate_vector.cpp:215)
    if ((MODE & MF_HOLDIF) && interior(I, si, sj) && gmem(I.imap)[cs] >= 0) Wg = ex_load<R, MODE>(I, fg, si, sj, in_g);
    if (!(Wg.v[0] >= R(0))) flag(I, ST_UNPHYS);
    const bool own = (si == i && sj == j);
    if ((MODE & MF_OGFM) && own) {
        if (!ogfm_ghost<R>(I, fg, Wr, Wg)) flag(I, ST_UNPHYS);
        int* bb = sbb + 4 * fr;
        atomicMin(&bb[0], i - 2); atomicMax(&bb[1], i + 2);
        atomicMin(&bb[2], j - 2); atomicMax(&bb[3], j + 2);
    }
    extrap_put<R, MODE>(I, i, j, bout, extrap_out<R, MODE>(I, fr, Wr, Wg), fu, fv);
}

template <typename R, int MODE>
__global__ __launch_bounds__(256) void k_extrap(const Inst* __restrict__ insts, const int4* __restrict__ work,
                                                int bin, int bout) {
    const int4 w = work[blockIdx.x];
    const Inst& I = insts[w.x];
    if (!gmem(I.si)[SI_ACTIVE]) return;
    __shared__ int sbb[8];
    if (MODE & MF_OGFM) {
        if (threadIdx.x < 8) sbb[threadIdx.x] = (threadIdx.x & 1) ? INT_MIN : INT_MAX;
        __syncthreads();
    }
    const int i = w.y * 4 + (threadIdx.x >> 6);
    const int j = w.z * 64 + (threadIdx.x & 63);
    if (i < I.H && j < I.W) extrap_cell<R, MODE>(I, i, j, bin, bout, sbb);
    if (MODE & MF_OGFM) {
        __syncthreads();
        if (threadIdx.x < 8) {
            const int v = sbb[threadIdx.x];
            int* dst = I.si + SI_BBOX + threadIdx.x;
            if (threadIdx.x & 1) { if (v != INT_MIN) atomicMax(dst, v); }
            else if (v != INT_MAX) atomicMin(dst, v);
        }
    }
}

// ---------------------------------------------------------------------------
// NIT Jacobi iterations of k_extrap in one launch (temporal blocking).  A
// workgroup owns a TBX x TBY tile of interior cells and keeps the tile plus an
// NIT-cell halo of both fluids' primitives in LDS.  Iteration t is exact on
// the region shrunk by t cells, so after NIT iterations the tile holds what
// NIT launches of k_extrap produce, bit for bit (same arithmetic, same order).
// Ghost cells: k_extrap gives a ghost cell its BC source cell's new value,
// reflected.  With mirror boundaries the stencil reads only the first ghost
// layer, whose source is the adjacent interior cell, so that cell refreshes
// it after each iteration; the first iteration reads the arrays as they are,
// ghost frame included.  The last iteration writes every owned cell and each
// ghost cell whose source it is.  PY: the instances periodic in y (the
// others return at once).  Their region rows wrap to the interior rows they
// stand for and are computed like them, so no y ghost layer is refreshed;
// only the first iteration of a row on the y boundary whose upwind y
// neighbour lies across it reads that neighbour from the input's stored
// ghost row (kept in an LDS strip), as k_extrap does.  x-periodic contexts
// use k_extrap.
// MIN: read-side mode bits of the first iteration (MF_FIRST, MF_INJECT);
// MOUT: write-side bits of the last (MF_LAST, MF_OGFM; none = primitives to
// P[.][bout]).
// work: {inst, tile_y, tile_x, 0}, TBX x TBY interior tiles
// ---------------------------------------------------------------------------

// new ghost-fluid primitives of an LDS cell (extrap_cell's arithmetic) from
// its current ones c0 and those of its upwind x and y neighbours at LDS
// cells ox, oy
template <typename R>
__device__ __forceinline__ V4<R> tb_update(const V4A<R>* s, const V4<R>& c0, int ox, int oy, R wx, R wy) {
    return extrap_combine<R>(wx, wy, c0, tb_get<R>(s, ox), tb_get<R>(s, oy));
}

// the output of an owned cell (i, j), and the ghost cells it is the source of
// (inverse of src_row / src_col): rows {i, 2g-1-i, 2(Ny+g)-1-i} x columns
// {j, 2g-1-j, 2(Nx+g)-1-j} where those are ghost cells, which take oa
template <typename R, int MOUT, bool PY>
__device__ __forceinline__ void tb_put_owned(const Inst& I, int i, int j, int bout, const ExOut<R>& o,
                                             const ExOut<R>& oa, bool rfb, bool rft, bool rfx) {
    const int g = I.g;
    extrap_put<R, MOUT>(I, i, j, bout, o, false, false);
    if (i < 2 * g || i >= I.Ny || j < 2 * g || j >= I.Nx) {
#pragma unroll 1
        for (int m = 1; m < 9; ++m) {
            const int pr = m / 3, pc = m - 3 * (m / 3);
            int r = i, cc = j;
            bool fv = false, fu = false;
            if (pr == 1) { if (i >= 2 * g) continue; r = PY ? i + I.Ny : 2 * g - 1 - i; fv = !PY && rfb; }
            if (pr == 2) { if (i < I.Ny) continue; r = PY ? i - I.Ny : 2 * (I.Ny + g) - 1 - i; fv = !PY && rft; }
            if (pc == 1) { if (j >= 2 * g) continue; cc = 2 * g - 1 - j; fu = rfx; }
            if (pc == 2) { if (j < I.Nx) continue; cc = 2 * (I.Nx + g) - 1 - j; fu = rfx; }
            extrap_put<R, MOUT & MF_LAST>(I, r, cc, bout, oa, fu, fv);
        }
    }
}

// Whole FP32 steps with a ghost band: a chain that starts from conserved
// states writes its primitives only on tiles with a cell nearer the
// interface than Inst::band, so the chain after it reads the fill's input
// Q[.][0] at the cells farther than that.  A ghost-frame cell stands for its
// BC source cell, whose input is read with the BC's sign change (c: the
// source's index; fu, fv: the velocity component the BC negates).  True when
// the cell is such a cell.
template <typename R>
__device__ __forceinline__ bool far_src(const Inst& I, int i, int j, R bnd, long long& c, bool& fu, bool& fv) {
    const int si = src_row(I, i, fv), sj = src_col(I, j, fu);
    c = cidx(I, si, sj);
    return fabs(ld1<R>(I.phi, c)) >= bnd;
}

// primitives of both fluids from the fill's input at source cell c
template <typename R>
__device__ __forceinline__ void far_prims(const Inst& I, long long c, bool fu, bool fv, V4<R>& W1, V4<R>& W2) {
    W1 = ld4<R>(I.Q[0][0], I, c);
    W2 = ld4<R>(I.Q[1][0], I, c);
    if constexpr (sizeof(R) == 4) {
        W1 = cons2prims_f(fluid_f(I.rsf[0]), W1);
        W2 = cons2prims_f(fluid_f(I.rsf[1]), W2);
    }
    if (fu) { W1.v[1] = -W1.v[1]; W2.v[1] = -W2.v[1]; }
    if (fv) { W1.v[2] = -W1.v[2]; W2.v[2] = -W2.v[2]; }
}

template <typename R, int NIT, int MIN, int MOUT, bool PY, int SH>
__global__ __launch_bounds__(TB_SHAPES[SH][2]) void k_extrap_tb(const Inst* __restrict__ insts,
                                                                const int4* __restrict__ work, int bin, int bout,
                                                                int band) {
    static_assert(NIT >= 2, "the first iteration is not the last");
    constexpr int TBX = TB_SHAPES[SH][0], TBY = TB_SHAPES[SH][1], TB_THREADS = TB_SHAPES[SH][2];
    constexpr int RX = TBX + 2 * NIT, RY = TBY + 2 * NIT, NC = RX * RY;
    constexpr int SLOTS = (NC + TB_THREADS - 1) / TB_THREADS;
    constexpr int LM = MIN & (MF_FIRST | MF_INJECT);
    // s[f NC + cell]: the primitives of fluid f; PY: then the stored ghost
    // rows below and above the interior, s[2 NC + (2 side + f) RX + column]
    __shared__ V4A<R> s[2 * NC + (PY ? 4 * RX : 0)];
    __shared__ int sbb[8];
    const int4 w = work[blockIdx.x];
    const Inst& I = insts[w.x];
    if ((I.bc[2] == BC_P && I.bc[3] == BC_P) != PY || !gmem(I.si)[SI_ACTIVE]) return;
    const int g = I.g;
    const int ri = g + w.y * TBY - NIT, rj = g + w.z * TBX - NIT;   // region origin
    if (MOUT & MF_OGFM) {
        if (threadIdx.x < 8) sbb[threadIdx.x] = (threadIdx.x & 1) ? INT_MIN : INT_MAX;
    }
    const void* in0 = (MIN & MF_FIRST) ? I.Q[0][bin] : I.P[0][bin];
    const void* in1 = (MIN & MF_FIRST) ? I.Q[1][bin] : I.P[1][bin];
    const bool rfb = I.bc[2] == BC_R, rft = I.bc[3] == BC_R, rfx = I.bc[0] == BC_R;
    // FP32 whole steps with a ghost band (band != 0): a tile whose cells are
    // all farther than I.band (phi) from the interface holds real states the
    // fill does not change and ghost states the band leaves alone.  The first
    // chain of a fill writes nothing there (the next chain reads the fill's
    // input at such cells instead); the last chain passes the fill's input
    // through, the real fluid converted exactly as below and the ghost fluid
    // as it came in.  A chain that starts from primitives reads the fill's
    // input wherever |phi| >= I.band (sub_in).
    const bool subst = sizeof(R) == 4 && !(MOUT & MF_OGFM) && band && I.band > 0.0;
    const R bnd = (R)I.band;
    if constexpr (sizeof(R) == 4 && !(MOUT & MF_OGFM)) {
        if (subst) {
            const auto ph = gmem((const R*)I.phi);
            bool near = false;
            for (int q = threadIdx.x; q < TBX * TBY; q += TB_THREADS) {
                const int i = g + w.y * TBY + q / TBX, j = g + w.z * TBX + q % TBX;
                if (interior(I, i, j)) near = near || fabs(ph[cidx(I, i, j)]) < bnd;
            }
            if (!__syncthreads_or(near)) {
                if (!(MOUT & MF_LAST)) return;
                bool bad = false;
                for (int q = threadIdx.x; q < TBX * TBY; q += TB_THREADS) {
                    const int i = g + w.y * TBY + q / TBX, j = g + w.z * TBX + q % TBX;
                    if (!interior(I, i, j)) continue;
                    const long long c = cidx(I, i, j);
                    V4<R> W1, W2;
                    far_prims<R>(I, c, false, false, W1, W2);
                    const int fr = ph[c] <= R(0) ? 0 : 1;
                    const V4<R>& Wr = fr == 0 ? W1 : W2;
                    const V4<R>& Wg = fr == 0 ? W2 : W1;
                    if (!(Wg.v[0] >= R(0))) bad = true;
                    const ExOut<R> o = extrap_out<R, MOUT>(I, fr, Wr, Wg);
                    tb_put_owned<R, MOUT, PY>(I, i, j, bout, o, o, rfb, rft, rfx);
                }
                if (bad) flag(I, ST_UNPHYS);
                return;
            }
        }
    }
    const bool sub_in = subst && !(LM & MF_FIRST);
    const ExK<R> K = ex_k<R>(I);
    // per cell: the update weights, the ghost fluid's current primitives
    // (only this thread changes them), the LDS cells of the ghost fluid at
    // the cell and its upwind neighbours, and bits (1: interior with its
    // stencil inside the region, 2: fluid 2 real, 4: owned, 8: next to the
    // domain edge (its new state refreshes the first ghost layer), 16 and up:
    // distance to the region edge)
    R wxs[SLOTS], wys[SLOTS];
    V4<R> cur[SLOTS];
    int oc[SLOTS], ox[SLOTS], oy[SLOTS];
    int meta[SLOTS];
    const int ib = g, it = g + I.Ny - 1, jl = g, jr = g + I.Nx - 1;
    // in two passes, so that every load for the thread's cells is in flight
    // before any is used: the raw values (states, phi and its neighbours,
    // the interfacial index), then ex_load's conversion and injection,
    // normals, weights and the LDS writes
    V4<R> L1[SLOTS], L2[SLOTS];
    R pc[SLOTS], pe[SLOTS], pw[SLOTS], pn[SLOTS], ps[SLOTS];
    int im[SLOTS];
    unsigned sbits = 0;   // bit k: slot k's primitives come from the fill's input (sub_in)
    static_assert(SLOTS <= 32, "one bit per slot");
    const auto phig = gmem((const R*)I.phi);
    if constexpr (PY) {
        // the stored ghost rows g-1 and g+Ny as the first iteration reads them
        for (int q = threadIdx.x; q < 2 * RX; q += TB_THREADS) {
            const int side = q / RX, b = q - side * RX, j = rj + b;
            const int r = side ? g + I.Ny : g - 1;
            V4<R> W1 = {{R(0), R(0), R(0), R(0)}}, W2 = W1;
            long long cs = 0;
            bool su = false, sv = false;
            if (j >= 0 && j < I.W && sub_in && far_src<R>(I, r, j, bnd, cs, su, sv)) {
                far_prims<R>(I, cs, su, sv, W1, W2);
            } else if (j >= 0 && j < I.W) {
                const long long c = cidx(I, r, j);
                W1 = ld4<R>(in0, I, c);
                W2 = ld4<R>(in1, I, c);
                if (LM & MF_FIRST) {
                    if constexpr (sizeof(R) == 4) {
                        W1 = cons2prims_f(fluid_f(I.rsf[0]), W1);
                        W2 = cons2prims_f(fluid_f(I.rsf[1]), W2);
                    } else {
                        W1 = cons2prims<R>((R)I.gam[0], (R)I.pinf[0], W1);
                        W2 = cons2prims<R>((R)I.gam[1], (R)I.pinf[1], W2);
                    }
                }
                const int m = (LM & MF_INJECT) ? gmem(I.imap)[c] : -1;
                if (m >= 0) {
                    const auto sd = gmem(I.side) + 8 * (long long)m;
                    V4<R> Ws;
#pragma unroll
                    for (int qq = 0; qq < 4; ++qq) Ws.v[qq] = (R)sd[qq];
                    if (phig[c] <= R(0)) W1 = Ws; else W2 = Ws;
                }
            }
            tb_put<R>(s, 2 * NC + (2 * side) * RX + b, W1);
            tb_put<R>(s, 2 * NC + (2 * side + 1) * RX + b, W2);
        }
    }
#pragma unroll
    for (int k = 0; k < SLOTS; ++k) {
        const int idx = threadIdx.x + k * TB_THREADS;
        const int a = idx / RX, b = idx - (idx / RX) * RX;
        const int i = ri + a, j = rj + b;
        bool fl;
        const int is = PY ? src_row(I, i, fl) : i;   // the row the region row stands for
        const bool valid = idx < NC && is >= 0 && is < I.H && j >= 0 && j < I.W;
        const bool inner = valid && interior(I, is, j) && a > 0 && a < RY - 1 && b > 0 && b < RX - 1;
        const long long c = valid ? cidx(I, is, j) : 0;
        L1[k] = L2[k] = {{R(0), R(0), R(0), R(0)}};
        pc[k] = pe[k] = pw[k] = pn[k] = ps[k] = R(0);
        im[k] = -1;
        if (valid) {
            long long cs = 0;
            bool su = false, sv = false;
            if (sub_in && far_src<R>(I, is, j, bnd, cs, su, sv)) {
                far_prims<R>(I, cs, su, sv, L1[k], L2[k]);
                sbits |= 1u << k;
            } else {
                L1[k] = ld4<R>(in0, I, c);
                L2[k] = ld4<R>(in1, I, c);
            }
            if (LM & MF_INJECT) im[k] = gmem(I.imap)[c];
            if ((LM & MF_INJECT) || inner) pc[k] = phig[c];
        }
        if (inner) {
            pe[k] = phig[c + 1]; pw[k] = phig[c - 1];
            pn[k] = phig[c + I.pitch]; ps[k] = phig[c - I.pitch];
        }
    }
#pragma unroll
    for (int k = 0; k < SLOTS; ++k) {
        const int idx = threadIdx.x + k * TB_THREADS;
        wxs[k] = R(0); wys[k] = R(0); meta[k] = 0;
        oc[k] = ox[k] = oy[k] = idx;
        cur[k] = {{R(0), R(0), R(0), R(0)}};
        if (idx < NC) {
            const int a = idx / RX, b = idx - (idx / RX) * RX;
            const int i = ri + a, j = rj + b;
            bool fl;
            const int is = PY ? src_row(I, i, fl) : i;
            V4<R> W1 = {{R(0), R(0), R(0), R(0)}}, W2 = W1;
            if (is >= 0 && is < I.H && j >= 0 && j < I.W) {
                // ex_load<R, LM> of both fluids on the loaded values
                // (sub_in slots hold primitives already)
                W1 = L1[k]; W2 = L2[k];
                if ((LM & MF_FIRST) && !((sbits >> k) & 1u)) {
                    if constexpr (sizeof(R) == 4) {
                        W1 = cons2prims_f(fluid_f(I.rsf[0]), L1[k]);
                        W2 = cons2prims_f(fluid_f(I.rsf[1]), L2[k]);
                    } else {
                        W1 = cons2prims<R>((R)I.gam[0], (R)I.pinf[0], L1[k]);
                        W2 = cons2prims<R>((R)I.gam[1], (R)I.pinf[1], L2[k]);
                    }
                }
                if ((LM & MF_INJECT) && im[k] >= 0) {
                    const auto sd = gmem(I.side) + 8 * (long long)im[k];
                    V4<R> Ws;
#pragma unroll
                    for (int q = 0; q < 4; ++q) Ws.v[q] = (R)sd[q];
                    if (pc[k] <= R(0)) W1 = Ws; else W2 = Ws;
                }
                if (interior(I, is, j) && a > 0 && a < RY - 1 && b > 0 && b < RX - 1) {
                    R nx, ny;
                    unit_normal_v<R>(I, K, pe[k], pw[k], pn[k], ps[k], nx, ny);
                    extrap_weights<R>(K, nx, ny, wxs[k], wys[k]);
                    const bool own = a >= NIT && a < NIT + TBY && b >= NIT && b < NIT + TBX && interior(I, i, j);
                    const bool edge = (!PY && (i == ib || i == it)) || j == jl || j == jr;
                    const int dist = min(min(a, RY - 1 - a), min(b, RX - 1 - b));
                    const int fr = pc[k] <= R(0) ? 0 : 1;
                    meta[k] = 1 | (fr ? 2 : 0) | (own ? 4 : 0) | (edge ? 8 : 0) | (dist << 4);
                    int dj, di;
                    extrap_dirs<R>(1 - fr, nx, ny, dj, di);
                    // bit 12: the first iteration's upwind y neighbour is the
                    // stored ghost row, below (bit 13 clear) or above (set)
                    if (PY && ((is == ib && di < 0) || (is == it && di > 0)))
                        meta[k] |= 4096 | (di > 0 ? 8192 : 0);
                    oc[k] = (1 - fr) * NC + idx;
                    ox[k] = oc[k] + dj;
                    oy[k] = oc[k] + di * RX;
                    cur[k] = fr ? W1 : W2;
                }
            }
            tb_put<R>(s, idx, W1);
            tb_put<R>(s, NC + idx, W2);
        }
    }
    __syncthreads();
    bool bad = false;
#pragma unroll 1
    for (int t = 1; t < NIT; ++t) {
        int act = 0;
#pragma unroll
        for (int k = 0; k < SLOTS; ++k) {
            if ((meta[k] & 1) && ((meta[k] >> 4) & 255) >= t) {
                act |= 1 << k;
                int oyk = oy[k];
                if (PY && t == 1 && (meta[k] & 4096)) {
                    const int idx = threadIdx.x + k * TB_THREADS;
                    const int b = idx - (idx / RX) * RX;
                    oyk = 2 * NC + (2 * ((meta[k] & 8192) ? 1 : 0) + (oc[k] >= NC ? 1 : 0)) * RX + b;
                }
                cur[k] = tb_update<R>(s, cur[k], ox[k], oyk, wxs[k], wys[k]);
                if ((meta[k] & 4) && !(cur[k].v[0] >= R(0))) bad = true;
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < SLOTS; ++k) {
            if (!(act & (1 << k))) continue;
            tb_put<R>(s, oc[k], cur[k]);
            if (meta[k] & 8) {
                const int idx = threadIdx.x + k * TB_THREADS;
                const int a = idx / RX, b = idx - (idx / RX) * RX;
                const int i = ri + a, j = rj + b;
                // first ghost layer: this cell's new state, reflected
                const V4<R> W1 = tb_get<R>(s, idx), W2 = tb_get<R>(s, NC + idx);
                V4<R> Wv1 = W1, Wv2 = W2, Wu1 = W1, Wu2 = W2;
                Wv1.v[2] = -Wv1.v[2]; Wv2.v[2] = -Wv2.v[2];
                Wu1.v[1] = -Wu1.v[1]; Wu2.v[1] = -Wu2.v[1];
                if (!PY && i == ib) { tb_put<R>(s, idx - RX, rfb ? Wv1 : W1); tb_put<R>(s, NC + idx - RX, rfb ? Wv2 : W2); }
                if (!PY && i == it) { tb_put<R>(s, idx + RX, rft ? Wv1 : W1); tb_put<R>(s, NC + idx + RX, rft ? Wv2 : W2); }
                if (j == jl) { tb_put<R>(s, idx - 1, rfx ? Wu1 : W1); tb_put<R>(s, NC + idx - 1, rfx ? Wu2 : W2); }
                if (j == jr) { tb_put<R>(s, idx + 1, rfx ? Wu1 : W1); tb_put<R>(s, NC + idx + 1, rfx ? Wu2 : W2); }
            }
        }
        __syncthreads();
    }
    // last iteration: owned cells, written with the ghost cells they source
#pragma unroll
    for (int k = 0; k < SLOTS; ++k) {
        if (!(meta[k] & 4)) continue;
        const int idx = threadIdx.x + k * TB_THREADS;
        const int a = idx / RX, b = idx - (idx / RX) * RX;
        const int i = ri + a, j = rj + b;
        const int fr = (meta[k] & 2) ? 1 : 0, fg = 1 - fr;
        const V4<R> Wg = tb_update<R>(s, cur[k], ox[k], oy[k], wxs[k], wys[k]);
        if (!(Wg.v[0] >= R(0))) bad = true;
        const V4<R> Wr = tb_get<R>(s, fr * NC + idx);
        V4<R> Wo = Wg;
        if (MOUT & MF_OGFM) {
            if (!ogfm_ghost<R>(I, fg, Wr, Wo)) bad = true;
            int* bb = sbb + 4 * fr;
            atomicMin(&bb[0], i - 2); atomicMax(&bb[1], i + 2);
            atomicMin(&bb[2], j - 2); atomicMax(&bb[3], j + 2);
        }
        const ExOut<R> o = extrap_out<R, MOUT>(I, fr, Wr, Wo);
        ExOut<R> oa = o;
        if ((MOUT & MF_OGFM) && (i < 2 * g || i >= I.Ny || j < 2 * g || j >= I.Nx)) {
            // GFM_original's rule applies to interior cells only
            const V4<R> Ug = (MOUT & MF_LAST) ? to_cons<R>(I, fg, Wg) : Wg;
            if (fr == 0) oa.b = Ug; else oa.a = Ug;
        }
        tb_put_owned<R, MOUT, PY>(I, i, j, bout, o, oa, rfb, rft, rfx);
    }
    if (bad) flag(I, ST_UNPHYS);
    if (MOUT & MF_OGFM) {
        __syncthreads();
        if (threadIdx.x < 8) {
            const int v = sbb[threadIdx.x];
            int* dst = I.si + SI_BBOX + threadIdx.x;
            if (threadIdx.x & 1) { if (v != INT_MIN) atomicMax(dst, v); }
            else if (v != INT_MAX) atomicMin(dst, v);
        }
    }
}

// ---------------------------------------------------------------------------
// interfacial cells (external Levelset::is_interfacial_cell, restated as a
// 4-neighbour change of side with the side taken as phi <= 0).
// work: {inst, tile_y, tile_x, tile index}, 64x4 interior tiles
// ---------------------------------------------------------------------------

// The list is built in row-major cell order (the order of the reference's
// i, j loops), so that every interfacial cell an entry depends on has a
// smaller entry index: per row segment a count (k_iface_count, or k_dt with
// `count`), an exclusive scan of the counts per instance (k_iface_scan), then
// each segment that has interfacial cells writes them in column order from
// its offset (k_iface_mark).  imap holds the list index of each listed cell
// and -1 elsewhere; the previous list is erased entry by entry before the
// count.

// interfacial cells of a segment: bit q of the result is set for cell
// q*64 + lane
template <typename R>
__device__ __forceinline__ unsigned seg_iface(const Inst& I, const int4& w, int lane) {
    const int i = I.g + w.y, j0 = I.g + w.z * SEGW + lane, jend = I.g + I.Nx;
    const long long row = cidx(I, i, 0);
    R ph[SEG_Q];
#pragma unroll
    for (int k = 0; k < SEG_Q; ++k) ph[k] = j0 + 64 * k <= jend ? ld1<R>(I.phi, row + j0 + 64 * k) : R(0);
    return seg_iface_ph<R>(I, row, j0, jend, lane, ph);
}

template <typename R>
__global__ __launch_bounds__(256) void k_iface_count(const Inst* __restrict__ insts, const int4* __restrict__ work,
                                                     int nwork) {
    const int sw = blockIdx.x * SEG_PER_WG + (int)(threadIdx.x >> 6);
    if (sw >= nwork) return;
    const int4 w = work[sw];
    const Inst& I = insts[w.x];
    const int lane = threadIdx.x & 63;
    erase_prev_list(I, w.w, lane);
    if (!gmem(I.si)[SI_ACTIVE]) return;
    const unsigned m = seg_iface<R>(I, w, lane);
    int nc = 0;
#pragma unroll
    for (int k = 0; k < SEG_Q; ++k) nc += __popcll(__ballot((m >> k) & 1u));
    if (lane == 0) I.iseg[w.w] = nc;
}

// one workgroup per instance: exclusive scan of the segment counts in place;
// the total is the number of interfacial cells
__global__ __launch_bounds__(1024) void k_iface_scan(const Inst* __restrict__ insts, const int4* __restrict__ work) {
    __shared__ int wsum[16];   // inclusive prefix of the wavefront totals
    const Inst& I = insts[work[blockIdx.x].x];
    if (!gmem(I.si)[SI_ACTIVE]) return;
    const int n = I.nseg, tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int per = (n + 1023) / 1024;
    const int b0 = min(n, tid * per), b1 = min(n, b0 + per);
    // the thread's chunk of counts: in registers (all loads in flight at
    // once) when it has at most PC of them
    constexpr int PC = 32;
    int vals[PC];
    int sum = 0;
    if (per <= PC) {
#pragma unroll
        for (int u = 0; u < PC; ++u) vals[u] = b0 + u < b1 ? I.iseg[b0 + u] : 0;
#pragma unroll
        for (int u = 0; u < PC; ++u) sum += vals[u];
    } else {
        for (int q = b0; q < b1; ++q) sum += I.iseg[q];
    }
    // inclusive scan of the chunk sums: within each wavefront by shuffles,
    // then over the sixteen wavefront totals
    int x = sum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    if (wid == 0) {
        int t = lane < 16 ? wsum[lane] : 0;
#pragma unroll
        for (int d = 1; d < 16; d <<= 1) {
            const int y = __shfl_up(t, d);
            if (lane >= d) t += y;
        }
        if (lane < 16) wsum[lane] = t;
    }
    __syncthreads();
    int acc = x - sum + (wid > 0 ? wsum[wid - 1] : 0);
    if (per <= PC) {
#pragma unroll
        for (int u = 0; u < PC; ++u) {
            if (b0 + u < b1) I.iseg[b0 + u] = acc;
            acc += vals[u];
        }
    } else {
        for (int q = b0; q < b1; ++q) {
            const int c = I.iseg[q];
            I.iseg[q] = acc;
            acc += c;
        }
    }
    if (tid == 1023) {
        const int total = wsum[15];
        gmem(I.si)[SI_NIFACE] = total;
        gmem(I.si)[SI_NLIST] = min(total, I.icap);   // the entries k_iface_mark records
    }
}

template <typename R>
__global__ __launch_bounds__(256) void k_iface_mark(const Inst* __restrict__ insts, const int4* __restrict__ work,
                                                    int nwork) {
    const int sw = blockIdx.x * SEG_PER_WG + (int)(threadIdx.x >> 6);
    if (sw >= nwork) return;
    const int4 w = work[sw];
    const Inst& I = insts[w.x];
    if (!gmem(I.si)[SI_ACTIVE]) return;
    const int lane = threadIdx.x & 63;
    int k = I.iseg[w.w];
    const int next = w.w + 1 < I.nseg ? I.iseg[w.w + 1] : gmem(I.si)[SI_NIFACE];
    if (next == k) return;   // no interfacial cell in this segment
    const unsigned m = seg_iface<R>(I, w, lane);
    const int i = I.g + w.y;
    bool ovf = false, bad = false;
#pragma unroll
    for (int q = 0; q < SEG_Q; ++q) {
        const bool f = (m >> q) & 1u;
        const unsigned long long mask = __ballot(f);
        if (f) {
            const int kk = k + __popcll(mask & ((1ull << lane) - 1ull));
            const int j = I.g + w.z * SEGW + q * 64 + lane;
            if (kk >= next) {
                bad = true;   // more cells than the segment's count: never write past its range
            } else if (kk < I.icap) {
                I.ilist[kk] = i * I.W + j;
                I.imap[cidx(I, i, j)] = kk;
            } else {
                ovf = true;
            }
        }
        k += __popcll(mask);
    }
    if (k < next) bad = true;   // fewer: entries of the range would keep stale cells
    if (ovf) flag(I, ST_IOVF);
    if (bad) flag(I, ST_ILIST);
}

// ---------------------------------------------------------------------------
// Interfacial mixed Riemann problems (GFM_real.cpp:52-208,
// GFM_riemann.cpp:51-148, GFM_modified.cpp:65-200), one workgroup per
// instance, FP64.
//
// GFM_real and GFM_riemann overwrite the real fluid's primitives in place
// while they loop over rows i and columns j, so a cell whose probe reads an
// interfacial cell that comes earlier in that order sees the earlier cell's
// star state.  The kernel reproduces that order exactly: results go to a
// side buffer; a cell is computed once every earlier interfacial cell it
// reads has been computed (dependency rounds separated by barriers), and it
// reads those cells' star states and every other cell's original value.
// GFM_modified writes to separate arrays and has no such dependency.
// ---------------------------------------------------------------------------

// area fraction on the side n . (x - c) + phi <= 0 of a dx x dy cell, for the
// line through the cell centre c with phi's value and unit normal (nx, ny)
// there (Scardovelli and Zaleski, J. Comput. Phys. 164 (2000), 2D case): with
// a = |nx| dx <= b = |ny| dy (swapped if not) and m = -phi + (a + b)/2, the
// fraction is 0 (m <= 0), m^2 / (2ab) (m < a), (2m - a) / (2b) (m < b),
// 1 - (a + b - m)^2 / (2ab) (m < a + b), else 1
__host__ __device__ inline double area_fraction(double phi, double nx, double ny, double dx, double dy) {
#pragma clang fp contract(off)
    double a = fabs(nx) * dx, b = fabs(ny) * dy;
    if (a > b) { const double s = a; a = b; b = s; }
    const double m = -phi + 0.5 * (a + b);
    if (m <= 0.0) return 0.0;
    if (m >= a + b) return 1.0;
    if (m < a) return m * m / (2.0 * a * b);
    if (m < b) return (2.0 * m - a) / (2.0 * b);
    const double r = a + b - m;
    return 1.0 - r * r / (2.0 * a * b);
}


struct Probe { int i0, j0; double tx, ty; };

__device__ __forceinline__ Probe bilinear_stencil(const Inst& I, double x, double y) {
    // external grid_bilinear_interpolate, restated: cell-centred, clamped
    const double fj = (x - I.x0) / I.dx - 0.5 + I.g;
    const double fi = (y - I.y0) / I.dy - 0.5 + I.g;
    Probe p;
    p.j0 = (int)floor(fj);
    p.i0 = (int)floor(fi);
    p.j0 = min(max(p.j0, 0), I.W - 2);
    p.i0 = min(max(p.i0, 0), I.H - 2);
    p.tx = fmin(fmax(fj - p.j0, 0.0), 1.0);
    p.ty = fmin(fmax(fi - p.i0, 0.0), 1.0);
    return p;
}

// probe positions of GFM_base.cpp:4-33 (fluid 2 at +n -> pL, fluid 1 at -n -> pR)
template <typename R>
__device__ __forceinline__ void probes(const Inst& I, int i, int j, Probe& p2, Probe& p1) {
    const double cx = I.x0 + (j - I.g + 0.5) * I.dx, cy = I.y0 + (i - I.g + 0.5) * I.dy;
    double lsval = (double)ld1<R>(I.phi, cidx(I, i, j));
    double nx = ((double)ld1<R>(I.phi, cidx(I, i, j + 1)) - (double)ld1<R>(I.phi, cidx(I, i, j - 1))) / (2.0 * I.dx);
    double ny = ((double)ld1<R>(I.phi, cidx(I, i + 1, j)) - (double)ld1<R>(I.phi, cidx(I, i - 1, j))) / (2.0 * I.dy);
    double m = sqrt(nx * nx + ny * ny);
    if (m < 1e-12) { nx = 1.0; ny = 0.0; m = 1.0; }
    const double ds = fmin(I.dx, I.dy);
    if (fabs(lsval) > ds) lsval = copysign(ds, lsval);
    const double hx = nx / m, hy = ny / m;
    const double ix = cx - lsval * hx, iy = cy - lsval * hy;
    p2 = bilinear_stencil(I, ix + ds * hx, iy + ds * hy);
    p1 = bilinear_stencil(I, ix - ds * hx, iy - ds * hy);
}

// dependency of a read of fluid f at cell (ii,jj) by the cell with linear index lin
template <typename R>
__device__ __forceinline__ int dep_of(const Inst& I, int ii, int jj, int f, int lin) {
    const long long c = cidx(I, ii, jj);
    const int k = I.imap[c];
    if (k < 0) return -1;
    if (ii * I.W + jj >= lin) return -1;
    const int fr = ld1<R>(I.phi, c) <= R(0) ? 0 : 1;
    return fr == f ? k : -1;
}

// shared normal-frame mixed RP (e.g. GFM_riemann.cpp:76-121); p0 is the
// Newton starting pressure (see rs_star)
// FAST: pow_fast in place of pow (mixed-precision contexts)
template <bool FAST>
__device__ __forceinline__ unsigned iface_mrp(const Inst& I, const double* ip1, const double* ip2, double nx, double ny,
                                              Star& S, double* t1, double* t2, double p0) {
    const double u1n = ip1[1] * nx + ip1[2] * ny;
    const double u2n = ip2[1] * nx + ip2[2] * ny;
    t1[0] = ip1[1] - u1n * nx; t1[1] = ip1[2] - u1n * ny;
    t2[0] = ip2[1] - u2n * nx; t2[1] = ip2[2] - u2n * ny;
    // Lprims/Rprims with the 1e-6 pressure floor, through W->U->W as the
    // reference passes conserved states to the mixed solver
    V4<double> Lp, Rp;
    Lp.v[0] = ip1[0]; Lp.v[1] = u1n; Lp.v[2] = 0.0; Lp.v[3] = fmax(1e-6, ip1[3]);
    Rp.v[0] = ip2[0]; Rp.v[1] = u2n; Rp.v[2] = 0.0; Rp.v[3] = fmax(1e-6, ip2[3]);
    if constexpr (FAST) {
        // mixed precision: the floored primitives go to the solver directly
        // (the round trip only adds rounding); the same positivity test in
        // division-free form
        unsigned st = 0;
        const RSFluid& F1 = I.rsf[0];
        const RSFluid& F2 = I.rsf[1];
        if (!(Lp.v[0] > 0.0 && Lp.v[3] + F1.g * F1.pinf >= 0.0) || !(Rp.v[0] > 0.0 && Rp.v[3] + F2.g * F2.pinf >= 0.0))
            st |= ST_UNPHYS;
        RSSideM L, R;
        rs_side_m(L, F1, Lp.v[0], Lp.v[1], Lp.v[3]);
        rs_side_m(R, F2, Rp.v[0], Rp.v[1], Rp.v[3]);
        // tolerance 1e-7: with Newton's quadratic convergence the star
        // pressure is then within ~1e-13 relative, far below the FP32 storage
        // it goes to
        st |= rs_star_m2(F1, F2, L, R, S, p0, 1e-7);
        return st;
    }
    const V4<double> Ls = prims2cons(I.gam[0], I.pinf[0], Lp);
    const V4<double> Rs = prims2cons(I.gam[1], I.pinf[1], Rp);
    unsigned st = 0;
    if (!phys_state(Ls) || !phys_state(Rs)) st |= ST_UNPHYS;
    const V4<double> Lw = cons2prims(I.gam[0], I.pinf[0], Ls);
    const V4<double> Rw = cons2prims(I.gam[1], I.pinf[1], Rs);
    RSSide L, R;
    rs_side(L, I.gam[0], I.pinf[0], Lw.v[0], Lw.v[1], Lw.v[3]);
    rs_side(R, I.gam[1], I.pinf[1], Rw.v[0], Rw.v[1], Rw.v[3]);
    st |= rs_star<FAST>(L, R, S, p0, 1e-12);
    return st;
}

// original (unmodified) primitives of fluid f at a cell: the conversion of
// the conserved grid for GFM_real, the extrapolated primitives otherwise
template <typename R, int VAR>
__device__ __forceinline__ void orig_prims(const Inst& I, int f, int ii, int jj, int pbuf, double* W) {
    const long long c = cidx(I, ii, jj);
    if (VAR == V_REAL) {
        // GFM_real converts the stored state; in FP64 also for FP32 storage
        // (an interfacial star-state input, O(perimeter) work)
        const V4<R> U = ld4<R>(I.Q[f][0], I, c);
        V4<double> Ud;
        for (int k = 0; k < 4; ++k) Ud.v[k] = (double)U.v[k];
        const V4<double> P = cons2prims<double>(I.gam[f], I.pinf[f], Ud);
        W[0] = P.v[0]; W[1] = P.v[1]; W[2] = P.v[2]; W[3] = P.v[3];
        return;
    }
    const V4<R> P = ld4<R>(I.P[f][pbuf], I, c);
    W[0] = P.v[0]; W[1] = P.v[1]; W[2] = P.v[2]; W[3] = P.v[3];
}

// per interfacial entry: the normal, the corner values of both probes (index
// 0: fluid 1 at -n, 1: fluid 2 at +n; GFM_real uses corner 0 with weight 1)
// and which corners are earlier interfacial cells whose star state replaces
// the stored value
struct alignas(16) IfRec {
    double nx, ny, lsval, pspec, preal;
    double tx[2], ty[2];
    double cv[2][4][4];
    int dep[2][4];
};
static_assert(sizeof(IfRec) <= IFREC_BYTES, "interfacial record larger than its slot");

// a mixed cell of the VOF GFMs: its unit normal and z, its original
// primitives of both fluids, and its probes (index 0: fluid 1 at -n, 1:
// fluid 2 at +n) with the corner primitives and which corners are earlier
// mixed cells, whose new primitives replace the stored ones
struct alignas(16) VofRec {
    double nx, ny, z, pspec;
    double own[2][4];
    double tx[2], ty[2];
    double cv[2][4][4];
    int dep[2][4];
};
static_assert(sizeof(VofRec) <= VOFREC_BYTES, "mixed-cell record larger than its slot");

__device__ __forceinline__ void bilerp(const double (*cv)[4], double tx, double ty, double* out) {
    for (int k = 0; k < 4; ++k)
        out[k] = (1.0 - ty) * ((1.0 - tx) * cv[0][k] + tx * cv[1][k]) + ty * ((1.0 - tx) * cv[2][k] + tx * cv[3][k]);
}

// star state of one interfacial cell from its probe states; writes the
// side-buffer entry (GFM_riemann.cpp:97-145, GFM_real.cpp:152-205,
// GFM_modified.cpp:113-196)
template <int VAR, bool FAST, typename Rec>
__device__ __forceinline__ unsigned iface_finish(const Inst& I, const Rec& r, const double* ip1, const double* ip2,
                                                 double p0, double* out, double* pstar) {
    Star S;
    double t1[2], t2[2];
    unsigned st = iface_mrp<FAST>(I, ip1, ip2, r.nx, r.ny, S, t1, t2, p0);
    if (VAR == V_MODIFIED) {
        // the real fluid keeps its state except the density; the ghost takes
        // the star state with the real fluid's tangential velocity (no-slip)
        // or, with slip BCs, its own (GFM_modified.cpp:152-169)
        const double* tr = r.lsval <= 0.0 ? (I.use_slip ? t2 : t1) : (I.use_slip ? t1 : t2);
        const int fr = r.lsval <= 0.0 ? 0 : 1;
        const double rs_real = fr == 0 ? S.rl : S.rr;
        out[0] = rs_real * rpow<FAST>((r.preal + I.pinf[fr]) / (S.p + I.pinf[fr]), 1.0 / I.gam[fr]);
        out[1] = fr == 0 ? S.rr : S.rl;
        out[2] = S.u * r.nx + tr[0];
        out[3] = S.u * r.ny + tr[1];
        out[4] = S.p;
    } else if (r.lsval <= 0.0) {
        out[0] = S.rl; out[1] = S.u * r.nx + t1[0]; out[2] = S.u * r.ny + t1[1]; out[3] = S.p;
    } else {
        out[0] = S.rr; out[1] = S.u * r.nx + t2[0]; out[2] = S.u * r.ny + t2[1]; out[3] = S.p;
    }
    if (VAR != V_MODIFIED) out[4] = S.p;
    out[5] = S.u;   // u* n is the cell's extension velocity (e.g. GFM_riemann.cpp:145)
    if (S.p != S.p) st |= ST_RS;
    *pstar = S.p;
    return st;
}

// probe states from a record; dependent corners read the side buffer
template <int VAR>
__device__ __forceinline__ void probe_values(const Inst& I, const IfRec& r, bool use_side, double* ip1, double* ip2) {
    double cv[2][4][4];
    for (int p = 0; p < 2; ++p)
        for (int q = 0; q < 4; ++q) {
            const int d = r.dep[p][q];
            const double* src = (use_side && d >= 0) ? I.side + 8 * (long long)d : r.cv[p][q];
            for (int k = 0; k < 4; ++k) cv[p][q][k] = src[k];
        }
    if (VAR == V_REAL) {
        for (int k = 0; k < 4; ++k) { ip1[k] = cv[0][0][k]; ip2[k] = cv[1][0][k]; }
    } else {
        bilerp(cv[0], r.tx[0], r.ty[0], ip1);
        bilerp(cv[1], r.tx[1], r.ty[1], ip2);
    }
}

// probe states of a record in global memory, dependent corners from the
// side buffer
template <int VAR, typename Rec>
__device__ __forceinline__ void probe_values_side(const Inst& I, const Rec& r, double* ip1, double* ip2) {
    double cv[2][4][4];
    const auto side = gmem(I.side);
    for (int p = 0; p < 2; ++p)
        for (int q = 0; q < 4; ++q) {
            const int d = r.dep[p][q];
            const auto src = d >= 0 ? side + 8 * (long long)d : &r.cv[p][q][0];
            for (int k = 0; k < 4; ++k) cv[p][q][k] = src[k];
        }
    if (VAR == V_REAL) {
        for (int k = 0; k < 4; ++k) { ip1[k] = cv[0][0][k]; ip2[k] = cv[1][0][k]; }
    } else {
        bilerp(cv[0], r.tx[0], r.ty[0], ip1);
        bilerp(cv[1], r.tx[1], r.ty[1], ip2);
    }
}

// Pass 1, parallel over interfacial entries: build the records; entries
// that read no earlier interfacial cell are final here, the others get a
// speculative p* (from the unmodified states) as their Newton start.
// work: {inst, first entry of a 256-entry chunk, 0, 0}
template <typename R, int VAR>
__global__ __launch_bounds__(256) void k_iface_prep(const Inst* __restrict__ insts, const int4* __restrict__ work,
                                                    int pbuf) {
    const int4 w = work[blockIdx.x];
    const Inst& I = insts[w.x];
    if (!gmem(I.si)[SI_ACTIVE]) return;
    const int n = min(gmem(I.si)[SI_NIFACE], I.icap);
    const int k = w.y + threadIdx.x;
    if (k >= n) return;
    int lin = I.ilist[k];
    int i = lin / I.W, j = lin % I.W;
    if (lin < 0 || !interior(I, i, j)) {
        // only if the counts and the list disagreed (flagged): a valid cell keeps every access in bounds
        flag(I, ST_ILIST);
        i = I.g; j = I.g; lin = i * I.W + j;
    }
    IfRec r;
    r.lsval = (double)ld1<R>(I.phi, cidx(I, i, j));
    unit_normal<R>(I, i, j, r.nx, r.ny);
    r.preal = 0.0;
    for (int p = 0; p < 2; ++p) {
        r.tx[p] = r.ty[p] = 0.0;
        for (int q = 0; q < 4; ++q) r.dep[p][q] = -1;
    }
    unsigned st = 0;
    bool has_dep = false;
    if (VAR == V_REAL) {
        // GFM_real.cpp:77-128: neighbour across the interface whose normal is
        // closest in angle (first minimum in x-major order)
        double mind = 1e100;
        int io = -1000, jo = -1000;
        for (int xi = -1; xi <= 1; ++xi)
            for (int yi = -1; yi <= 1; ++yi) {
                const double o = (double)ld1<R>(I.phi, cidx(I, i + yi, j + xi));
                if (r.lsval * o <= 0.0) {
                    double ox, oy;
                    unit_normal<R>(I, i + yi, j + xi, ox, oy);
                    double dp = r.nx * ox + r.ny * oy;
                    dp = fmin(fmax(dp, -1.0), 1.0);
                    const double ad = acos(dp);
                    if (ad < mind) { io = i + yi; jo = j + xi; mind = ad; }
                }
            }
        if (io == -1000) { io = i; jo = j; st |= ST_UNPHYS; }
        const int fown = r.lsval <= 0.0 ? 0 : 1, fo = 1 - fown;
        orig_prims<R, VAR>(I, fown, i, j, pbuf, r.cv[fown][0]);
        orig_prims<R, VAR>(I, fo, io, jo, pbuf, r.cv[fo][0]);
        r.dep[fo][0] = dep_of<R>(I, io, jo, fo, lin);
        has_dep = r.dep[fo][0] >= 0;
    } else {
        Probe pr[2];
        probes<R>(I, i, j, pr[1], pr[0]);
        for (int p = 0; p < 2; ++p) {
            r.tx[p] = pr[p].tx;
            r.ty[p] = pr[p].ty;
            for (int q = 0; q < 4; ++q) {
                const int ii = pr[p].i0 + (q >> 1), jj = pr[p].j0 + (q & 1);
                orig_prims<R, VAR>(I, p, ii, jj, pbuf, r.cv[p][q]);
                if (VAR == V_RIEMANN) {
                    r.dep[p][q] = dep_of<R>(I, ii, jj, p, lin);
                    has_dep |= r.dep[p][q] >= 0;
                }
            }
        }
        if (VAR == V_MODIFIED) {
            const int fr = r.lsval <= 0.0 ? 0 : 1;
            r.preal = (double)ld4<R>(I.P[fr][pbuf], I, cidx(I, i, j)).v[3];
        }
    }
    double ip1[4], ip2[4], out[8], ps;
    probe_values<VAR>(I, r, false, ip1, ip2);
    st |= iface_finish<VAR, (sizeof(R) == 4)>(I, r, ip1, ip2, -1e300, out, &ps);
    r.pspec = ps;
    const auto recs = gmem((IfRec*)I.irec);
    recs[k] = r;
    gmem(I.idone)[k] = has_dep ? 0 : 1;
    if (!has_dep) {
        const auto o = gmem(I.side) + 8 * (long long)k;
        for (int q = 0; q < 6; ++q) o[q] = out[q];
    }
    flag(I, st);
}

// star state of pending entry k once every earlier interfacial cell it reads
// has its own (the side buffer holds those)
template <bool OG, bool FAST>
__device__ __forceinline__ unsigned vof_solve_pending(const Inst& I, const VofRec& r, int k);

template <int VAR, bool FAST, typename Rec>
__device__ __forceinline__ unsigned iface_solve_pending(const Inst& I, const Rec& r, int k) {
    if constexpr (VAR == V_VOF || VAR == V_OGFM_VOF) {
        return vof_solve_pending<VAR == V_OGFM_VOF, FAST>(I, r, k);
    } else {
        double ip1[4], ip2[4], out[8], ps;
        probe_values_side<VAR>(I, r, ip1, ip2);
        const unsigned st = iface_finish<VAR, FAST>(I, r, ip1, ip2, r.pspec, out, &ps);
        const auto o = gmem(I.side) + 8 * (long long)k;
        for (int q = 0; q < 6; ++q) o[q] = out[q];
        return st;
    }
}

// Pass 2, one workgroup per instance.  Entries are in row-major cell order
// and every entry an entry depends on comes earlier.  Each lane holds one
// entry, solves it as soon as the entries it reads are final (LDS flags,
// acquire / release fences for the side buffer) and then claims the next
// unclaimed entry.  Entries are claimed in increasing order, so the earliest
// unfinished entry is always held by a lane and has its inputs final: the
// pass cannot stall; a spin limit guards it anyway (ST_IROUNDS).  With `plain`
// set it runs plain rounds instead: solve whatever is ready, barrier,
// repeat.  Both give the same bits.
constexpr unsigned IFACE_SPIN_LIMIT = 1u << 24;

template <int VAR> struct RecOf { using type = IfRec; };
template <> struct RecOf<V_VOF> { using type = VofRec; };
template <> struct RecOf<V_OGFM_VOF> { using type = VofRec; };

// fin[k] != 0: entry k's star state is final (LDS bytes, or for lists longer
// than I.icap_lds pass 1's flags in I.idone, in global memory)
template <typename R, int VAR, typename FL>
__device__ __forceinline__ void iface_rounds_run(const Inst& I, int n, int plain, FL* fin, int (&nunk)[2],
                                                 int& nclaim) {
    constexpr bool FAST = sizeof(R) == 4;
    using Rec = typename RecOf<VAR>::type;
    const auto recs = gmem((const Rec*)I.irec);
    const int tid = threadIdx.x, nt = blockDim.x;
    if (tid == 0) nclaim = nt;
    __syncthreads();
    unsigned st = 0;
    int rounds = 0;
    if (!plain) {
        // every lane holds one entry at a time and claims the next unclaimed
        // one (in increasing order) once its entry is final
        unsigned spins = 0;
        int k = tid;
        bool have = k < n;
        int dep[8];
        auto claim = [&]() {
            while (have && fin[k]) {   // final after pass 1: take the next
                k = atomicAdd(&nclaim, 1);
                have = k < n;
            }
            if (have) {
#pragma unroll
                for (int q = 0; q < 8; ++q) dep[q] = recs[k].dep[q >> 2][q & 3];
            }
        };
        claim();
        while (true) {
            if (have) {
                bool ready = true;
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    if (dep[q] >= 0 && __atomic_load_n(&fin[dep[q]], __ATOMIC_RELAXED) == 0) ready = false;
                if (ready) {
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                    st |= iface_solve_pending<VAR, FAST>(I, recs[k], k);
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                    __atomic_store_n(&fin[k], (FL)1, __ATOMIC_RELAXED);
                    k = atomicAdd(&nclaim, 1);
                    have = k < n;
                    claim();
                }
            }
            if (__all(!have)) break;
            if (++spins > IFACE_SPIN_LIMIT) { st |= ST_IROUNDS; break; }
        }
        rounds = (int)min(spins, 1u << 30);
    } else {
        for (rounds = 0;; ++rounds) {
            if (tid == 0) nunk[rounds & 1] = 0;
            __syncthreads();
            for (int k = tid; k < n; k += nt) {
                if (fin[k]) continue;
                bool ready = true;
                for (int a = 0; a < 2; ++a)
                    for (int q = 0; q < 4; ++q) {
                        const int d = recs[k].dep[a][q];
                        if (d >= 0 && fin[d] != 1) ready = false;
                    }
                if (!ready) { atomicAdd(&nunk[rounds & 1], 1); continue; }
                st |= iface_solve_pending<VAR, FAST>(I, recs[k], k);
                fin[k] = 2;
            }
            __syncthreads();
            for (int k = tid; k < n; k += nt)
                if (fin[k] == 2) fin[k] = 1;
            __syncthreads();
            if (nunk[rounds & 1] == 0) break;
            if (rounds > n + 1) { if (tid == 0) flag(I, ST_IROUNDS); break; }
        }
    }
    flag(I, st);
    if (tid == 0) gmem(I.si)[SI_ROUNDS] = rounds + 1;
}

template <typename R, int VAR>
__global__ __launch_bounds__(512) void k_iface_rounds(const Inst* __restrict__ insts, const int4* __restrict__ work,
                                                      int plain) {
    __shared__ unsigned char fin[ICAP_LDS];   // 1: star state final
    __shared__ int nunk[2];                   // per round parity (plain rounds)
    __shared__ int nclaim;                    // next unclaimed entry
    const int4 w = work[blockIdx.x];
    const Inst& I = insts[w.x];
    if (!gmem(I.si)[SI_ACTIVE]) return;
    const int n = min(gmem(I.si)[SI_NIFACE], I.icap);
    if (n <= I.icap_lds) {
        for (int k = threadIdx.x; k < n; k += blockDim.x) fin[k] = I.idone[k] ? 1 : 0;
        iface_rounds_run<R, VAR>(I, n, plain, fin, nunk, nclaim);
    } else {
        // a longer list keeps its flags where pass 1 wrote them (each entry's
        // flag is written by the lane that solves it; no later kernel reads them)
        iface_rounds_run<R, VAR>(I, n, plain, I.idone, nunk, nclaim);
    }
}


// GFM_modified write-back (GFM_modified.cpp:37-47, :230-239): the grids are
// W->U of the unextrapolated primitives with the interfacial changes; the
// velocity field is the real fluid's (the 6-iteration extrapolation at :202
// only changes ghost values and its result is not written back).
// work: full-array 64x4 tiles
// PRIMS: the new states go to the primitives P[.][qout] (for the extrapolation
// that follows) and the velocity field is left to that extrapolation's last
// iteration
template <typename R, bool PRIMS>
__global__ __launch_bounds__(256) void k_mgfm_final(const Inst* __restrict__ insts, const int4* __restrict__ work,
                                                    int qin, int qout) {
    const int4 w = work[blockIdx.x];
    const Inst& I = insts[w.x];
    if (!gmem(I.si)[SI_ACTIVE]) return;
    const int i = w.y * 4 + (threadIdx.x >> 6);
    const int j = w.z * 64 + (threadIdx.x & 63);
    if (i >= I.H || j >= I.W) return;
    const long long c = cidx(I, i, j);
    const V4<R> U1 = ld4<R>(I.Q[0][qin], I, c), U2 = ld4<R>(I.Q[1][qin], I, c);
    V4<R> W1 = cons2prims<R>((R)I.gam[0], (R)I.pinf[0], U1);
    V4<R> W2 = cons2prims<R>((R)I.gam[1], (R)I.pinf[1], U2);
    const int fr = ld1<R>(I.phi, c) <= R(0) ? 0 : 1;
    if (!PRIMS) {
        const auto vel = gmem((R*)I.vel);
        vel[c] = fr == 0 ? W1.v[1] : W2.v[1];
        vel[c + I.plane] = fr == 0 ? W1.v[2] : W2.v[2];
    }
    if (interior(I, i, j)) {
        const int k = I.imap[c];
        if (k >= 0) {
            const double* s = I.side + 8 * (long long)k;
            if (fr == 0) {
                W1.v[0] = (R)s[0];
                W2.v[0] = (R)s[1]; W2.v[1] = (R)s[2]; W2.v[2] = (R)s[3]; W2.v[3] = (R)s[4];
            } else {
                W2.v[0] = (R)s[0];
                W1.v[0] = (R)s[1]; W1.v[1] = (R)s[2]; W1.v[2] = (R)s[3]; W1.v[3] = (R)s[4];
            }
        }
    }
    if (PRIMS) {
        st4<R>(I.P[0][qout], I, c, W1);
        st4<R>(I.P[1][qout], I, c, W2);
        return;
    }
    if (sizeof(R) == 4 && !(interior(I, i, j) && I.imap[c] >= 0)) {
        // FP32 storage: both fluids of a cell that is not interfacial are
        // unchanged, so their stored states go out as they came in (the
        // W->U of :230-239 would return them up to an FP32 round trip)
        st4<R>(I.Q[0][qout], I, c, U1);
        st4<R>(I.Q[1][qout], I, c, U2);
        return;
    }
    st4<R>(I.Q[0][qout], I, c, prims2cons<R>((R)I.gam[0], (R)I.pinf[0], W1));
    st4<R>(I.Q[1][qout], I, c, prims2cons<R>((R)I.gam[1], (R)I.pinf[1], W2));
}

// ---------------------------------------------------------------------------
// VOF-based GFMs: GFM_VOF (GFM_VOF.cpp:8-325) and OGFM_VOF (OGFM_VOF.cpp:8-274).
// Their interface (GFM_VOF_interface: get_z, get_mixedcell_normal,
// get_interfacelocation; the reference's trackers are external) is the level
// set's PLIC reconstruction: in each cell the line through its centre with
// phi's value and unit normal there.  z is the area fraction on fluid 1's side
// (area_fraction), the normal points into fluid 2 and the interface location
// is the midpoint of the line's chord through the cell.
//
// The mixed cells (0 < z < 1) within three cells of the interior are listed in
// row-major order, the order of the reference's loops (GFM_VOF.cpp:73-216,
// OGFM_VOF.cpp:108-182), and solved in that order: a mixed cell's probes read
// the primitives of earlier mixed cells as those modified them in place (both
// fluids for GFM_VOF, the ghost fluid for OGFM_VOF), through the dataflow pass
// of the interfacial solve (k_iface_rounds).  A listed cell's side-buffer
// entry holds VSIDE doubles: its MRP states of fluid 1 and fluid 2 (GFM_VOF:
// the star states; OGFM_VOF: the probe states), then its new primitives of
// fluid 1 and fluid 2.  The final pass writes every cell: a listed cell's new
// primitives; a pure interior cell within three cells of listed ones the
// inverse-distance average of their MRP states of its ghost fluid (OGFM_VOF:
// isentropic to the real fluid's pressure, with its velocity); any other cell
// its own state.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool vof_range(const Inst& I, int i, int j) {
    return i >= I.g - 3 && i < I.g + I.Ny + 3 && j >= I.g - 3 && j < I.g + I.Nx + 3;
}

// z of a cell; the outermost ring has no centred gradient and takes 1 or 0 by
// phi's sign.  With |phi| >= (dx + dy)/2 >= (a + b)/2 the line misses the
// cell and area_fraction returns exactly 0 or 1 (m <= 0 or m >= a + b in
// floating point too: a <= |nx| dx and b <= |ny| dy round monotonically), so
// the normal is computed only nearer the interface.
template <typename R>
__device__ __forceinline__ double vof_z(const Inst& I, int i, int j) {
    const double ph = (double)ld1<R>(I.phi, cidx(I, i, j));
    if (i == 0 || j == 0 || i == I.H - 1 || j == I.W - 1) return ph <= 0.0 ? 1.0 : 0.0;
    if (fabs(ph) >= 0.5 * (I.dx + I.dy)) return ph <= 0.0 ? 1.0 : 0.0;
    double nx, ny;
    unit_normal<R>(I, i, j, nx, ny);
    return area_fraction(ph, nx, ny, I.dx, I.dy);
}

struct VofGeom { double z, nx, ny, lx, ly; };

// z, unit normal and interface location of a cell: the chord midpoint is the
// foot point c - phi n moved along the tangent (-ny, nx) to the middle of the
// parameter interval that the cell's sides clip (the foot point if the line
// misses the cell)
template <typename R>
__device__ __forceinline__ VofGeom vof_geom(const Inst& I, int i, int j) {
#pragma clang fp contract(off)
    VofGeom gm;
    const double p = (double)ld1<R>(I.phi, cidx(I, i, j));
    const double cx = I.x0 + (j - I.g + 0.5) * I.dx, cy = I.y0 + (i - I.g + 0.5) * I.dy;
    if (i == 0 || j == 0 || i == I.H - 1 || j == I.W - 1) {
        gm.z = p <= 0.0 ? 1.0 : 0.0;
        gm.nx = 1.0; gm.ny = 0.0; gm.lx = cx; gm.ly = cy;
        return gm;
    }
    unit_normal<R>(I, i, j, gm.nx, gm.ny);
    gm.z = area_fraction(p, gm.nx, gm.ny, I.dx, I.dy);
    const double fx = -p * gm.nx, fy = -p * gm.ny;
    double lo = -1e300, hi = 1e300;
    if (gm.ny != 0.0) {
        const double s1 = (fx - 0.5 * I.dx) / gm.ny, s2 = (fx + 0.5 * I.dx) / gm.ny;
        lo = fmax(lo, fmin(s1, s2));
        hi = fmin(hi, fmax(s1, s2));
    }
    if (gm.nx != 0.0) {
        const double s1 = (-0.5 * I.dy - fy) / gm.nx, s2 = (0.5 * I.dy - fy) / gm.nx;
        lo = fmax(lo, fmin(s1, s2));
        hi = fmin(hi, fmax(s1, s2));
    }
    const double sm = lo <= hi ? 0.5 * (lo + hi) : 0.0;
    gm.lx = cx + (fx - sm * gm.ny);
    gm.ly = cy + (fy + sm * gm.nx);
    return gm;
}

// FP64 primitives of fluid f from its conserved state U
template <typename R>
__device__ __forceinline__ void vof_prims_of(const Inst& I, int f, const V4<R>& U, double* W) {
    V4<double> Ud;
    for (int k = 0; k < 4; ++k) Ud.v[k] = (double)U.v[k];
    const V4<double> P = cons2prims<double>(I.gam[f], I.pinf[f], Ud);
    for (int k = 0; k < 4; ++k) W[k] = P.v[k];
}

// FP64 primitives of fluid f at a cell, from the stored conserved state
template <typename R>
__device__ __forceinline__ void vof_prims(const Inst& I, int f, int i, int j, double* W) {
    const V4<R> U = ld4<R>(I.Q[f][0], I, cidx(I, i, j));
    V4<double> Ud;
    for (int k = 0; k < 4; ++k) Ud.v[k] = (double)U.v[k];
    const V4<double> P = cons2prims<double>(I.gam[f], I.pinf[f], Ud);
    for (int k = 0; k < 4; ++k) W[k] = P.v[k];
}

// isentropic_extrapolation (OGFM_VOF.cpp:156, GFM_VOF.cpp:198)
template <bool FAST>
__device__ __forceinline__ double vof_isen(const Inst& I, int f, double rho_old, double p_old, double p_new) {
    return rho_old * rpow<FAST>((p_new + I.pinf[f]) / (p_old + I.pinf[f]), 1.0 / I.gam[f]);
}

// the mixed cell's MRP states and new primitives from its probe states
// (GFM_VOF.cpp:112-213; OGFM_VOF.cpp:148-167); out: VSIDE doubles
template <bool OG, bool FAST>
__device__ __forceinline__ unsigned vof_finish(const Inst& I, const VofRec& r, const double* ip1, const double* ip2,
                                               double p0, double* out, double* pstar) {
#pragma clang fp contract(off)
    double* M1 = out;
    double* M2 = out + 4;
    double* N1 = out + 8;
    double* N2 = out + 12;
    for (int k = 0; k < 4; ++k) { N1[k] = r.own[0][k]; N2[k] = r.own[1][k]; }
    if (OG) {
        for (int k = 0; k < 4; ++k) { M1[k] = ip1[k]; M2[k] = ip2[k]; }
        // the ghost fluid: the real fluid's velocity and pressure, its own
        // probe's density made isentropic to that pressure
        if (r.z > 0.5) {
            N2[0] = vof_isen<FAST>(I, 1, ip2[0], ip2[3], r.own[0][3]);
            N2[1] = r.own[0][1]; N2[2] = r.own[0][2]; N2[3] = r.own[0][3];
        } else {
            N1[0] = vof_isen<FAST>(I, 0, ip1[0], ip1[3], r.own[1][3]);
            N1[1] = r.own[1][1]; N1[2] = r.own[1][2]; N1[3] = r.own[1][3];
        }
        *pstar = 0.0;
        return 0;
    }
    Star S;
    double t1[2], t2[2];
    unsigned st = iface_mrp<FAST>(I, ip1, ip2, r.nx, r.ny, S, t1, t2, p0);
    // slip: each fluid keeps its tangential velocity; no-slip: both take the
    // majority fluid's (GFM_VOF.cpp:161-178)
    const double* ta = I.use_slip ? t1 : (r.z > 0.5 ? t1 : t2);
    const double* tb = I.use_slip ? t2 : (r.z > 0.5 ? t1 : t2);
    M1[0] = S.rl; M1[1] = S.u * r.nx + ta[0]; M1[2] = S.u * r.ny + ta[1]; M1[3] = S.p;
    M2[0] = S.rr; M2[1] = S.u * r.nx + tb[0]; M2[2] = S.u * r.ny + tb[1]; M2[3] = S.p;
    // the majority fluid keeps its state but for the density, isentropic
    // from its star state to its own pressure; the other takes its star state
    if (r.z > 0.5) {
        N1[0] = vof_isen<FAST>(I, 0, S.rl, S.p, r.own[0][3]);
        for (int k = 0; k < 4; ++k) N2[k] = M2[k];
    } else {
        N2[0] = vof_isen<FAST>(I, 1, S.rr, S.p, r.own[1][3]);
        for (int k = 0; k < 4; ++k) N1[k] = M1[k];
    }
    if (S.p != S.p) st |= ST_RS;
    *pstar = S.p;
    return st;
}

// probe states of a record: a dependent corner reads the new primitives of
// the earlier mixed cell from the side buffer (GFM_base.cpp:36-56)
template <bool SIDE>
__device__ __forceinline__ void vof_probe_values(const Inst& I, const VofRec& r, double* ip1, double* ip2) {
    double cv[2][4][4];
    const auto side = gmem(I.side);
    for (int p = 0; p < 2; ++p)
        for (int q = 0; q < 4; ++q) {
            const int d = SIDE ? r.dep[p][q] : -1;
            for (int k = 0; k < 4; ++k) cv[p][q][k] = d >= 0 ? side[(long long)VSIDE * d + 8 + 4 * p + k] : r.cv[p][q][k];
        }
    bilerp(cv[0], r.tx[0], r.ty[0], ip1);
    bilerp(cv[1], r.tx[1], r.ty[1], ip2);
}

template <bool OG, bool FAST>
__device__ __forceinline__ unsigned vof_solve_pending(const Inst& I, const VofRec& r, int k) {
    double ip1[4], ip2[4], out[VSIDE], ps;
    vof_probe_values<true>(I, r, ip1, ip2);
    const unsigned st = vof_finish<OG, FAST>(I, r, ip1, ip2, r.pspec, out, &ps);
    const auto o = gmem(I.side) + (long long)VSIDE * k;
    for (int q = 0; q < VSIDE; ++q) o[q] = out[q];
    return st;
}

// mixed cells of a row segment of the grown interior: bit q for cell q*64 + lane
template <typename R>
__device__ __forceinline__ unsigned vof_seg_mixed(const Inst& I, const int4& w, int lane) {
    const int i = I.g - 3 + w.y, jend = I.g + I.Nx + 3;
    unsigned m = 0;
#pragma unroll
    for (int q = 0; q < SEG_Q; ++q) {
        const int j = I.g - 3 + w.z * SEGW + q * 64 + lane;
        if (j < jend) {
            const double z = vof_z<R>(I, i, j);
            if (0.0 < z && z < 1.0) m |= 1u << q;
        }
    }
    return m;
}

// erases the previous mixed-cell list from imap (entries [s SEGW, (s+1) SEGW)
// for segment s, which covers the list: at most one entry per cell of the
// grown interior) and counts the segment's mixed cells
// work: {inst, row of the grown interior, segment column, segment index}
template <typename R>
__global__ __launch_bounds__(256) void k_vof_count(const Inst* __restrict__ insts, const int4* __restrict__ work,
                                                   int nwork) {
    const int sw = blockIdx.x * SEG_PER_WG + (int)(threadIdx.x >> 6);
    if (sw >= nwork) return;
    const int4 w = work[sw];
    const Inst& I = insts[w.x];
    const int lane = threadIdx.x & 63;
    {
        const int n = gmem(I.si)[SI_NLIST];
#pragma unroll
        for (int q = 0; q < SEG_Q; ++q) {
            const int k = w.w * SEGW + q * 64 + lane;
            if (k < n) {
                const int v = I.ilist[k];
                const int i = v / I.W, j = v - (v / I.W) * I.W;
                if (v >= 0 && vof_range(I, i, j)) I.imap[cidx(I, i, j)] = -1;
            }
        }
    }
    if (!gmem(I.si)[SI_ACTIVE]) return;
    const unsigned m = vof_seg_mixed<R>(I, w, lane);
    int nc = 0;
#pragma unroll
    for (int k = 0; k < SEG_Q; ++k) nc += __popcll(__ballot((m >> k) & 1u));
    if (lane == 0) I.iseg[w.w] = nc;
}

// each segment with mixed cells lists them in column order from its offset
template <typename R>
__global__ __launch_bounds__(256) void k_vof_mark(const Inst* __restrict__ insts, const int4* __restrict__ work,
                                                  int nwork) {
    const int sw = blockIdx.x * SEG_PER_WG + (int)(threadIdx.x >> 6);
    if (sw >= nwork) return;
    const int4 w = work[sw];
    const Inst& I = insts[w.x];
    if (!gmem(I.si)[SI_ACTIVE]) return;
    const int lane = threadIdx.x & 63;
    int k = I.iseg[w.w];
    const int next = w.w + 1 < I.nseg ? I.iseg[w.w + 1] : gmem(I.si)[SI_NIFACE];
    if (next == k) return;
    const unsigned m = vof_seg_mixed<R>(I, w, lane);
    const int i = I.g - 3 + w.y;
    bool ovf = false, bad = false;
#pragma unroll
    for (int q = 0; q < SEG_Q; ++q) {
        const bool f = (m >> q) & 1u;
        const unsigned long long mask = __ballot(f);
        if (f) {
            const int kk = k + __popcll(mask & ((1ull << lane) - 1ull));
            const int j = I.g - 3 + w.z * SEGW + q * 64 + lane;
            if (kk >= next) {
                bad = true;
            } else if (kk < I.icap) {
                I.ilist[kk] = i * I.W + j;
                I.imap[cidx(I, i, j)] = kk;
            } else {
                ovf = true;
            }
        }
        k += __popcll(mask);
    }
    if (k < next) bad = true;
    if (ovf) flag(I, ST_IOVF);
    if (bad) flag(I, ST_ILIST);
}

// a read of fluid f at corner (ii, jj) by the mixed cell with linear index
// lin depends on an earlier mixed cell if that cell changed fluid f: both
// fluids for GFM_VOF, the ghost fluid (fluid 2 where z > 1/2) for OGFM_VOF
template <typename R, bool OG>
__device__ __forceinline__ int vof_dep(const Inst& I, int ii, int jj, int f, int lin) {
    const int k = I.imap[cidx(I, ii, jj)];
    if (k < 0 || ii * I.W + jj >= lin) return -1;
    if (OG) {
        const int fg = vof_z<R>(I, ii, jj) > 0.5 ? 1 : 0;
        if (fg != f) return -1;
    }
    return k;
}

// Pass 1, parallel over the mixed cells: records; a cell that reads no
// earlier mixed cell is final here, the others get a speculative p* (GFM_VOF)
// work: {inst, first entry of a 256-entry chunk, 0, 0}
template <typename R, bool OG>
__global__ __launch_bounds__(256) void k_vof_prep(const Inst* __restrict__ insts, const int4* __restrict__ work) {
#pragma clang fp contract(off)
    const int4 w = work[blockIdx.x];
    const Inst& I = insts[w.x];
    if (!gmem(I.si)[SI_ACTIVE]) return;
    const int n = min(gmem(I.si)[SI_NIFACE], I.icap);
    const int k = w.y + threadIdx.x;
    if (k >= n) return;
    int lin = I.ilist[k];
    int i = lin / I.W, j = lin % I.W;
    if (lin < 0 || !vof_range(I, i, j)) {
        flag(I, ST_ILIST);   // only if the counts and the list disagreed
        i = I.g; j = I.g; lin = i * I.W + j;
    }
    VofRec r;
    const VofGeom gm = vof_geom<R>(I, i, j);
    r.nx = gm.nx; r.ny = gm.ny; r.z = gm.z;
    vof_prims<R>(I, 0, i, j, r.own[0]);
    vof_prims<R>(I, 1, i, j, r.own[1]);
    const double ds = fmin(I.dx, I.dy);
    Probe pr[2];
    pr[0] = bilinear_stencil(I, gm.lx - ds * gm.nx, gm.ly - ds * gm.ny);   // fluid 1 at -n
    pr[1] = bilinear_stencil(I, gm.lx + ds * gm.nx, gm.ly + ds * gm.ny);   // fluid 2 at +n
    bool has_dep = false;
    for (int p = 0; p < 2; ++p) {
        r.tx[p] = pr[p].tx;
        r.ty[p] = pr[p].ty;
        for (int q = 0; q < 4; ++q) {
            const int ii = pr[p].i0 + (q >> 1), jj = pr[p].j0 + (q & 1);
            vof_prims<R>(I, p, ii, jj, r.cv[p][q]);
            r.dep[p][q] = vof_dep<R, OG>(I, ii, jj, p, lin);
            has_dep |= r.dep[p][q] >= 0;
        }
    }
    double ip1[4], ip2[4], out[VSIDE], ps;
    vof_probe_values<false>(I, r, ip1, ip2);
    const unsigned st = vof_finish<OG, (sizeof(R) == 4)>(I, r, ip1, ip2, -1e300, out, &ps);
    r.pspec = ps;
    gmem((VofRec*)I.irec)[k] = r;
    gmem(I.idone)[k] = has_dep ? 0 : 1;
    if (!has_dep) {
        const auto o = gmem(I.side) + (long long)VSIDE * k;
        for (int q = 0; q < VSIDE; ++q) o[q] = out[q];
    }
    flag(I, st);
}

// Final pass (GFM_VOF.cpp:221-319, OGFM_VOF.cpp:55-93, :171-180, :188-268):
// the new primitives of every cell, W->U into Q[.][1], the velocity field
// and, for OGFM_VOF, the update boxes.  A 64x4 tile keeps the list indices of
// its cells and a three-cell halo in LDS for the inverse-distance fill.
// work: full-array 64x4 tiles
template <typename R, bool OG>
__global__ __launch_bounds__(256) void k_vof_final(const Inst* __restrict__ insts, const int4* __restrict__ work) {
#pragma clang fp contract(off)
    constexpr int TW = 64 + 6, TH = 4 + 6;
    __shared__ int smap[TH][TW];
    __shared__ int sbb[8];
    const int4 w = work[blockIdx.x];
    const Inst& I = insts[w.x];
    if (!gmem(I.si)[SI_ACTIVE]) return;
    const int ti = w.y * 4, tj = w.z * 64;
    bool any = false;
    for (int q = threadIdx.x; q < TH * TW; q += 256) {
        const int a = q / TW, b = q - a * TW;
        const int ii = ti - 3 + a, jj = tj - 3 + b;
        int v = -1;
        if (ii >= 0 && ii < I.H && jj >= 0 && jj < I.W) v = I.imap[cidx(I, ii, jj)];
        smap[a][b] = v;
        any |= v >= 0;
    }
    if (OG && threadIdx.x < 8) sbb[threadIdx.x] = (threadIdx.x & 1) ? INT_MIN : INT_MAX;
    any = __syncthreads_or(any);
    const int li = threadIdx.x >> 6, lj = threadIdx.x & 63;
    const int i = ti + li, j = tj + lj;
    unsigned st = 0;
    if (i < I.H && j < I.W) {
        const long long c = cidx(I, i, j);
        const V4<R> U[2] = {ld4<R>(I.Q[0][0], I, c), ld4<R>(I.Q[1][0], I, c)};
        const double z = vof_z<R>(I, i, j);
        // GFM_VOF.cpp:59-66, OGFM_VOF.cpp:64-81: the real fluid's state
        const int fr = z >= 0.5 ? 0 : 1;
        if (!phys_state(U[fr])) st |= ST_UNPHYS;
        // W[f]: fluid f's new primitives where changed[f], else converted on demand
        double W[2][4];
        bool have[2] = {false, false}, changed[2] = {false, false};
        const int self = smap[li + 3][lj + 3];
        if (self >= 0) {
            const double* s = gmem(I.side) + (long long)VSIDE * self;
            for (int k = 0; k < 4; ++k) { W[0][k] = s[8 + k]; W[1][k] = s[12 + k]; }
            have[0] = have[1] = changed[0] = changed[1] = true;
        } else if (any && interior(I, i, j) && (z == 0.0 || z == 1.0)) {
            // the fluid whose ghost state is set: fluid 1 where z = 0 (flag 1)
            const int fg = z == 0.0 ? 0 : 1;
            double sw = 0.0, sp[4] = {0.0, 0.0, 0.0, 0.0};
            bool near = false;
            const double cx = I.x0 + (j - I.g + 0.5) * I.dx, cy = I.y0 + (i - I.g + 0.5) * I.dy;
            for (int xi = -3; xi <= 3; ++xi)
                for (int yi = -3; yi <= 3; ++yi) {
                    const int kk = smap[li + 3 + yi][lj + 3 + xi];
                    if (kk < 0) continue;
                    near = true;
                    const double ox = I.x0 + (j + xi - I.g + 0.5) * I.dx, oy = I.y0 + (i + yi - I.g + 0.5) * I.dy;
                    const double ex = ox - cx, ey = oy - cy;
                    const double wt = 1.0 / sqrt(ex * ex + ey * ey);
                    const double* s = gmem(I.side) + (long long)VSIDE * kk + 4 * fg;
                    sw += wt;
                    for (int k = 0; k < 4; ++k) sp[k] += wt * s[k];
                }
            if (near) {
                for (int k = 0; k < 4; ++k) sp[k] /= sw;
                double* G = W[fg];
                if (OG) {
                    const int fo = 1 - fg;
                    vof_prims_of<R>(I, fo, U[fo], W[fo]);
                    have[fo] = true;
                    G[0] = vof_isen<(sizeof(R) == 4)>(I, fg, sp[0], sp[3], W[fo][3]);
                    G[1] = W[fo][1]; G[2] = W[fo][2]; G[3] = W[fo][3];
                } else {
                    for (int k = 0; k < 4; ++k) G[k] = sp[k];
                }
                have[fg] = changed[fg] = true;
            }
        }
        // GFM_VOF converts every cell, OGFM_VOF the interior (:263-266); the
        // ghost frame is overwritten by the BCs that follow either way.  FP32
        // storage: an unchanged state goes out as it came in (the W->U would
        // return it up to an FP32 round trip)
        const bool conv_cell = !OG || interior(I, i, j);
        for (int f = 0; f < 2; ++f) {
            const bool convert = conv_cell && (sizeof(R) == 8 || changed[f]);
            if ((convert || f == fr) && !have[f]) {
                vof_prims_of<R>(I, f, U[f], W[f]);
                have[f] = true;
            }
            V4<R> Uo = U[f];
            if (convert) {
                V4<double> Wd;
                for (int k = 0; k < 4; ++k) Wd.v[k] = W[f][k];
                const V4<double> Ud = prims2cons<double>(I.gam[f], I.pinf[f], Wd);
                if (!phys_state(Ud)) st |= ST_UNPHYS;
                for (int k = 0; k < 4; ++k) Uo.v[k] = (R)Ud.v[k];
            }
            st4<R>(I.Q[f][1], I, c, Uo);
        }
        // the real fluid's velocity (GFM_VOF.cpp:306-319; OGFM_VOF.cpp:171-180,
        // within three cells of the interior only, zero beyond)
        const auto vel = gmem((R*)I.vel);
        const bool vin = !OG || vof_range(I, i, j);
        vel[c] = vin ? (R)W[fr][1] : R(0);
        vel[c + I.plane] = vin ? (R)W[fr][2] : R(0);
        if (OG) {
            // OGFM_VOF.cpp:64-81: each fluid's real cells grown by two
            int* bb = sbb + 4 * fr;
            atomicMin(&bb[0], i - 2); atomicMax(&bb[1], i + 2);
            atomicMin(&bb[2], j - 2); atomicMax(&bb[3], j + 2);
        }
    }
    flag(I, st);
    if (OG) {
        __syncthreads();
        if (threadIdx.x < 8) {
            const int v = sbb[threadIdx.x];
            int* dst = I.si + SI_BBOX + threadIdx.x;
            if (threadIdx.x & 1) { if (v != INT_MIN) atomicMax(dst, v); }
            else if (v != INT_MAX) atomicMin(dst, v);
        }
    }
}

// ---------------------------------------------------------------------------
// Extension velocity field (GFM_base::use_extension_vfield, GFM_base.hpp:20;
// extrapolate_extension_vfield.cpp:4-122, called with 20 iterations at
// GFM_original.cpp:141-144, GFM_real.cpp:212-215, GFM_riemann.cpp:152-155,
// GFM_modified.cpp:204-207).  The field starts at zero except in interior
// interfacial cells, which hold u* n from the interfacial solve
// (GFM_real.cpp:205, GFM_riemann.cpp:145, GFM_modified.cpp:196) or, for
// GFM_original, the real fluid's velocity (GFM_original.cpp:67-71, :93-97).
// Each Jacobi iteration moves every other interior cell's value along the
// normal, away from the interface on both sides; a cell whose phi gradient
// vanishes keeps its value.  BCs are folded in as a gather: a ghost cell takes
// its BC source cell's new value with the component normal to a reflective
// side negated (the external apply_gridVector2dtype_BCs, restated with the
// map of apply_BCs_euler).  The two planes ping-pong through P[0][0] and
// P[0][1], free once the fill has written the conserved grids; the last
// iteration writes the velocity field.  work: full-array 64x4 tiles
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ bool iface_cell(const Inst& I, int i, int j) {
    const auto ph = gmem((const R*)I.phi);
    const long long c = cidx(I, i, j);
    const bool s0 = ph[c] <= R(0);
    return ((ph[c + 1] <= R(0)) != s0) || ((ph[c - 1] <= R(0)) != s0) || ((ph[c + I.pitch] <= R(0)) != s0) ||
           ((ph[c - I.pitch] <= R(0)) != s0);
}

template <typename R, bool OGFM>
__global__ __launch_bounds__(256) void k_vext_init(const Inst* __restrict__ insts, const int4* __restrict__ work) {
    const int4 w = work[blockIdx.x];
    const Inst& I = insts[w.x];
    if (!gmem(I.si)[SI_ACTIVE]) return;
    const int i = w.y * 4 + (threadIdx.x >> 6);
    const int j = w.z * 64 + (threadIdx.x & 63);
    if (i >= I.H || j >= I.W) return;
    const long long c = cidx(I, i, j);
    R vx = R(0), vy = R(0);
    if (interior(I, i, j) && iface_cell<R>(I, i, j)) {
        if (OGFM) {
            const int fr = ld1<R>(I.phi, c) <= R(0) ? 0 : 1;
            V4<R> W;
            if constexpr (sizeof(R) == 4) W = cons2prims_f(fluid_f(I.rsf[fr]), ld4<R>(I.Q[fr][0], I, c));
            else W = cons2prims<R>((R)I.gam[fr], (R)I.pinf[fr], ld4<R>(I.Q[fr][0], I, c));
            vx = W.v[1];
            vy = W.v[2];
        } else {
            const int k = gmem(I.imap)[c];
            if (k >= 0 && k < I.icap) {
                const double us = gmem(I.side)[8 * (long long)k + 5];
                const auto r = gmem((const IfRec*)I.irec) + k;
                vx = (R)(us * r->nx);
                vy = (R)(us * r->ny);
            } else {
                flag(I, ST_ILIST);
            }
        }
    }
    const auto out = gmem((R*)I.P[0][0]);
    out[c] = vx;
    out[c + I.plane] = vy;
}

template <typename R>
__global__ __launch_bounds__(256) void k_vext_iter(const Inst* __restrict__ insts, const int4* __restrict__ work,
                                                   int vin, int last) {
#pragma clang fp contract(off)
    const int4 w = work[blockIdx.x];
    const Inst& I = insts[w.x];
    if (!gmem(I.si)[SI_ACTIVE]) return;
    const int i = w.y * 4 + (threadIdx.x >> 6);
    const int j = w.z * 64 + (threadIdx.x & 63);
    if (i >= I.H || j >= I.W) return;
    bool fv, fu;
    const int si = src_row(I, i, fv);
    const int sj = src_col(I, j, fu);
    const auto v = gmem((const R*)I.P[0][vin]);
    const long long cs = cidx(I, si, sj), P2 = I.plane;
    R vx = v[cs], vy = v[cs + P2];
    if (!iface_cell<R>(I, si, sj)) {   // extrapolate_extension_vfield.cpp:46
        const auto ph = gmem((const R*)I.phi);
        const R pe = ph[cs + 1], pw = ph[cs - 1], pn = ph[cs + I.pitch], ps = ph[cs - I.pitch];
        R nx, ny;
        bool ok;
        if constexpr (sizeof(R) == 8) {
            nx = (pe - pw) / (2.0 * I.dx);
            ny = (pn - ps) / (2.0 * I.dy);
            const double m = sqrt(nx * nx + ny * ny);
            ok = m > 1e-12;   // :52
            if (ok) { nx /= m; ny /= m; }
        } else {
            nx = (pe - pw) * (float)(0.5 / I.dx);
            ny = (pn - ps) * (float)(0.5 / I.dy);
            const float m2 = nx * nx + ny * ny;
            ok = m2 > 1e-24f;
            if (ok) { const float r = rsqrtf(m2); nx *= r; ny *= r; }
        }
        if (ok) {
            const double pdt = 0.5 * fmin(I.dx, I.dy);   // :15
            const R ax = (R)(pdt / I.dx) * nx, ay = (R)(pdt / I.dy) * ny;
            const R ve[2] = {v[cs + 1], v[cs + 1 + P2]}, vw[2] = {v[cs - 1], v[cs - 1 + P2]};
            const R vn[2] = {v[cs + I.pitch], v[cs + I.pitch + P2]}, vs[2] = {v[cs - I.pitch], v[cs - I.pitch + P2]};
            R o[2];
            const R c0[2] = {vx, vy};
            if (ph[cs] <= R(0)) {
                // information along -n (:59-83)
                for (int k = 0; k < 2; ++k) {
                    const R dx = nx > R(0) ? ve[k] - c0[k] : c0[k] - vw[k];
                    const R dy = ny > R(0) ? vn[k] - c0[k] : c0[k] - vs[k];
                    o[k] = (c0[k] + ax * dx) + ay * dy;
                }
            } else {
                // information along +n (:85-109)
                for (int k = 0; k < 2; ++k) {
                    const R dx = nx > R(0) ? c0[k] - vw[k] : ve[k] - c0[k];
                    const R dy = ny > R(0) ? c0[k] - vs[k] : vn[k] - c0[k];
                    o[k] = (c0[k] - ax * dx) - ay * dy;
                }
            }
            vx = o[0];
            vy = o[1];
        }
    }
    if (fu) vx = -vx;
    if (fv) vy = -vy;
    const auto out = gmem((R*)(last ? I.vel : I.P[0][1 - vin]));
    const long long c = cidx(I, i, j);
    out[c] = vx;
    out[c + P2] = vy;
}

// ---------------------------------------------------------------------------
// Initial conditions on the device (the host library gfm_ic.cpp, restated
// per cell): sim_onefluid.cpp:196-422 (TTC1-5, circular explosion with 10x10
// sub-cell mixing), sim_twofluid.cpp:774-859 (shocked bubble), :1245-1347
// (RMI interface), :675-742 (planar tubes).  FP64, every operation rounded on
// its own as the host code is compiled; work: full-array 64x4 tiles.
// ---------------------------------------------------------------------------
__constant__ double c_ttc[5][2][4] = {
    {{1.0, 0.0, 0.0, 1.0}, {0.125, 0.0, 0.0, 0.1}},
    {{1.0, -2.0, 0.0, 0.4}, {1.0, 2.0, 0.0, 0.4}},
    {{1.0, 0.0, 0.0, 1000.0}, {1.0, 0.0, 0.0, 0.01}},
    {{1.0, 0.0, 0.0, 0.1}, {1.0, 0.0, 0.0, 100.0}},
    {{5.99924, 19.5975, 0.0, 460.894}, {5.99242, -6.19633, 0.0, 46.0950}},
};

// euler_misc.hpp:89-108 with the stiffened-gas energy
__device__ __forceinline__ V4<double> ic_cons(double g, double pinf, const double* W) {
#pragma clang fp contract(off)
    const double rho = W[0], u = W[1], v = W[2], p = W[3];
    const double e = (p + g * pinf) / ((g - 1.0) * rho);
    V4<double> U;
    U.v[0] = rho;
    U.v[1] = rho * u;
    U.v[2] = rho * v;
    U.v[3] = rho * e + 0.5 * rho * (u * u + v * v);
    return U;
}

template <typename R>
__device__ __forceinline__ void ic_put(void* base, const Inst& I, long long c, const V4<double>& U) {
    V4<R> a;
    for (int k = 0; k < 4; ++k) a.v[k] = (R)U.v[k];
    st4<R>(base, I, c, a);
}

template <typename R>
__global__ __launch_bounds__(256) void k_ic(const Inst* __restrict__ insts, const int4* __restrict__ work,
                                            const IcDesc* __restrict__ descs, const int* __restrict__ map, int q,
                                            int twofluid) {
#pragma clang fp contract(off)
    const int4 w = work[blockIdx.x];
    const int d = map[w.x];
    if (d < 0) return;
    const Inst& I = insts[w.x];
    const IcDesc& D = descs[d];
    const int i = w.y * 4 + (threadIdx.x >> 6);
    const int j = w.z * 64 + (threadIdx.x & 63);
    if (i >= I.H || j >= I.W) return;
    const long long c = cidx(I, i, j);
    const double x = I.x0 + (j - I.g + 0.5) * I.dx, y = I.y0 + (i - I.g + 0.5) * I.dy;   // cellcentre_coord
    const double g1 = I.gam[0], p1 = I.pinf[0], g2 = I.gam[1], p2 = I.pinf[1];
    double phi = 0.0;
    switch (D.kind) {
        case 1: {   // TTC: set_planar_IC (sim_onefluid.cpp:426-449), a x + b y <= 0.5 -> left
            const int k = D.test_case % 10 - 1;
            const bool ydir = D.test_case > 10;
            double W[4];
            const int side = (ydir ? y : x) <= 0.5 ? 0 : 1;
            for (int m = 0; m < 4; ++m) W[m] = c_ttc[k][side][m];
            if (ydir) { const double t = W[1]; W[1] = W[2]; W[2] = t; }
            ic_put<R>(I.Q[0][q], I, c, ic_cons(g1, p1, W));
            break;
        }
        case 2: {   // circular explosion: 10x10 sub-cell samples inside r = 0.4 about (1, 1)
            const double L[4] = {1.0, 0.0, 0.0, 1.0}, Rr[4] = {0.125, 0.0, 0.0, 0.1};
            const int N = 10;
            int inside = 0;
            const double bx = x - 0.5 * I.dx, by = y - 0.5 * I.dy, ddx = I.dx / N, ddy = I.dy / N;
            for (int a = 0; a < N; ++a)
                for (int b = 0; b < N; ++b) {
                    const double sx = bx + (a + 0.5) * ddx - 1.0;
                    const double sy = by + (b + 0.5) * ddy - 1.0;
                    if (sqrt(sx * sx + sy * sy) <= 0.4) inside++;
                }
            const double f = (double)inside / (N * N);
            double W[4];
            for (int m = 0; m < 4; ++m) W[m] = f * L[m] + (1.0 - f) * Rr[m];
            ic_put<R>(I.Q[0][q], I, c, ic_cons(g1, p1, W));
            break;
        }
        case 3: {   // shocked bubble (sim_twofluid.cpp:774-859): fluid 1 = bubble, phi < 0 inside
            ic_put<R>(I.Q[0][q], I, c, ic_cons(g1, p1, D.w1));
            ic_put<R>(I.Q[1][q], I, c, ic_cons(g2, p2, x < D.shock_x ? D.w2 : D.w3));
            const double ddx = x - D.xc, ddy = y - D.yc;
            phi = sqrt(ddx * ddx + ddy * ddy) - D.R;
            break;
        }
        case 4: {   // RMI interface x_I(y) = x_if - amp cos(2 pi modes (y - y0) / Ly)
            const double Ly = I.Ny * I.dy;
            const double kk = 2.0 * 3.14159265358979323846 * D.modes / Ly;
            const double arg = kk * (y - I.y0);
            const double xi = D.x_if - D.amp * cos(arg);
            const double dxi = D.amp * kk * sin(arg);
            ic_put<R>(I.Q[0][q], I, c, ic_cons(g1, p1, x < D.shock_x ? D.w1 : D.w2));
            ic_put<R>(I.Q[1][q], I, c, ic_cons(g2, p2, D.w3));
            phi = (xi - x) / sqrt(1.0 + dxi * dxi);
            break;
        }
        default: {  // planar two-material tube (sim_twofluid.cpp:675-742)
            ic_put<R>(I.Q[0][q], I, c, ic_cons(g1, p1, D.w1));
            ic_put<R>(I.Q[1][q], I, c, ic_cons(g2, p2, D.w2));
            phi = D.axis == 0 ? x - D.pos : y - D.pos;
            break;
        }
    }
    if (twofluid) gmem((R*)I.phi)[c] = (R)phi;
}

template <typename R>
void launch_ic(const LaunchCtx& L, const IcDesc* descs, const int* map, int q, int twofluid) {
    hipLaunchKernelGGL((k_ic<R>), dim3((unsigned)L.n_full), dim3(256), 0, L.stream, L.insts, L.w_full, descs, map, q,
                       twofluid);
}

// ---------------------------------------------------------------------------
// A15 level set: TVD-RK3 stages of HOUC advection with the stored velocity
// (update_interface_coupled, sim_twofluid.cpp:72), BCs folded in as a gather.
//   stage 1: A = phi + dt L(phi)
//   stage 2: B = 3/4 phi + 1/4 (A + dt L(A))
//   stage 3: A = 1/3 phi + 2/3 (B + dt L(B))
// work: full-array 64x4 tiles
// ---------------------------------------------------------------------------
// -(u d/dx + v d/dy) psi by HOUC upwinding; psi points at the cell, sx / sy
// are the strides of a column / row step.  Every operation rounds on its own
// (fixed order, no contraction) so that each kernel using it gets the same
// bits.
// (idx, idy: 1/dx, 1/dy in the storage precision)
template <typename R>
__device__ __forceinline__ R houc_rhs(const Houc& H, int n, const R* psi, long long sx, long long sy, R u, R v,
                                      R idx, R idy) {
#pragma clang fp contract(off)
    R px = R(0), py = R(0);
    if (u > R(0)) { for (int m = 0; m < n; ++m) px += (R)H.c[m] * psi[(H.o0 + m) * sx]; }
    else { for (int m = 0; m < n; ++m) px -= (R)H.c[m] * psi[(-H.o0 - m) * sx]; }
    if (v > R(0)) { for (int m = 0; m < n; ++m) py += (R)H.c[m] * psi[(H.o0 + m) * sy]; }
    else { for (int m = 0; m < n; ++m) py -= (R)H.c[m] * psi[(-H.o0 - m) * sy]; }
    return -(u * (px * idx) + v * (py * idy));
}

template <typename R>
__device__ __forceinline__ R ls_rhs(const Inst& I, const Houc& H, const R* psi, int i, int j) {
    const long long c = cidx(I, i, j);
    const auto vel = gmem((const R*)I.vel);
    return houc_rhs<R>(H, H.n, psi + c, 1, I.pitch, vel[c], vel[c + I.plane], R(1.0 / I.dx), R(1.0 / I.dy));
}

// the TVD-RK3 combinations of the three stages
template <typename R, int STAGE>
__device__ __forceinline__ R rk3_combine(R phi0, R prev, R dt, R rhs) {
#pragma clang fp contract(off)
    if (STAGE == 1) return phi0 + dt * rhs;
    if (STAGE == 2) return R(0.75) * phi0 + R(0.25) * (prev + dt * rhs);
    return R(1.0 / 3.0) * phi0 + R(2.0 / 3.0) * (prev + dt * rhs);
}

template <typename R, int STAGE>
__global__ __launch_bounds__(256) void k_ls_stage(const Inst* __restrict__ insts, const int4* __restrict__ work,
                                                  Houc H) {
    const int4 w = work[blockIdx.x];
    const Inst& I = insts[w.x];
    if (!gmem(I.si)[SI_ACTIVE]) return;
    const int i = w.y * 4 + (threadIdx.x >> 6);
    const int j = w.z * 64 + (threadIdx.x & 63);
    if (i >= I.H || j >= I.W) return;
    bool f0, f1;
    const int si = src_row(I, i, f0), sj = src_col(I, j, f1);
    const R dt = (R)gmem(I.sd)[SD_DT];
    const auto phi = gmem((const R*)I.phi);
    const auto A = gmem((const R*)I.phiA);
    const auto B = gmem((const R*)I.phiB);
    const long long cs = cidx(I, si, sj);
    R val;
    if (STAGE == 1) val = rk3_combine<R, 1>(phi[cs], R(0), dt, ls_rhs<R>(I, H, phi, si, sj));
    else if (STAGE == 2) val = rk3_combine<R, 2>(phi[cs], A[cs], dt, ls_rhs<R>(I, H, A, si, sj));
    else val = rk3_combine<R, 3>(phi[cs], B[cs], dt, ls_rhs<R>(I, H, B, si, sj));
    R* out = (R*)(STAGE == 2 ? I.phiB : I.phiA);
    out[cidx(I, i, j)] = val;
}

// ---------------------------------------------------------------------------
// The three TVD-RK3 stages in one launch (temporal blocking, as k_extrap_tb):
// a LTX x LTY tile plus a 3 r halo (r = HOUC stencil radius) in LDS, stage t
// exact on the region shrunk by t r cells, the r ghost layers the stencil
// reads refreshed from their mirror sources after stages 1 and 2, the result
// written to A for the owned cells and the ghost cells they source.  Same
// arithmetic as k_ls_stage, bit for bit.  Mirror BCs in x; in y mirror or
// periodic (region rows wrapped to their interior rows).
// work: the RTX x RTY interior tiles of the reinitialisation
// ---------------------------------------------------------------------------
constexpr int LTX = 64, LTY = 32;

// PY: the instances periodic in y (the others return at once), so that
// their extra reads cost the mirror-bounded instances no registers
template <typename R, int ORDER, bool PY>
__global__ __launch_bounds__(1024) void k_ls_rk3_tb(const Inst* __restrict__ insts, const int4* __restrict__ work,
                                                    Houc H) {
    constexpr int N = ORDER + 1, RAD = N / 2, HALO = 3 * RAD;
    constexpr int RX = LTX + 2 * HALO, RY = LTY + 2 * HALO, NC = RX * RY;
    constexpr int SLOTS = (NC + 1023) / 1024;
    __shared__ R s0[NC], s1[NC];
    const int4 w = work[blockIdx.x];
    const Inst& I = insts[w.x];
    if ((I.bc[2] == BC_P && I.bc[3] == BC_P) != PY || !gmem(I.si)[SI_ACTIVE]) return;
    const int g = I.g;
    const int ri = g + w.y * LTY - HALO, rj = g + w.z * LTX - HALO;
    const auto phi = gmem((const R*)I.phi);
    const auto vel = gmem((const R*)I.vel);
    const R dt = (R)gmem(I.sd)[SD_DT];
    const R idxr = R(1.0 / I.dx), idyr = R(1.0 / I.dy);
    // periodic in y (x then has mirror boundaries): a region row outside the
    // interior is the interior row it wraps to (src_row), computed like any
    // interior row, and no y ghost layer is refreshed.  Stage 1 of a row
    // within RAD of the y boundary reads phi's stored ghost rows, as
    // k_ls_stage does (they need not be periodic images: after an upload they
    // hold the initial condition's own values); the stages after it read
    // stage values, whose ghost rows are images
    // per cell: phi, velocity; meta bits 0-7: distance to the region edge,
    // 8: interior, 9: owned, 10: within RAD of the domain edge (its new
    // value refreshes mirror ghost cells), 11: periodic in y and within RAD
    // of the y boundary (stage 1 from global phi)
    constexpr bool py = PY;
    R p0[SLOTS], us[SLOTS], vs[SLOTS];
    int meta[SLOTS];
#pragma unroll
    for (int k = 0; k < SLOTS; ++k) {
        const int idx = threadIdx.x + k * 1024;
        p0[k] = R(0); us[k] = R(0); vs[k] = R(0); meta[k] = 0;
        if (idx < NC) {
            const int a = idx / RX, b = idx - (idx / RX) * RX;
            const int i = ri + a, j = rj + b;
            bool fl;
            const int is = py ? src_row(I, i, fl) : i;
            R val = R(0);
            if (is >= 0 && is < I.H && j >= 0 && j < I.W) {
                const long long c = cidx(I, is, j);
                val = phi[c];
                if (interior(I, is, j)) {
                    us[k] = vel[c];
                    vs[k] = vel[c + I.plane];
                    const int d = min(min(a, RY - 1 - a), min(b, RX - 1 - b));
                    const bool own = a >= HALO && a < HALO + LTY && b >= HALO && b < HALO + LTX && interior(I, i, j);
                    const bool edge = (!py && (i - g < RAD || g + I.Ny - 1 - i < RAD)) || j - g < RAD ||
                                      g + I.Nx - 1 - j < RAD;
                    const bool yb = py && (is - g < RAD || g + I.Ny - 1 - is < RAD);
                    meta[k] = d | 256 | (own ? 512 : 0) | (edge ? 1024 : 0) | (yb ? 2048 : 0);
                }
            }
            p0[k] = val;
            s0[idx] = val;
        }
    }
    __syncthreads();
    const int ib = g, it = g + I.Ny - 1, jl = g, jr = g + I.Nx - 1;
#pragma unroll
    for (int st = 1; st <= 2; ++st) {
        const R* in = st == 1 ? s0 : s1;
        R* out = st == 1 ? s1 : s0;
#pragma unroll
        for (int k = 0; k < SLOTS; ++k) {
            if (!(meta[k] & 256) || (meta[k] & 255) < st * RAD) continue;
            const int idx = threadIdx.x + k * 1024;
            R rhs;
            if (PY && st == 1 && (meta[k] & 2048)) {
                const int a = idx / RX, b = idx - (idx / RX) * RX;
                bool fl;
                const int is = src_row(I, ri + a, fl);
                rhs = houc_rhs<R>(H, N, (const R*)I.phi + cidx(I, is, rj + b), 1, I.pitch, us[k], vs[k], idxr, idyr);
            } else {
                rhs = houc_rhs<R>(H, N, in + idx, 1, RX, us[k], vs[k], idxr, idyr);
            }
            const R nv = st == 1 ? rk3_combine<R, 1>(p0[k], R(0), dt, rhs) : rk3_combine<R, 2>(p0[k], in[idx], dt, rhs);
            out[idx] = nv;
            if (meta[k] & 1024) {
                // mirror ghosts within the stencil radius of the boundary
                const int a = idx / RX, b = idx - (idx / RX) * RX;
                const int i = ri + a, j = rj + b;
                if (!py && i - ib < RAD) { const int ga = a - 2 * (i - ib) - 1; if (ga >= 0) out[ga * RX + b] = nv; }
                if (!py && it - i < RAD) { const int ga = a + 2 * (it - i) + 1; if (ga < RY) out[ga * RX + b] = nv; }
                if (j - jl < RAD) { const int gb = b - 2 * (j - jl) - 1; if (gb >= 0) out[a * RX + gb] = nv; }
                if (jr - j < RAD) { const int gb = b + 2 * (jr - j) + 1; if (gb < RX) out[a * RX + gb] = nv; }
            }
        }
        __syncthreads();
    }
    const auto A = gmem((R*)I.phiA);
#pragma unroll
    for (int k = 0; k < SLOTS; ++k) {
        if (!(meta[k] & 512)) continue;
        const int idx = threadIdx.x + k * 1024;
        const R rhs = houc_rhs<R>(H, N, s0 + idx, 1, RX, us[k], vs[k], idxr, idyr);
        const R nv = rk3_combine<R, 3>(p0[k], s0[idx], dt, rhs);
        const int a = idx / RX, b = idx - (idx / RX) * RX;
        const int i = ri + a, j = rj + b;
        A[cidx(I, i, j)] = nv;
        if (i < 2 * g || i >= I.Ny || j < 2 * g || j >= I.Nx) {
#pragma unroll 1
            for (int m = 1; m < 9; ++m) {
                const int pr = m / 3, pc = m - 3 * (m / 3);
                int r = i, cc = j;
                if (pr == 1) { if (i >= 2 * g) continue; r = py ? i + I.Ny : 2 * g - 1 - i; }
                if (pr == 2) { if (i < I.Ny) continue; r = py ? i - I.Ny : 2 * (I.Ny + g) - 1 - i; }
                if (pc == 1) { if (j >= 2 * g) continue; cc = 2 * g - 1 - j; }
                if (pc == 2) { if (j < I.Nx) continue; cc = 2 * (I.Nx + g) - 1 - j; }
                A[cidx(I, r, cc)] = nv;
            }
        }
    }
}

__device__ __forceinline__ bool reinit_step(const Inst& I, int every) {
    return every > 0 && ((gmem(I.si)[SI_NSTEP] + 1) % every) == 0;
}

// copy A -> phi on steps without reinitialisation
template <typename R>
__global__ __launch_bounds__(256) void k_ls_copy(const Inst* __restrict__ insts, const int4* __restrict__ work,
                                                 int every) {
    const int4 w = work[blockIdx.x];
    const Inst& I = insts[w.x];
    if (!gmem(I.si)[SI_ACTIVE] || reinit_step(I, every)) return;
    const int i = w.y * 4 + (threadIdx.x >> 6);
    const int j = w.z * 64 + (threadIdx.x & 63);
    if (i >= I.H || j >= I.W) return;
    const long long c = cidx(I, i, j);
    ((R*)I.phi)[c] = ((const R*)I.phiA)[c];
}

// Sussman-Smereka-Osher reinitialisation, first-order Godunov, Russo-Smereka
// subcell fix near the interface; phi0 frozen in A.  Every operation rounds
// on its own so that the per-iteration and the blocked kernels agree bit for
// bit.  reinit_setup: the per-cell constants from phi0 at the cell and its
// four neighbours (mode 1: interface cell, 2: phi0 > 0, 4: phi0 < 0;
// cst: phi0 / |grad phi0| at interface cells, the smoothed sign otherwise).
template <typename R>
__device__ __forceinline__ void reinit_setup(R f0v, R f0e, R f0w, R f0n, R f0s, double dx, double dy, double h,
                                             int& mode, R& cst) {
#pragma clang fp contract(off)
    const bool s0 = f0v <= R(0);
    const bool ifc = ((f0e <= R(0)) != s0) || ((f0w <= R(0)) != s0) || ((f0n <= R(0)) != s0) || ((f0s <= R(0)) != s0);
    mode = (ifc ? 1 : 0) | (f0v > R(0) ? 2 : 0) | (f0v < R(0) ? 4 : 0);
    if (ifc) {
        const R gx = (f0e - f0w) / R(2.0 * dx), gy = (f0n - f0s) / R(2.0 * dy);
        R gm = sqrt(gx * gx + gy * gy);
        gm = fmax(gm, fabs(f0e - f0v) / R(dx));
        gm = fmax(gm, fabs(f0v - f0w) / R(dx));
        gm = fmax(gm, fabs(f0n - f0v) / R(dy));
        gm = fmax(gm, fabs(f0v - f0s) / R(dy));
        gm = fmax(gm, R(1e-12));
        cst = f0v / gm;
    } else {
        cst = f0v / sqrt(f0v * f0v + R(h * h));
    }
}

// one pseudo-time step of the cell value f from its neighbours' values; the
// one-sided differences are squared before the 1/dx^2, 1/dy^2 scaling
// (max((a/dx)^2, (b/dx)^2) = max(a^2, b^2) / dx^2)
template <typename R>
__device__ __forceinline__ R reinit_update(int mode, R cst, R f, R fw, R fe, R fs, R fn, R idx2, R idy2, R dtau) {
#pragma clang fp contract(off)
    if (mode & 1) {
        const R sg = (mode & 2) ? R(1) : ((mode & 4) ? R(-1) : R(0));
        return f - R(0.5) * (sg * fabs(f) - cst);
    }
    const R a = f - fw, b = fe - f, c = f - fs, d = fn - f;
    R gx, gy;
    if (mode & 2) {
        const R ap = fmax(a, R(0)), bm = fmin(b, R(0)), cp = fmax(c, R(0)), dm = fmin(d, R(0));
        gx = fmax(ap * ap, bm * bm);
        gy = fmax(cp * cp, dm * dm);
    } else {
        const R am = fmin(a, R(0)), bp = fmax(b, R(0)), cm = fmin(c, R(0)), dp = fmax(d, R(0));
        gx = fmax(am * am, bp * bp);
        gy = fmax(cm * cm, dp * dp);
    }
    return f - dtau * cst * (sqrt(gx * idx2 + gy * idy2) - R(1));
}

// the constants reinit_update takes: 1/dx^2, 1/dy^2, dtau = h/2
template <typename R>
struct ReinitK { R idx2, idy2, dtau; };
template <typename R>
__device__ __forceinline__ ReinitK<R> reinit_k(const Inst& I) {
    ReinitK<R> k;
    k.idx2 = R(1.0 / (I.dx * I.dx));
    k.idy2 = R(1.0 / (I.dy * I.dy));
    k.dtau = R(0.5 * fmin(I.dx, I.dy));
    return k;
}

template <typename R>
__device__ __forceinline__ void reinit_cell_setup(const Inst& I, const R* P0g, int si, int sj, int& mode, R& cst) {
    const auto P0 = gmem(P0g);
    const double h = fmin(I.dx, I.dy);
    reinit_setup<R>(P0[cidx(I, si, sj)], P0[cidx(I, si, sj + 1)], P0[cidx(I, si, sj - 1)], P0[cidx(I, si + 1, sj)],
                    P0[cidx(I, si - 1, sj)], I.dx, I.dy, h, mode, cst);
}

// Iteration k of K reads (k==0 ? A : previous output) and writes phi when
// K-1-k is even, else B.  gate: 0 = always (explicit reinit), otherwise only
// on reinit steps.
template <typename R>
__global__ __launch_bounds__(256) void k_ls_reinit(const Inst* __restrict__ insts, const int4* __restrict__ work,
                                                   int k, int K, int gate_every, int force) {
    const int4 w = work[blockIdx.x];
    const Inst& I = insts[w.x];
    if (!force && (!gmem(I.si)[SI_ACTIVE] || !reinit_step(I, gate_every))) return;
    const int i = w.y * 4 + (threadIdx.x >> 6);
    const int j = w.z * 64 + (threadIdx.x & 63);
    if (i >= I.H || j >= I.W) return;
    bool f0, f1;
    const int si = src_row(I, i, f0), sj = src_col(I, j, f1);
    const auto P0 = gmem((const R*)I.phiA);
    const bool out_phi = ((K - 1 - k) % 2) == 0;
    const R* in = k == 0 ? P0 : (const R*)(out_phi ? I.phiB : I.phi);
    R* out = (R*)(out_phi ? I.phi : I.phiB);
    int mode;
    R cst;
    reinit_cell_setup<R>(I, P0, si, sj, mode, cst);
    R fw = R(0), fe = R(0), fs = R(0), fn = R(0);
    if (!(mode & 1)) {
        fw = in[cidx(I, si, sj - 1)]; fe = in[cidx(I, si, sj + 1)];
        fs = in[cidx(I, si - 1, sj)]; fn = in[cidx(I, si + 1, sj)];
    }
    const ReinitK<R> rk = reinit_k<R>(I);
    out[cidx(I, i, j)] = reinit_update<R>(mode, cst, in[cidx(I, si, sj)], fw, fe, fs, fn, rk.idx2, rk.idy2, rk.dtau);
}

// ---------------------------------------------------------------------------
// Up to RT_MAXIT iterations of k_ls_reinit in one launch (temporal blocking,
// as k_extrap_tb): a RTX x RTY tile of interior cells plus an RT_MAXIT-cell
// halo of the iterate in LDS; iteration t is exact on the region shrunk by t
// cells; the first ghost layer (the only one the stencil reads) is refreshed
// from its interior source after each iteration; the last iteration writes
// the owned cells and the ghost cells they source.  Mirror BCs in x; in y
// mirror or periodic (PY: wrapped region rows, as k_ls_rk3_tb).
// in_sel / out_sel: 0 = A, 1 = phi, 2 = B.
// work: {inst, tile_y, tile_x, 0}, RTX x RTY interior tiles
// ---------------------------------------------------------------------------
constexpr int RTX = 64, RTY = 32, RT_THREADS = 1024, RT_MAXIT = 10;

template <typename R>
__device__ __forceinline__ R* ls_buf(const Inst& I, int sel) {
    return (R*)(sel == 0 ? I.phiA : (sel == 1 ? I.phi : I.phiB));
}

// minimum waves per SIMD: FP32 8, i.e. two workgroups of 1024 per CU with
// at most 64 VGPRs; FP64 fits one workgroup
template <typename R>
struct RtWaves { static constexpr int value = sizeof(R) == 4 ? 8 : 1; };

// PY: the instances periodic in y, as in k_ls_rk3_tb
template <typename R, bool PY>
__global__ __launch_bounds__(RT_THREADS, RtWaves<R>::value) void k_ls_reinit_tb(const Inst* __restrict__ insts,
                                                             const int4* __restrict__ work, int nit, int in_sel,
                                                             int out_sel, int gate_every, int force, int copy_idle) {
    constexpr int RX = RTX + 2 * RT_MAXIT, RY = RTY + 2 * RT_MAXIT, NC = RX * RY;
    constexpr int SLOTS = (NC + RT_THREADS - 1) / RT_THREADS;
    __shared__ R s[NC];
    __shared__ R scs[NC];   // reinit_setup's constant per cell (in LDS: it would not fit the registers)
    const int4 w = work[blockIdx.x];
    const Inst& I = insts[w.x];
    if ((I.bc[2] == BC_P && I.bc[3] == BC_P) != PY) return;
    const int g = I.g;
    // periodic in y, as in k_ls_rk3_tb: wrapped region rows, no y ghost
    // refresh; the first iteration of a row on the y boundary reads its
    // outer neighbour from the input's stored ghost row, as k_ls_reinit does
    constexpr bool py = PY;
    if (!force) {
        if (!gmem(I.si)[SI_ACTIVE]) return;
        if (!reinit_step(I, gate_every)) {
            // a step without reinitialisation: phi = A (k_ls_copy's work)
            if (copy_idle) {
                const auto A = gmem((const R*)I.phiA);
                const auto phi = gmem((R*)I.phi);
                for (int q = threadIdx.x; q < RTX * RTY; q += RT_THREADS) {
                    const int i = g + w.y * RTY + q / RTX, j = g + w.z * RTX + q % RTX;
                    if (!interior(I, i, j)) continue;
                    const long long c0 = cidx(I, i, j);
                    phi[c0] = A[c0];
                    if (!(i < 2 * g || i >= I.Ny || j < 2 * g || j >= I.Nx)) continue;   // no ghost image
                    for (int m = 1; m < 9; ++m) {
                        const int pr = m / 3, pc = m - 3 * (m / 3);
                        int r = i, cc = j;
                        if (pr == 1) { if (i >= 2 * g) continue; r = py ? i + I.Ny : 2 * g - 1 - i; }
                        if (pr == 2) { if (i < I.Ny) continue; r = py ? i - I.Ny : 2 * (I.Ny + g) - 1 - i; }
                        if (pc == 1) { if (j >= 2 * g) continue; cc = 2 * g - 1 - j; }
                        if (pc == 2) { if (j < I.Nx) continue; cc = 2 * (I.Nx + g) - 1 - j; }
                        const long long c = cidx(I, r, cc);
                        phi[c] = A[c];
                    }
                }
            }
            return;
        }
    }
    const int ri = g + w.y * RTY - RT_MAXIT, rj = g + w.z * RTX - RT_MAXIT;
    const auto P0 = gmem((const R*)I.phiA);
    const auto in = gmem((const R*)ls_buf<R>(I, in_sel));
    const auto out = gmem(ls_buf<R>(I, out_sel));
    // per cell: reinit_setup's mode and constant; bits 3 / 4: periodic in y
    // and on the bottom / top interior row, bit 8: interior with its stencil
    // inside the region, bit 9: owned, bit 10: next to the domain edge
    // (refreshes the first ghost layer), bits 11+: distance to the region
    // edge
    const int ib = g, it = g + I.Ny - 1, jl = g, jr = g + I.Nx - 1;
    int meta[SLOTS];
#pragma unroll
    for (int k = 0; k < SLOTS; ++k) {
        const int idx = threadIdx.x + k * RT_THREADS;
        meta[k] = 0;
        if (idx < NC) {
            const int a = idx / RX, b = idx - (idx / RX) * RX;
            const int i = ri + a, j = rj + b;
            bool fl;
            const int is = py ? src_row(I, i, fl) : i;
            R v = R(0);
            if (is >= 0 && is < I.H && j >= 0 && j < I.W) {
                v = in[cidx(I, is, j)];
                if (interior(I, is, j) && a > 0 && a < RY - 1 && b > 0 && b < RX - 1) {
                    int mode;
                    R cs;
                    reinit_cell_setup<R>(I, P0, is, j, mode, cs);
                    scs[idx] = cs;
                    const bool own = a >= RT_MAXIT && a < RT_MAXIT + RTY && b >= RT_MAXIT && b < RT_MAXIT + RTX &&
                                     interior(I, i, j);
                    const bool edge = (!py && (i == ib || i == it)) || j == jl || j == jr;
                    const int dist = min(min(a, RY - 1 - a), min(b, RX - 1 - b));
                    const int yb = py ? ((is == ib ? 8 : 0) | (is == it ? 16 : 0)) : 0;
                    meta[k] = mode | yb | 256 | (own ? 512 : 0) | (edge ? 1024 : 0) | (dist << 11);
                }
            }
            s[idx] = v;
        }
    }
    __syncthreads();
    const ReinitK<R> rk = reinit_k<R>(I);
    for (int t = 1; t <= nit; ++t) {
        R nv[SLOTS];
        int act = 0;
#pragma unroll
        for (int k = 0; k < SLOTS; ++k) {
            const int idx = threadIdx.x + k * RT_THREADS;
            if ((meta[k] & 256) && (meta[k] >> 11) >= t) {
                act |= 1 << k;
                R fs = s[idx - RX], fn = s[idx + RX];
                if (PY && t == 1 && (meta[k] & 24)) {
                    const int a = idx / RX, b = idx - (idx / RX) * RX;
                    bool fl;
                    const int is = src_row(I, ri + a, fl);
                    if (meta[k] & 8) fs = in[cidx(I, is - 1, rj + b)];
                    if (meta[k] & 16) fn = in[cidx(I, is + 1, rj + b)];
                }
                nv[k] = reinit_update<R>(meta[k] & 7, scs[idx], s[idx], s[idx - 1], s[idx + 1], fs, fn, rk.idx2,
                                         rk.idy2, rk.dtau);
            }
        }
        if (t == nit) {
            // owned cells and their ghost images (inverse of src_row / src_col)
#pragma unroll
            for (int k = 0; k < SLOTS; ++k) {
                if (!(meta[k] & 512)) continue;
                const int idx = threadIdx.x + k * RT_THREADS;
                const int a = idx / RX, b = idx - (idx / RX) * RX;
                const int i = ri + a, j = rj + b;
                out[cidx(I, i, j)] = nv[k];
                if (i < 2 * g || i >= I.Ny || j < 2 * g || j >= I.Nx) {
#pragma unroll 1
                    for (int m = 1; m < 9; ++m) {
                        const int pr = m / 3, pc = m - 3 * (m / 3);
                        int r = i, cc = j;
                        if (pr == 1) { if (i >= 2 * g) continue; r = py ? i + I.Ny : 2 * g - 1 - i; }
                        if (pr == 2) { if (i < I.Ny) continue; r = py ? i - I.Ny : 2 * (I.Ny + g) - 1 - i; }
                        if (pc == 1) { if (j >= 2 * g) continue; cc = 2 * g - 1 - j; }
                        if (pc == 2) { if (j < I.Nx) continue; cc = 2 * (I.Nx + g) - 1 - j; }
                        out[cidx(I, r, cc)] = nv[k];
                    }
                }
            }
            break;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < SLOTS; ++k) {
            if (!(act & (1 << k))) continue;
            const int idx = threadIdx.x + k * RT_THREADS;
            s[idx] = nv[k];
            if (meta[k] & 1024) {
                const int a = idx / RX, b = idx - (idx / RX) * RX;
                const int i = ri + a, j = rj + b;
                if (!py && i == ib) s[idx - RX] = nv[k];
                if (!py && i == it) s[idx + RX] = nv[k];
                if (j == jl) s[idx - 1] = nv[k];
                if (j == jr) s[idx + 1] = nv[k];
            }
        }
        __syncthreads();
    }
}

// copy phi -> A (frozen phi0 for an explicit reinitialisation)
template <typename R>
__global__ __launch_bounds__(256) void k_phi_to_A(const Inst* __restrict__ insts, const int4* __restrict__ work) {
    const int4 w = work[blockIdx.x];
    const Inst& I = insts[w.x];
    const int i = w.y * 4 + (threadIdx.x >> 6);
    const int j = w.z * 64 + (threadIdx.x & 63);
    if (i >= I.H || j >= I.W) return;
    const long long c = cidx(I, i, j);
    ((R*)I.phiA)[c] = ((const R*)I.phi)[c];
}

// ---------------------------------------------------------------------------
// A13 at the end of the step (sim_twofluid.cpp:74-75): ghost frame of the
// current conserved grids, then t += dt and the step count (:77-78).
// work: {inst, strip, chunk start, 0}; strips 0 bottom, 1 top, 2 left, 3 right
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool frame_cell(const Inst& I, int strip, int idx, int& i, int& j) {
    const int g = I.g;
    if (strip <= 1) {
        if (idx >= g * I.W) return false;
        i = idx / I.W; j = idx % I.W;
        if (strip == 1) i += g + I.Ny;
    } else {
        if (idx >= g * I.Ny) return false;
        i = g + idx / g; j = idx % g;
        if (strip == 3) j += g + I.Nx;
    }
    return true;
}

// velocity at (x, y): the bilinear stencil of the probes on the two planes of
// the velocity field (the solved velocity field's get_velocity, external)
template <typename R>
__device__ __forceinline__ void vel_at(const Inst& I, double x, double y, double& u, double& v) {
#pragma clang fp contract(off)
    const Probe p = bilinear_stencil(I, x, y);
    const auto vel = gmem((const R*)I.vel);
    const long long a = cidx(I, p.i0, p.j0), c = a + I.pitch;
    const double sx = 1.0 - p.tx, sy = 1.0 - p.ty;
    u = sy * (sx * (double)vel[a] + p.tx * (double)vel[a + 1]) +
        p.ty * (sx * (double)vel[c] + p.tx * (double)vel[c + 1]);
    v = sy * (sx * (double)vel[a + I.plane] + p.tx * (double)vel[a + 1 + I.plane]) +
        p.ty * (sx * (double)vel[c + I.plane] + p.tx * (double)vel[c + 1 + I.plane]);
}

// the passive tracers of sim_twofluid.cpp:1861-1885 after step n (t, dt of
// that step, the velocity field its ghost fill stored): record t and each
// tracer's position and velocity, then move it by one classical RK4 step of
// dt in that field (time_step_RK4, external), and put it back on its initial
// y if asked (:1882-1884).  One thread.
template <typename R>
__device__ void advance_tracers(const Inst& I) {
#pragma clang fp contract(off)
    const int n = gmem(I.si)[SI_NSTEP];
    const double t = gmem(I.sd)[SD_T], dt = gmem(I.sd)[SD_DT];
    double* pos = I.trk;
    const int stride = 1 + 4 * I.ntr;
    double* h = n < I.hist_cap ? I.trk + TR_HDR + (long long)n * stride : nullptr;
    if (h) h[0] = t;
    for (int k = 0; k < I.ntr; ++k) {
        const double x = pos[2 * k], y = pos[2 * k + 1];
        double u1, v1, u2, v2, u3, v3, u4, v4;
        vel_at<R>(I, x, y, u1, v1);
        if (h) { h[1 + 4 * k] = x; h[2 + 4 * k] = y; h[3 + 4 * k] = u1; h[4 + 4 * k] = v1; }
        vel_at<R>(I, x + 0.5 * dt * u1, y + 0.5 * dt * v1, u2, v2);
        vel_at<R>(I, x + 0.5 * dt * u2, y + 0.5 * dt * v2, u3, v3);
        vel_at<R>(I, x + dt * u3, y + dt * v3, u4, v4);
        pos[2 * k] = x + dt / 6.0 * (u1 + 2.0 * u2 + 2.0 * u3 + u4);
        pos[2 * k + 1] = I.tr_keep_y ? pos[2 * TR_MAX + 2 * k + 1] : y + dt / 6.0 * (v1 + 2.0 * v2 + 2.0 * v3 + v4);
    }
}

template <typename R>
__global__ __launch_bounds__(256) void k_bc(const Inst* __restrict__ insts, const int4* __restrict__ work,
                                            int q, int nfluid, int advance) {
    const int4 w = work[blockIdx.x];
    const Inst& I = insts[w.x];
    const bool active = gmem(I.si)[SI_ACTIVE] != 0;
    int i, j;
    if (frame_cell(I, w.y, w.z + threadIdx.x, i, j)) {
        bool fv, fu;
        const int si = src_row(I, i, fv), sj = src_col(I, j, fu);
        for (int f = 0; f < nfluid; ++f) {
            V4<R> U = ld4<R>(I.Q[f][q], I, cidx(I, si, sj));
            if (fu) U.v[1] = -U.v[1];
            if (fv) U.v[2] = -U.v[2];
            st4<R>(I.Q[f][q], I, cidx(I, i, j), U);
        }
    }
    if (active && advance && w.y == 0 && w.z == 0 && threadIdx.x == 0) {
        gmem(I.sd)[SD_T] += gmem(I.sd)[SD_DT];
        gmem(I.si)[SI_NSTEP] += 1;
        if (I.ntr > 0) advance_tracers<R>(I);
    }
}

// ---------------------------------------------------------------------------
// batched Riemann solves for the pure_RS_base / mixed_RS_base drop-ins
// ---------------------------------------------------------------------------
__global__ void k_pure_rs(int n, int px, double g, double pinf, const double* UL, const double* UR, double* F,
                          unsigned* status) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    V4<double> L, Rr;
    for (int k = 0; k < 4; ++k) { L.v[k] = UL[4 * t + k]; Rr.v[k] = UR[4 * t + k]; }
    unsigned st = 0;
    const V4<double> f = pure_rs_flux<double>(px != 0, g, pinf, L, Rr, st);
    for (int k = 0; k < 4; ++k) F[4 * t + k] = f.v[k];
    if (st) atomicOr(status, st);
}

// the face solve of k_fs<float> (FP32 storage) for one face: FP32 conserved
// states -> primitives -> trivial or full FP32 solve -> FP32 flux
__global__ void k_pure_rs_f32(int n, int px, RSFluid F, const double* UL, const double* UR, double* Fl,
                              unsigned* status) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    V4<float> a, b;
    for (int k = 0; k < 4; ++k) { a.v[k] = (float)UL[4 * t + k]; b.v[k] = (float)UR[4 * t + k]; }
    unsigned st = 0;
    if (!phys_state_f(a) || !phys_state_f(b)) st |= ST_UNPHYS;
    const FluidF Ff = fluid_f(F);
    const V4<float> WL = cons2prims_f(Ff, a), WR = cons2prims_f(Ff, b);
    const V4<float> f = rs_trivial(px != 0, WL, WR) ? rs_flux_trivial_f(px != 0, Ff, WL, WR)
                                                     : rs_flux_full_f(F, Ff, px != 0, WL, WR, st);
    for (int k = 0; k < 4; ++k) Fl[4 * t + k] = (double)f.v[k];
    if (st) atomicOr(status, st);
}

__global__ void k_mixed_rs(int n, double gL, double pL, double gR, double pR, const double* UL, const double* UR,
                           double* out, unsigned* status) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    V4<double> L, Rr;
    for (int k = 0; k < 4; ++k) { L.v[k] = UL[4 * t + k]; Rr.v[k] = UR[4 * t + k]; }
    unsigned st = 0;
    if (!phys_state(L) || !phys_state(Rr)) st |= ST_UNPHYS;
    const V4<double> WL = cons2prims(gL, pL, L), WR = cons2prims(gR, pR, Rr);
    RSSide a, b;
    rs_side(a, gL, pL, WL.v[0], WL.v[1], WL.v[3]);
    rs_side(b, gR, pR, WR.v[0], WR.v[1], WR.v[3]);
    Star S;
    st |= rs_star(a, b, S, -1e300, 1e-12);
    out[4 * t] = S.p; out[4 * t + 1] = S.u; out[4 * t + 2] = S.rl; out[4 * t + 3] = S.rr;
    if (st) atomicOr(status, st);
}

// per-step counters for phase-by-phase use (the step path resets them in k_dt)
__global__ void k_reset(const Inst* __restrict__ insts, const int4* __restrict__ work, int narrow_boxes) {
    const Inst& I = insts[work[blockIdx.x].x];
    if (threadIdx.x != 0) return;
    gmem(I.si)[SI_ACTIVE] = 1;
    gmem(I.si)[SI_NIFACE] = 0;
    gmem(I.si)[SI_DTCOUNT] = 0;
    reset_boxes(I, narrow_boxes);
}

// ---------------------------------------------------------------------------
// launchers (explicit instantiation per storage type)
// ---------------------------------------------------------------------------
#define GRID(n) dim3((unsigned)(n))

// ---------------------------------------------------------------------------
// CLSVOF tracker (settings ITM "CLSVOF", sim_twofluid.cpp:571-574).  The
// reference's tracker is a class of an external library, so this is the
// published method, not a restatement of its code: a coupled level set and
// volume of fluid (Sussman and Puckett 2000).  Inst::zv holds three FP64
// planes: the volume fraction z of fluid 1, the first sweep's result, and
// the divergence term's flag of the step's input.  Each step, between the
// level set's advection (phi -> phiA) and its reinitialisation:
//  - two directionally split sweeps (x first on even steps, y first on odd
//    ones) move z by geometric PLIC fluxes with the cell-centred divergence
//    term of Weymouth and Yue (2010); a donor cell is reconstructed with the
//    start-of-step level set's unit normal and the line that leaves its z;
// and after it:
//  - every cell's phi takes the PLIC line of its z along its own normal
//    wherever the two disagree (k_clsvof_fix, into phiA, then k_cls_copy).
// Full-array kernels: a ghost-frame cell takes the result of its BC source
// cell (the BCs of a scalar).
// ---------------------------------------------------------------------------
constexpr double CLS_ZEPS = 1e-12;   // z this near 0 or 1 counts as a pure cell

// inverse of area_fraction in phi for a dx x dy cell (Scardovelli and
// Zaleski 2000, 2D): m = sqrt(2abz) | bz + a/2 | a + b - sqrt(2ab(1-z)), the
// line value (a + b)/2 - m
__device__ __forceinline__ double plic_line(double z, double nx, double ny, double dx, double dy) {
#pragma clang fp contract(off)
    double a = fabs(nx) * dx, b = fabs(ny) * dy;
    if (a > b) { const double t = a; a = b; b = t; }
    double m;
    if (z <= 0.0) m = 0.0;
    else if (z >= 1.0) m = a + b;
    else if (2.0 * b * z <= a) m = sqrt(2.0 * a * b * z);
    else if (2.0 * b * (1.0 - z) <= a) m = a + b - sqrt(2.0 * a * b * (1.0 - z));
    else m = b * z + 0.5 * a;
    return 0.5 * (a + b) - m;
}

// signed volume fraction of fluid 1 through the face between cell (iA, jA)
// and the next cell along d, at face velocity u over dt, in units of one
// cell: the donor's strip of width |u| dt next to the face
template <typename R>
__device__ __forceinline__ double cls_face_flux(const Inst& I, int d, int iA, int jA, double u, double dt,
                                                const double* zs) {
#pragma clang fp contract(off)
    const double h = d == 0 ? I.dx : I.dy, w = fabs(u) * dt;
    if (w == 0.0) return 0.0;
    const int i = u > 0.0 ? iA : iA + (d == 1 ? 1 : 0), j = u > 0.0 ? jA : jA + (d == 0 ? 1 : 0);
    const double z = zs[cidx(I, i, j)];
    double f;
    if (z <= CLS_ZEPS || z >= 1.0 - CLS_ZEPS) {
        f = z;   // a pure donor: its strip holds z of its area
    } else {
        double nx, ny;
        unit_normal<R>(I, i, j, nx, ny);
        const double sl = plic_line(z, nx, ny, I.dx, I.dy);
        const double off = (u > 0.0 ? 0.5 : -0.5) * (h - w);   // strip centre - donor centre along d
        const double sc = sl + (d == 0 ? nx : ny) * off;
        f = d == 0 ? area_fraction(sc, nx, ny, w, I.dy) : area_fraction(sc, nx, ny, I.dx, w);
    }
    return (u > 0.0 ? 1.0 : -1.0) * (w / h) * f;
}

// rows per thread of the tracker's full-array kernels (64 x 4 CLS_ROWS tiles,
// LaunchCtx::w_p16): a workgroup's chain of dependent loads (work item,
// descriptor, scalars) is paid once for CLS_ROWS cells per thread
constexpr int CLS_ROWS = 4;

// one sweep: the first reads plane 0 (the step's input) and writes plane 1
// and the flags (plane 2), the second reads plane 1 and the flags and
// writes plane 0.  Face velocities are the means of the cells' velocity
// components along d (the fill's velocity field, ghost frame included);
// results clipped to [0, 1].  work: 64x16 tiles over the padded array
template <typename R>
__global__ __launch_bounds__(256) void k_cls_sweep(const Inst* __restrict__ insts, const int4* __restrict__ work,
                                                   int second) {
#pragma clang fp contract(off)
    const int4 w = work[blockIdx.x];
    const Inst& I = insts[w.x];
    if (!gmem(I.si)[SI_ACTIVE]) return;
    const int j = w.z * 64 + (threadIdx.x & 63);
    if (j >= I.W) return;
    const int odd = gmem(I.si)[SI_NSTEP] & 1;
    const int d = second ? 1 - odd : odd;
    const double dt = gmem(I.sd)[SD_DT];
    double* zv = gmem(I.zv);
    const double* zs = zv + (second ? I.plane : 0);
    double* out = zv + (second ? 0 : I.plane);
    double* flags = zv + 2 * I.plane;
    const R* vel = gmem((const R*)I.vel) + (d == 0 ? 0 : I.plane);
    const int di = d == 1 ? 1 : 0, dj = d == 0 ? 1 : 0;
    const double h = d == 0 ? I.dx : I.dy;
    bool fu;
    const int sj = src_col(I, j, fu);
#pragma unroll
    for (int r = 0; r < CLS_ROWS; ++r) {
        const int i = w.y * (4 * CLS_ROWS) + (threadIdx.x >> 6) + 4 * r;
        if (i >= I.H) break;
        bool fv;
        const int si = src_row(I, i, fv);
        const long long c = cidx(I, i, j), q = cidx(I, si, sj);
        double cf;
        if (second) {
            cf = flags[c];
        } else {
            cf = zs[q] > 0.5 ? 1.0 : 0.0;
            flags[c] = cf;
        }
        const double uc = (double)vel[q];
        const double uL = 0.5 * ((double)vel[cidx(I, si - di, sj - dj)] + uc);
        const double uR = 0.5 * (uc + (double)vel[cidx(I, si + di, sj + dj)]);
        const double fL = cls_face_flux<R>(I, d, si - di, sj - dj, uL, dt, zs);
        const double fR = cls_face_flux<R>(I, d, si, sj, uR, dt, zs);
        const double v = zs[q] - (fR - fL) + cf * ((uR - uL) * dt / h);
        out[c] = fmin(fmax(v, 0.0), 1.0);
    }
}

// the coupling: phi at a cell takes the PLIC line of its z along phi's unit
// normal where the two disagree (a mixed cell always; a pure one where phi's
// line cuts it or has the other sign: phi keeps its magnitude but at least
// half the cell's extent along n, with the sign of z's fluid).  Written to
// phiA (the copy follows).  work: 64x16 tiles over the padded array
template <typename R>
__global__ __launch_bounds__(256) void k_clsvof_fix(const Inst* __restrict__ insts, const int4* __restrict__ work) {
#pragma clang fp contract(off)
    const int4 w = work[blockIdx.x];
    const Inst& I = insts[w.x];
    if (!gmem(I.si)[SI_ACTIVE]) return;
    const int j = w.z * 64 + (threadIdx.x & 63);
    if (j >= I.W) return;
    bool fu;
    const int sj = src_col(I, j, fu);
#pragma unroll
    for (int r = 0; r < CLS_ROWS; ++r) {
        const int i = w.y * (4 * CLS_ROWS) + (threadIdx.x >> 6) + 4 * r;
        if (i >= I.H) break;
        bool fv;
        const int si = src_row(I, i, fv);
        const long long q = cidx(I, si, sj);
        double nx, ny;
        unit_normal<R>(I, si, sj, nx, ny);
        const double zc = gmem(I.zv)[q], p = (double)ld1<R>(I.phi, q);
        double out = p;
        if (zc > CLS_ZEPS && zc < 1.0 - CLS_ZEPS) {
            out = plic_line(zc, nx, ny, I.dx, I.dy);
        } else {
            const double zp = zc >= 0.5 ? 1.0 : 0.0;
            if (area_fraction(p, nx, ny, I.dx, I.dy) != zp) {
                const double half = 0.5 * (fabs(nx) * I.dx + fabs(ny) * I.dy);
                const double mag = fmax(fabs(p), half);
                out = zp == 1.0 ? -mag : mag;
            }
        }
        gmem((R*)I.phiA)[cidx(I, i, j)] = (R)out;
    }
}

// phi = phiA after the coupling (k_ls_copy's work on 64x16 tiles)
template <typename R>
__global__ __launch_bounds__(256) void k_cls_copy(const Inst* __restrict__ insts, const int4* __restrict__ work) {
    const int4 w = work[blockIdx.x];
    const Inst& I = insts[w.x];
    if (!gmem(I.si)[SI_ACTIVE]) return;
    const int j = w.z * 64 + (threadIdx.x & 63);
    if (j >= I.W) return;
#pragma unroll
    for (int r = 0; r < CLS_ROWS; ++r) {
        const int i = w.y * (4 * CLS_ROWS) + (threadIdx.x >> 6) + 4 * r;
        if (i >= I.H) break;
        const long long c = cidx(I, i, j);
        gmem((R*)I.phi)[c] = gmem((const R*)I.phiA)[c];
    }
}

// initialisation(): z from the level set's PLIC reconstruction (vof_z of
// the cell, of its BC source in the ghost frame); only >= 0: that instance
template <typename R>
__global__ __launch_bounds__(256) void k_clsvof_init(const Inst* __restrict__ insts, const int4* __restrict__ work,
                                                     int only) {
    const int4 w = work[blockIdx.x];
    if (only >= 0 && w.x != only) return;
    const Inst& I = insts[w.x];
    const int i = w.y * 4 + (threadIdx.x >> 6);
    const int j = w.z * 64 + (threadIdx.x & 63);
    if (i >= I.H || j >= I.W) return;
    bool fv, fu;
    const int si = src_row(I, i, fv), sj = src_col(I, j, fu);
    gmem(I.zv)[cidx(I, i, j)] = vof_z<R>(I, si, sj);
}

template <typename R>
void launch_clsvof_init(const LaunchCtx& L, int only) {
    hipLaunchKernelGGL((k_clsvof_init<R>), GRID(L.n_full), dim3(256), 0, L.stream, L.insts, L.w_full, only);
}


template <typename R>
void launch_dt(const LaunchCtx& L, int twofluid, int q, int narrow_boxes, int count_iface) {
    hipLaunchKernelGGL((k_dt<R>), GRID((L.n_seg + SEG_PER_WG - 1) / SEG_PER_WG), dim3(256), 0, L.stream, L.insts,
                       L.w_seg, L.n_seg, twofluid, q, count_iface);
    hipLaunchKernelGGL(k_dt_final, GRID(L.n_inst), dim3(256), 0, L.stream, L.insts, L.w_inst, twofluid, narrow_boxes);
}

template <typename R, int TX, int TY>
void launch_fs_shape(const LaunchCtx& L, bool muscl, int qin, int qout, int copy_inactive, int qorig) {
    if (L.fs_persist > 0 && L.fs_next) {
        // eight workgroups per CU, more than fit at once: the ones that do
        // not fit start as the first ones leave, find the queues empty and
        // leave too, so the grid needs no occupancy figure
        const int grid = std::min(8 * L.fs_persist, (L.n_fs + 7) / 8 * 8);
        launch_fill_u32(L.stream, L.fs_next, 8, 0u);
        if (muscl)
            hipLaunchKernelGGL((k_fs_persist<R, true, TX, TY>), GRID(grid), dim3(TX * TY), 0, L.stream, L.insts,
                               L.w_fs, L.n_fs, L.fs_next, qin, qout, copy_inactive, qorig);
        else
            hipLaunchKernelGGL((k_fs_persist<R, false, TX, TY>), GRID(grid), dim3(TX * TY), 0, L.stream, L.insts,
                               L.w_fs, L.n_fs, L.fs_next, qin, qout, copy_inactive, qorig);
        return;
    }
    if (muscl)
        hipLaunchKernelGGL((k_fs<R, true, TX, TY>), GRID(L.n_fs), dim3(TX * TY), 0, L.stream, L.insts, L.w_fs, qin,
                           qout, copy_inactive, qorig);
    else
        hipLaunchKernelGGL((k_fs<R, false, TX, TY>), GRID(L.n_fs), dim3(TX * TY), 0, L.stream, L.insts, L.w_fs, qin,
                           qout, copy_inactive, qorig);
}

// the context's flux tile shape (FS_SHAPES[L.fs_shape]; its work table w_fs
// was built for that shape)
template <typename R>
void launch_fs(const LaunchCtx& L, bool muscl, int qin, int qout, int copy_inactive, int qorig) {
    static_assert(FS_NSHAPE == 4, "one case per shape");
    switch (L.fs_shape) {
        case 1: launch_fs_shape<R, FS_SHAPES[1][0], FS_SHAPES[1][1]>(L, muscl, qin, qout, copy_inactive, qorig); break;
        case 2: launch_fs_shape<R, FS_SHAPES[2][0], FS_SHAPES[2][1]>(L, muscl, qin, qout, copy_inactive, qorig); break;
        case 3: launch_fs_shape<R, FS_SHAPES[3][0], FS_SHAPES[3][1]>(L, muscl, qin, qout, copy_inactive, qorig); break;
        default: launch_fs_shape<R, FS_SHAPES[0][0], FS_SHAPES[0][1]>(L, muscl, qin, qout, copy_inactive, qorig); break;
    }
}

template <typename R>
void launch_extrap(const LaunchCtx& L, int mode, int bin, int bout) {
#define EXC(M)                                                                                              \
    case M:                                                                                                 \
        hipLaunchKernelGGL((k_extrap<R, M>), GRID(L.n_full), dim3(256), 0, L.stream, L.insts, L.w_full, bin, \
                           bout);                                                                           \
        break;
    switch (mode) {
        EXC(0)
        EXC(MF_FIRST)
        EXC(MF_INJECT)
        EXC(MF_LAST)
        EXC(MF_FIRST | MF_INJECT)
        EXC(MF_FIRST | MF_LAST)
        EXC(MF_LAST | MF_OGFM)
        EXC(MF_FIRST | MF_LAST | MF_OGFM)
        EXC(MF_INJECT | MF_LAST)
        EXC(MF_FIRST | MF_INJECT | MF_LAST)
        EXC(MF_RAW)
        EXC(MF_HOLDIF)
        EXC(MF_HOLDIF | MF_LAST)
        default: break;
    }
#undef EXC
}

// the blocked extrapolation on tile shape TB_SHAPES[SH] (the context's w_tb
// was built for it): the mirror-y and the periodic-y variant, each for the
// instances it serves
template <typename R, int N, int MI, int MO, int SH>
void launch_extrap_tb_shape(const LaunchCtx& L, int bin, int bout, int band) {
    if (L.mirror_y)
        hipLaunchKernelGGL((k_extrap_tb<R, N, MI, MO, false, SH>), GRID(L.n_tb), dim3(TB_SHAPES[SH][2]), 0, L.stream,
                           L.insts, L.w_tb, bin, bout, band);
    if (L.periodic_y)
        hipLaunchKernelGGL((k_extrap_tb<R, N, MI, MO, true, SH>), GRID(L.n_tb), dim3(TB_SHAPES[SH][2]), 0, L.stream,
                           L.insts, L.w_tb, bin, bout, band);
}

// NIT iterations of launch_extrap in one launch; false when the combination
// has no temporally blocked instance (the caller then launches per iteration)
template <typename R>
bool launch_extrap_tb(const LaunchCtx& L, int nit, int min, int mout, int bin, int bout, int band) {
#define EXT(N, MI, MO)                                                                  \
    if (nit == N && min == (MI) && mout == (MO)) {                                      \
        switch (L.tb_shape) {                                                           \
            case 1: launch_extrap_tb_shape<R, N, MI, MO, 1>(L, bin, bout, band); break; \
            case 2: launch_extrap_tb_shape<R, N, MI, MO, 2>(L, bin, bout, band); break; \
            default: launch_extrap_tb_shape<R, N, MI, MO, 0>(L, bin, bout, band); break;\
        }                                                                               \
        return true;                                                                    \
    }
    static_assert(TB_NSHAPE == 3, "one case per shape");
    EXT(6, MF_FIRST, MF_LAST | MF_OGFM)           // GFM_original
    EXT(6, MF_FIRST | MF_INJECT, MF_LAST)         // GFM_real
    EXT(2, MF_FIRST, 0)                           // GFM_riemann / GFM_modified, before the star states
    EXT(6, MF_INJECT, MF_LAST)                    // GFM_riemann, after the star states
#undef EXT
    return false;
}

template <typename R>
void launch_iface(const LaunchCtx& L, int variant, int pbuf, int counted) {
    const unsigned nsb = (unsigned)((L.n_seg + SEG_PER_WG - 1) / SEG_PER_WG);
    if (!counted)
        hipLaunchKernelGGL((k_iface_count<R>), dim3(nsb), dim3(256), 0, L.stream, L.insts, L.w_seg, L.n_seg);
    hipLaunchKernelGGL(k_iface_scan, GRID(L.n_inst), dim3(1024), 0, L.stream, L.insts, L.w_inst);
    hipLaunchKernelGGL((k_iface_mark<R>), dim3(nsb), dim3(256), 0, L.stream, L.insts, L.w_seg, L.n_seg);
    switch (variant) {
        case V_REAL:
            hipLaunchKernelGGL((k_iface_prep<R, V_REAL>), GRID(L.n_ich), dim3(256), 0, L.stream, L.insts, L.w_ich, pbuf);
            hipLaunchKernelGGL((k_iface_rounds<R, V_REAL>), GRID(L.n_inst), dim3(512), 0, L.stream, L.insts, L.w_inst,
                               L.iface_plain);
            break;
        case V_RIEMANN:
            hipLaunchKernelGGL((k_iface_prep<R, V_RIEMANN>), GRID(L.n_ich), dim3(256), 0, L.stream, L.insts, L.w_ich, pbuf);
            hipLaunchKernelGGL((k_iface_rounds<R, V_RIEMANN>), GRID(L.n_inst), dim3(512), 0, L.stream, L.insts, L.w_inst,
                               L.iface_plain);
            break;
        case V_MODIFIED:   // no dependencies: the parallel pass is final
            hipLaunchKernelGGL((k_iface_prep<R, V_MODIFIED>), GRID(L.n_ich), dim3(256), 0, L.stream, L.insts, L.w_ich, pbuf);
            break;
        default: break;
    }
}

template <typename R>
void launch_vof(const LaunchCtx& L, bool ogfm) {
    const int nsb = (L.n_vseg + SEG_PER_WG - 1) / SEG_PER_WG;
    hipLaunchKernelGGL((k_vof_count<R>), dim3(nsb), dim3(256), 0, L.stream, L.insts, L.w_vseg, L.n_vseg);
    hipLaunchKernelGGL(k_iface_scan, GRID(L.n_inst), dim3(1024), 0, L.stream, L.insts, L.w_inst);
    hipLaunchKernelGGL((k_vof_mark<R>), dim3(nsb), dim3(256), 0, L.stream, L.insts, L.w_vseg, L.n_vseg);
    if (ogfm) {
        hipLaunchKernelGGL((k_vof_prep<R, true>), GRID(L.n_ich), dim3(256), 0, L.stream, L.insts, L.w_ich);
        hipLaunchKernelGGL((k_iface_rounds<R, V_OGFM_VOF>), GRID(L.n_inst), dim3(512), 0, L.stream, L.insts,
                           L.w_inst, L.iface_plain);
        hipLaunchKernelGGL((k_vof_final<R, true>), GRID(L.n_full), dim3(256), 0, L.stream, L.insts, L.w_full);
    } else {
        hipLaunchKernelGGL((k_vof_prep<R, false>), GRID(L.n_ich), dim3(256), 0, L.stream, L.insts, L.w_ich);
        hipLaunchKernelGGL((k_iface_rounds<R, V_VOF>), GRID(L.n_inst), dim3(512), 0, L.stream, L.insts,
                           L.w_inst, L.iface_plain);
        hipLaunchKernelGGL((k_vof_final<R, false>), GRID(L.n_full), dim3(256), 0, L.stream, L.insts, L.w_full);
    }
}


template <typename R>
void launch_mgfm_final(const LaunchCtx& L, int qin, int qout, bool prims_out) {
    if (prims_out)
        hipLaunchKernelGGL((k_mgfm_final<R, true>), GRID(L.n_full), dim3(256), 0, L.stream, L.insts, L.w_full, qin, qout);
    else
        hipLaunchKernelGGL((k_mgfm_final<R, false>), GRID(L.n_full), dim3(256), 0, L.stream, L.insts, L.w_full, qin, qout);
}

template <typename R>
void launch_vext(const LaunchCtx& L, bool ogfm) {
    if (ogfm) hipLaunchKernelGGL((k_vext_init<R, true>), GRID(L.n_full), dim3(256), 0, L.stream, L.insts, L.w_full);
    else hipLaunchKernelGGL((k_vext_init<R, false>), GRID(L.n_full), dim3(256), 0, L.stream, L.insts, L.w_full);
    constexpr int N = 20;   // extrapolate_extension_vfield(params, ls, 20, newvelocities)
    for (int k = 0; k < N; ++k)
        hipLaunchKernelGGL((k_vext_iter<R>), GRID(L.n_full), dim3(256), 0, L.stream, L.insts, L.w_full, k & 1,
                           k == N - 1 ? 1 : 0);
}

// K reinitialisation iterations from A into phi: blocked launches of up to
// RT_MAXIT iterations each, or one launch per iteration (x-periodic contexts)
template <typename R>
void launch_reinit_chain(const LaunchCtx& L, int K, int every, int force, int copy_idle) {
    if (L.tb) {
        const int m = (K + RT_MAXIT - 1) / RT_MAXIT;
        int in = 0, done = 0;
        for (int c = 0; c < m; ++c) {
            const int nit = (K - done + (m - c) - 1) / (m - c);   // even split
            const int out = ((m - 1 - c) % 2 == 0) ? 1 : 2;
            if (L.mirror_y)
                hipLaunchKernelGGL((k_ls_reinit_tb<R, false>), GRID(L.n_rt), dim3(RT_THREADS), 0, L.stream, L.insts,
                                   L.w_rt, nit, in, out, every, force, c == 0 ? copy_idle : 0);
            if (L.periodic_y)
                hipLaunchKernelGGL((k_ls_reinit_tb<R, true>), GRID(L.n_rt), dim3(RT_THREADS), 0, L.stream, L.insts,
                                   L.w_rt, nit, in, out, every, force, c == 0 ? copy_idle : 0);
            in = out;
            done += nit;
        }
        return;
    }
    for (int k = 0; k < K; ++k)
        hipLaunchKernelGGL((k_ls_reinit<R>), GRID(L.n_full), dim3(256), 0, L.stream, L.insts, L.w_full, k, K, every,
                           force);
}

template <typename R>
void launch_ls(const LaunchCtx& L, const Houc& H, int every, int iters, bool clsvof) {
    if (iters <= 0) every = 0;   // no reinitialisation iterations: phi = A every step
    // CLSVOF: z's sweeps read the level set at the start of the step (phi,
    // before the copy or reinitialisation from A); the coupling follows
    auto sweeps = [&]() {
        if (!clsvof) return;
        hipLaunchKernelGGL((k_cls_sweep<R>), GRID(L.n_p16), dim3(256), 0, L.stream, L.insts, L.w_p16, 0);
        hipLaunchKernelGGL((k_cls_sweep<R>), GRID(L.n_p16), dim3(256), 0, L.stream, L.insts, L.w_p16, 1);
    };
    auto couple = [&]() {
        if (!clsvof) return;
        hipLaunchKernelGGL((k_clsvof_fix<R>), GRID(L.n_p16), dim3(256), 0, L.stream, L.insts, L.w_p16);
        hipLaunchKernelGGL((k_cls_copy<R>), GRID(L.n_p16), dim3(256), 0, L.stream, L.insts, L.w_p16);
    };
    if (L.tb) {
        static_assert(LTX == RTX && LTY == RTY, "the fused level-set kernel uses the reinitialisation tiles");
#define RK3(O)                                                                                                         \
    do {                                                                                                               \
        if (L.mirror_y)                                                                                             \
            hipLaunchKernelGGL((k_ls_rk3_tb<R, O, false>), GRID(L.n_rt), dim3(1024), 0, L.stream, L.insts, L.w_rt, H); \
        if (L.periodic_y)                                                                                           \
            hipLaunchKernelGGL((k_ls_rk3_tb<R, O, true>), GRID(L.n_rt), dim3(1024), 0, L.stream, L.insts, L.w_rt, H);  \
    } while (0)
        switch (H.n) {
            case 2: RK3(1); break;
            case 4: RK3(3); break;
            case 6: RK3(5); break;
            case 8: RK3(7); break;
            default: RK3(9); break;
        }
#undef RK3
        sweeps();
        if (every > 0) launch_reinit_chain<R>(L, iters, every, 0, 1);
        else hipLaunchKernelGGL((k_ls_copy<R>), GRID(L.n_full), dim3(256), 0, L.stream, L.insts, L.w_full, every);
        couple();
        return;
    }
    hipLaunchKernelGGL((k_ls_stage<R, 1>), GRID(L.n_full), dim3(256), 0, L.stream, L.insts, L.w_full, H);
    hipLaunchKernelGGL((k_ls_stage<R, 2>), GRID(L.n_full), dim3(256), 0, L.stream, L.insts, L.w_full, H);
    hipLaunchKernelGGL((k_ls_stage<R, 3>), GRID(L.n_full), dim3(256), 0, L.stream, L.insts, L.w_full, H);
    sweeps();
    hipLaunchKernelGGL((k_ls_copy<R>), GRID(L.n_full), dim3(256), 0, L.stream, L.insts, L.w_full, every);
    if (every > 0) launch_reinit_chain<R>(L, iters, every, 0, 0);
    couple();
}

template <typename R>
void launch_reinit_now(const LaunchCtx& L, int iters) {
    hipLaunchKernelGGL((k_phi_to_A<R>), GRID(L.n_full), dim3(256), 0, L.stream, L.insts, L.w_full);
    launch_reinit_chain<R>(L, iters, 0, 1, 0);
}

template <typename R>
void launch_reset(const LaunchCtx& L, int narrow_boxes) {
    hipLaunchKernelGGL(k_reset, GRID(L.n_inst), dim3(64), 0, L.stream, L.insts, L.w_inst, narrow_boxes);
}

template <typename R>
void launch_bc(const LaunchCtx& L, int q, int nfluid, int advance) {
    hipLaunchKernelGGL((k_bc<R>), GRID(L.n_frame), dim3(256), 0, L.stream, L.insts, L.w_frame, q, nfluid, advance);
}

void launch_pure_rs(hipStream_t s, int n, int px, double g, double pinf, const double* UL, const double* UR, double* F,
                    unsigned* st) {
    hipLaunchKernelGGL(k_pure_rs, dim3((n + 255) / 256), dim3(256), 0, s, n, px, g, pinf, UL, UR, F, st);
}

// ---------------------------------------------------------------------------
// z: the fraction of each cell's area on fluid 1's side (phi <= 0), from the
// plane through the cell centre with phi's value there and the level set's
// unit normal (central differences, the (1,0) fallback of unit_normal).  The
// area under a line in a rectangle (Scardovelli & Zaleski, J. Comput. Phys.
// 164 (2000), 2D case): with a = |nx| dx <= b = |ny| dy (swapped if not) and
// m = -phi + (a + b)/2, the fraction is 0 (m <= 0), m^2 / (2ab) (m < a),
// (2m - a) / (2b) (m < b), 1 - (a + b - m)^2 / (2ab) (m < a + b), else 1.
// The outermost ring of the padded grid takes 1 or 0 by phi's sign.
// Written to phiA.  work: {inst, tile_y, tile_x, 0}, 64x4 tiles over the
// padded array
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void k_zfrac(const Inst* __restrict__ insts, const int4* __restrict__ work) {
    const int4 w = work[blockIdx.x];
    const Inst& I = insts[w.x];
    const int i = w.y * 4 + (threadIdx.x >> 6);
    const int j = w.z * 64 + (threadIdx.x & 63);
    if (i >= I.H || j >= I.W) return;
    const long long c = cidx(I, i, j);
    const double ph = (double)ld1<R>(I.phi, c);
    double z;
    if (i == 0 || j == 0 || i == I.H - 1 || j == I.W - 1) {
        z = ph <= 0.0 ? 1.0 : 0.0;
    } else {
        double nx, ny;
        unit_normal<R>(I, i, j, nx, ny);
        z = area_fraction(ph, nx, ny, I.dx, I.dy);
    }
    gmem((R*)I.phiA)[c] = (R)z;
}

template <typename R>
void launch_zfrac(const LaunchCtx& L) {
    hipLaunchKernelGGL((k_zfrac<R>), GRID(L.n_full), dim3(256), 0, L.stream, L.insts, L.w_full);
}

// FP32 storage on the FP64 path (analysis): the conserved grids Q[.][q] of
// both fluids and phi rounded to FP32 and back, every padded cell.  Applied
// after each step, it separates what FP32 storage alone does to a job from
// what the FP32 arithmetic of the mixed path adds.
// work: {inst, tile_y, tile_x, 0}, 64x4 tiles over the padded array
__global__ __launch_bounds__(256) void k_round_state_f32(const Inst* __restrict__ insts,
                                                         const int4* __restrict__ work, int q) {
    const int4 w = work[blockIdx.x];
    const Inst& I = insts[w.x];
    const int i = w.y * 4 + (threadIdx.x >> 6);
    const int j = w.z * 64 + (threadIdx.x & 63);
    if (i >= I.H || j >= I.W) return;
    const long long c = cidx(I, i, j);
    for (int f = 0; f < 2; ++f) {
        V4<double> U = ld4<double>(I.Q[f][q], I, c);
#pragma unroll
        for (int k = 0; k < 4; ++k) U.v[k] = (double)(float)U.v[k];
        st4<double>(I.Q[f][q], I, c, U);
    }
    double* ph = (double*)I.phi;
    ph[c] = (double)(float)ph[c];
}

void launch_round_state_f32(const LaunchCtx& L, int q) {
    hipLaunchKernelGGL(k_round_state_f32, GRID(L.n_full), dim3(256), 0, L.stream, L.insts, L.w_full, q);
}

__global__ void k_fill_u32(unsigned* p, int n, unsigned v) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

void launch_fill_u32(hipStream_t s, unsigned* p, int n, unsigned v) {
    hipLaunchKernelGGL(k_fill_u32, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
}

void launch_pure_rs_f32(hipStream_t s, int n, int px, double g, double pinf, const double* UL, const double* UR,
                        double* F, unsigned* st) {
    RSFluid rf;
    rs_fluid(rf, g, pinf);
    hipLaunchKernelGGL(k_pure_rs_f32, dim3((n + 255) / 256), dim3(256), 0, s, n, px, rf, UL, UR, F, st);
}

void launch_mixed_rs(hipStream_t s, int n, double gL, double pL, double gR, double pR, const double* UL,
                     const double* UR, double* out, unsigned* st) {
    hipLaunchKernelGGL(k_mixed_rs, dim3((n + 255) / 256), dim3(256), 0, s, n, gL, pL, gR, pR, UL, UR, out, st);
}

#define INST(R)                                                               \
    template void launch_dt<R>(const LaunchCtx&, int, int, int, int);              \
    template void launch_fs<R>(const LaunchCtx&, bool, int, int, int, int);        \
    template void launch_extrap<R>(const LaunchCtx&, int, int, int);          \
    template bool launch_extrap_tb<R>(const LaunchCtx&, int, int, int, int, int, int); \
    template void launch_iface<R>(const LaunchCtx&, int, int, int);           \
    template void launch_mgfm_final<R>(const LaunchCtx&, int, int, bool);     \
    template void launch_vext<R>(const LaunchCtx&, bool);                       \
    template void launch_ls<R>(const LaunchCtx&, const Houc&, int, int, bool); \
    template void launch_clsvof_init<R>(const LaunchCtx&, int);                 \
    template void launch_reinit_now<R>(const LaunchCtx&, int);                \
    template void launch_bc<R>(const LaunchCtx&, int, int, int);               \
    template void launch_vof<R>(const LaunchCtx&, bool);                       \
    template void launch_reset<R>(const LaunchCtx&, int);                      \
    template void launch_ic<R>(const LaunchCtx&, const IcDesc*, const int*, int, int); \
    template void launch_zfrac<R>(const LaunchCtx&);
INST(double)
INST(float)

}  // namespace gfm
