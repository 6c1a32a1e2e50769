// gfm_capi.hip -- host side of the C ABI declared in include/gfm_hip.h:
// context and instance management, device memory layout, work tables,
// step sequencing and hipGraph capture.  No CPU fallback: every numerical
// phase runs on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "gfm_hip.h"
#include "gfm_kernels.hpp"

using gfm::Inst;

namespace {

thread_local std::string g_err;

int fail(const char* what, hipError_t e) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return GFM_E_HIP;
}
int fail_arg(const char* what) {
    g_err = what;
    return GFM_E_ARG;
}

#define HC(call)                                   \
    do {                                           \
        hipError_t e_ = (call);                    \
        if (e_ != hipSuccess) return fail(#call, e_); \
    } while (0)

struct InstHost {
    gfm_sim_info info;
    gfm_eos eos;
    double cfl, T;
    int max_steps;
    int ntr = 0, tr_keep_y = 0;          // passive tracers (gfm_set_tracers)
    double tr_hdr[gfm::TR_HDR] = {};     // positions, then initial positions
    void* mem = nullptr;   // one allocation per instance
    Inst d;                // descriptor as uploaded (device pointers)
};

void houc_coeffs(int order, gfm::Houc& H) {
    static const int c1[] = {-1, 1};
    static const int c3[] = {1, -6, 3, 2};
    static const int c5[] = {-2, 15, -60, 20, 30, -3};
    static const int c7[] = {3, -28, 126, -420, 105, 252, -42, 4};
    static const int c9[] = {-4, 45, -240, 840, -2520, 504, 1680, -360, 60, -5};
    const int* src = c5;
    double den = 60.0;
    switch (order) {
        case 1: src = c1; den = 1.0; break;
        case 3: src = c3; den = 6.0; break;
        case 7: src = c7; den = 420.0; break;
        case 9: src = c9; den = 2520.0; break;
        default: order = 5; break;
    }
    H.n = order + 1;
    for (int m = 0; m < H.n; ++m) H.c[m] = src[m] / den;
    for (int m = H.n; m < 10; ++m) H.c[m] = 0.0;
    H.o0 = -(H.n / 2);
}

size_t round_up(size_t a, size_t b) { return (a + b - 1) / b * b; }

// The mixed path's ghost band in cells.  The states the interface reads
// (the ghost fluid within two or three cells of it) come from a fill that
// extrapolates from the real fluid and keeps part of each ghost state's own
// value; the flux update moves values two cells per step, so a ghost state
// far out reaches them only over many steps, each of which scales its weight
// down.
constexpr int GHOST_BAND_DEFAULT = 24;

// index of the tile shape named by environment variable `var` ("XxY") in a
// table of n shapes of `stride` ints each ({x, y, ...}); unset: shape 0
int tile_shape(const char* var, const int* table, int stride, int n, int* out) {
    *out = 0;
    const char* v = std::getenv(var);
    if (!v || !v[0]) return GFM_OK;
    int x = 0, y = 0;
    char tail = 0;
    if (std::sscanf(v, "%dx%d%c", &x, &y, &tail) == 2)
        for (int k = 0; k < n; ++k)
            if (table[k * stride] == x && table[k * stride + 1] == y) { *out = k; return GFM_OK; }
    std::string shapes;
    for (int k = 0; k < n; ++k)
        shapes += (k ? ", " : "") + std::to_string(table[k * stride]) + "x" + std::to_string(table[k * stride + 1]);
    return fail_arg((std::string(var) + "=" + v + " is not one of the compiled tile shapes: " + shapes).c_str());
}

}  // namespace

struct gfm_ctx {
    gfm_methods m;
    std::vector<InstHost> inst;
    bool committed = false;
    hipStream_t stream = nullptr;
    Inst* d_insts = nullptr;
    int4* d_work = nullptr;
    gfm::LaunchCtx L{};
    gfm::Houc houc{};
    int cur = 0;                       // current conserved buffer (one-fluid step path)
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    bool tb = false;                   // temporally blocked kernels (no x-periodic instance)
    bool mgfm_extrapolate = false;     // GFM_MGFM_EXTRAPOLATE=1 (GFM_modified only; not the reference)
    bool use_extension_vfield = false; // GFM_base::use_extension_vfield (sim_twofluid.cpp:551 sets it false)
    bool use_slip_BCs = false;         // GFM_base::use_slip_BCs
    int fs_shape = 0;                  // flux tile gfm::FS_SHAPES[fs_shape] (GFM_FS_TILE, e.g. "32x16")
    int tb_shape = 0;                  // blocked extrapolation tile gfm::TB_SHAPES[tb_shape] (GFM_TB_TILE)
    int fs_persist = 0;                // GFM_FS_PERSISTENT=1: the persistent flux kernel (the GPU's CU count)
    unsigned* d_fs_next = nullptr;     // its queue counters, 8 per launch context
    int ghost_band = 0;                // mixed path: ghost states updated within this many cells of their
                                       // fluid's real region (GFM_GHOST_BAND)
    bool round_f32 = false;            // GFM_ROUND_STATE_F32=1 (FP64 two-fluid only; analysis, not the reference)
    // two-fluid step path: the instances in groups, each group's step on its
    // own stream (group 0 on `stream`), forked from and joined back to
    // `stream`, so that the latency-bound phases of one group (the ordered
    // interfacial solve: one workgroup per instance) overlap the
    // throughput-bound phases of the others
    std::vector<gfm::LaunchCtx> GL;
    std::vector<hipStream_t> gstream;  // gstream[0] == stream
    hipEvent_t ev_fork = nullptr;
    std::vector<hipEvent_t> ev_join, ev_stag;
    hipGraphExec_t gexec_multi = nullptr;   // graph_steps two-fluid steps of all groups
    gfm::IcDesc* d_ic = nullptr;       // initial-condition descriptors, one slot per instance
    int* d_icmap = nullptr;            // per instance: its descriptor, or -1
    int graph_steps = 0;
    bool mixed() const { return m.precision == GFM_MIXED; }
    bool twofluid() const { return m.sim_type == GFM_SIM_TWOFLUID; }
    bool clsvof() const { return twofluid() && m.itm == GFM_ITM_CLSVOF; }
    int nfluid() const { return twofluid() ? 2 : 1; }
    size_t es() const { return mixed() ? 4 : 8; }
    // OGFM and OGFM_VOF narrow the update boxes to each fluid's real cells
    int narrow() const { return twofluid() && (m.gfm == GFM_ORIGINAL || m.gfm == GFM_OGFMVOF) ? 1 : 0; }
    bool vof() const { return twofluid() && (m.gfm == GFM_VOFGFM || m.gfm == GFM_OGFMVOF); }
    // the variants whose interfacial list the dt pass counts
    int counted() const { return twofluid() && m.gfm != GFM_ORIGINAL && !vof() ? 1 : 0; }
};

namespace {

template <typename R>
void dispatch_dt(gfm_ctx* c, int q) { gfm::launch_dt<R>(c->L, c->twofluid() ? 1 : 0, q, c->narrow(), 0); }

// N Jacobi iterations of the extrapolation.  Source: conserved Q[.][qin]
// (from_q) or primitives P[.][pstart]; destination: conserved Q[.][qout]
// with W->U and the velocity field (to_q) or the P ping-pong.  Returns the
// P buffer holding the result when !to_q.  all_extra: mode bits of every
// iteration (32 = MF_HOLDIF: interfacial cells keep their ghost state), for
// which there is no blocked kernel.
template <typename R>
int extrap_chain(gfm_ctx* c, const gfm::LaunchCtx& L, int N, bool from_q, int qin, int pstart, bool inject_first,
                 bool to_q, int qout, int last_extra, int all_extra = 0, int band = 0) {
    if (L.tb && N >= 2 && all_extra == 0) {
        // one launch for all N iterations; it writes only the last one's
        // destination, which is the same buffer as in the loop below
        int cur_p = pstart;
        for (int k = 0; k < N && !(k == N - 1 && to_q); ++k) cur_p = (k == 0 && from_q) ? 0 : 1 - cur_p;
        const int min = (from_q ? 1 : 0) | (inject_first ? 2 : 0);
        const int mout = to_q ? (4 | last_extra) : 0;
        if (gfm::launch_extrap_tb<R>(L, N, min, mout, from_q ? qin : pstart, to_q ? qout : cur_p, band))
            return cur_p;
    }
    int cur_p = pstart;
    for (int k = 0; k < N; ++k) {
        int mode = all_extra, bin, bout;
        if (k == 0) {
            bin = from_q ? qin : pstart;
            if (from_q) mode |= 1;            // MF_FIRST
            if (inject_first) mode |= 2;      // MF_INJECT
        } else {
            bin = cur_p;
        }
        if (k == N - 1 && to_q) {
            mode |= 4 | last_extra;           // MF_LAST (| MF_OGFM)
            bout = qout;
        } else {
            bout = (k == 0 && from_q) ? 0 : 1 - cur_p;
            cur_p = bout;
        }
        gfm::launch_extrap<R>(L, mode, bin, bout);
    }
    return cur_p;
}

// GFM_base::set_ghost_states for the context's variant: Q[.][0] -> Q[.][1].
// `mark`, if given, is recorded on L's stream once the interfacial solve (or,
// for GFM_original, the whole chain) is launched.  `band`: the fill of a
// whole mixed-precision step, whose blocked extrapolation skips the tiles
// beyond the ghost band in its first chain and passes them through in its
// last (the flux kernel does not update their ghost states)
template <typename R>
int ghost_chain(gfm_ctx* c, const gfm::LaunchCtx& L, int counted = 0, hipEvent_t mark = nullptr, int band = 0) {
    switch (c->m.gfm) {
        case GFM_ORIGINAL:   // GFM_original.cpp:34-165
            extrap_chain<R>(c, L, 6, true, 0, 0, false, true, 1, 8);
            if (mark) HC(hipEventRecord(mark, L.stream));
            break;
        case GFM_REAL: {     // GFM_real.cpp:32-247
            gfm::launch_iface<R>(L, gfm::V_REAL, 0, counted);
            if (mark) HC(hipEventRecord(mark, L.stream));
            extrap_chain<R>(c, L, 6, true, 0, 0, true, true, 1, 0, 0, band);
            break;
        }
        case GFM_RIEMANN: {  // GFM_riemann.cpp:33-187
            const int p = extrap_chain<R>(c, L, 2, true, 0, 0, false, false, 0, 0, 0, band);
            gfm::launch_iface<R>(L, gfm::V_RIEMANN, p, counted);
            if (mark) HC(hipEventRecord(mark, L.stream));
            extrap_chain<R>(c, L, 6, false, 0, p, true, true, 1, 0, 0, band);
            break;
        }
        case GFM_MODIFIED: { // GFM_modified.cpp:37-239
            const int p = extrap_chain<R>(c, L, 2, true, 0, 0, false, false, 0, 0, 0, band);
            gfm::launch_iface<R>(L, gfm::V_MODIFIED, p, counted);
            if (mark) HC(hipEventRecord(mark, L.stream));
            if (c->mgfm_extrapolate) {
                // not the reference (GFM_MGFM_EXTRAPOLATE=1): the extrapolation of
                // GFM_modified.cpp:202 applied to the new states, which :230-239
                // otherwise discard; the interfacial cells keep theirs
                gfm::launch_mgfm_final<R>(L, 0, p, true);
                extrap_chain<R>(c, L, 6, false, 0, p, false, true, 1, 0, 32);
            } else {
                gfm::launch_mgfm_final<R>(L, 0, 1);
            }
            break;
        }
        case GFM_VOFGFM:     // GFM_VOF.cpp:8-325
        case GFM_OGFMVOF:   // OGFM_VOF.cpp:8-274
            gfm::launch_vof<R>(L, c->m.gfm == GFM_OGFMVOF);
            if (mark) HC(hipEventRecord(mark, L.stream));
            // both apply the BCs to their result (GFM_VOF.cpp:302-303,
            // OGFM_VOF.cpp:271-272) and store the real fluid's velocity,
            // whatever use_extension_vfield says
            gfm::launch_bc<R>(L, 1, 2, 0);
            return GFM_OK;
        default: break;
    }
    // the extension velocity field replaces the real fluid's velocity the
    // chain wrote (e.g. GFM_riemann.cpp:152-174)
    if (c->use_extension_vfield) gfm::launch_vext<R>(L, c->m.gfm == GFM_ORIGINAL);
    return GFM_OK;
}

// one two-fluid step of the instances of L (sim_twofluid.cpp:64-88)
template <typename R>
int twofluid_step(gfm_ctx* c, const gfm::LaunchCtx& L, hipEvent_t mark) {
    const bool muscl = c->m.fs == GFM_FS_MUSCL;
    // the dt pass also counts the interfacial cells when the variant lists them
    const int count = c->counted();
    gfm::launch_dt<R>(L, 1, 0, c->narrow(), count);
    int r = ghost_chain<R>(c, L, count, mark, c->mixed() ? 1 : 0);
    if (r) return r;
    gfm::launch_fs<R>(L, muscl, 1, 0, 0, 0);   // the fill read Q[.][0]
    gfm::launch_ls<R>(L, c->houc, c->m.reinit_every, c->m.reinit_iters, c->clsvof());
    gfm::launch_bc<R>(L, 0, 2, 1);
    // FP32 storage of the step's result on the FP64 path: what storing the
    // state in FP32 alone does to a job, without the mixed path's FP32 sweeps
    if (c->round_f32) gfm::launch_round_state_f32(L, 0);
    return GFM_OK;
}

// nsteps two-fluid steps.  With several instance groups, each group's steps
// go to its own stream with no synchronisation between the groups until all
// nsteps are launched; for more than one step, group g starts once group g-1
// has reached its first interfacial solve, so that the groups'
// latency-bound phases fall at different times.  Instances are independent,
// so the result is the same as with one stream.
template <typename R>
int twofluid_steps(gfm_ctx* c, int nsteps) {
    const size_t ng = c->GL.size();
    if (ng <= 1) {
        for (int s = 0; s < nsteps; ++s) {
            int r = twofluid_step<R>(c, c->L, nullptr);
            if (r) return r;
        }
        return GFM_OK;
    }
    const bool stagger = nsteps > 1;
    HC(hipEventRecord(c->ev_fork, c->stream));
    for (int s = 0; s < nsteps; ++s)
        for (size_t g = 0; g < ng; ++g) {
            if (s == 0 && g > 0) HC(hipStreamWaitEvent(c->gstream[g], stagger ? c->ev_stag[g - 1] : c->ev_fork, 0));
            int r = twofluid_step<R>(c, c->GL[g], (stagger && s == 0 && g + 1 < ng) ? c->ev_stag[g] : nullptr);
            if (r) return r;
        }
    for (size_t g = 1; g < ng; ++g) {
        HC(hipEventRecord(c->ev_join[g], c->gstream[g]));
        HC(hipStreamWaitEvent(c->stream, c->ev_join[g], 0));
    }
    return GFM_OK;
}

// nsteps steps of sim_twofluid::run_sim (sim_twofluid.cpp:64-88) or one step
// of sim_onefluid::run_sim (sim_onefluid.cpp:40-54) on conserved buffer cur
template <typename R>
int step_launches(gfm_ctx* c, int cur, int nsteps = 1) {
    if (c->twofluid()) return twofluid_steps<R>(c, nsteps);
    const bool muscl = c->m.fs == GFM_FS_MUSCL;
    gfm::launch_dt<R>(c->L, 0, cur, 0, 0);
    gfm::launch_fs<R>(c->L, muscl, cur, 1 - cur, 1);
    gfm::launch_bc<R>(c->L, 1 - cur, 1, 1);
    return GFM_OK;
}

int copy_q(gfm_ctx* c, int from, int to) {
    for (auto& h : c->inst)
        for (int f = 0; f < c->nfluid(); ++f)
            HC(hipMemcpyAsync(h.d.Q[f][to], h.d.Q[f][from], 4 * h.d.plane * c->es(), hipMemcpyDeviceToDevice,
                              c->stream));
    return GFM_OK;
}

// XCD-aware order: MI355X dispatches workgroup k to XCD k % 8, so
// workgroup p takes tile (p % 8) * ceil(n/8) + p / 8 and each XCD walks a
// contiguous run of tiles whose halos share its L2.  `group` consecutive
// entries form one workgroup's work.
void xcd_order(std::vector<int4>& v, size_t begin, size_t group = 1) {
    const size_t ne = v.size() - begin;
    const size_t n = (ne + group - 1) / group;
    if (n < 16) return;
    const size_t per = (n + 7) / 8;
    std::vector<int4> src(v.begin() + begin, v.end()), out;
    out.reserve(ne);
    for (size_t p = 0; out.size() < ne; ++p) {
        const size_t t = (p % 8) * per + p / 8;
        if (t < n)
            for (size_t q = t * group; q < std::min(ne, (t + 1) * group); ++q) out.push_back(src[q]);
        if (p > 16 * n) break;
    }
    std::copy(out.begin(), out.end(), v.begin() + begin);
}

}  // namespace

static int check_inst(gfm_ctx* c, int inst) {
    if (!c) return fail_arg("null ctx");
    if (!c->committed) return fail_arg("context not committed");
    if (inst < 0 || inst >= (int)c->inst.size()) return fail_arg("instance out of range");
    return GFM_OK;
}

template <typename R>
static int upload_planes(gfm_ctx* c, const InstHost& h, void* dst, const double* src, int ncomp) {
    const Inst& d = h.d;
    std::vector<R> buf((size_t)ncomp * d.plane, R(0));
    for (int i = 0; i < d.H; ++i)
        for (int j = 0; j < d.W; ++j)
            for (int k = 0; k < ncomp; ++k)
                buf[(size_t)k * d.plane + (size_t)i * d.pitch + j] = (R)src[((size_t)i * d.W + j) * ncomp + k];
    HC(hipMemcpyAsync(dst, buf.data(), buf.size() * sizeof(R), hipMemcpyHostToDevice, c->stream));
    HC(hipStreamSynchronize(c->stream));
    return GFM_OK;
}

template <typename R>
static int download_planes(gfm_ctx* c, const InstHost& h, const void* src, double* dst, int ncomp) {
    const Inst& d = h.d;
    std::vector<R> buf((size_t)ncomp * d.plane);
    HC(hipMemcpyAsync(buf.data(), src, buf.size() * sizeof(R), hipMemcpyDeviceToHost, c->stream));
    HC(hipStreamSynchronize(c->stream));
    for (int i = 0; i < d.H; ++i)
        for (int j = 0; j < d.W; ++j)
            for (int k = 0; k < ncomp; ++k)
                dst[((size_t)i * d.W + j) * ncomp + k] = (double)buf[(size_t)k * d.plane + (size_t)i * d.pitch + j];
    return GFM_OK;
}

static int xfer(gfm_ctx* c, int inst, void* dev, double* host, int ncomp, bool up) {
    const InstHost& h = c->inst[inst];
    if (c->mixed())
        return up ? upload_planes<float>(c, h, dev, host, ncomp) : download_planes<float>(c, h, dev, host, ncomp);
    return up ? upload_planes<double>(c, h, dev, host, ncomp) : download_planes<double>(c, h, dev, host, ncomp);
}

static int read_si(gfm_ctx* c, int inst, int* si) {
    HC(hipMemcpyAsync(si, c->inst[inst].d.si, gfm::SI_N * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HC(hipStreamSynchronize(c->stream));
    return GFM_OK;
}

static int ready(gfm_ctx* c) {
    if (!c) return fail_arg("null ctx");
    if (!c->committed) return fail_arg("context not committed");
    return GFM_OK;
}

static int post_launch(gfm_ctx* c) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail("kernel launch", e);
    return GFM_OK;
}

// the CLSVOF tracker's initialisation() (z from the level set) for one
// instance, or every instance with inst < 0: whenever the host sets or
// reinitialises phi, z follows it
static int clsvof_init(gfm_ctx* c, int inst) {
    if (!c->clsvof()) return GFM_OK;
    if (c->mixed()) gfm::launch_clsvof_init<float>(c->L, inst);
    else gfm::launch_clsvof_init<double>(c->L, inst);
    return post_launch(c);
}

// kind: 0 pure (FP64), 1 mixed, 2 pure with the mixed-precision face solver
static int rs_common(int n, const double* UL, const double* UR, double* out, unsigned* status, int kind,
                     int px, double a, double b, double cc, double d) {
    if (n < 0 || (n > 0 && (!UL || !UR || !out))) return fail_arg("bad arguments");
    if (n == 0) { if (status) *status = 0; return GFM_OK; }
    double *dL = nullptr, *dR = nullptr, *dO = nullptr;
    unsigned* dS = nullptr;
    const size_t bytes = (size_t)n * 4 * sizeof(double);
    HC(hipMalloc((void**)&dL, bytes));
    HC(hipMalloc((void**)&dR, bytes));
    HC(hipMalloc((void**)&dO, bytes));
    HC(hipMalloc((void**)&dS, sizeof(unsigned)));
    HC(hipMemcpy(dL, UL, bytes, hipMemcpyHostToDevice));
    HC(hipMemcpy(dR, UR, bytes, hipMemcpyHostToDevice));
    HC(hipMemset(dS, 0, sizeof(unsigned)));
    if (kind == 1) gfm::launch_mixed_rs(nullptr, n, a, b, cc, d, dL, dR, dO, dS);
    else if (kind == 2) gfm::launch_pure_rs_f32(nullptr, n, px, a, b, dL, dR, dO, dS);
    else gfm::launch_pure_rs(nullptr, n, px, a, b, dL, dR, dO, dS);
    HC(hipGetLastError());
    HC(hipDeviceSynchronize());
    unsigned st = 0;
    HC(hipMemcpy(out, dO, bytes, hipMemcpyDeviceToHost));
    HC(hipMemcpy(&st, dS, sizeof(unsigned), hipMemcpyDeviceToHost));
    (void)hipFree(dL); (void)hipFree(dR); (void)hipFree(dO); (void)hipFree(dS);
    if (status) *status = st;
    return GFM_OK;
}

// Staging of the allocation-free face solver (gfm_pure_rs_solve_noalloc):
// the status word, then the left states, right states and fluxes of up to
// NOALLOC_MAX_FACES faces.  Every GPU the library runs on holds its own copy
// from the load of the code object until the process ends; host code reaches
// it only through the per-device address from hipGetSymbolAddress.
constexpr int NOALLOC_MAX_FACES = 1 << 18;
constexpr size_t NOALLOC_FACE_BYTES = (size_t)NOALLOC_MAX_FACES * 4 * sizeof(double);
constexpr size_t NOALLOC_BYTES = 256 + 3 * NOALLOC_FACE_BYTES;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch[NOALLOC_MAX_DEVICES] = {};

// the current device's copy of `scratch`, looked up once per device
int scratch_of_current_device(unsigned char** out) {
    int dev = 0;
    HC(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail_arg("device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        HC(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail_arg("no device address for the scratch array");
        g_scratch[dev] = p;
    }
    *out = (unsigned char*)g_scratch[dev];
    return GFM_OK;
}

}  // namespace

namespace {

// work tables of a set of instances, appended to `all`
struct TableOffsets {
    size_t o[11];
    int n[11];
};

TableOffsets append_tables(const gfm_ctx* c, const std::vector<Inst>& descs, const std::vector<int>& members,
                           std::vector<int4>& all) {
    std::vector<int4> w_full, w_int, w_fs, w_frame, w_inst, w_ich, w_tb, w_rt, w_seg, w_vseg, w_p16;
    const int nf = c->nfluid();
    for (int ki : members) {
        const Inst& d = descs[ki];
        const gfm_sim_info& s = c->inst[ki].info;
        const int fx = (d.W + 63) / 64, fy = (d.H + 3) / 4;
        for (int ty = 0; ty < fy; ++ty)
            for (int tx = 0; tx < fx; ++tx) w_full.push_back(make_int4(ki, ty, tx, 0));
        if (c->clsvof())   // the tracker's full-array kernels: 64x16 tiles, four rows per thread
            for (int ty = 0; ty < (d.H + 15) / 16; ++ty)
                for (int tx = 0; tx < fx; ++tx) w_p16.push_back(make_int4(ki, ty, tx, 0));
        const int tix = (s.Nx + 63) / 64, tiy = (s.Ny + 3) / 4;
        int li = 0;
        for (int ty = 0; ty < tiy; ++ty)
            for (int tx = 0; tx < tix; ++tx) w_int.push_back(make_int4(ki, ty, tx, li++));
        const int fsx = gfm::FS_SHAPES[c->fs_shape][0], fsy = gfm::FS_SHAPES[c->fs_shape][1];
        const int sx = (s.Nx + fsx - 1) / fsx, sy = (s.Ny + fsy - 1) / fsy;
        for (int f = 0; f < nf; ++f)
            for (int ty = 0; ty < sy; ++ty)
                for (int tx = 0; tx < sx; ++tx) w_fs.push_back(make_int4(ki, ty, tx, f));
        const int strip_n[4] = {d.g * d.W, d.g * d.W, d.g * d.Ny, d.g * d.Ny};
        for (int st = 0; st < 4; ++st)
            for (int o = 0; o < strip_n[st]; o += 256) w_frame.push_back(make_int4(ki, st, o, 0));
        w_inst.push_back(make_int4(ki, 0, 0, 0));
        if (c->twofluid()) {
            const int tbx = gfm::TB_SHAPES[c->tb_shape][0], tby = gfm::TB_SHAPES[c->tb_shape][1];
            for (int ty = 0; ty < (s.Ny + tby - 1) / tby; ++ty)
                for (int tx = 0; tx < (s.Nx + tbx - 1) / tbx; ++tx) w_tb.push_back(make_int4(ki, ty, tx, 0));
        }
        {
            const int nsx = (s.Nx + 255) / 256;
            for (int r = 0; r < s.Ny; ++r)
                for (int x = 0; x < nsx; ++x) w_seg.push_back(make_int4(ki, r, x, r * nsx + x));
        }
        if (c->twofluid())
            for (int ty = 0; ty < (s.Ny + 31) / 32; ++ty)
                for (int tx = 0; tx < (s.Nx + 63) / 64; ++tx) w_rt.push_back(make_int4(ki, ty, tx, 0));
        if (c->twofluid())
            for (int o = 0; o < d.icap; o += 256) w_ich.push_back(make_int4(ki, o, 0, 0));
        if (c->vof()) {
            const int nsx = (s.Nx + 6 + 255) / 256;
            for (int r = 0; r < s.Ny + 6; ++r)
                for (int x = 0; x < nsx; ++x) w_vseg.push_back(make_int4(ki, r, x, r * nsx + x));
        }
    }
    xcd_order(w_full, 0);
    xcd_order(w_p16, 0);
    xcd_order(w_fs, 0);
    xcd_order(w_tb, 0);
    xcd_order(w_rt, 0);
    xcd_order(w_seg, 0, gfm::SEG_PER_WG);
    xcd_order(w_vseg, 0, gfm::SEG_PER_WG);
    if (w_ich.empty()) w_ich.push_back(make_int4(members.empty() ? 0 : members[0], 1 << 30, 0, 0));
    const std::vector<int4>* t[11] = {&w_full, &w_int, &w_fs, &w_frame, &w_inst, &w_ich, &w_tb, &w_rt, &w_seg,
                                      &w_vseg, &w_p16};
    TableOffsets r{};
    for (int q = 0; q < 11; ++q) {
        r.o[q] = all.size();
        r.n[q] = (int)t[q]->size();
        all.insert(all.end(), t[q]->begin(), t[q]->end());
    }
    return r;
}

gfm::LaunchCtx launch_ctx(const gfm_ctx* c, const TableOffsets& t, hipStream_t stream, int iface_plain,
                          const std::vector<int>& members) {
    gfm::LaunchCtx L{};
    L.stream = stream;
    L.insts = c->d_insts;
    int4* w = c->d_work;
    L.w_full = w + t.o[0]; L.n_full = t.n[0];
    L.w_int = w + t.o[1]; L.n_int = t.n[1];
    L.w_fs = w + t.o[2]; L.n_fs = t.n[2];
    L.w_frame = w + t.o[3]; L.n_frame = t.n[3];
    L.w_inst = w + t.o[4]; L.n_inst = t.n[4];
    L.w_ich = w + t.o[5]; L.n_ich = t.n[5];
    L.w_tb = w + t.o[6]; L.n_tb = t.n[6];
    L.w_rt = w + t.o[7]; L.n_rt = t.n[7];
    L.w_seg = w + t.o[8]; L.n_seg = t.n[8];
    L.w_vseg = w + t.o[9]; L.n_vseg = t.n[9];
    L.w_p16 = w + t.o[10]; L.n_p16 = t.n[10];
    L.tb = c->tb ? 1 : 0;
    for (int k : members) {
        if (c->inst[k].info.BC_B == GFM_BC_PERIODIC) L.periodic_y = 1;
        else L.mirror_y = 1;
    }
    L.iface_plain = iface_plain;
    L.fs_shape = c->fs_shape;
    L.fs_persist = c->fs_persist;
    L.fs_next = nullptr;   // set by the caller
    L.tb_shape = c->tb_shape;
    return L;
}

}  // namespace

namespace {

// graph of `launch`'s work, captured on `stream`; `forked` are the streams
// the work forks to, each of which must have joined back before the capture
// ends.  *out is set only to a checked, instantiated graph.
template <typename F>
int capture(hipStream_t stream, const std::vector<hipStream_t>& forked, hipGraphExec_t* out, F launch) {
    hipGraph_t g = nullptr;
    HC(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    int r = launch();
    hipError_t ec = hipStreamEndCapture(stream, &g);
    if (r) { if (g) (void)hipGraphDestroy(g); return r; }
    if (ec != hipSuccess) return fail("hipStreamEndCapture", ec);
    if (!g) return fail_arg("capture produced no graph");
    for (hipStream_t s : forked) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        hipError_t e = hipStreamIsCapturing(s, &cs);
        if (e != hipSuccess || cs != hipStreamCaptureStatusNone) {
            (void)hipGraphDestroy(g);
            return e != hipSuccess ? fail("hipStreamIsCapturing", e) : fail_arg("a forked stream did not join the capture");
        }
    }
    size_t nodes = 0;
    hipError_t en = hipGraphGetNodes(g, nullptr, &nodes);
    if (en != hipSuccess || nodes == 0) {
        (void)hipGraphDestroy(g);
        return en != hipSuccess ? fail("hipGraphGetNodes", en) : fail_arg("captured graph is empty");
    }
    hipGraphExec_t x = nullptr;
    hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess || !x) return e != hipSuccess ? fail("hipGraphInstantiate", e) : fail_arg("no graph executable");
    *out = x;
    return GFM_OK;
}

}  // namespace

extern "C" {

const char* gfm_last_error(void) { return g_err.c_str(); }

int gfm_ctx_create(const gfm_methods* methods, gfm_ctx** out) {
    if (!methods || !out) return fail_arg("null argument");
    if (methods->sim_type != GFM_SIM_ONEFLUID && methods->sim_type != GFM_SIM_TWOFLUID) return fail_arg("sim_type");
    if (methods->fs != GFM_FS_MUSCL && methods->fs != GFM_FS_GODUNOV) return fail_arg("fs");
    if (methods->gfm < GFM_ORIGINAL || methods->gfm > GFM_OGFMVOF) return fail_arg("gfm");
    if (methods->precision != GFM_FP64 && methods->precision != GFM_MIXED) return fail_arg("precision");
    const int o = methods->ls_order;
    if (o != 1 && o != 3 && o != 5 && o != 7 && o != 9) return fail_arg("ls_order must be 1,3,5,7 or 9");
    if (methods->reinit_every < 0 || methods->reinit_iters < 0) return fail_arg("reinit");
    if (methods->itm != GFM_ITM_LS && methods->itm != GFM_ITM_CLSVOF) return fail_arg("itm");
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) return fail("hipGetDeviceCount (no GPU)", e == hipSuccess ? hipErrorNoDevice : e);
    gfm_ctx* c = new gfm_ctx();
    c->m = *methods;
    houc_coeffs(o, c->houc);
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete c; return fail("hipStreamCreate", e); }
    *out = c;
    return GFM_OK;
}

int gfm_ctx_destroy(gfm_ctx* c) {
    if (!c) return GFM_OK;
    // teardown: errors here leave nothing to recover, so they are not reported
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& g : c->gexec)
        if (g) (void)hipGraphExecDestroy(g);
    if (c->gexec_multi) (void)hipGraphExecDestroy(c->gexec_multi);
    for (auto& h : c->inst)
        if (h.mem) (void)hipFree(h.mem);
    if (c->d_insts) (void)hipFree(c->d_insts);
    if (c->d_ic) (void)hipFree(c->d_ic);
    if (c->d_icmap) (void)hipFree(c->d_icmap);
    if (c->d_work) (void)hipFree(c->d_work);
    if (c->d_fs_next) (void)hipFree(c->d_fs_next);
    for (size_t g = 1; g < c->gstream.size(); ++g) (void)hipStreamDestroy(c->gstream[g]);
    for (size_t g = 1; g < c->ev_join.size(); ++g) (void)hipEventDestroy(c->ev_join[g]);
    for (auto ev : c->ev_stag) (void)hipEventDestroy(ev);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return GFM_OK;
}

int gfm_instance_add(gfm_ctx* c, const gfm_sim_info* info, const gfm_eos* eos, double cfl, double T, int max_steps,
                     int* inst_id) {
    if (!c || !info || !eos) return fail_arg("null argument");
    if (c->committed) return fail_arg("context already committed");
    if (info->Nx < 1 || info->Ny < 1) return fail_arg("Nx, Ny must be positive");
    const int gmin = c->twofluid() ? 6 : 2;   // MUSCL stencil 2; HOUC-9 stencil 5 + normal 1
    if (info->numGC < gmin) return fail_arg("numGC too small for the stencils");
    if (c->twofluid() && info->numGC < (c->m.ls_order + 1) / 2 + 1) return fail_arg("numGC too small for HOUC order");
    if (!(info->dx > 0.0) || !(info->dy > 0.0)) return fail_arg("dx, dy must be positive");
    const int bcs[4] = {info->BC_L, info->BC_R, info->BC_B, info->BC_T};
    for (int b : bcs)
        if (b < GFM_BC_TRANSMISSIVE || b > GFM_BC_PERIODIC) return fail_arg("bad BC");
    if ((info->BC_L == GFM_BC_PERIODIC) != (info->BC_R == GFM_BC_PERIODIC)) return fail_arg("periodic BCs come in L/R pairs");
    if ((info->BC_B == GFM_BC_PERIODIC) != (info->BC_T == GFM_BC_PERIODIC)) return fail_arg("periodic BCs come in B/T pairs");
    if (info->BC_L == GFM_BC_PERIODIC && info->Nx < info->numGC) return fail_arg("periodic x needs Nx >= numGC");
    if (info->BC_B == GFM_BC_PERIODIC && info->Ny < info->numGC) return fail_arg("periodic y needs Ny >= numGC");
    if (info->Nx < info->numGC || info->Ny < info->numGC) return fail_arg("mirror BCs need Nx, Ny >= numGC");
    InstHost h;
    h.info = *info;
    h.eos = *eos;
    h.cfl = cfl;
    h.T = T;
    h.max_steps = max_steps;
    c->inst.push_back(h);
    if (inst_id) *inst_id = (int)c->inst.size() - 1;
    return GFM_OK;
}

int gfm_num_instances(const gfm_ctx* c) { return c ? (int)c->inst.size() : 0; }

int gfm_set_gfm_flags(gfm_ctx* c, int use_extension_vfield, int use_slip_BCs) {
    if (!c) return fail_arg("null ctx");
    if (c->committed) return fail_arg("the GFM flags are set before gfm_ctx_commit");
    if (!c->twofluid() && (use_extension_vfield || use_slip_BCs)) return fail_arg("GFM flags need a two-fluid context");
    if (use_slip_BCs && c->m.gfm == GFM_ORIGINAL) return fail_arg("GFM_original has no slip BCs (GFM_original.cpp:21)");
    if (use_slip_BCs && c->m.gfm == GFM_OGFMVOF) return fail_arg("OGFM_VOF has no slip BCs (OGFM_VOF.cpp:20-21)");
    c->use_extension_vfield = use_extension_vfield != 0;
    c->use_slip_BCs = use_slip_BCs != 0;
    return GFM_OK;
}

int gfm_ctx_commit(gfm_ctx* c) {
    if (!c) return fail_arg("null ctx");
    if (c->committed) return GFM_OK;
    if (c->inst.empty()) return fail_arg("no instances");
    // tile shapes (the autotuner's choice, or the defaults): "XxY" from the
    // shape tables; anything else is an error rather than a silent default
    {
        // the mixed path's ghost band, in cells (GFM_GHOST_BAND; 0: update every ghost state)
        const char* gb = std::getenv("GFM_GHOST_BAND");
        c->ghost_band = gb ? std::atoi(gb) : GHOST_BAND_DEFAULT;
        if (c->ghost_band < 0) return fail_arg("GFM_GHOST_BAND must be >= 0");
    }
    int r = tile_shape("GFM_FS_TILE", &gfm::FS_SHAPES[0][0], 2, gfm::FS_NSHAPE, &c->fs_shape);
    if (!r) r = tile_shape("GFM_TB_TILE", &gfm::TB_SHAPES[0][0], 3, gfm::TB_NSHAPE, &c->tb_shape);
    if (r) return r;
    const size_t es = c->es();
    std::vector<Inst> descs;
    bool periodic_x = false;
    for (size_t k = 0; k < c->inst.size(); ++k) {
        InstHost& h = c->inst[k];
        const gfm_sim_info& s = h.info;
        Inst d{};
        d.Nx = s.Nx; d.Ny = s.Ny; d.g = s.numGC;
        d.W = s.Nx + 2 * s.numGC;
        d.H = s.Ny + 2 * s.numGC;
        d.pitch = (int)round_up(d.W, 64);
        d.plane = (long long)round_up((size_t)d.pitch * d.H, 64);
        d.x0 = s.x0; d.y0 = s.y0; d.dx = s.dx; d.dy = s.dy;
        d.bc[0] = s.BC_L; d.bc[1] = s.BC_R; d.bc[2] = s.BC_B; d.bc[3] = s.BC_T;
        d.gam[0] = h.eos.gamma1; d.pinf[0] = h.eos.pinf1;
        d.gam[1] = c->twofluid() ? h.eos.gamma2 : h.eos.gamma1;
        d.pinf[1] = c->twofluid() ? h.eos.pinf2 : h.eos.pinf1;
        for (int f = 0; f < 2; ++f) gfm::rs_fluid(d.rsf[f], d.gam[f], d.pinf[f]);
        d.cfl = h.cfl; d.T = h.T; d.max_steps = h.max_steps;
        d.use_bbox = 1;   // boxes are reset to the full domain each step unless OGFM narrows them
        d.use_slip = c->use_slip_BCs ? 1 : 0;
        d.band = c->mixed() && c->twofluid() ? c->ghost_band * std::max(s.dx, s.dy) : 0.0;
        d.ntile_dt = s.Ny * ((s.Nx + 255) / 256);   // dt partials: one per 256-cell row segment
        // the VOF GFMs list the mixed cells of the interior grown by three cells
        const int lnx = c->vof() ? s.Nx + 6 : s.Nx, lny = c->vof() ? s.Ny + 6 : s.Ny;
        d.icap = c->twofluid() ? std::min(lnx * lny, 16 * (lnx + lny) + 1024) : 1;
        // GFM_IFACE_LDS_ENTRIES lowers the longest list whose ordered solve keeps
        // its flags in LDS (the tests run the global-memory form on small grids)
        const char* il = std::getenv("GFM_IFACE_LDS_ENTRIES");
        d.icap_lds = (il && il[0]) ? std::max(0, std::min(std::atoi(il), gfm::ICAP_LDS)) : gfm::ICAP_LDS;
        // layout of the single allocation
        const size_t P = (size_t)d.plane;
        const int nf = c->nfluid();
        size_t off = 0;
        auto take = [&](size_t bytes) { size_t o = off; off = round_up(off + bytes, 256); return o; };
        size_t oQ[2][2] = {{0, 0}, {0, 0}}, oP[2][2] = {{0, 0}, {0, 0}};
        for (int f = 0; f < nf; ++f)
            for (int b = 0; b < 2; ++b) oQ[f][b] = take(4 * P * es);
        size_t ophi = 0, oA = 0, oB = 0, ovel = 0, oimap = 0, oilist = 0, oside = 0, orec = 0, odone = 0, oseg = 0;
        d.nseg = lny * ((lnx + 255) / 256);
        if (c->twofluid()) {
            for (int f = 0; f < 2; ++f)
                for (int b = 0; b < 2; ++b) oP[f][b] = take(4 * P * es);
            ophi = take(P * es); oA = take(P * es); oB = take(P * es);
            ovel = take(2 * P * es);
            oimap = take(P * sizeof(int));
            oilist = take((size_t)d.icap * sizeof(int));
            oside = take((size_t)d.icap * (c->vof() ? gfm::VSIDE : 8) * sizeof(double));
            orec = take((size_t)d.icap * (c->vof() ? gfm::VOFREC_BYTES : gfm::IFREC_BYTES));
            odone = take((size_t)d.icap * sizeof(int));
            oseg = take((size_t)d.nseg * sizeof(int));   // interfacial counts per segment
        }
        const size_t opart = take((size_t)d.ntile_dt * 4 * sizeof(double));
        d.hist_cap = c->twofluid() ? (int)std::min<long long>((long long)h.max_steps + 1, 1 << 20) : 0;
        if (d.hist_cap < 0) d.hist_cap = 0;
        const size_t ohist = take((size_t)d.hist_cap * 3 * sizeof(double) + 8);
        d.ntr = h.ntr;
        d.tr_keep_y = h.tr_keep_y;
        const size_t otrk = h.ntr ? take((gfm::TR_HDR + (size_t)d.hist_cap * (1 + 4 * h.ntr)) * sizeof(double)) : 0;
        const size_t ozv = c->clsvof() ? take(3 * P * sizeof(double)) : 0;   // CLSVOF: z, first sweep, flags
        const size_t osd = take(gfm::SD_N * sizeof(double));
        const size_t osi = take(gfm::SI_N * sizeof(int));
        hipError_t e = hipMalloc(&h.mem, off);
        if (e != hipSuccess) return fail("hipMalloc (instance state)", e);
        HC(hipMemsetAsync(h.mem, 0, off, c->stream));
        char* base = (char*)h.mem;
        for (int f = 0; f < 2; ++f)
            for (int b = 0; b < 2; ++b) {
                d.Q[f][b] = base + oQ[f < nf ? f : 0][b];
                d.P[f][b] = c->twofluid() ? (void*)(base + oP[f][b]) : nullptr;
            }
        if (c->twofluid()) {
            d.phi = base + ophi; d.phiA = base + oA; d.phiB = base + oB; d.vel = base + ovel;
            d.imap = (int*)(base + oimap); d.ilist = (int*)(base + oilist); d.side = (double*)(base + oside);
            d.irec = base + orec; d.idone = (int*)(base + odone); d.iseg = (int*)(base + oseg);
            HC(hipMemsetAsync(d.imap, 0xff, P * sizeof(int), c->stream));   // -1 everywhere
        }
        d.partials = (double*)(base + opart);
        d.mass_hist = (double*)(base + ohist);
        d.trk = h.ntr ? (double*)(base + otrk) : nullptr;
        d.zv = c->clsvof() ? (double*)(base + ozv) : nullptr;
        if (h.ntr) HC(hipMemcpyAsync(d.trk, h.tr_hdr, sizeof(h.tr_hdr), hipMemcpyHostToDevice, c->stream));
        d.sd = (double*)(base + osd);
        d.si = (int*)(base + osi);
        h.d = d;
        descs.push_back(d);
        if (s.BC_L == GFM_BC_PERIODIC || s.BC_R == GFM_BC_PERIODIC) periodic_x = true;
    }
    const char* mx = std::getenv("GFM_MGFM_EXTRAPOLATE");
    c->mgfm_extrapolate = c->m.gfm == GFM_MODIFIED && mx && mx[0] == '1';
    const char* rf = std::getenv("GFM_ROUND_STATE_F32");
    c->round_f32 = c->twofluid() && !c->mixed() && rf && rf[0] == '1';
    const char* no_tb = std::getenv("GFM_NO_TEMPORAL_BLOCKING");
    // the blocked kernels wrap the rows of y-periodic instances themselves
    c->tb = c->twofluid() && !periodic_x && !(no_tb && no_tb[0] == '1');
    // GFM_NO_TEMPORAL_BLOCKING=1 / GFM_IFACE_PLAIN_ROUNDS=1 select the simpler
    // schedules that the tests compare the default ones against
    const char* plain = std::getenv("GFM_IFACE_PLAIN_ROUNDS");
    const int iface_plain = (plain && plain[0] == '1') ? 1 : 0;
    HC(hipMalloc((void**)&c->d_ic, descs.size() * sizeof(gfm::IcDesc)));
    HC(hipMalloc((void**)&c->d_icmap, descs.size() * sizeof(int)));
    HC(hipMalloc((void**)&c->d_insts, descs.size() * sizeof(Inst)));
    HC(hipMemcpyAsync(c->d_insts, descs.data(), descs.size() * sizeof(Inst), hipMemcpyHostToDevice, c->stream));
    // step groups: GFM_STEP_GROUPS (maximum 4, the hardware queues a
    // process gets) groups of instances with about equal cell counts,
    // largest instances first.  Default 2 in mixed precision, 4 in FP64: on
    // 64 shocked bubbles the mixed path measured 3.80 ms per step with two
    // groups, 3.92 with three, 3.94 with four and 4.51 on one stream; the
    // FP64 path, whose kernels take longer next to the same interfacial
    // solve, 12.7 ms with four groups and 13.9 with two (6 and 8 groups were
    // slower for both).
    const int ninst = (int)c->inst.size();
    int ng = 1;
    if (c->twofluid()) {
        const char* eg = std::getenv("GFM_STEP_GROUPS");
        ng = eg ? std::atoi(eg) : (c->mixed() ? 2 : 4);
        ng = std::max(1, std::min(ng, std::min(ninst, 4)));
    }
    std::vector<std::vector<int>> members(ng);
    {
        std::vector<int> order(ninst);
        for (int k = 0; k < ninst; ++k) order[k] = k;
        auto cells = [&](int k) { return (long long)c->inst[k].info.Nx * c->inst[k].info.Ny; };
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cells(a) > cells(b); });
        std::vector<long long> load(ng, 0);
        for (int k : order) {
            const int g = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            members[g].push_back(k);
            load[g] += cells(k);
        }
        for (auto& m : members) std::sort(m.begin(), m.end());
    }
    std::vector<int> everyone(ninst);
    for (int k = 0; k < ninst; ++k) everyone[k] = k;
    // device work tables: the whole context's, then each group's (if more than one)
    std::vector<int4> all;
    std::vector<TableOffsets> offs;
    offs.push_back(append_tables(c, descs, everyone, all));
    if (ng > 1)
        for (int g = 0; g < ng; ++g) offs.push_back(append_tables(c, descs, members[g], all));
    HC(hipMalloc((void**)&c->d_work, all.size() * sizeof(int4)));
    HC(hipMemcpyAsync(c->d_work, all.data(), all.size() * sizeof(int4), hipMemcpyHostToDevice, c->stream));
    {
        // the persistent flux kernel (GFM_FS_PERSISTENT=1): its grid spans the
        // GPU's CUs; queue counters for the context and for each group
        const char* fp = std::getenv("GFM_FS_PERSISTENT");
        c->fs_persist = 0;
        if (fp && fp[0] == '1') {
            int dev = 0, ncu = 0;
            HC(hipGetDevice(&dev));
            HC(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
            c->fs_persist = std::max(ncu, 1);
        }
        HC(hipMalloc((void**)&c->d_fs_next, 8 * sizeof(unsigned) * (ng + 1)));
    }
    c->L = launch_ctx(c, offs[0], c->stream, iface_plain, everyone);
    c->L.fs_next = c->d_fs_next;
    {
        // steps per graph of a grouped context (GFM_GRAPH_STEPS, default 100)
        const char* gs = std::getenv("GFM_GRAPH_STEPS");
        c->graph_steps = std::max(2, gs ? std::atoi(gs) : 100);
    }
    c->GL.clear();
    c->gstream.assign(1, c->stream);
    c->ev_join.assign(1, nullptr);
    if (ng > 1) {
        HC(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
        c->ev_stag.assign(ng - 1, nullptr);
        for (auto& ev : c->ev_stag) HC(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        for (int g = 0; g < ng; ++g) {
            if (g > 0) {
                hipStream_t st = nullptr;
                hipEvent_t ev = nullptr;
                HC(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
                c->gstream.push_back(st);
                HC(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
                c->ev_join.push_back(ev);
            }
            c->GL.push_back(launch_ctx(c, offs[1 + g], c->gstream[g], iface_plain, members[g]));
            c->GL.back().fs_next = c->d_fs_next + 8 * (1 + g);
        }
    }
    // initial scalars: t = 0, step 0, active
    for (auto& h : c->inst) {
        int si[gfm::SI_N] = {0};
        si[gfm::SI_ACTIVE] = 1;
        const int g = h.info.numGC;
        const int full[4] = {g, g + h.info.Ny, g, g + h.info.Nx};
        for (int f = 0; f < 2; ++f)
            for (int q = 0; q < 4; ++q) si[gfm::SI_BBOX + 4 * f + q] = full[q];
        HC(hipMemcpyAsync(h.d.si, si, sizeof(si), hipMemcpyHostToDevice, c->stream));
    }
    HC(hipStreamSynchronize(c->stream));
    c->committed = true;
    return GFM_OK;
}


int gfm_upload_state(gfm_ctx* c, int inst, const double* g1, const double* g2, const double* phi) {
    int r = check_inst(c, inst);
    if (r) return r;
    if (!g1) return fail_arg("grid1 is NULL");
    InstHost& h = c->inst[inst];
    const int q = c->twofluid() ? 0 : c->cur;
    if ((r = xfer(c, inst, h.d.Q[0][q], (double*)g1, 4, true))) return r;
    if (c->twofluid()) {
        if (!g2 || !phi) return fail_arg("two-fluid upload needs grid2 and phi");
        if ((r = xfer(c, inst, h.d.Q[1][0], (double*)g2, 4, true))) return r;
        if ((r = xfer(c, inst, h.d.phi, (double*)phi, 1, true))) return r;
        if ((r = clsvof_init(c, inst))) return r;
    }
    return GFM_OK;
}

int gfm_download_state(gfm_ctx* c, int inst, double* g1, double* g2, double* phi) {
    int r = check_inst(c, inst);
    if (r) return r;
    InstHost& h = c->inst[inst];
    const int q = c->twofluid() ? 0 : c->cur;
    if (g1 && (r = xfer(c, inst, h.d.Q[0][q], g1, 4, false))) return r;
    if (c->twofluid()) {
        if (g2 && (r = xfer(c, inst, h.d.Q[1][0], g2, 4, false))) return r;
        if (phi && (r = xfer(c, inst, h.d.phi, phi, 1, false))) return r;
    }
    return GFM_OK;
}

int gfm_ic_device(gfm_ctx* c, int n, const int* inst, const gfm_ic_desc* desc) {
    int r = ready(c);
    if (r) return r;
    if (n < 0 || (n > 0 && (!inst || !desc))) return fail_arg("bad arguments");
    static_assert(sizeof(gfm::IcDesc) == sizeof(gfm_ic_desc), "descriptor layouts differ");
    const int ninst = (int)c->inst.size();
    std::vector<int> map(ninst, -1);
    for (int k = 0; k < n; ++k) {
        if (inst[k] < 0 || inst[k] >= ninst) return fail_arg("instance out of range");
        const int kind = desc[k].kind;
        if (c->twofluid() ? (kind < GFM_IC_BUBBLE || kind > GFM_IC_PLANAR) : (kind != GFM_IC_TTC && kind != GFM_IC_EXPLOSION))
            return fail_arg("initial condition kind does not fit the context");
        if (kind == GFM_IC_TTC && (desc[k].test_case % 10 < 1 || desc[k].test_case % 10 > 5 || desc[k].test_case > 15))
            return fail_arg("TTC test case");
        if (map[inst[k]] >= 0) return fail_arg("instance listed twice");
        map[inst[k]] = k;
    }
    if (n == 0) return GFM_OK;
    HC(hipMemcpyAsync(c->d_ic, desc, (size_t)n * sizeof(gfm_ic_desc), hipMemcpyHostToDevice, c->stream));
    HC(hipMemcpyAsync(c->d_icmap, map.data(), (size_t)ninst * sizeof(int), hipMemcpyHostToDevice, c->stream));
    const int q = c->twofluid() ? 0 : c->cur;
    if (c->mixed()) gfm::launch_ic<float>(c->L, c->d_ic, c->d_icmap, q, c->twofluid() ? 1 : 0);
    else gfm::launch_ic<double>(c->L, c->d_ic, c->d_icmap, q, c->twofluid() ? 1 : 0);
    if ((r = post_launch(c))) return r;
    for (int k = 0; k < n; ++k)
        if ((r = clsvof_init(c, inst[k]))) return r;
    HC(hipStreamSynchronize(c->stream));
    for (int k = 0; k < n; ++k)
        if ((r = gfm_set_time(c, inst[k], 0.0, 0))) return r;
    return GFM_OK;
}

int gfm_upload_vfield(gfm_ctx* c, int inst, const double* vf) {
    int r = check_inst(c, inst);
    if (r) return r;
    if (!c->twofluid() || !vf) return fail_arg("velocity field needs a two-fluid context");
    return xfer(c, inst, c->inst[inst].d.vel, (double*)vf, 2, true);
}

int gfm_download_vfield(gfm_ctx* c, int inst, double* vf) {
    int r = check_inst(c, inst);
    if (r) return r;
    if (!c->twofluid() || !vf) return fail_arg("velocity field needs a two-fluid context");
    return xfer(c, inst, c->inst[inst].d.vel, vf, 2, false);
}

int gfm_set_time(gfm_ctx* c, int inst, double t, int nstep) {
    int r = check_inst(c, inst);
    if (r) return r;
    const Inst& d = c->inst[inst].d;
    HC(hipMemcpyAsync(d.sd + gfm::SD_T, &t, sizeof(double), hipMemcpyHostToDevice, c->stream));
    HC(hipMemcpyAsync(d.si + gfm::SI_NSTEP, &nstep, sizeof(int), hipMemcpyHostToDevice, c->stream));
    HC(hipStreamSynchronize(c->stream));
    return GFM_OK;
}

int gfm_set_dt(gfm_ctx* c, int inst, double dt) {
    int r = check_inst(c, inst);
    if (r) return r;
    const Inst& d = c->inst[inst].d;
    HC(hipMemcpyAsync(d.sd + gfm::SD_DT, &dt, sizeof(double), hipMemcpyHostToDevice, c->stream));
    HC(hipStreamSynchronize(c->stream));
    return GFM_OK;
}

int gfm_get_time(gfm_ctx* c, int inst, double* t, int* nstep, double* dt) {
    int r = check_inst(c, inst);
    if (r) return r;
    const Inst& d = c->inst[inst].d;
    double sd[gfm::SD_N];
    int si[gfm::SI_N];
    HC(hipMemcpyAsync(sd, d.sd, sizeof(sd), hipMemcpyDeviceToHost, c->stream));
    HC(hipMemcpyAsync(si, d.si, sizeof(si), hipMemcpyDeviceToHost, c->stream));
    HC(hipStreamSynchronize(c->stream));
    if (t) *t = sd[gfm::SD_T];
    if (nstep) *nstep = si[gfm::SI_NSTEP];
    if (dt) *dt = sd[gfm::SD_DT];
    return GFM_OK;
}


int gfm_status(gfm_ctx* c, int inst, unsigned* flags) {
    int r = check_inst(c, inst);
    if (r) return r;
    int si[gfm::SI_N];
    if ((r = read_si(c, inst, si))) return r;
    if (flags) *flags = (unsigned)si[gfm::SI_STATUS];
    return GFM_OK;
}

int gfm_iface_info(gfm_ctx* c, int inst, int* ncells, int* rounds) {
    int r = check_inst(c, inst);
    if (r) return r;
    int si[gfm::SI_N];
    if ((r = read_si(c, inst, si))) return r;
    if (ncells) *ncells = si[gfm::SI_NIFACE];
    if (rounds) *rounds = si[gfm::SI_ROUNDS];
    return GFM_OK;
}

int gfm_set_bbox(gfm_ctx* c, int inst, const int* b1, const int* b2) {
    int r = check_inst(c, inst);
    if (r) return r;
    if (!b1) return fail_arg("bbox1 is NULL");
    int v[8];
    for (int k = 0; k < 4; ++k) { v[k] = b1[k]; v[4 + k] = b2 ? b2[k] : b1[k]; }
    HC(hipMemcpyAsync(c->inst[inst].d.si + gfm::SI_BBOX, v, sizeof(v), hipMemcpyHostToDevice, c->stream));
    HC(hipStreamSynchronize(c->stream));
    return GFM_OK;
}

int gfm_get_bbox(gfm_ctx* c, int inst, int* b1, int* b2) {
    int r = check_inst(c, inst);
    if (r) return r;
    int si[gfm::SI_N];
    if ((r = read_si(c, inst, si))) return r;
    const gfm_sim_info& s = c->inst[inst].info;
    const int g = s.numGC;
    for (int f = 0; f < 2; ++f) {
        int* b = f == 0 ? b1 : b2;
        if (!b) continue;
        const int* v = si + gfm::SI_BBOX + 4 * f;
        b[0] = std::max(v[0], g); b[1] = std::min(v[1], g + s.Ny);
        b[2] = std::max(v[2], g); b[3] = std::min(v[3], g + s.Nx);
    }
    return GFM_OK;
}

int gfm_mass_history(gfm_ctx* c, int inst, int n, double* out3) {
    int r = check_inst(c, inst);
    if (r) return r;
    const Inst& d = c->inst[inst].d;
    if (n < 0 || n > d.hist_cap || (n > 0 && !out3)) return fail_arg("history length beyond capacity");
    if (n == 0) return GFM_OK;
    HC(hipMemcpyAsync(out3, d.mass_hist, (size_t)n * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HC(hipStreamSynchronize(c->stream));
    return GFM_OK;
}

int gfm_set_tracers(gfm_ctx* c, int inst, int n, const double* xy, int keep_y) {
    if (!c) return fail_arg("null ctx");
    if (inst < 0 || inst >= (int)c->inst.size()) return fail_arg("bad instance index");
    if (!c->twofluid()) return fail_arg("tracers need a two-fluid context (they move in its velocity field)");
    if (n < 0 || n > gfm::TR_MAX || (n > 0 && !xy)) return fail_arg("0 to 4 tracers, positions as x0 y0 x1 y1 ...");
    InstHost& h = c->inst[inst];
    if (c->committed && n != h.ntr) return fail_arg("the number of tracers is fixed at gfm_ctx_commit");
    for (int k = 0; k < 2 * n; ++k) {
        h.tr_hdr[k] = xy[k];
        h.tr_hdr[2 * gfm::TR_MAX + k] = xy[k];
    }
    h.ntr = n;
    h.tr_keep_y = keep_y != 0;
    if (!c->committed || n == 0) return GFM_OK;
    if (!keep_y != !h.d.tr_keep_y) return fail_arg("keep_y is fixed at gfm_ctx_commit");
    HC(hipMemcpyAsync(h.d.trk, h.tr_hdr, sizeof(h.tr_hdr), hipMemcpyHostToDevice, c->stream));
    HC(hipStreamSynchronize(c->stream));
    return GFM_OK;
}

int gfm_tracer_history(gfm_ctx* c, int inst, int n, double* out) {
    int r = check_inst(c, inst);
    if (r) return r;
    const Inst& d = c->inst[inst].d;
    if (d.ntr == 0) return fail_arg("instance has no tracers");
    if (n < 0 || n > d.hist_cap || (n > 0 && !out)) return fail_arg("history length beyond capacity");
    if (n == 0) return GFM_OK;
    HC(hipMemcpyAsync(out, d.trk + gfm::TR_HDR, (size_t)n * (1 + 4 * d.ntr) * sizeof(double),
                      hipMemcpyDeviceToHost, c->stream));
    HC(hipStreamSynchronize(c->stream));
    return GFM_OK;
}

int gfm_set_output_hold(gfm_ctx* c, int inst, double last, double interval) {
    int r = check_inst(c, inst);
    if (r) return r;
    const Inst& d = c->inst[inst].d;
    const double v[2] = {last, interval};
    const int on = interval >= 0.0 ? 1 : 0;
    HC(hipMemcpyAsync(d.sd + gfm::SD_HOLD_LAST, v, sizeof(v), hipMemcpyHostToDevice, c->stream));
    HC(hipMemcpyAsync(d.si + gfm::SI_HOLD, &on, sizeof(int), hipMemcpyHostToDevice, c->stream));
    HC(hipStreamSynchronize(c->stream));
    return GFM_OK;
}

int gfm_download_z(gfm_ctx* c, int inst, double* z) {
    int r = check_inst(c, inst);
    if (r) return r;
    if (!c->twofluid() || !z) return fail_arg("z needs a two-fluid context and an output array");
    // the CLSVOF tracker's own volume fraction (FP64 in both precisions)
    if (c->clsvof()) return download_planes<double>(c, c->inst[inst], c->inst[inst].d.zv, z, 1);
    if (c->mixed()) gfm::launch_zfrac<float>(c->L);
    else gfm::launch_zfrac<double>(c->L);
    HC(hipGetLastError());
    return xfer(c, inst, c->inst[inst].d.phiA, z, 1, false);
}

int gfm_upload_z(gfm_ctx* c, int inst, const double* z) {
    int r = check_inst(c, inst);
    if (r) return r;
    if (!c->clsvof() || !z) return fail_arg("z upload needs a CLSVOF context and an input array");
    return upload_planes<double>(c, c->inst[inst], c->inst[inst].d.zv, z, 1);
}

int gfm_mass_sums(gfm_ctx* c, int inst, double* out2) {
    int r = check_inst(c, inst);
    if (r) return r;
    double sd[gfm::SD_N];
    HC(hipMemcpyAsync(sd, c->inst[inst].d.sd, sizeof(sd), hipMemcpyDeviceToHost, c->stream));
    HC(hipStreamSynchronize(c->stream));
    out2[0] = sd[gfm::SD_MASS1];
    out2[1] = sd[gfm::SD_MASS2];
    return GFM_OK;
}


int gfm_compute_dt(gfm_ctx* c) {
    int r = ready(c);
    if (r) return r;
    const int q = c->twofluid() ? 0 : c->cur;
    if (c->mixed()) dispatch_dt<float>(c, q);
    else dispatch_dt<double>(c, q);
    return post_launch(c);
}

int gfm_set_ghost_states(gfm_ctx* c) {
    int r = ready(c);
    if (r) return r;
    if (!c->twofluid()) return fail_arg("ghost states need a two-fluid context");
    if (c->mixed()) { gfm::launch_reset<float>(c->L, c->narrow()); ghost_chain<float>(c, c->L); }
    else { gfm::launch_reset<double>(c->L, c->narrow()); ghost_chain<double>(c, c->L); }
    if ((r = post_launch(c))) return r;
    // keep the phase-by-phase state in buffer 0 (but not the OGFM boxes,
    // which live in the scalars)
    return copy_q(c, 1, 0);
}

int gfm_fs_update(gfm_ctx* c) {
    int r = ready(c);
    if (r) return r;
    const bool muscl = c->m.fs == GFM_FS_MUSCL;
    const int q = c->twofluid() ? 0 : c->cur;
    // the update writes the cells of the realcells boxes only; every other
    // cell keeps its input value (FS_MUSCL.cpp:162-169, :296-302)
    if ((r = copy_q(c, q, 1 - q))) return r;
    if (c->mixed()) gfm::launch_fs<float>(c->L, muscl, q, 1 - q, 0);
    else gfm::launch_fs<double>(c->L, muscl, q, 1 - q, 0);
    if ((r = post_launch(c))) return r;
    return copy_q(c, 1 - q, q);
}

int gfm_ls_update(gfm_ctx* c) {
    int r = ready(c);
    if (r) return r;
    if (!c->twofluid()) return fail_arg("level set needs a two-fluid context");
    if (c->mixed()) gfm::launch_ls<float>(c->L, c->houc, c->m.reinit_every, c->m.reinit_iters, c->clsvof());
    else gfm::launch_ls<double>(c->L, c->houc, c->m.reinit_every, c->m.reinit_iters, c->clsvof());
    return post_launch(c);
}

int gfm_ls_reinit(gfm_ctx* c, int iters) {
    int r = ready(c);
    if (r) return r;
    if (!c->twofluid()) return fail_arg("level set needs a two-fluid context");
    if (iters <= 0) return GFM_OK;
    if (c->mixed()) gfm::launch_reinit_now<float>(c->L, iters);
    else gfm::launch_reinit_now<double>(c->L, iters);
    if ((r = post_launch(c))) return r;
    return clsvof_init(c, -1);
}

int gfm_apply_bcs(gfm_ctx* c) {
    int r = ready(c);
    if (r) return r;
    const int q = c->twofluid() ? 0 : c->cur;
    if (c->mixed()) gfm::launch_bc<float>(c->L, q, c->nfluid(), 1);
    else gfm::launch_bc<double>(c->L, q, c->nfluid(), 1);
    return post_launch(c);
}

int gfm_extrapolate(gfm_ctx* c, int N) {
    int r = ready(c);
    if (r) return r;
    if (!c->twofluid()) return fail_arg("extrapolation needs a two-fluid context");
    if (N <= 0) return GFM_OK;
    int q = 0;
    for (int k = 0; k < N; ++k) {
        if (c->mixed()) { if (k == 0) gfm::launch_reset<float>(c->L, c->narrow()); gfm::launch_extrap<float>(c->L, 16, q, 1 - q); }
        else { if (k == 0) gfm::launch_reset<double>(c->L, c->narrow()); gfm::launch_extrap<double>(c->L, 16, q, 1 - q); }
        q = 1 - q;
    }
    if ((r = post_launch(c))) return r;
    if (q != 0) return copy_q(c, 1, 0);
    return GFM_OK;
}

int gfm_step(gfm_ctx* c, int nsteps) {
    int r = ready(c);
    if (r) return r;
    if (c->twofluid()) {
        r = c->mixed() ? step_launches<float>(c, 0, nsteps) : step_launches<double>(c, 0, nsteps);
        if (r) return r;
        return post_launch(c);
    }
    for (int s = 0; s < nsteps; ++s) {
        r = c->mixed() ? step_launches<float>(c, c->cur) : step_launches<double>(c, c->cur);
        if (r) return r;
        c->cur = 1 - c->cur;
    }
    return post_launch(c);
}

int gfm_step_graph(gfm_ctx* c, int nsteps) {
    int r = ready(c);
    if (r) return r;
    if (nsteps <= 0) return GFM_OK;
    const int ng = c->twofluid() ? 1 : 2;
    for (int k = 0; k < ng; ++k)
        if (!c->gexec[k] && (r = capture(c->stream, c->gstream, &c->gexec[k], [&]() {
                return c->mixed() ? step_launches<float>(c, k) : step_launches<double>(c, k);
            })))
            return r;
    int s = 0;
    if (c->twofluid() && c->GL.size() > 1 && nsteps >= c->graph_steps) {
        // instance groups: blocks of graph_steps steps in one graph, in which
        // the groups run staggered and independent (twofluid_steps)
        if (!c->gexec_multi && (r = capture(c->stream, c->gstream, &c->gexec_multi, [&]() {
                return c->mixed() ? step_launches<float>(c, 0, c->graph_steps)
                                  : step_launches<double>(c, 0, c->graph_steps);
            })))
            return r;
        for (; s + c->graph_steps <= nsteps; s += c->graph_steps) HC(hipGraphLaunch(c->gexec_multi, c->stream));
    }
    for (; s < nsteps; ++s) {
        const int k = c->twofluid() ? 0 : c->cur;
        HC(hipGraphLaunch(c->gexec[k], c->stream));
        if (!c->twofluid()) c->cur = 1 - c->cur;
    }
    return GFM_OK;
}

int gfm_profile_steps(gfm_ctx* c, int nsteps, double* ms) {
    int r = ready(c);
    if (r) return r;
    if (!ms || nsteps < 0) return fail_arg("bad arguments");
    for (int k = 0; k < 5; ++k) ms[k] = 0.0;
    hipEvent_t ev[6];
    for (auto& e : ev) HC(hipEventCreate(&e));
    const bool muscl = c->m.fs == GFM_FS_MUSCL;
    for (int s = 0; s < nsteps; ++s) {
        const int cur = c->cur;
        HC(hipEventRecord(ev[0], c->stream));
        if (c->twofluid()) {
            if (c->mixed()) {
                gfm::launch_dt<float>(c->L, 1, 0, c->narrow(), c->counted());
                HC(hipEventRecord(ev[1], c->stream));
                ghost_chain<float>(c, c->L, c->counted(), nullptr, 1);
                HC(hipEventRecord(ev[2], c->stream));
                gfm::launch_fs<float>(c->L, muscl, 1, 0, 0, 0);
                HC(hipEventRecord(ev[3], c->stream));
                gfm::launch_ls<float>(c->L, c->houc, c->m.reinit_every, c->m.reinit_iters, c->clsvof());
                HC(hipEventRecord(ev[4], c->stream));
                gfm::launch_bc<float>(c->L, 0, 2, 1);
            } else {
                gfm::launch_dt<double>(c->L, 1, 0, c->narrow(), c->counted());
                HC(hipEventRecord(ev[1], c->stream));
                ghost_chain<double>(c, c->L, c->counted());
                HC(hipEventRecord(ev[2], c->stream));
                gfm::launch_fs<double>(c->L, muscl, 1, 0, 0, 0);
                HC(hipEventRecord(ev[3], c->stream));
                gfm::launch_ls<double>(c->L, c->houc, c->m.reinit_every, c->m.reinit_iters, c->clsvof());
                HC(hipEventRecord(ev[4], c->stream));
                gfm::launch_bc<double>(c->L, 0, 2, 1);
            }
        } else {
            if (c->mixed()) gfm::launch_dt<float>(c->L, 0, cur, 0, 0);
            else gfm::launch_dt<double>(c->L, 0, cur, 0, 0);
            HC(hipEventRecord(ev[1], c->stream));
            HC(hipEventRecord(ev[2], c->stream));
            if (c->mixed()) gfm::launch_fs<float>(c->L, muscl, cur, 1 - cur, 1);
            else gfm::launch_fs<double>(c->L, muscl, cur, 1 - cur, 1);
            HC(hipEventRecord(ev[3], c->stream));
            HC(hipEventRecord(ev[4], c->stream));
            if (c->mixed()) gfm::launch_bc<float>(c->L, 1 - cur, 1, 1);
            else gfm::launch_bc<double>(c->L, 1 - cur, 1, 1);
            c->cur = 1 - cur;
        }
        HC(hipEventRecord(ev[5], c->stream));
        HC(hipEventSynchronize(ev[5]));
        for (int k = 0; k < 5; ++k) {
            float t = 0.f;
            HC(hipEventElapsedTime(&t, ev[k], ev[k + 1]));
            ms[k] += t;
        }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    return post_launch(c);
}

int gfm_sync(gfm_ctx* c) {
    if (!c) return fail_arg("null ctx");
    HC(hipStreamSynchronize(c->stream));
    return GFM_OK;
}


int gfm_pure_rs_solve(int n, int paralleltox, double gamma, double pinf, const double* UL, const double* UR,
                      double* flux, unsigned* status) {
    return rs_common(n, UL, UR, flux, status, 0, paralleltox, gamma, pinf, 0.0, 0.0);
}

int gfm_pure_rs_solve_f32(int n, int paralleltox, double gamma, double pinf, const double* UL, const double* UR,
                          double* flux, unsigned* status) {
    return rs_common(n, UL, UR, flux, status, 2, paralleltox, gamma, pinf, 0.0, 0.0);
}

int gfm_mixed_rs_solve(int n, double gL, double pL, double gR, double pR, const double* UL, const double* UR,
                       double* out4, unsigned* status) {
    return rs_common(n, UL, UR, out4, status, 1, 0, gL, pL, gR, pR);
}

int gfm_pure_rs_solve_noalloc(int n, int paralleltox, double gamma, double pinf, const double* UL, const double* UR,
                              double* flux, unsigned* status) {
    if (n < 0 || (n > 0 && (!UL || !UR || !flux))) return fail_arg("bad arguments");
    if (n > NOALLOC_MAX_FACES) return fail_arg("more faces than the allocation-free solver stages (262144)");
    if (n == 0) { if (status) *status = 0; return GFM_OK; }
    unsigned char* base = nullptr;
    int r = scratch_of_current_device(&base);
    if (r) return r;
    unsigned* dS = (unsigned*)base;
    double* dL = (double*)(base + 256);
    double* dR = (double*)(base + 256 + NOALLOC_FACE_BYTES);
    double* dO = (double*)(base + 256 + 2 * NOALLOC_FACE_BYTES);
    const size_t bytes = (size_t)n * 4 * sizeof(double);
    const hipStream_t s = nullptr;
    gfm::launch_fill_u32(s, dS, 1, 0u);   // the previous call's status stays in the array
    HC(hipMemcpyAsync(dL, UL, bytes, hipMemcpyHostToDevice, s));
    HC(hipMemcpyAsync(dR, UR, bytes, hipMemcpyHostToDevice, s));
    gfm::launch_pure_rs(s, n, paralleltox, gamma, pinf, dL, dR, dO, dS);
    HC(hipGetLastError());
    unsigned st = 0;
    HC(hipMemcpyAsync(flux, dO, bytes, hipMemcpyDeviceToHost, s));
    HC(hipMemcpyAsync(&st, dS, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    HC(hipStreamSynchronize(s));
    if (status) *status = st;
    return GFM_OK;
}

}  // extern "C"
