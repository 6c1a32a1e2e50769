"""ctypes front end of libgfm_hip.so (C ABI in include/gfm_hip.h).

Host arrays use the reference's grid layout: numpy float64 of shape
(Ny+2g, Nx+2g, 4) for a fluid state, (Ny+2g, Nx+2g) for phi and
(Ny+2g, Nx+2g, 2) for the velocity field.  There is no CPU fallback: if the
library is not built, or no GPU is present, the calls raise.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# GFM_HIP_LIB selects another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get("GFM_HIP_LIB", os.path.join(HERE, "libgfm_hip.so"))

BC_TRANSMISSIVE, BC_REFLECTIVE, BC_PERIODIC = 0, 1, 2
SIM_ONEFLUID, SIM_TWOFLUID = 0, 1
FS_MUSCL, FS_GODUNOV = 0, 1
GFM_ORIGINAL, GFM_REAL, GFM_RIEMANN, GFM_MODIFIED, GFM_VOFGFM, GFM_OGFMVOF = 0, 1, 2, 3, 4, 5
FP64, MIXED = 0, 1
ITM_LS, ITM_CLSVOF = 0, 1
ST_UNPHYSICAL, ST_RS_FAIL, ST_IFACE_OVERFLOW, ST_IFACE_ROUNDS, ST_IFACE_LIST = 1, 2, 4, 8, 16

GFM_NAMES = {"OGFM": GFM_ORIGINAL, "realGFM": GFM_REAL, "riemannGFM": GFM_RIEMANN, "MGFM": GFM_MODIFIED,
             "VOFGFM": GFM_VOFGFM, "OGFMVOF": GFM_OGFMVOF}

# every entry point declared in include/gfm_hip.h
EXPORTS = [
    "gfm_ctx_create", "gfm_ctx_destroy", "gfm_instance_add", "gfm_ctx_commit", "gfm_num_instances",
    "gfm_set_gfm_flags",
    "gfm_upload_state", "gfm_download_state", "gfm_upload_vfield", "gfm_download_vfield",
    "gfm_set_time", "gfm_get_time", "gfm_set_dt",
    "gfm_compute_dt", "gfm_set_ghost_states", "gfm_fs_update", "gfm_ls_update", "gfm_apply_bcs",
    "gfm_extrapolate", "gfm_ls_reinit", "gfm_get_bbox", "gfm_set_bbox",
    "gfm_step", "gfm_step_graph", "gfm_sync", "gfm_profile_steps", "gfm_status", "gfm_mass_sums", "gfm_mass_history", "gfm_iface_info",
    "gfm_set_output_hold", "gfm_set_tracers", "gfm_tracer_history", "gfm_download_z", "gfm_upload_z",
    "gfm_pure_rs_solve", "gfm_pure_rs_solve_f32", "gfm_mixed_rs_solve", "gfm_pure_rs_solve_noalloc",
    "gfm_ic_onefluid", "gfm_ic_bubble", "gfm_ic_rmi", "gfm_ic_planar", "gfm_shock_state", "gfm_ic_device",
    "gfm_last_error",
]


class SimInfo(C.Structure):
    _fields_ = [("Nx", C.c_int), ("Ny", C.c_int), ("numGC", C.c_int),
                ("x0", C.c_double), ("y0", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
                ("BC_L", C.c_int), ("BC_R", C.c_int), ("BC_B", C.c_int), ("BC_T", C.c_int)]

    @property
    def shape(self):
        return (self.Ny + 2 * self.numGC, self.Nx + 2 * self.numGC)


class Eos(C.Structure):
    _fields_ = [("gamma1", C.c_double), ("pinf1", C.c_double),
                ("gamma2", C.c_double), ("pinf2", C.c_double)]


class Methods(C.Structure):
    _fields_ = [("sim_type", C.c_int), ("fs", C.c_int), ("gfm", C.c_int), ("ls_order", C.c_int),
                ("reinit_every", C.c_int), ("reinit_iters", C.c_int), ("precision", C.c_int), ("itm", C.c_int)]


IC_TTC, IC_EXPLOSION, IC_BUBBLE, IC_RMI, IC_PLANAR = 1, 2, 3, 4, 5


class IcDesc(C.Structure):
    """gfm_ic_desc: an initial condition computed on the device (gfm_ic_device)"""
    _fields_ = [("kind", C.c_int), ("test_case", C.c_int), ("axis", C.c_int), ("modes", C.c_int),
                ("xc", C.c_double), ("yc", C.c_double), ("R", C.c_double), ("x_if", C.c_double),
                ("amp", C.c_double), ("pos", C.c_double), ("shock_x", C.c_double),
                ("w1", C.c_double * 4), ("w2", C.c_double * 4), ("w3", C.c_double * 4)]


def ic_desc(kind, w1=(0, 0, 0, 0), w2=(0, 0, 0, 0), w3=(0, 0, 0, 0), **kw):
    d = IcDesc()
    d.kind = kind
    for k, v in kw.items():
        setattr(d, k, v)
    for name, w in (("w1", w1), ("w2", w2), ("w3", w3)):
        getattr(d, name)[:] = [float(x) for x in w]
    return d


class GfmError(RuntimeError):
    pass


_lib = None
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_U = C.POINTER(C.c_uint)


def load():
    """Load libgfm_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GfmError("libgfm_hip.so is not built (run __graft_entry__.build() or make -C phd-2d-gfm_amd)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "gfm_ctx_create": [C.POINTER(Methods), C.POINTER(vp)],
        "gfm_ctx_destroy": [vp],
        "gfm_instance_add": [vp, C.POINTER(SimInfo), C.POINTER(Eos), C.c_double, C.c_double, C.c_int, _I],
        "gfm_ctx_commit": [vp],
        "gfm_num_instances": [vp],
        "gfm_set_gfm_flags": [vp, C.c_int, C.c_int],
        "gfm_upload_state": [vp, C.c_int, _D, _D, _D],
        "gfm_download_state": [vp, C.c_int, _D, _D, _D],
        "gfm_upload_vfield": [vp, C.c_int, _D],
        "gfm_download_vfield": [vp, C.c_int, _D],
        "gfm_set_time": [vp, C.c_int, C.c_double, C.c_int],
        "gfm_get_time": [vp, C.c_int, _D, _I, _D],
        "gfm_set_dt": [vp, C.c_int, C.c_double],
        "gfm_compute_dt": [vp], "gfm_set_ghost_states": [vp], "gfm_fs_update": [vp],
        "gfm_ls_update": [vp], "gfm_apply_bcs": [vp],
        "gfm_extrapolate": [vp, C.c_int], "gfm_ls_reinit": [vp, C.c_int],
        "gfm_get_bbox": [vp, C.c_int, _I, _I],
        "gfm_set_bbox": [vp, C.c_int, _I, _I],
        "gfm_step": [vp, C.c_int], "gfm_step_graph": [vp, C.c_int], "gfm_sync": [vp],
        "gfm_profile_steps": [vp, C.c_int, _D],
        "gfm_status": [vp, C.c_int, _U], "gfm_mass_sums": [vp, C.c_int, _D],
        "gfm_mass_history": [vp, C.c_int, C.c_int, _D],
        "gfm_iface_info": [vp, C.c_int, _I, _I],
        "gfm_set_output_hold": [vp, C.c_int, C.c_double, C.c_double],
        "gfm_set_tracers": [vp, C.c_int, C.c_int, _D, C.c_int],
        "gfm_tracer_history": [vp, C.c_int, C.c_int, _D],
        "gfm_download_z": [vp, C.c_int, _D],
        "gfm_upload_z": [vp, C.c_int, _D],
        "gfm_pure_rs_solve": [C.c_int, C.c_int, C.c_double, C.c_double, _D, _D, _D, _U],
        "gfm_pure_rs_solve_f32": [C.c_int, C.c_int, C.c_double, C.c_double, _D, _D, _D, _U],
        "gfm_pure_rs_solve_noalloc": [C.c_int, C.c_int, C.c_double, C.c_double, _D, _D, _D, _U],
        "gfm_mixed_rs_solve": [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, _D, _D, _D, _U],
        "gfm_ic_onefluid": [C.POINTER(SimInfo), C.POINTER(Eos), C.c_int, _D],
        "gfm_ic_bubble": [C.POINTER(SimInfo), C.POINTER(Eos), C.c_double, C.c_double, C.c_double, _D, _D, _D,
                          C.c_double, _D, _D, _D],
        "gfm_ic_rmi": [C.POINTER(SimInfo), C.POINTER(Eos), C.c_double, C.c_double, C.c_int, _D, _D, _D,
                       C.c_double, _D, _D, _D],
        "gfm_ic_planar": [C.POINTER(SimInfo), C.POINTER(Eos), C.c_int, C.c_double, _D, _D, _D, _D, _D],
        "gfm_shock_state": [C.c_double, C.c_double, _D, C.c_double, C.c_int, _D],
        "gfm_ic_device": [vp, C.c_int, _I, C.POINTER(IcDesc)],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.gfm_last_error.restype = C.c_char_p
    L.gfm_last_error.argtypes = []
    _lib = L
    return L


def _check(rc, what):
    if rc != 0:
        raise GfmError("%s failed (%d): %s" % (what, rc, load().gfm_last_error().decode()))


def _dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_D)


def methods(sim_type=SIM_TWOFLUID, fs=FS_MUSCL, gfm=GFM_RIEMANN, ls_order=5, reinit_every=5,
            reinit_iters=10, precision=FP64, itm=ITM_LS):
    return Methods(sim_type, fs, gfm, ls_order, reinit_every, reinit_iters, precision, itm)


class Context:
    """One batch of instances sharing a gfm_methods (a gfm_ctx)."""

    def __init__(self, m):
        L = load()
        self.m = m
        self.h = C.c_void_p()
        _check(L.gfm_ctx_create(C.byref(m), C.byref(self.h)), "gfm_ctx_create")
        self.infos = []
        self.ntr = {}

    def close(self):
        if self.h:
            load().gfm_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def twofluid(self):
        return self.m.sim_type == SIM_TWOFLUID

    def add(self, info, eos, cfl, T=1e300, max_steps=1 << 30):
        k = C.c_int(0)
        _check(load().gfm_instance_add(self.h, C.byref(info), C.byref(eos), cfl, T, max_steps, C.byref(k)),
               "gfm_instance_add")
        self.infos.append(info)
        return k.value

    def set_gfm_flags(self, use_extension_vfield=False, use_slip_BCs=False):
        """GFM_base::use_extension_vfield / use_slip_BCs, before commit"""
        _check(load().gfm_set_gfm_flags(self.h, int(use_extension_vfield), int(use_slip_BCs)), "gfm_set_gfm_flags")

    def commit(self):
        _check(load().gfm_ctx_commit(self.h), "gfm_ctx_commit")

    def upload(self, k, grid1, grid2=None, phi=None):
        _check(load().gfm_upload_state(self.h, k, _dp(grid1), _dp(grid2), _dp(phi)), "gfm_upload_state")

    def download(self, k):
        H, W = self.infos[k].shape
        g1 = np.zeros((H, W, 4))
        g2 = np.zeros((H, W, 4)) if self.twofluid else None
        phi = np.zeros((H, W)) if self.twofluid else None
        _check(load().gfm_download_state(self.h, k, _dp(g1), _dp(g2), _dp(phi)), "gfm_download_state")
        return (g1, g2, phi) if self.twofluid else g1

    def ic_device(self, items):
        """initial conditions computed on the device: items = [(instance, IcDesc), ...]; t = 0, step 0"""
        n = len(items)
        inst = (C.c_int * max(n, 1))(*[k for k, _ in items])
        descs = (IcDesc * max(n, 1))(*[d for _, d in items])
        _check(load().gfm_ic_device(self.h, n, inst, descs), "gfm_ic_device")

    def upload_vfield(self, k, vf):
        _check(load().gfm_upload_vfield(self.h, k, _dp(vf)), "gfm_upload_vfield")

    def download_vfield(self, k):
        H, W = self.infos[k].shape
        vf = np.zeros((H, W, 2))
        _check(load().gfm_download_vfield(self.h, k, _dp(vf)), "gfm_download_vfield")
        return vf

    def set_time(self, k, t, nstep=0):
        _check(load().gfm_set_time(self.h, k, t, nstep), "gfm_set_time")

    def set_dt(self, k, dt):
        _check(load().gfm_set_dt(self.h, k, dt), "gfm_set_dt")

    def get_time(self, k):
        t, n, dt = C.c_double(0), C.c_int(0), C.c_double(0)
        _check(load().gfm_get_time(self.h, k, C.byref(t), C.byref(n), C.byref(dt)), "gfm_get_time")
        return t.value, n.value, dt.value

    def _call(self, name, *a):
        _check(getattr(load(), name)(self.h, *a), name)

    def compute_dt(self):
        self._call("gfm_compute_dt")

    def set_ghost_states(self):
        self._call("gfm_set_ghost_states")

    def fs_update(self):
        self._call("gfm_fs_update")

    def ls_update(self):
        self._call("gfm_ls_update")

    def apply_bcs(self):
        self._call("gfm_apply_bcs")

    def extrapolate(self, N):
        self._call("gfm_extrapolate", N)

    def ls_reinit(self, iters):
        self._call("gfm_ls_reinit", iters)

    def step(self, n=1):
        self._call("gfm_step", n)

    def step_graph(self, n=1):
        self._call("gfm_step_graph", n)

    def sync(self):
        self._call("gfm_sync")

    def profile_steps(self, n):
        """per-phase milliseconds summed over n eager steps: dt, ghost, fs, ls, bc"""
        ms = np.zeros(5)
        _check(load().gfm_profile_steps(self.h, n, _dp(ms)), "gfm_profile_steps")
        return ms

    def status(self, k):
        s = C.c_uint(0)
        _check(load().gfm_status(self.h, k, C.byref(s)), "gfm_status")
        return s.value

    def mass(self, k):
        out = np.zeros(2)
        _check(load().gfm_mass_sums(self.h, k, _dp(out)), "gfm_mass_sums")
        return out

    def mass_history(self, k, n):
        out = np.zeros((n, 3))
        _check(load().gfm_mass_history(self.h, k, n, _dp(out)), "gfm_mass_history")
        return out

    def set_output_hold(self, k, last, interval):
        """instance k takes no step while t - last > interval (interval < 0: off)"""
        _check(load().gfm_set_output_hold(self.h, k, last, interval), "gfm_set_output_hold")

    def set_tracers(self, k, xy, keep_y=False):
        """passive RK4 tracers of instance k; xy: (n, 2) positions, n <= 4"""
        xy = np.ascontiguousarray(np.asarray(xy, dtype=np.float64).reshape(-1, 2))
        _check(load().gfm_set_tracers(self.h, k, len(xy), _dp(xy), int(keep_y)), "gfm_set_tracers")
        self.ntr[k] = len(xy)

    def tracer_history(self, k, n):
        """(n, 1 + 4 ntr): t, then x, y, u, v per tracer, for steps 0 .. n-1"""
        out = np.zeros((n, 1 + 4 * self.ntr.get(k, 0)))
        _check(load().gfm_tracer_history(self.h, k, n, _dp(out)), "gfm_tracer_history")
        return out

    def download_z(self, k):
        H, W = self.infos[k].shape
        z = np.zeros((H, W))
        _check(load().gfm_download_z(self.h, k, _dp(z)), "gfm_download_z")
        return z

    def upload_z(self, k, z):
        """the CLSVOF tracker's z (after upload, which derives it from phi)"""
        _check(load().gfm_upload_z(self.h, k, _dp(np.ascontiguousarray(z, dtype=np.float64))), "gfm_upload_z")

    def iface_info(self, k):
        n, r = C.c_int(0), C.c_int(0)
        _check(load().gfm_iface_info(self.h, k, C.byref(n), C.byref(r)), "gfm_iface_info")
        return n.value, r.value

    def bbox(self, k):
        b1 = (C.c_int * 4)()
        b2 = (C.c_int * 4)()
        _check(load().gfm_get_bbox(self.h, k, b1, b2), "gfm_get_bbox")
        return list(b1), list(b2)


def pure_rs_solve(paralleltox, gamma, pinf, UL, UR, fp32=False, noalloc=False):
    """face fluxes of the FP64 solver, or (fp32) of the mixed-precision
    path's FP32 face solver on the states rounded to FP32, or (noalloc) of
    the FP64 solver through the entry point that allocates nothing"""
    UL = np.ascontiguousarray(UL, dtype=np.float64).reshape(-1, 4)
    UR = np.ascontiguousarray(UR, dtype=np.float64).reshape(-1, 4)
    F = np.zeros_like(UL)
    st = C.c_uint(0)
    name = "gfm_pure_rs_solve_f32" if fp32 else ("gfm_pure_rs_solve_noalloc" if noalloc else "gfm_pure_rs_solve")
    _check(getattr(load(), name)(len(UL), int(paralleltox), gamma, pinf, _dp(UL), _dp(UR), _dp(F), C.byref(st)),
           name)
    return F, st.value


def mixed_rs_solve(gL, pinfL, gR, pinfR, UL, UR):
    UL = np.ascontiguousarray(UL, dtype=np.float64).reshape(-1, 4)
    UR = np.ascontiguousarray(UR, dtype=np.float64).reshape(-1, 4)
    out = np.zeros_like(UL)
    st = C.c_uint(0)
    _check(load().gfm_mixed_rs_solve(len(UL), gL, pinfL, gR, pinfR, _dp(UL), _dp(UR), _dp(out), C.byref(st)),
           "gfm_mixed_rs_solve")
    return out, st.value


def _v(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def ic_onefluid(info, eos, test_case):
    H, W = info.shape
    g = np.zeros((H, W, 4))
    _check(load().gfm_ic_onefluid(C.byref(info), C.byref(eos), test_case, _dp(g)), "gfm_ic_onefluid")
    return g


def ic_bubble(info, eos, xc, yc, R, bubble, pre, post, shock_x):
    H, W = info.shape
    g1, g2, phi = np.zeros((H, W, 4)), np.zeros((H, W, 4)), np.zeros((H, W))
    _check(load().gfm_ic_bubble(C.byref(info), C.byref(eos), xc, yc, R, _dp(_v(bubble)), _dp(_v(pre)),
                                _dp(_v(post)), shock_x, _dp(g1), _dp(g2), _dp(phi)), "gfm_ic_bubble")
    return g1, g2, phi


def ic_rmi(info, eos, x_if, amp, modes, light, post, heavy, shock_x):
    H, W = info.shape
    g1, g2, phi = np.zeros((H, W, 4)), np.zeros((H, W, 4)), np.zeros((H, W))
    _check(load().gfm_ic_rmi(C.byref(info), C.byref(eos), x_if, amp, modes, _dp(_v(light)), _dp(_v(post)),
                             _dp(_v(heavy)), shock_x, _dp(g1), _dp(g2), _dp(phi)), "gfm_ic_rmi")
    return g1, g2, phi


def ic_planar(info, eos, axis, pos, prims1, prims2):
    H, W = info.shape
    g1, g2, phi = np.zeros((H, W, 4)), np.zeros((H, W, 4)), np.zeros((H, W))
    _check(load().gfm_ic_planar(C.byref(info), C.byref(eos), axis, pos, _dp(_v(prims1)), _dp(_v(prims2)),
                                _dp(g1), _dp(g2), _dp(phi)), "gfm_ic_planar")
    return g1, g2, phi


def shock_state(gamma, pinf, pre, mach, direction):
    out = np.zeros(4)
    _check(load().gfm_shock_state(gamma, pinf, _dp(_v(pre)), mach, int(direction), _dp(out)), "gfm_shock_state")
    return out
