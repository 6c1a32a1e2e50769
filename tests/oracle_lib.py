"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  Grids are numpy arrays in the oracle's
cell-major layout: shape (Ny+2g, Nx+2g, 4) float64 for fluid states,
(Ny+2g, Nx+2g) for phi, (Ny+2g, Nx+2g, 2) for the velocity field.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

BC_T, BC_R, BC_P = 0, 1, 2
GFM_ORIGINAL, GFM_REAL, GFM_RIEMANN, GFM_MODIFIED, GFM_VOF, GFM_OGFM_VOF = 0, 1, 2, 3, 4, 5
FS_MUSCL, FS_GODUNOV = 0, 1
ITM_LS, ITM_CLSVOF = 0, 1


class Grid(C.Structure):
    _fields_ = [("Nx", C.c_int), ("Ny", C.c_int), ("g", C.c_int),
                ("x0", C.c_double), ("y0", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
                ("bcL", C.c_int), ("bcR", C.c_int), ("bcB", C.c_int), ("bcT", C.c_int)]

    @property
    def shape(self):
        return (self.Ny + 2 * self.g, self.Nx + 2 * self.g)


class Eos(C.Structure):
    _fields_ = [("gamma1", C.c_double), ("pinf1", C.c_double),
                ("gamma2", C.c_double), ("pinf2", C.c_double)]


class Run(C.Structure):
    _fields_ = [("gfm", C.c_int), ("fs", C.c_int), ("ls_order", C.c_int),
                ("reinit_every", C.c_int), ("reinit_iters", C.c_int),
                ("cfl", C.c_double), ("T", C.c_double), ("max_steps", C.c_int),
                ("use_extension_vfield", C.c_int), ("use_slip_BCs", C.c_int), ("itm", C.c_int)]


_lib = None
_other = {}
D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    if a.dtype == np.float64:
        return a.ctypes.data_as(D)
    if a.dtype == np.int32:
        return a.ctypes.data_as(I)
    raise TypeError(a.dtype)


def build():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def lib(path=None):
    """the oracle (or, with path, another build of the same source, e.g. liboracle_mgfm_new.so)"""
    global _lib
    if path is not None and path != LIB_PATH:
        if path not in _other:
            if not os.path.exists(path):
                build()
            _other[path] = _bind(C.CDLL(path))
        return _other[path]
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _bind(C.CDLL(LIB_PATH))
    return _lib


def _bind(L):
    dbl, i32 = C.c_double, C.c_int
    sig = {
        "or_eos_pressure": (dbl, [dbl, dbl, dbl, dbl]),
        "or_eos_soundspeed": (dbl, [dbl, dbl, dbl, dbl]),
        "or_eos_isentropic": (dbl, [dbl, dbl, dbl, dbl, dbl]),
        "or_cons_to_prims": (None, [dbl, dbl, D, D]),
        "or_prims_to_cons": (None, [dbl, dbl, D, D]),
        "or_flux_prims": (None, [dbl, dbl, D, D]),
        "or_exact_rs_star": (i32, [dbl, dbl, dbl, dbl, D, D, D, I]),
        "or_exact_rs_sample": (None, [dbl, dbl, dbl, dbl, D, D, D, dbl, D]),
        "or_pure_rs_flux": (i32, [i32, dbl, dbl, D, D, D]),
        "or_mixed_rs": (i32, [dbl, dbl, dbl, dbl, D, D, D]),
        "or_apply_bcs": (None, [C.POINTER(Grid), D]),
        "or_apply_bcs_scalar": (None, [C.POINTER(Grid), D]),
        "or_apply_bcs_vector": (None, [C.POINTER(Grid), D]),
        "or_extrapolate_extension_vfield": (None, [C.POINTER(Grid), D, i32, D]),
        "or_set_ghost_states_flags": (i32, [C.POINTER(Grid), C.POINTER(Eos), i32, D, D, D, I, I, D, i32, i32]),
        "or_fs_muscl": (i32, [C.POINTER(Grid), dbl, dbl, dbl, I, D, D]),
        "or_fs_godunov": (i32, [C.POINTER(Grid), dbl, dbl, dbl, I, D, D]),
        "or_compute_dt_onefluid": (dbl, [C.POINTER(Grid), dbl, dbl, dbl, D, dbl, dbl]),
        "or_compute_dt_twofluid": (dbl, [C.POINTER(Grid), C.POINTER(Eos), dbl, D, D, D, dbl, dbl]),
        "or_fluid_mass": (dbl, [C.POINTER(Grid), dbl, D, D]),
        "or_ls_normal": (None, [C.POINTER(Grid), D, i32, i32, D]),
        "or_ls_is_interfacial": (i32, [C.POINTER(Grid), D, i32, i32]),
        "or_bilinear": (None, [C.POINTER(Grid), D, dbl, dbl, D]),
        "or_tracers_step": (None, [C.POINTER(Grid), D, i32, D, D, i32, dbl, dbl, D]),
        "or_volume_fraction": (None, [C.POINTER(Grid), D, D]),
        "or_vof_interface": (None, [C.POINTER(Grid), D, D, D, D]),
        "or_houc_coeffs": (None, [i32, D, I]),
        "or_ls_advect": (None, [C.POINTER(Grid), i32, D, D, dbl]),
        "or_ls_reinit": (None, [C.POINTER(Grid), D, i32]),
        "or_extrapolate_vector": (i32, [C.POINTER(Grid), D, i32, D, D, i32]),
        "or_set_ghost_states": (i32, [C.POINTER(Grid), C.POINTER(Eos), i32, D, D, D, I, I, D]),
        "or_onefluid_step": (i32, [C.POINTER(Grid), dbl, dbl, i32, dbl, dbl, D, D, D, D]),
        "or_twofluid_step": (i32, [C.POINTER(Grid), C.POINTER(Eos), C.POINTER(Run), i32, D, D, D, D, D, D, D, D]),
        "or_twofluid_run": (i32, [C.POINTER(Grid), C.POINTER(Eos), C.POINTER(Run), D, D, D, I, D]),
        "or_plic_line": (dbl, [dbl, dbl, dbl, dbl, dbl]),
        "or_plic_area": (dbl, [dbl, dbl, dbl, dbl, dbl]),
        "or_vof_advect": (None, [C.POINTER(Grid), i32, D, D, D, dbl]),
        "or_clsvof_correct": (None, [C.POINTER(Grid), D, D]),
        "or_clsvof_init": (None, [C.POINTER(Grid), D, D]),
        "or_twofluid_step_z": (i32, [C.POINTER(Grid), C.POINTER(Eos), C.POINTER(Run), i32, D, D, D, D, D, D, D, D,
                                     D]),
        "or_twofluid_run_z": (i32, [C.POINTER(Grid), C.POINTER(Eos), C.POINTER(Run), D, D, D, D, I, D]),
        "or_onefluid_run": (i32, [C.POINTER(Grid), dbl, dbl, i32, dbl, dbl, i32, D, I, D]),
        "or_cellcentre": (None, [C.POINTER(Grid), i32, i32, D, D]),
        "or_ic_onefluid": (i32, [C.POINTER(Grid), dbl, dbl, i32, D]),
        "or_ic_bubble": (None, [C.POINTER(Grid), C.POINTER(Eos), dbl, dbl, dbl, D, D, D, dbl, D, D, D]),
        "or_ic_rmi": (None, [C.POINTER(Grid), C.POINTER(Eos), dbl, dbl, i32, D, D, D, dbl, D, D, D]),
        "or_ic_planar_twofluid": (None, [C.POINTER(Grid), C.POINTER(Eos), i32, dbl, D, D, D, D, D]),
        "or_shock_state": (None, [dbl, dbl, D, dbl, i32, D]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def arr(*v):
    return np.ascontiguousarray(np.array(v, dtype=np.float64))


def rs_star(gL, pinfL, gR, pinfR, WL, WR):
    out = np.zeros(4)
    it = C.c_int(0)
    st = lib().or_exact_rs_star(gL, pinfL, gR, pinfR, _p(arr(*WL)), _p(arr(*WR)), _p(out), C.byref(it))
    return out, st, it.value


def rs_sample(gL, pinfL, gR, pinfR, WL, WR, star, S):
    out = np.zeros(3)
    lib().or_exact_rs_sample(gL, pinfL, gR, pinfR, _p(arr(*WL)), _p(arr(*WR)), _p(np.ascontiguousarray(star, dtype=np.float64)), S, _p(out))
    return out


def prims_to_cons(g, pinf, W):
    U = np.zeros(4)
    lib().or_prims_to_cons(g, pinf, _p(arr(*W)), _p(U))
    return U


def cons_to_prims(g, pinf, U):
    W = np.zeros(4)
    lib().or_cons_to_prims(g, pinf, _p(arr(*U)), _p(W))
    return W


def pure_rs_flux(parallel_x, g, pinf, UL, UR):
    F = np.zeros(4)
    st = lib().or_pure_rs_flux(int(parallel_x), g, pinf, _p(arr(*UL)), _p(arr(*UR)), _p(F))
    return F, st


def mixed_rs(gL, pinfL, gR, pinfR, UL, UR):
    out = np.zeros(4)
    st = lib().or_mixed_rs(gL, pinfL, gR, pinfR, _p(arr(*UL)), _p(arr(*UR)), _p(out))
    return out, st


def shock_state(g, pinf, pre, mach, direction):
    out = np.zeros(4)
    lib().or_shock_state(g, pinf, _p(arr(*pre)), mach, int(direction), _p(out))
    return out


def make_grid(Nx, Ny, g, x0, y0, Lx, Ly, bcs=(BC_T, BC_T, BC_T, BC_T)):
    return Grid(Nx, Ny, g, x0, y0, Lx / Nx, Ly / Ny, *bcs)


def new_state(G):
    H, W = G.shape
    return np.zeros((H, W, 4)), np.zeros((H, W))


def twofluid_run(info, eos, state, variant, steps, fs=FS_MUSCL, ls_order=5, reinit_every=5, reinit_iters=10,
                 cfl=0.6, init_reinit=0):
    """`steps` oracle steps of one two-fluid instance from `state` (g1, g2, phi; copied) after
    `init_reinit` reinitialisation iterations of phi; returns (g1, g2, phi, status, t)"""
    g1, g2, phi = (np.ascontiguousarray(a, dtype=np.float64).copy() for a in state)
    Gr = Grid(info.Nx, info.Ny, info.numGC, info.x0, info.y0, info.dx, info.dy,
              info.BC_L, info.BC_R, info.BC_B, info.BC_T)
    E = Eos(eos.gamma1, eos.pinf1, eos.gamma2, eos.pinf2)
    if init_reinit:
        lib().or_ls_reinit(Gr, _p(phi), init_reinit)
    run = Run(variant, fs, ls_order, reinit_every, reinit_iters, cfl, 1e300, steps)
    s = np.zeros(1, dtype=np.int32)
    t = np.zeros(1)
    st = lib().or_twofluid_run(Gr, E, run, _p(g1), _p(g2), _p(phi), _p(s), _p(t))
    assert int(s[0]) == steps
    return g1, g2, phi, st, float(t[0])


def run_many(fn, jobs, threads=None):
    """fn(*job) for every job on host threads (the oracle's scratch is thread-local and ctypes
    releases the GIL while a call runs)"""
    from concurrent.futures import ThreadPoolExecutor
    lib()
    n = max(1, min(len(jobs), threads or min(16, os.cpu_count() or 1)))
    with ThreadPoolExecutor(n) as ex:
        return list(ex.map(lambda j: fn(*j), jobs))
