"""The allocation-free face solver (gfm_pure_rs_solve_noalloc) against gfm_pure_rs_solve through the C
ABI: the same fluxes and status, bit for bit, on two calls in a row, the second with fewer faces and
without the unphysical face of the first (its status bit must not carry over through the shared
staging array); more faces than the staging holds is refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import gfm_hip as G

MAX_FACES = 1 << 18


def states(rng, n, g=1.4):
    rho = rng.uniform(0.5, 5.0, (n, 2))
    u = rng.normal(0.0, 0.3, (n, 2))      # no face opens a vacuum
    v = rng.normal(0.0, 0.3, (n, 2))
    p = rng.uniform(0.5, 5.0, (n, 2))
    E = p / (g - 1.0) + 0.5 * rho * (u * u + v * v)
    U = np.stack([rho, rho * u, rho * v, E], -1)
    return np.ascontiguousarray(U[:, 0]), np.ascontiguousarray(U[:, 1])


@pytest.mark.parametrize("px", [1, 0])
def test_noalloc_matches_allocating_solver(px):
    print("CASE noalloc px=%d" % px, flush=True)
    rng = np.random.default_rng(11 + px)
    for n, bad in ((MAX_FACES, True), (12345, False)):
        UL, UR = states(rng, n)
        if bad:
            UL[7, 3] = -1.0            # negative energy: an unphysical state sets a status bit
        ref, st_ref = G.pure_rs_solve(px, 1.4, 0.0, UL, UR)
        got, st = G.pure_rs_solve(px, 1.4, 0.0, UL, UR, noalloc=True)
        assert (st_ref != 0) == bad                         # the inputs are as intended
        assert st == st_ref
        assert np.array_equal(got, ref, equal_nan=True)     # the unphysical face's flux is NaN in both


def test_noalloc_refuses_more_faces_than_staged():
    L = G.load()
    n = MAX_FACES + 1
    UL = np.zeros((n, 4))
    F = np.zeros((n, 4))
    st = C.c_uint(0)
    D = C.POINTER(C.c_double)
    rc = L.gfm_pure_rs_solve_noalloc(n, 1, 1.4, 0.0, UL.ctypes.data_as(D), UL.ctypes.data_as(D),
                                     F.ctypes.data_as(D), C.byref(st))
    assert rc == -1 and b"262144" in L.gfm_last_error()
