"""Workload definitions built on the library's own initial conditions.

- circular_explosion: single-fluid circular explosion (sim_onefluid.cpp:99-141,
  :359-414), 101x101 on [0,2]^2, ghost width 2, transmissive, T = 0.25.
- shocked_bubble: two-fluid shocked bubble in SI units, after the reference's
  shocked_helium_bubble (sim_twofluid.cpp:394-404, :774-859) with the bubble
  gas, density and Mach number as parameters; domain 0.325 x 0.089, bubble
  R = 0.025 at (0.175, 0.0445), shock at x = 0.225 moving left into air at
  rest (rho 1, p 1e5, gamma 1.4); reflective top/bottom, transmissive
  left/right; ghost width 8.
- bubble_batch: seeded uneven shocked bubbles (grids s*(256x72), s ~ U[1,8]).
- rmi / rmi_batch: single-mode Richtmyer-Meshkov interfaces after
  sim_twofluid.cpp:1245-1347, periodic in y.
"""
import numpy as np

import gfm_hip as G


def circular_explosion(N=101):
    info = G.SimInfo(N, N, 2, 0.0, 0.0, 2.0 / N, 2.0 / N,
                     G.BC_TRANSMISSIVE, G.BC_TRANSMISSIVE, G.BC_TRANSMISSIVE, G.BC_TRANSMISSIVE)
    eos = G.Eos(1.4, 0.0, 1.4, 0.0)
    grid = G.ic_onefluid(info, eos, 100)
    return info, eos, grid


def shocked_bubble(Nx=1024, Ny=288, bubble_gamma=1.648, bubble_rho=0.182, mach=1.22, R=0.025,
                   p0=1e5, numGC=8):
    Lx, Ly = 0.325, 0.089
    info = G.SimInfo(Nx, Ny, numGC, 0.0, 0.0, Lx / Nx, Ly / Ny,
                     G.BC_TRANSMISSIVE, G.BC_TRANSMISSIVE, G.BC_REFLECTIVE, G.BC_REFLECTIVE)
    eos = G.Eos(bubble_gamma, 0.0, 1.4, 0.0)
    pre = [1.0, 0.0, 0.0, p0]
    post = G.shock_state(1.4, 0.0, pre, mach, -1)
    bub = [bubble_rho, 0.0, 0.0, p0]
    g1, g2, phi = G.ic_bubble(info, eos, 0.175, 0.0445, R, bub, pre, post, 0.225)
    return info, eos, (g1, g2, phi)


def bubble_batch(n=64, seed=20261203, smin=1.0, smax=8.0, base=(256, 72)):
    """Seeded instances: grid s*base with s ~ U[smin, smax]; Mach ~ U[1.1,2.5],
    R ~ U[0.015,0.035], bubble/ambient density ratio ~ U[0.1,5],
    bubble gamma ~ U[1.09,1.67]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        s = rng.uniform(smin, smax)
        Nx, Ny = int(round(base[0] * s)), int(round(base[1] * s))
        mach = rng.uniform(1.1, 2.5)
        R = rng.uniform(0.015, 0.035)
        ratio = rng.uniform(0.1, 5.0)
        gam = rng.uniform(1.09, 1.67)
        out.append(dict(Nx=Nx, Ny=Ny, bubble_gamma=gam, bubble_rho=ratio, mach=mach, R=R))
    return out


def rmi(Nx=512, Ny=128, amp_over_lambda=0.05, modes=1, mach=1.2, numGC=8):
    """Air (fluid 1, gamma 1.4, rho 1, p 1) right of the interface at x = 2.9
    on [0,4]x[0,1], heavy gas (fluid 2, gamma 1.093, rho 5.04) left of it,
    shock at x = 3.2 moving left; periodic in y."""
    Lx, Ly = 4.0, 1.0
    info = G.SimInfo(Nx, Ny, numGC, 0.0, 0.0, Lx / Nx, Ly / Ny,
                     G.BC_TRANSMISSIVE, G.BC_TRANSMISSIVE, G.BC_PERIODIC, G.BC_PERIODIC)
    eos = G.Eos(1.4, 0.0, 1.093, 0.0)
    light = [1.0, 0.0, 0.0, 1.0]
    post = G.shock_state(1.4, 0.0, light, mach, -1)
    heavy = [5.04, 0.0, 0.0, 1.0]
    amp = amp_over_lambda * Ly / modes
    g1, g2, phi = G.ic_rmi(info, eos, 2.9, amp, modes, light, post, heavy, 3.2)
    return info, eos, (g1, g2, phi)


def rmi_batch(n=256, seed=20261204, smin=1.0, smax=4.0):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        s = rng.uniform(smin, smax)
        out.append(dict(Nx=int(round(512 * s)), Ny=int(round(128 * s)),
                        amp_over_lambda=rng.uniform(0.01, 0.1), modes=int(rng.integers(1, 5)),
                        mach=rng.uniform(1.2, 3.0)))
    return out


def shocked_bubble_ic(Nx=1024, Ny=288, bubble_gamma=1.648, bubble_rho=0.182, mach=1.22, R=0.025, p0=1e5,
                      numGC=8):
    """shocked_bubble's instance as (info, eos, gfm_ic_desc) for the device initial conditions"""
    Lx, Ly = 0.325, 0.089
    info = G.SimInfo(Nx, Ny, numGC, 0.0, 0.0, Lx / Nx, Ly / Ny,
                     G.BC_TRANSMISSIVE, G.BC_TRANSMISSIVE, G.BC_REFLECTIVE, G.BC_REFLECTIVE)
    eos = G.Eos(bubble_gamma, 0.0, 1.4, 0.0)
    pre = [1.0, 0.0, 0.0, p0]
    post = G.shock_state(1.4, 0.0, pre, mach, -1)
    desc = G.ic_desc(G.IC_BUBBLE, [bubble_rho, 0.0, 0.0, p0], pre, post, xc=0.175, yc=0.0445, R=R, shock_x=0.225)
    return info, eos, desc


def rmi_ic(Nx=512, Ny=128, amp_over_lambda=0.05, modes=1, mach=1.2, numGC=8):
    """rmi's instance as (info, eos, gfm_ic_desc) for the device initial conditions"""
    Lx, Ly = 4.0, 1.0
    info = G.SimInfo(Nx, Ny, numGC, 0.0, 0.0, Lx / Nx, Ly / Ny,
                     G.BC_TRANSMISSIVE, G.BC_TRANSMISSIVE, G.BC_PERIODIC, G.BC_PERIODIC)
    eos = G.Eos(1.4, 0.0, 1.093, 0.0)
    light = [1.0, 0.0, 0.0, 1.0]
    post = G.shock_state(1.4, 0.0, light, mach, -1)
    desc = G.ic_desc(G.IC_RMI, light, post, [5.04, 0.0, 0.0, 1.0], x_if=2.9, amp=amp_over_lambda * Ly / modes,
                     modes=modes, shock_x=3.2)
    return info, eos, desc


def interior(info, a):
    g = info.numGC
    return a[g:g + info.Ny, g:g + info.Nx]
