"""Converged device solutions against the closed forms the mathematics
fixes (GPU): the attached oblique shock of a supersonic compression ramp
(theta-beta-M relation and the Rankine-Hugoniot jumps) and the laminar
flat-plate skin friction (Blasius, Cf = 0.664 / sqrt(Re_x)).  These pin the
whole converged path (fluxes, reconstruction, limiter, boundary conditions,
viscous terms, implicit and explicit steppers) independently of the
oracle."""
import math

import numpy as np
import pytest

from csu_amd import cases, native

pytestmark = pytest.mark.gpu


def oblique_shock(M1: float, theta_deg: float, g: float):
    """Weak-shock solution of tan(theta) = 2 cot(beta) (M^2 sin^2 beta - 1) /
    (M^2 (g + cos 2 beta) + 2) by bisection, and the jumps across it:
    (beta in degrees, p2/p1, rho2/rho1, M2)."""
    th = math.radians(theta_deg)

    def f(b):
        return 2 / math.tan(b) * (M1 ** 2 * math.sin(b) ** 2 - 1) / (M1 ** 2 * (g + math.cos(2 * b)) + 2) - math.tan(th)

    lo, hi = math.asin(1 / M1) + 1e-12, math.radians(64.0)  # below the maximum-deflection angle
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(lo) * f(mid) <= 0:
            hi = mid
        else:
            lo = mid
    b = 0.5 * (lo + hi)
    mn1 = M1 * math.sin(b)
    p21 = 1 + 2 * g / (g + 1) * (mn1 ** 2 - 1)
    r21 = (g + 1) * mn1 ** 2 / ((g - 1) * mn1 ** 2 + 2)
    mn2 = math.sqrt((1 + (g - 1) / 2 * mn1 ** 2) / (g * mn1 ** 2 - (g - 1) / 2))
    return math.degrees(b), p21, r21, mn2 / math.sin(b - th)


def test_oblique_shock_closed_form_values():
    """The closed-form routine reproduces the textbook values for M 2.5 over
    15 degrees (gamma 1.4): beta 36.9449 deg, p2/p1 2.4675, rho2/rho1
    1.86655, M2 1.87353."""
    b, p21, r21, m2 = oblique_shock(2.5, 15.0, 1.4)
    assert abs(b - 36.944900) < 1e-5
    assert abs(p21 - 2.46750013) < 1e-7 and abs(r21 - 1.86654863) < 1e-7 and abs(m2 - 1.87352601) < 1e-7


def _centres(c):
    A = c.arrays()
    xy, cv = A["node_xy"], A["cell_vert"]
    n = (cv >= 0).sum(1)
    xs = np.where(cv >= 0, xy[np.maximum(cv, 0), 0], 0).sum(1) / n
    ys = np.where(cv >= 0, xy[np.maximum(cv, 0), 1], 0).sum(1) / n
    return xs, ys


def _converge(dev, drop, cap, chunk):
    peak, it = 0.0, 0
    while it < cap:
        r, _ = dev.iterate(chunk)
        it += chunk
        assert np.isfinite(r).all()
        peak = max(peak, float(r[:, 0].max()))
        if r[-1, 0] <= drop * peak:
            return it
    raise AssertionError(f"no {drop:g} drop in {cap} iterations")


def test_ramp_oblique_shock(tmpcase):
    """Config 0's scheme (Roe, second order, Venkatakrishnan K=5, RK5 at CFL
    0.5, local time steps) on the 200x100 jittered-triangle ramp (15 deg at
    x = 0.3, M 2.5) with the domain 0.8 high, so that the shock leaves
    through the outlet instead of reflecting off the top boundary, run to a
    1e-10 drop.  Between the ramp and the shock (cells 0.05 clear of the
    closed-form shock line): pressure, density and Mach number within 1e-3
    of the closed-form jumps in the median and 1 % in every cell; above the
    shock line (0.05 clear) the free stream within 1e-3, which also places
    the shock at the closed-form angle."""
    d = tmpcase("rs")
    cases.write_case(d, cases.ramp_mesh(nx=200, ny=100, height=0.8, seed=5), cases.ramp_input(threads=1, nmax=1))
    c = native.HostCase(d)
    dev = native.Device(c)
    it = _converge(dev, 1e-10, 20000, 500)
    g = c.params.gamma
    U = dev.get_state()
    dev.close()
    rho = U[0]
    u, v = U[1] / rho, U[2] / rho
    p = (g - 1) * (U[3] - 0.5 * rho * (u * u + v * v))
    M = np.sqrt(u * u + v * v) / np.sqrt(g * p / rho)
    U0 = c.initial_state()
    r1 = U0[0, 0]
    p1 = float(c.params.pin[3])
    beta, p21, r21, m2 = oblique_shock(2.5, 15.0, g)
    xc, yc = _centres(c)
    tb, ta = math.tan(math.radians(beta)), math.tan(math.radians(15.0))
    below = (xc > 0.4) & (xc < 0.95) & (yc > (xc - 0.3) * ta + 0.01) & (yc < (xc - 0.3) * tb - 0.05)
    above = (xc > 0.4) & (xc < 0.95) & (yc > (xc - 0.3) * tb + 0.05) & (yc < 0.75)
    assert below.sum() > 3000 and above.sum() > 3000
    print("iterations", it, "cells below / above", below.sum(), above.sum(), flush=True)
    for name, q, ex in (("p2/p1", p / p1, p21), ("rho2/rho1", rho / r1, r21), ("M2", M, m2)):
        e = q[below] / ex - 1
        print(name, "median", np.median(e), "range", e.min(), e.max(), flush=True)
        assert abs(np.median(e)) < 1e-3 and np.abs(e).max() < 1e-2
    for name, q, ex in (("p/p1", p / p1, 1.0), ("rho/rho1", rho / r1, 1.0), ("M", M, 2.5)):
        e = q[above] / ex - 1
        print(name, "free stream", np.abs(e).max(), flush=True)
        assert np.abs(e).max() < 1e-3


def test_laminar_plate_blasius(tmpcase):
    """Config 1's mesh and scheme (500x100 hybrid plate, AUSM, Venkatakrishnan
    K=5, DPLUR) at CFL 200, run to a 1e-8 drop from the density-residual
    peak: the wall skin friction within 3 % of Blasius' 0.664/sqrt(Re_x)
    from x = 0.05 to the end of the plate (M 0.3, Re_L 1e5: compressibility
    and the finite Reynolds number move Cf by about 2 %)."""
    d = tmpcase("bl")
    cases.write_case(d, cases.plate_mesh(nx=500, ny=100, growth=1.08, tri_frac=0.2, seed=1),
                     cases.plate_input(threads=1, nmax=1, timeScheme="DPLUR", CFL="200"))
    c = native.HostCase(d)
    dev = native.Device(c)
    it = _converge(dev, 1e-8, 40000, 1000)
    wt = dev.wall_tensor()
    dev.close()
    A = c.arrays()
    xy, bf = A["node_xy"], A["bface"]
    wall = A["mark_type"][A["bface_mark"]] == 3  # CSU_BC_WALL
    xm = 0.5 * (xy[bf[:, 1], 0] + xy[bf[:, 2], 0])
    U0 = c.initial_state()
    r_inf, u_inf = U0[0, 0], U0[1, 0] / U0[0, 0]
    T_inf = float(c.params.pin[4])
    mu = 1.458e-6 * T_inf ** 1.5 / (T_inf + 110.4)
    sel = wall & (xm >= 0.05)
    assert sel.sum() > 400
    cf = wt[sel, 1] / (0.5 * r_inf * u_inf ** 2)
    blasius = 0.664 / np.sqrt(r_inf * u_inf * xm[sel] / mu)
    ratio = cf / blasius
    print("iterations", it, "Cf / Blasius", ratio.min(), ratio.max(), flush=True)
    assert np.abs(ratio - 1).max() < 0.03
