"""Exact solution of the 1D Riemann problem for a calorically perfect gas
(test infrastructure): the closed form the reference's Sod cases are
checked against (caseSod/characteristics.py, caseSod/analisys.py:169-170).

Follows the exact Riemann solver of E. F. Toro, "Riemann Solvers and
Numerical Methods for Fluid Dynamics", 3rd ed., ch. 4: Newton iteration on
the star pressure f(p) = f_L(p) + f_R(p) + (u_R - u_L) = 0 (eqs. 4.5-4.7),
then sampling of the wave pattern at x/t (section 4.5).
"""
from __future__ import annotations

import numpy as np


def _f(p, rk, pk, ck, g):
    """Toro eq. 4.6/4.7 and its derivative."""
    if p > pk:   # shock
        A, B = 2.0 / ((g + 1) * rk), (g - 1) / (g + 1) * pk
        q = np.sqrt(A / (p + B))
        return (p - pk) * q, q * (1 - 0.5 * (p - pk) / (B + p))
    # rarefaction
    pr = p / pk
    f = 2 * ck / (g - 1) * (pr ** ((g - 1) / (2 * g)) - 1)
    return f, 1.0 / (rk * ck) * pr ** (-(g + 1) / (2 * g))


def star_state(L, R, g=1.4):
    """(p*, u*) of left/right states (rho, u, p)."""
    rl, ul, pl = L
    rr, ur, pr = R
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    p = max(1e-8, 0.5 * (pl + pr))
    for _ in range(100):
        fl, dl = _f(p, rl, pl, cl, g)
        fr, dr = _f(p, rr, pr, cr, g)
        dp = (fl + fr + ur - ul) / (dl + dr)
        p_new = max(1e-8 * p, p - dp)
        if abs(p_new - p) <= 1e-14 * (p_new + p):
            p = p_new
            break
        p = p_new
    fl, _ = _f(p, rl, pl, cl, g)
    fr, _ = _f(p, rr, pr, cr, g)
    return p, 0.5 * (ul + ur) + 0.5 * (fr - fl)


def sample(L, R, s, g=1.4):
    """rho, u, p at similarity coordinates s = (x - x0)/t (array)."""
    rl, ul, pl = L
    rr, ur, pr = R
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    ps, us = star_state(L, R, g)
    s = np.asarray(s, dtype=float)
    out = np.empty((3, s.size))
    gm = (g - 1) / (g + 1)
    for k, x in enumerate(s.ravel()):
        if x <= us:   # left of the contact
            if ps > pl:   # left shock
                sl = ul - cl * np.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
                if x <= sl:
                    out[:, k] = rl, ul, pl
                else:
                    out[:, k] = rl * (ps / pl + gm) / (gm * ps / pl + 1), us, ps
            else:         # left rarefaction
                cs = cl * (ps / pl) ** ((g - 1) / (2 * g))
                if x <= ul - cl:
                    out[:, k] = rl, ul, pl
                elif x >= us - cs:
                    out[:, k] = rl * (ps / pl) ** (1 / g), us, ps
                else:
                    c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - x))
                    u = 2 / (g + 1) * (cl + (g - 1) / 2 * ul + x)
                    r = rl * (c / cl) ** (2 / (g - 1))
                    out[:, k] = r, u, pl * (c / cl) ** (2 * g / (g - 1))
        else:         # right of the contact
            if ps > pr:   # right shock
                sr = ur + cr * np.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
                if x >= sr:
                    out[:, k] = rr, ur, pr
                else:
                    out[:, k] = rr * (ps / pr + gm) / (gm * ps / pr + 1), us, ps
            else:         # right rarefaction
                cs = cr * (ps / pr) ** ((g - 1) / (2 * g))
                if x >= ur + cr:
                    out[:, k] = rr, ur, pr
                elif x <= us + cs:
                    out[:, k] = rr * (ps / pr) ** (1 / g), us, ps
                else:
                    c = 2 / (g + 1) * (cr - (g - 1) / 2 * (ur - x))
                    u = 2 / (g + 1) * (-cr + (g - 1) / 2 * ur + x)
                    r = rr * (c / cr) ** (2 / (g - 1))
                    out[:, k] = r, u, pr * (c / cr) ** (2 * g / (g - 1))
    return out.reshape((3,) + s.shape)
