"""Exact solution of the Riemann problem of an ideal gas (test helper).

Textbook method (E. F. Toro, *Riemann Solvers and Numerical Methods for Fluid
Dynamics*, 3rd ed., ch. 4): the star-region pressure is the root of
f_L(p) + f_R(p) + (u_R - u_L) = 0, with f_K the shock (Rankine-Hugoniot) or
rarefaction (isentropic) branch, found by Newton's method; the solution at
x/t is then sampled wave by wave.  Independent of the solver under test.
"""
from __future__ import annotations

import numpy as np


def _fk(p, rho, pk, ck, g):
    if p > pk:                                            # shock
        A = 2.0 / ((g + 1.0) * rho)
        B = (g - 1.0) / (g + 1.0) * pk
        q = np.sqrt(A / (p + B))
        return (p - pk) * q, q * (1.0 - 0.5 * (p - pk) / (p + B))
    r = p / pk                                            # rarefaction
    f = 2.0 * ck / (g - 1.0) * (r ** ((g - 1.0) / (2.0 * g)) - 1.0)
    return f, 1.0 / (rho * ck) * r ** (-(g + 1.0) / (2.0 * g))


def star_state(rl, ul, pl, rr, ur, pr, g):
    """(p*, u*) of the Riemann problem."""
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    p = max(1e-12, 0.5 * (pl + pr) - 0.125 * (ur - ul) * (rl + rr) * (cl + cr))
    for _ in range(100):
        fl, dl = _fk(p, rl, pl, cl, g)
        fr, dr = _fk(p, rr, pr, cr, g)
        step = (fl + fr + ur - ul) / (dl + dr)
        pn = max(1e-14, p - step)
        if abs(pn - p) <= 1e-15 * (pn + p):
            p = pn
            break
        p = pn
    fl, _ = _fk(p, rl, pl, cl, g)
    fr, _ = _fk(p, rr, pr, cr, g)
    return p, 0.5 * (ul + ur) + 0.5 * (fr - fl)


def sample(x, t, x0, left, right, g):
    """Density, velocity and pressure at positions ``x`` and time ``t`` > 0
    for left/right states (rho, u, p) separated at ``x0``."""
    rl, ul, pl = left
    rr, ur, pr = right
    ps, us = star_state(rl, ul, pl, rr, ur, pr, g)
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    gm = (g - 1.0) / (g + 1.0)
    x = np.asarray(x, np.float64)
    s = (x - x0) / t
    rho, u, p = np.empty_like(s), np.empty_like(s), np.empty_like(s)
    for i, si in enumerate(s):
        if si <= us:                                      # left of the contact
            if ps > pl:                                   # left shock
                sl = ul - cl * np.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
                if si <= sl:
                    rho[i], u[i], p[i] = rl, ul, pl
                else:
                    rho[i] = rl * (ps / pl + gm) / (gm * ps / pl + 1.0)
                    u[i], p[i] = us, ps
            else:                                         # left rarefaction
                csl = cl * (ps / pl) ** ((g - 1) / (2 * g))
                if si <= ul - cl:
                    rho[i], u[i], p[i] = rl, ul, pl
                elif si >= us - csl:
                    rho[i] = rl * (ps / pl) ** (1.0 / g)
                    u[i], p[i] = us, ps
                else:
                    c = 2.0 / (g + 1) * (cl + (g - 1) / 2 * (ul - si))
                    u[i] = 2.0 / (g + 1) * (cl + (g - 1) / 2 * ul + si)
                    rho[i] = rl * (c / cl) ** (2.0 / (g - 1))
                    p[i] = pl * (c / cl) ** (2.0 * g / (g - 1))
        else:                                             # right of the contact
            if ps > pr:                                   # right shock
                sr = ur + cr * np.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
                if si >= sr:
                    rho[i], u[i], p[i] = rr, ur, pr
                else:
                    rho[i] = rr * (ps / pr + gm) / (gm * ps / pr + 1.0)
                    u[i], p[i] = us, ps
            else:                                         # right rarefaction
                csr = cr * (ps / pr) ** ((g - 1) / (2 * g))
                if si >= ur + cr:
                    rho[i], u[i], p[i] = rr, ur, pr
                elif si <= us + csr:
                    rho[i] = rr * (ps / pr) ** (1.0 / g)
                    u[i], p[i] = us, ps
                else:
                    c = 2.0 / (g + 1) * (cr - (g - 1) / 2 * (ur - si))
                    u[i] = 2.0 / (g + 1) * (-cr + (g - 1) / 2 * ur + si)
                    rho[i] = rr * (c / cr) ** (2.0 / (g - 1))
                    p[i] = pr * (c / cr) ** (2.0 * g / (g - 1))
    return rho, u, p
