"""Exact solution of the 1-D Riemann problem for an ideal gas (Toro, "Riemann
Solvers and Numerical Methods for Fluid Dynamics", ch. 4): Newton iteration for
the star pressure, then sampling of the self-similar solution.  Test helper."""
from __future__ import annotations

import math

import numpy as np


def _f(p, rho, pk, ck, g):
    if p > pk:  # shock
        A = 2.0 / ((g + 1.0) * rho)
        B = (g - 1.0) / (g + 1.0) * pk
        s = math.sqrt(A / (p + B))
        return (p - pk) * s, s * (1.0 - 0.5 * (p - pk) / (B + p))
    pr = p / pk  # rarefaction
    f = 2.0 * ck / (g - 1.0) * (pr ** ((g - 1.0) / (2.0 * g)) - 1.0)
    return f, 1.0 / (rho * ck) * pr ** (-(g + 1.0) / (2.0 * g))


def star_state(rl, ul, pl, rr, ur, pr, g):
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    p = max(1e-12, 0.5 * (pl + pr))
    for _ in range(100):
        fl, dl = _f(p, rl, pl, cl, g)
        fr, dr = _f(p, rr, pr, cr, g)
        dp = (fl + fr + ur - ul) / (dl + dr)
        pn = max(1e-14, p - dp)
        if abs(pn - p) < 1e-15 * (pn + p):
            p = pn
            break
        p = pn
    fl, _ = _f(p, rl, pl, cl, g)
    fr, _ = _f(p, rr, pr, cr, g)
    return p, 0.5 * (ul + ur) + 0.5 * (fr - fl)


def sample(x, t, x0, left, right, g):
    """rho, u, p at positions x (array) and time t; left/right = (rho, u, p)."""
    rl, ul, pl = left
    rr, ur, pr = right
    ps, us = star_state(rl, ul, pl, rr, ur, pr, g)
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    out = np.zeros((len(x), 3))
    for i, xi in enumerate(np.asarray(x, float)):
        s = (xi - x0) / t
        if s <= us:  # left of contact
            if ps > pl:  # left shock
                rs = rl * ((ps / pl + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pl + 1))
                sl = ul - cl * math.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
                out[i] = (rl, ul, pl) if s <= sl else (rs, us, ps)
            else:  # left rarefaction
                rs = rl * (ps / pl) ** (1 / g)
                cs = cl * (ps / pl) ** ((g - 1) / (2 * g))
                if s <= ul - cl:
                    out[i] = (rl, ul, pl)
                elif s >= us - cs:
                    out[i] = (rs, us, ps)
                else:
                    u = 2 / (g + 1) * (cl + (g - 1) / 2 * ul + s)
                    c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - s))
                    r = rl * (c / cl) ** (2 / (g - 1))
                    out[i] = (r, u, pl * (c / cl) ** (2 * g / (g - 1)))
        else:  # right of contact
            if ps > pr:  # right shock
                rs = rr * ((ps / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pr + 1))
                sr = ur + cr * math.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
                out[i] = (rs, us, ps) if s <= sr else (rr, ur, pr)
            else:  # right rarefaction
                rs = rr * (ps / pr) ** (1 / g)
                cs = cr * (ps / pr) ** ((g - 1) / (2 * g))
                if s >= ur + cr:
                    out[i] = (rr, ur, pr)
                elif s <= us + cs:
                    out[i] = (rs, us, ps)
                else:
                    u = 2 / (g + 1) * (-cr + (g - 1) / 2 * ur + s)
                    c = 2 / (g + 1) * (cr - (g - 1) / 2 * (ur - s))
                    r = rr * (c / cr) ** (2 / (g - 1))
                    out[i] = (r, u, pr * (c / cr) ** (2 * g / (g - 1)))
    return out
