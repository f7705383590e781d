"""Exact solution of the 1-D Euler Riemann problem for an ideal gas (Toro,
Riemann Solvers and Numerical Methods for Fluid Dynamics, ch. 4): Newton
iteration on the star pressure, then sampling of the self-similar solution.
Independent of the code under test; used to pin the full step."""
import numpy as np


def _fk(p, rho, pk, ck, g):
    if p > pk:  # shock
        A = 2 / ((g + 1) * rho)
        B = (g - 1) / (g + 1) * pk
        f = (p - pk) * np.sqrt(A / (p + B))
        df = np.sqrt(A / (B + p)) * (1 - (p - pk) / (2 * (B + p)))
    else:  # rarefaction
        f = 2 * ck / (g - 1) * ((p / pk) ** ((g - 1) / (2 * g)) - 1)
        df = 1 / (rho * ck) * (p / pk) ** (-(g + 1) / (2 * g))
    return f, df


def star_state(L, R, g):
    rl, ul, pl = L
    rr, ur, pr = R
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    p = max(1e-12, 0.5 * (pl + pr))
    for _ in range(200):
        fl, dfl = _fk(p, rl, pl, cl, g)
        fr, dfr = _fk(p, rr, pr, cr, g)
        dp = (fl + fr + ur - ul) / (dfl + dfr)
        pn = max(1e-14, p - dp)
        if abs(pn - p) < 1e-15 * (pn + p):
            p = pn
            break
        p = pn
    fl, _ = _fk(p, rl, pl, cl, g)
    fr, _ = _fk(p, rr, pr, cr, g)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def sample(L, R, g, xi):
    """(rho, u, p) at the similarity coordinates xi = (x - x0) / t."""
    rl, ul, pl = L
    rr, ur, pr = R
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    ps, us = star_state(L, R, g)
    out = np.zeros((len(xi), 3))
    for n, s in enumerate(xi):
        if s <= us:  # left of contact
            if ps > pl:
                rs = rl * ((ps / pl + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pl + 1))
                S = ul - cl * np.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
                out[n] = (rl, ul, pl) if s <= S else (rs, us, ps)
            else:
                rs = rl * (ps / pl) ** (1 / g)
                cs = cl * (ps / pl) ** ((g - 1) / (2 * g))
                sh, st = ul - cl, us - cs
                if s <= sh:
                    out[n] = (rl, ul, pl)
                elif s >= st:
                    out[n] = (rs, us, ps)
                else:
                    u = 2 / (g + 1) * (cl + (g - 1) / 2 * ul + s)
                    c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - s))
                    out[n] = (rl * (c / cl) ** (2 / (g - 1)), u, pl * (c / cl) ** (2 * g / (g - 1)))
        else:
            if ps > pr:
                rs = rr * ((ps / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pr + 1))
                S = ur + cr * np.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
                out[n] = (rr, ur, pr) if s >= S else (rs, us, ps)
            else:
                rs = rr * (ps / pr) ** (1 / g)
                cs = cr * (ps / pr) ** ((g - 1) / (2 * g))
                sh, st = ur + cr, us + cs
                if s >= sh:
                    out[n] = (rr, ur, pr)
                elif s <= st:
                    out[n] = (rs, us, ps)
                else:
                    u = 2 / (g + 1) * (-cr + (g - 1) / 2 * ur + s)
                    c = 2 / (g + 1) * (cr - (g - 1) / 2 * (ur - s))
                    out[n] = (rr * (c / cr) ** (2 / (g - 1)), u, pr * (c / cr) ** (2 * g / (g - 1)))
    return out
