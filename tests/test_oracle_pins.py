"""Closed forms and literal arrays that pin the parts of the oracle no
reference-compiled golden vector reaches: the Runge-Kutta combine weights,
the outflow and reflecting guard fill, and the monopole scale.

Each check is chosen so that a plausible slip fails it: a wrong RK weight
(b3 = {1, 1/4, 1/3} or b2 = {1, 1/4}), a reflecting guard without the sign
flip, an outflow guard copied from the wrong layer, a monopole off by a
factor.  None of them restates the oracle's own formula.
"""
import numpy as np
import pytest

import oracle
from mara2_amd import abi, problems


# ---- Runge-Kutta combine weights (src/SolutionSchemes.cpp:119-131, :232-266) -----------------
@pytest.mark.parametrize("rk_order", [1, 2, 3])
def test_rk_weights_on_the_cooling_ode(rk_order):
    """A uniform NewtonianMHD state with a cooling rate has zero flux
    divergence, so every stage is L = -rate * eps on the thermal energy
    eps = p / (gamma - 1) and L = +rate * eps on the auxiliary field
    (src/ConservationLaws.cpp:489-501).  The Shu-Osher schemes applied to
    deps/dt = -rate eps give eps' = eps * T(z), z = -rate dt, with T the
    degree-s Taylor polynomial of exp(z): 1 + z, 1 + z + z^2/2,
    1 + z + z^2/2 + z^3/6.  The auxiliary field takes the remainder.
    z = -0.035 makes every weight visible: b2 = {1, 1/4} moves the RK2
    result by ~1e-2, b3 = {1, 1/4, 1/3} the RK3 result by ~3e-1, and the
    z^3/6 term alone is 7e-6 of p, far above the 1e-13 tolerance."""
    gamma, rate, dt = 5.0 / 3.0, 0.7, 0.05
    z = -rate * dt
    taylor = {1: 1 + z, 2: 1 + z + z * z / 2, 3: 1 + z + z * z / 2 + z ** 3 / 6}[rk_order]
    d = abi.make_desc(law="mhd", recon="plm", riemann="hlle", bc="periodic", cells=(8, 6, 1),
                      rk_order=rk_order, gamma=gamma, cooling_rate=rate)
    P = problems.empty(d)
    p0, aux0 = 0.9, 0.25
    P[..., 0] = 1.3
    P[..., 1:4] = (0.2, -0.1, 0.05)
    P[..., 4] = p0
    P[..., 5:8] = (0.3, 0.4, -0.2)
    P[..., 8] = aux0
    P0 = P.copy()
    err, _ = oracle.advance(d, P, 0.0, dt)
    assert err == 0
    g = d.guards
    inner = P[g[0]:g[0] + 8, g[1]:g[1] + 6, 0]
    np.testing.assert_allclose(inner[..., 4], p0 * taylor, rtol=1e-13, atol=0)
    np.testing.assert_allclose(inner[..., 8], aux0 + (p0 - p0 * taylor) / (gamma - 1), rtol=1e-13, atol=0)
    for q in (0, 1, 2, 3, 5, 6, 7):
        np.testing.assert_allclose(inner[..., q], P0[g[0], g[1], 0, q], rtol=1e-14, atol=1e-15)


# ---- guard fill, written out from src/BoundaryConditions.cpp:24-176 --------------------------
# Source cell of every position along one axis with guards, and the sign the
# normal velocity takes there.  These are typed from the Region arithmetic of
# the reference (left guard [n, n+1) <- [g, g+1) for outflow, <- [2g-n-1, 2g-n)
# for reflecting, [0, g) <- [-2g, -g) for periodic; the right side mirrored),
# not computed.
GUARD_MAPS = {
    # g = 2, three interior cells (indices 2, 3, 4), N = 7
    (2, "periodic"): ([3, 4, 2, 3, 4, 2, 3], [1, 1, 1, 1, 1, 1, 1]),
    (2, "outflow"): ([2, 2, 2, 3, 4, 4, 4], [1, 1, 1, 1, 1, 1, 1]),
    (2, "reflecting"): ([3, 2, 2, 3, 4, 4, 3], [-1, -1, 1, 1, 1, -1, -1]),
    # g = 3, four interior cells (indices 3 .. 6), N = 10
    (3, "periodic"): ([4, 5, 6, 3, 4, 5, 6, 3, 4, 5], [1] * 10),
    (3, "outflow"): ([3, 3, 3, 3, 4, 5, 6, 6, 6, 6], [1] * 10),
    (3, "reflecting"): ([5, 4, 3, 3, 4, 5, 6, 6, 5, 4], [-1, -1, -1, 1, 1, 1, 1, -1, -1, -1]),
}


def test_guard_fill_1d_literal_arrays():
    """Outflow copies the first/last interior layer into every guard layer;
    reflecting mirrors about the wall and negates the normal velocity only;
    periodic copies the opposite interior slab.  The g = 2 arrays are spelled
    out in full; the g = 3 ones through the maps above."""
    d = abi.make_desc(law="hydro", recon="plm", bc="outflow", cells=(3, 1, 1))
    P = np.full(abi.array_shape(d), np.nan)
    for i in range(2, 5):
        P[i, 0, 0, :] = 10 * i + np.arange(1, 6)
    z2, z3, z4 = ([21, 22, 23, 24, 25], [31, 32, 33, 34, 35], [41, 42, 43, 44, 45])
    r2, r3, r4 = ([21, -22, 23, 24, 25], [31, -32, 33, 34, 35], [41, -42, 43, 44, 45])
    expect = {
        "periodic": [z3, z4, z2, z3, z4, z2, z3],
        "outflow": [z2, z2, z2, z3, z4, z4, z4],
        "reflecting": [r3, r2, z2, z3, z4, r4, r3],
    }
    for bc, rows in expect.items():
        d.bc = abi.BCS[bc]
        Q = P.copy()
        assert oracle.apply_bc(d, Q) == 0
        assert np.array_equal(Q[:, 0, 0, :], np.array(rows, dtype=float)), bc

    for (g, bc), (src, sgn) in GUARD_MAPS.items():
        recon = "plm" if g == 2 else "weno5"
        n = 3 if g == 2 else 4
        d = abi.make_desc(law="hydro", recon=recon, bc=bc, cells=(n, 1, 1))
        P = np.full(abi.array_shape(d), np.nan)
        for i in range(g, g + n):
            P[i, 0, 0, :] = 10 * i + np.arange(1, 6)
        assert oracle.apply_bc(d, P) == 0
        for pos, (s, m) in enumerate(zip(src, sgn)):
            want = 10 * s + np.arange(1, 6, dtype=float)
            want[1] *= m
            assert np.array_equal(P[pos, 0, 0, :], want), (g, bc, pos)


@pytest.mark.parametrize("bc", ["periodic", "outflow", "reflecting"])
def test_guard_fill_2d_corners_are_axis_sequential(bc):
    """applySimple fills x guards over the full y extent (guard rows
    included), then y guards over the full x extent, so a corner guard is the
    y image of an x image: P[i, j] = cell (map[i], map[j]) with the x
    velocity signed by sign[i] and the y velocity by sign[j].  The guards
    start as NaN, so a fill that skipped the corners would leave NaN."""
    src, sgn = GUARD_MAPS[(2, bc)]
    d = abi.make_desc(law="mhd", recon="plm", bc=bc, cells=(3, 3, 1))
    P = np.full(abi.array_shape(d), np.nan)
    for i in range(2, 5):
        for j in range(2, 5):
            P[i, j, 0, :] = 100 * i + 10 * j + np.arange(8)
    assert oracle.apply_bc(d, P) == 0
    assert not np.isnan(P).any()
    for i in range(7):
        for j in range(7):
            want = 100 * src[i] + 10 * src[j] + np.arange(8, dtype=float)
            want[1] *= sgn[i]
            want[2] *= sgn[j]
            assert np.array_equal(P[i, j, 0, :], want), (bc, i, j)


def test_reflecting_wall_is_the_mirror_half_of_a_periodic_run():
    """Data even about x = 0 and x = 1/2 on a periodic [0, 1] (velocity odd)
    stays mirror-symmetric, and its left half is what a reflecting box on
    [0, 1/2] computes: the reflecting guards hold exactly the mirrored cells
    the periodic run has beside each wall.  PLM + HLLE, RK2, fixed dt, a
    density jump included; the two runs agree bit for bit."""
    m = 48
    dper = abi.make_desc(law="hydro", recon="plm", riemann="hlle", bc="periodic", cells=(2 * m, 1, 1),
                         rk_order=2, gamma=1.4)
    dref = abi.make_desc(law="hydro", recon="plm", riemann="hlle", bc="reflecting", cells=(m, 1, 1),
                         rk_order=2, gamma=1.4, upper=(0.5, 1.0, 1.0))
    g = dper.guards[0]
    Pp = problems.empty(dper)
    x = problems.cell_centres(dper, 0)[g:g + 2 * m]
    # the left half is set, the right half is its mirror image (velocity negated)
    xl = x[:m]
    rho = 1.0 + 0.5 * np.cos(2 * np.pi * xl) + np.where(np.abs(xl - 0.25) < 0.1, 1.5, 0.0)
    pre = 1.0 + 0.3 * np.cos(4 * np.pi * xl)
    vel = 0.4 * np.sin(2 * np.pi * xl)
    Pp[g:g + 2 * m, 0, 0, 0] = np.concatenate([rho, rho[::-1]])
    Pp[g:g + 2 * m, 0, 0, 1] = np.concatenate([vel, -vel[::-1]])
    Pp[g:g + 2 * m, 0, 0, 4] = np.concatenate([pre, pre[::-1]])
    assert np.array_equal(Pp[g:g + m, 0, 0, 0], Pp[g + 2 * m - 1:g + m - 1:-1, 0, 0, 0])
    assert np.array_equal(Pp[g:g + m, 0, 0, 1], -Pp[g + 2 * m - 1:g + m - 1:-1, 0, 0, 1])
    # the reflecting box's cells are the left half of the periodic one
    assert np.array_equal(problems.cell_centres(dref, 0)[g:g + m], x[:m])
    Pr = problems.empty(dref)
    Pr[g:g + m] = Pp[g:g + m]
    assert oracle.apply_bc(dper, Pp) == 0 and oracle.apply_bc(dref, Pr) == 0
    for _ in range(12):
        for desc, P in ((dper, Pp), (dref, Pr)):
            err, _ = oracle.advance(desc, P, 0.0, 2e-3)
            assert err == 0
    assert np.array_equal(Pr[g:g + m], Pp[g:g + m])
    # the reflecting run moved: the wall is not a no-op
    assert not np.array_equal(Pr[g:g + m, 0, 0, 0], rho)
    # and the periodic run kept its symmetry
    assert np.array_equal(Pp[g:g + m, 0, 0, 1], -Pp[g + 2 * m - 1:g + m - 1:-1, 0, 0, 1])


# ---- monopole scale (src/CellCenteredFieldCT.cpp:159-229) ------------------------------------
def _linear_field(d, a, b, c):
    P = problems.empty(d)
    P[..., 0] = 1.0
    P[..., 4] = 1.0
    x = problems.cell_centres(d, 0)[:, None, None]
    y = problems.cell_centres(d, 1)[None, :, None]
    zc = problems.cell_centres(d, 2)[None, None, :]
    P[..., 5] = a * x + 0 * y + 0 * zc
    P[..., 6] = b * y + 0 * x + 0 * zc
    P[..., 7] = c * zc + 0 * x + 0 * y
    return P


def test_monopole_of_a_linear_field_is_its_closed_form():
    """B = (a x, b y, c z) at cell centres.  Each vertex value is
    0.25 * (sum of four x differences + four y + four z) / meshSpacing, and
    meshSpacing stays at its default 1 (src/CellCenteredFieldCT.cpp:11; the
    regression programs never set it), so every vertex holds
    a dx + b dy + c dz.  The cell value averages the six vertex pairs:
    in 3-D the same number; in 2-D the vertex array's k = 0 layer is never
    written, so a cell gets one sixth of it."""
    a, b, c = 0.7, -1.9, 2.3
    d3 = abi.make_desc(law="mhd", recon="plm", cells=(6, 5, 4), lower=(0, 0, 0), upper=(1.2, 0.5, 2.0))
    dx, dy, dz = 1.2 / 6, 0.5 / 5, 2.0 / 4
    err, m = oracle.monopole_max(d3, _linear_field(d3, a, b, c))
    assert err == 0
    np.testing.assert_allclose(m, abs(a * dx + b * dy + c * dz), rtol=1e-12)

    d2 = abi.make_desc(law="mhd", recon="plm", cells=(8, 6, 1), upper=(1.0, 3.0, 1.0))
    err, m = oracle.monopole_max(d2, _linear_field(d2, a, b, 0.0))
    assert err == 0
    np.testing.assert_allclose(m, abs(a / 8 + b * 3.0 / 6) / 6, rtol=1e-12)

    # every interior cell carries the value, not only the maximum
    N = abi.array_shape(d3)[:3]
    B = np.ascontiguousarray(_linear_field(d3, a, b, c)[..., 5:8])
    Mc = oracle.monopole_cell(B)
    g = d3.guards
    inner = Mc[g[0]:N[0] - g[0], g[1]:N[1] - g[1], g[2]:N[2] - g[2]]
    np.testing.assert_allclose(inner, a * dx + b * dy + c * dz, rtol=1e-12)
