"""Toro's five shock tubes (Riemann Solvers and Numerical Methods for Fluid
Dynamics, 3rd ed., table 6.2) through the GPU step: PLM (theta 1.5) + HLLC,
RK2, CFL 0.4, outflow, gamma 1.4, run to each test's final time with the
device loop.  Each run is bit-identical to the oracle, and its density error
against the exact solution (tests/exact_riemann.py, independent of the code
under test) falls as the mesh is refined."""
import numpy as np
import pytest

from mara2_amd import abi, problems
from mara2_amd.runtime import Block
from exact_riemann import sample

pytestmark = pytest.mark.gpu

# (left rho, u, p), (right rho, u, p), diaphragm x0, final time
TORO = [((1.0, 0.75, 1.0), (0.125, 0.0, 0.1), 0.3, 0.2),
        ((1.0, -2.0, 0.4), (1.0, 2.0, 0.4), 0.5, 0.15),
        ((1.0, 0.0, 1000.0), (1.0, 0.0, 0.01), 0.5, 0.012),
        ((5.99924, 19.5975, 460.894), (5.99242, -6.19633, 46.0950), 0.4, 0.035),
        ((1.0, -19.59745, 1000.0), (1.0, -19.59745, 0.01), 0.8, 0.012)]


def tube(n, L, R, x0):
    d = abi.make_desc(law="hydro", gamma=1.4, recon="plm", plm_theta=1.5, riemann="hllc", bc="outflow",
                      cells=(n, 1, 1), rk_order=2)
    P = problems.empty(d)
    x = problems.cell_centres(d, 0)
    for q, (a, b) in enumerate(zip(L, R)):
        P[:, 0, 0, (0, 1, 4)[q]] = np.where(x < x0, a, b)
    return d, problems.apply_bc(d, P)


@pytest.mark.parametrize("num", [1, 2, 3, 4, 5])
def test_toro_tube(num):
    import oracle
    L, R, x0, tf = TORO[num - 1]
    errs = []
    for n in (200, 400):
        d, P0 = tube(n, L, R, x0)
        with Block(d) as b:
            b.upload(P0)
            t, steps, _ = b.run(0.4, tf, 10**6)
            P = b.download()
        Po = P0.copy()
        err, to, so = oracle.run(d, Po, 0.4, tf)
        assert err == 0 and (t, steps) == (to, so)
        assert np.array_equal(P, Po)
        x = problems.cell_centres(d, 0)[2:-2]
        ex = sample(L, R, 1.4, (x - x0) / tf)
        errs.append(np.mean(np.abs(P[2:-2, 0, 0, 0] - ex[:, 0])) / np.mean(np.abs(ex[:, 0])))
    print(f"Toro test {num}: relative L1(rho) {errs[0]:.3e} (200), {errs[1]:.3e} (400)")
    assert errs[0] < 0.1
    assert errs[1] < errs[0] / 1.3
