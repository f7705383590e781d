"""The mixed form of the relativistic MHD configuration (WENO5 + HLLE,
flux-CT, RK3): reconstruction, face states and Riemann fluxes in binary32,
the cons-to-prim root solve, the eigenvalue quartic, RK accumulation and
flux-CT in binary64.  Held to relative L1 <= 1e-5 of the conserved fields
against the FP64 oracle after 10 steps with identical time steps, no
inversion failures, and the corner monopole at binary64 roundoff.  32^3,
max W = 3, modes up to |k| = 2: the FP64 parity test's state (at 24^3 the
reference scheme itself loses an inversion in step 6, in the oracle and in
both GPU forms alike)."""
import numpy as np
import pytest

import oracle
from mara2_amd import problems
from mara2_amd.runtime import Block
from helpers import interior, rel_l1

pytestmark = pytest.mark.gpu

REL_L1_MIXED = 1e-5


def _rmhd_cons(desc, P):
    """U of RelativisticMHD::fromPrimitive (src/ConservationLaws.cpp:711-732)
    for every zone, via the oracle's point function."""
    out = np.empty(P.shape[:3] + (8,))
    flat = P.reshape(-1, 8)
    res = out.reshape(-1, 8)
    for n in range(flat.shape[0]):
        err, U, _, _ = oracle.from_primitive(desc.law, desc.gamma, flat[n], (1.0, 0.0, 0.0))
        assert err == 0
        res[n] = U
    return out


@pytest.mark.parametrize("n", [32])
def test_mixed_rmhd_close_to_fp64_oracle(n):
    d64 = problems.rmhd_desc(n)
    dmx = problems.rmhd_desc(n, "mixed")
    P0 = problems.rmhd_3d(d64, kmax=2, wmax=3.0)
    Po = P0.copy()
    failures = 0
    with Block(dmx) as blk:
        blk.upload(P0)
        for step in range(10):
            err, dto = oracle.courant_dt(d64, Po)
            assert err == 0
            st, rep = blk.advance(0.3 * dto, check=False)
            failures += rep.c2p_failures
            assert st == 0, (step, rep.as_dict())
            err, _ = oracle.advance(d64, Po, 0.0, 0.3 * dto)
            assert err == 0
        Pg = blk.download()
        m = blk.monopole_max()
    assert failures == 0
    err = rel_l1(_rmhd_cons(d64, interior(d64, Pg)), _rmhd_cons(d64, interior(d64, Po)))
    print(f"mixed RMHD {n}^3 rel L1 vs FP64 oracle after 10 steps: {err:.3e}; monopole {m:.3e}")
    assert err <= REL_L1_MIXED, err
    assert m <= 1e-12 * np.abs(Pg[..., 5:8]).max(), m
