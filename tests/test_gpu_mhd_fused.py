"""Parity of the fused 2-D MHD stage kernel (k_mhd2d: fluxes, flux-CT,
divergence, update and ghost images in one pass) with the oracle, bit for
bit in the FP64 form: tile-ragged shapes (extents that are not multiples of
the 64 x 4 / 64 x 8 tiles), every reconstruction, boundary condition, RK
order of two or three stages, and flux-CT on and off."""
import numpy as np
import pytest

from mara2_amd import abi, problems
from mara2_amd.runtime import Block
from helpers import run_pair, rel_l1

pytestmark = pytest.mark.gpu


def _mhd_state(desc, seed, amp=0.3):
    P = problems.empty(desc)
    rng = np.random.default_rng(seed)
    shp = P.shape[:3]
    P[..., 0] = 1.0 + 0.5 * rng.random(shp)
    P[..., 4] = 1.0 + 0.5 * rng.random(shp)
    P[..., 1:4] = amp * rng.standard_normal(shp + (3,))
    P[..., 5:8] = 0.5 + 0.2 * rng.standard_normal(shp + (3,))
    return problems.apply_bc(desc, P)


@pytest.mark.parametrize("rk", [2, 3])
@pytest.mark.parametrize("recon", ["pcm", "plm", "weno5"])
@pytest.mark.parametrize("ct", [1, 0])
def test_mhd2d_ragged(recon, rk, ct):
    desc = abi.make_desc(law="mhd", recon=recon, riemann="hlle", bc="periodic", cells=(70, 37, 1),
                         upper=(1.0, 0.6, 1.0), rk_order=rk, field_ct=ct)
    P0 = _mhd_state(desc, 51)
    Pg, Po, dg, do = run_pair(desc, P0, 4, 0.3, Block)
    assert dg == do
    diff = np.argwhere(Pg != Po)
    assert diff.size == 0, f"{len(diff)} differ, first {diff[0]}"


@pytest.mark.parametrize("bc", ["periodic", "outflow", "reflecting"])
def test_mhd2d_boundaries(bc):
    desc = abi.make_desc(law="mhd", recon="plm", riemann="hlle", bc=bc, cells=(130, 9, 1), field_ct=1)
    P0 = _mhd_state(desc, 52)
    Pg, Po, dg, do = run_pair(desc, P0, 5, 0.4, Block)
    assert dg == do
    assert np.array_equal(Pg, Po)


def test_mhd2d_mixed_close_to_fp64_and_divergence_free():
    d64 = problems.mhd_desc(96)
    dmx = problems.mhd_desc(96, "mixed")
    P0 = problems.mhd_2d(d64)
    out = {}
    for d in (d64, dmx):
        with Block(d) as blk:
            blk.upload(P0)
            blk.run(0.4, 1e300, 20)
            out[d.precision] = (blk.download(), blk.monopole_max())
    (P64, m64), (Pmx, mmx) = out[0], out[1]
    g = d64.guards
    inner = (slice(g[0], g[0] + 96), slice(g[1], g[1] + 96), slice(0, 1))
    assert rel_l1(Pmx[inner], P64[inner]) < 1e-5
    assert m64 < 1e-12 and mmx < 1e-12
