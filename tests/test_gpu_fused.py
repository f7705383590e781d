"""Parity of the fused hydro stage kernel (2-D and 3-D hydro blocks) with the
oracle, bit for bit in the FP64 form, on tile-ragged shapes (extents that
are not multiples of the 64 x 8 x 16 tile) and every scheme combination."""
import numpy as np
import pytest

from mara2_amd import abi, problems
from mara2_amd.runtime import Block
from helpers import interior, rel_l1, prim_to_cons, run_pair

pytestmark = pytest.mark.gpu


def _random_state(desc, seed, amp=0.3):
    P = problems.empty(desc)
    rng = np.random.default_rng(seed)
    shp = P.shape[:3]
    P[..., 0] = 1.0 + 0.5 * rng.random(shp)
    P[..., 4] = 1.0 + 0.5 * rng.random(shp)
    P[..., 1:4] = amp * rng.standard_normal(shp + (3,))
    if desc.cells[2] == 1:
        P[..., 3] = 0.1
    return problems.apply_bc(desc, P)


@pytest.mark.parametrize("rk", [1, 2, 3])
@pytest.mark.parametrize("recon,riemann", [("pcm", "hlle"), ("plm", "hllc"), ("plm", "hlle"), ("weno5", "hllc")])
def test_fused_3d_ragged(recon, riemann, rk):
    desc = abi.make_desc(law="hydro", recon=recon, riemann=riemann, bc="periodic",
                         cells=(70, 12, 19), upper=(1.0, 0.5, 0.75), rk_order=rk)
    P0 = _random_state(desc, 11)
    Pg, Po, dg, do = run_pair(desc, P0, 4, 0.3, Block)
    assert dg == do
    diff = np.argwhere(Pg != Po)
    assert diff.size == 0, f"{len(diff)} differ, first {diff[0]}"


@pytest.mark.parametrize("bc", ["periodic", "outflow", "reflecting"])
@pytest.mark.parametrize("rk", [1, 3])
def test_fused_2d(bc, rk):
    desc = abi.make_desc(law="hydro", recon="plm", riemann="hllc", bc=bc, cells=(130, 9, 1), rk_order=rk)
    P0 = _random_state(desc, 12)
    Pg, Po, dg, do = run_pair(desc, P0, 5, 0.4, Block)
    assert dg == do
    assert np.array_equal(Pg, Po)


def test_fused_turbulence_64():
    desc = problems.turbulence_desc(64)
    P0 = problems.turbulence_3d(desc)
    Pg, Po, dg, do = run_pair(desc, P0, 3, 0.3, Block)
    assert dg == do
    assert np.array_equal(Pg, Po)


@pytest.mark.parametrize("rk", [2, 3])
def test_mixed_turbulence_close_to_fp64(rk):
    """Mixed form (binary32 fluxes, binary64 RK accumulation) stays within
    relative L1 1e-5 of the FP64 oracle on the turbulence box after 10 steps
    with identical time steps (RK3: the tile's intermediate stage outputs
    carry unfilled guards, read through their interior images)."""
    import oracle
    desc = problems.turbulence_desc(48, "mixed")
    d64 = problems.turbulence_desc(48, "fp64")
    desc.rk_order = d64.rk_order = rk
    P0 = problems.turbulence_3d(d64)
    Po = P0.copy()
    with Block(desc) as blk:
        blk.upload(P0)
        for _ in range(10):
            err, dto = oracle.courant_dt(d64, Po)
            blk.advance(0.3 * dto)
            oracle.advance(d64, Po, 0.0, 0.3 * dto)
        Pg = blk.download()
    err = rel_l1(prim_to_cons(d64, interior(d64, Pg)), prim_to_cons(d64, interior(d64, Po)))
    assert err <= 1e-5, err


def test_mixed_mhd_monopole_roundoff():
    """Mixed form keeps flux-CT in binary64: the corner monopole stays at
    binary64 roundoff."""
    desc = problems.mhd_desc(64, "mixed")
    P0 = problems.mhd_2d(desc)
    with Block(desc) as blk:
        blk.upload(P0)
        blk.run(0.4, 1e30, 10)
        m = blk.monopole_max()
    assert m < 1e-12, m


@pytest.mark.parametrize("case", ["turbulence3d", "hydro2d", "mhd2d", "hydro3d_rk3"])
def test_device_loop_matches_oracle_loop(case):
    """mara_block_run (device dt, clock and stop flag; each step's CFL reduced
    by the previous step's last stage) against the oracle's maraMainLoop."""
    import oracle
    if case == "turbulence3d":
        desc = problems.turbulence_desc(32)
        P0, cfl = problems.turbulence_3d(desc), 0.3
    elif case == "hydro2d":
        desc = abi.make_desc(law="hydro", recon="plm", riemann="hllc", bc="outflow", cells=(100, 37, 1))
        P0, cfl = _random_state(desc, 21), 0.4
    elif case == "mhd2d":
        desc = problems.mhd_desc(48)
        P0, cfl = problems.mhd_2d(desc), 0.4
    else:
        desc = abi.make_desc(law="hydro", recon="plm", riemann="hlle", bc="periodic", cells=(40, 24, 20), rk_order=3)
        P0, cfl = _random_state(desc, 22), 0.3
    Po = P0.copy()
    err, to, so = oracle.run(desc, Po, cfl, 0.05, max_steps=12)
    assert err == 0
    with Block(desc) as blk:
        blk.upload(P0)
        tg, sg, rep = blk.run(cfl, 0.05, 12)
        Pg = blk.download()
    assert sg == so and tg == to
    assert np.array_equal(Pg, Po)
