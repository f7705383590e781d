"""NewtonianMHD with a cooling rate (mara_desc.cooling_rate; setCoolingRate,
src/ConservationLaws.cpp:354-358, addSourceTerms :489-501): the auxiliary
ninth field and the cooling terms on the flux-array path, bit for bit against
the oracle in FP64 -- 1-D, 2-D and 3-D, flux-CT on and off, RK 1 to 3, the
device loop --, the mixed form within 1e-5 of it, and the configurations it
is refused in."""
import numpy as np
import pytest

from mara2_amd import abi, problems, runtime
from mara2_amd.runtime import Batch, Block
from helpers import run_pair

pytestmark = pytest.mark.gpu


def _state(desc, seed, amp=0.3):
    P = problems.empty(desc)
    assert P.shape[3] == 9
    rng = np.random.default_rng(seed)
    shp = P.shape[:3]
    P[..., 0] = 1.0 + 0.5 * rng.random(shp)
    P[..., 4] = 1.0 + 0.5 * rng.random(shp)
    P[..., 1:4] = amp * rng.standard_normal(shp + (3,))
    P[..., 5:8] = 0.5 + 0.2 * rng.standard_normal(shp + (3,))
    P[..., 8] = rng.random(shp)
    return problems.apply_bc(desc, P)


CASES = {
    "1d_plm_rk2": dict(cells=(200, 1, 1), recon="plm", bc="outflow", rk_order=2, field_ct=1),
    "2d_plm_ct_rk2": dict(cells=(70, 37, 1), recon="plm", bc="periodic", rk_order=2, field_ct=1),
    "2d_weno_noct_rk3": dict(cells=(40, 24, 1), recon="weno5", bc="periodic", rk_order=3, field_ct=0),
    "2d_pcm_ct_rk1": dict(cells=(48, 30, 1), recon="pcm", bc="outflow", rk_order=1, field_ct=1),
    "3d_plm_ct_rk2": dict(cells=(16, 12, 10), recon="plm", bc="periodic", rk_order=2, field_ct=1),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_cooling_matches_oracle(case):
    desc = abi.make_desc(law="mhd", riemann="hlle", cooling_rate=0.8, **CASES[case])
    P0 = _state(desc, 81)
    Pg, Po, dg, do = run_pair(desc, P0, 4, 0.3, Block)
    assert dg == do
    diff = np.argwhere(Pg != Po)
    assert diff.size == 0, f"{len(diff)} differ, first {diff[0]}"


def test_device_loop_with_cooling_matches_oracle_loop():
    import oracle
    desc = abi.make_desc(law="mhd", riemann="hlle", cooling_rate=2.0, **CASES["2d_plm_ct_rk2"])
    P0 = _state(desc, 82)
    Po = P0.copy()
    err, to, so = oracle.run(desc, Po, 0.4, 0.05, max_steps=40)
    assert err == 0
    with Block(desc) as blk:
        blk.upload(P0)
        tg, sg, rep = blk.run(0.4, 0.05, 40)
        Pg = blk.download()
        mono = blk.monopole_max()
    assert (sg, tg) == (so, to)
    assert np.array_equal(Pg, Po)
    e, mo = oracle.monopole_max(desc, Po)
    assert e == 0 and mono == mo


@pytest.mark.parametrize("case", ["2d_plm_ct_rk2", "1d_plm_rk2"])
def test_mixed_cooling_close_to_the_fp64_oracle(case):
    """The mixed form (binary32 fluxes; c2p, RK, CT and the cooling terms in
    binary64) within 1e-5 relative L1 of the FP64 oracle after 10 steps."""
    desc = abi.make_desc(law="mhd", riemann="hlle", cooling_rate=0.8, precision="mixed", **CASES[case])
    d64 = abi.make_desc(law="mhd", riemann="hlle", cooling_rate=0.8, **CASES[case])
    P0 = _state(desc, 84)
    import oracle
    Po = P0.copy()
    with Block(desc) as blk:
        blk.upload(P0)
        for _ in range(10):
            err, do = oracle.courant_dt(d64, Po)
            assert err == 0
            blk.advance(0.3 * do)
            err, _ = oracle.advance(d64, Po, 0.0, 0.3 * do)
            assert err == 0
        Pm = blk.download()
    g = desc.guards
    inner = tuple(slice(g[a], g[a] + desc.cells[a]) for a in range(3))
    assert np.sum(np.abs(Pm[inner] - Po[inner])) / np.sum(np.abs(Po[inner])) < 1e-5


def test_refused_configurations():
    with pytest.raises(runtime.MaraError):
        Block(abi.make_desc(law="hydro", cells=(32, 32, 1), cooling_rate=0.5))


def test_batched_cooled_solve_matches_oracle():
    """A cooled solve in a batch (the batch's flux-array kernels with the
    cooling terms) against the oracle, bit for bit."""
    import oracle
    desc = abi.make_desc(law="mhd", riemann="hlle", cooling_rate=0.5, **CASES["2d_plm_ct_rk2"])
    P0 = _state(desc, 83)
    with Batch(1 << 28, rk_order=2, precision="fp64", max_solves=4) as b:
        s = b.admit(desc, P0, 0.4, 1e300, 5)
        while not b.finished():
            b.step(5)
        done, t, steps, st, _ = b.query(s)
        assert st == 0 and steps == 5
        Pg = b.retire(s)
    Po = P0.copy()
    err, to, so = oracle.run(desc, Po, 0.4, 1e300, max_steps=5)
    assert err == 0 and (to, so) == (t, steps)
    assert np.array_equal(Pg, Po)


def test_cooling_with_viscous_fluxes_matches_oracle():
    """Both hooks of the stage on one block (flux-CT off): the viscous fluxes
    go in before the divergence, the cooling terms after it, as in the
    reference's advance (src/SolutionSchemes.cpp:183-207)."""
    import os
    import oracle
    src = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mara2_amd", "sources")
    desc = abi.make_desc(law="mhd", riemann="hlle", cooling_rate=0.6, **CASES["2d_weno_noct_rk3"])
    P0 = _state(desc, 85)
    Po = P0.copy()
    with Block(desc) as blk:
        blk.set_viscous_flux(os.path.join(src, "viscous_constant.hsaco"), "mara_viscous_constant", (0.01,))
        blk.upload(P0)
        for _ in range(4):
            dg = 0.3 * blk.courant_dt()
            err, do = oracle.courant_dt(desc, Po)
            assert err == 0 and dg == 0.3 * do
            blk.advance(dg)
            err, _ = oracle.advance(desc, Po, 0.0, dg, None, ("constant", [0.01]))
            assert err == 0
        Pg = blk.download()
    assert np.array_equal(Pg, Po)
