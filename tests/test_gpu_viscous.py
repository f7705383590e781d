"""Viscous fluxes (mara_block_set_viscous_flux): the per-face function in a
gfx950 code object (mara2_amd/sources/viscous_constant.hsaco, written with
include/mara_source.h) against the same function in the oracle
(oracle_advance_visc), bit for bit in FP64.  A block with the hook takes the
flux-array path whatever its own kernels are (fused 2-D/3-D hydro, 1-D
pencils, fused 2-D MHD); with flux-CT the hook sees the CT-corrected fluxes,
as the reference's fluxCorrection hands them over (binary64 staging,
k_visc_stage); in the mixed form it sees binary64 copies of the binary32
fluxes and the update reads their binary32 rounding (held to 1e-5 of the
oracle).  Removing it restores the block's own kernels."""
import os

import numpy as np
import pytest

from mara2_amd import abi, problems, runtime
from mara2_amd.runtime import Block

pytestmark = pytest.mark.gpu

SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mara2_amd", "sources")
VISC = (os.path.join(SRC, "viscous_constant.hsaco"), "mara_viscous_constant")


def _state(desc, seed, amp=0.3):
    P = problems.empty(desc)
    rng = np.random.default_rng(seed)
    shp = P.shape[:3]
    P[..., 0] = 1.0 + 0.5 * rng.random(shp)
    P[..., 4] = 1.0 + 0.5 * rng.random(shp)
    P[..., 1:4] = amp * rng.standard_normal(shp + (3,))
    if desc.law != 0:
        P[..., 5:8] = 0.5 + 0.2 * rng.standard_normal(shp + (3,))
    return problems.apply_bc(desc, P)


CASES = {
    "hydro2d_fused_to_flux": dict(law="hydro", recon="plm", riemann="hllc", bc="periodic", cells=(70, 37, 1)),
    "hydro3d_fused_to_flux": dict(law="hydro", recon="plm", riemann="hllc", bc="periodic", cells=(20, 12, 10)),
    "hydro1d_pencil_to_flux": dict(law="hydro", recon="plm", riemann="hlle", bc="outflow", cells=(300, 1, 1)),
    "hydro2d_weno_rk3": dict(law="hydro", recon="weno5", riemann="hlle", bc="reflecting", cells=(40, 24, 1),
                             rk_order=3),
    "mhd2d_noct_to_flux": dict(law="mhd", recon="plm", riemann="hlle", bc="periodic", cells=(70, 37, 1), field_ct=0),
    "mhd2d_ct_to_flux": dict(law="mhd", recon="plm", riemann="hlle", bc="periodic", cells=(70, 37, 1), field_ct=1),
    "mhd3d_ct": dict(law="mhd", recon="plm", riemann="hlle", bc="outflow", cells=(20, 12, 10), field_ct=1),
    "mhd2d_ct_cooled": dict(law="mhd", recon="plm", riemann="hlle", bc="periodic", cells=(36, 28, 1), field_ct=1,
                            cooling_rate=0.3),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_constant_viscosity_matches_oracle(case):
    import oracle
    desc = abi.make_desc(**CASES[case])
    P0 = _state(desc, 71)
    nu = 0.01
    Po = P0.copy()
    with Block(desc) as blk:
        blk.set_viscous_flux(*VISC, (nu,))
        blk.upload(P0)
        for _ in range(4):
            dg = 0.3 * blk.courant_dt()
            err, do = oracle.courant_dt(desc, Po)
            assert err == 0 and dg == 0.3 * do
            blk.advance(dg)
            err, _ = oracle.advance(desc, Po, 0.0, dg, None, ("constant", [nu]))
            assert err == 0
        Pg = blk.download()
    diff = np.argwhere(Pg != Po)
    assert diff.size == 0, f"{len(diff)} differ, first {diff[0]}"


def test_device_loop_with_viscosity_matches_oracle_loop():
    """mara_block_run (graph replays) with the hook against the oracle's loop."""
    import oracle
    desc = abi.make_desc(**CASES["hydro2d_fused_to_flux"])
    P0 = _state(desc, 72)
    Po = P0.copy()
    err, to, so = oracle.run(desc, Po, 0.3, 0.05, max_steps=30, viscous=("constant", [0.02]))
    assert err == 0
    with Block(desc) as blk:
        blk.set_viscous_flux(*VISC, (0.02,))
        blk.upload(P0)
        tg, sg, rep = blk.run(0.3, 0.05, 30)
        Pg = blk.download()
    assert (sg, tg) == (so, to)
    assert np.array_equal(Pg, Po)


@pytest.mark.parametrize("case", ["hydro2d_fused_to_flux", "hydro1d_pencil_to_flux", "mhd2d_noct_to_flux"])
def test_removing_the_hook_restores_the_blocks_own_kernels(case):
    desc = abi.make_desc(**CASES[case])
    P0 = _state(desc, 73)
    with Block(desc) as plain:
        plain.upload(P0)
        plain.run(0.3, 1e300, 5)
        ref = plain.download()
    with Block(desc) as blk:
        blk.set_viscous_flux(*VISC, (0.05,))
        blk.set_viscous_flux(None)
        blk.upload(P0)
        blk.run(0.3, 1e300, 5)
        assert np.array_equal(blk.download(), ref)


def test_zero_viscosity_gives_the_plain_flux_array_step():
    """nu = 0 subtracts exact zeros: the result is the block's plain step
    (the flux-array path and the fused kernel agree bit for bit)."""
    desc = abi.make_desc(**CASES["hydro2d_fused_to_flux"])
    P0 = _state(desc, 74)
    out = []
    for nu in (None, 0.0):
        with Block(desc) as blk:
            if nu is not None:
                blk.set_viscous_flux(*VISC, (nu,))
            blk.upload(P0)
            blk.run(0.3, 1e300, 5)
            out.append(blk.download())
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("case", ["hydro2d_fused_to_flux", "hydro3d_fused_to_flux", "mhd2d_ct_to_flux"])
def test_mixed_form_with_viscosity_close_to_oracle(case):
    """The mixed form with the hook: binary32 fluxes, the hook on their
    binary64 copies (CT-corrected under flux-CT), within the mixed form's 1e-5
    of the binary64 oracle."""
    import oracle
    from helpers import interior, prim_to_cons, rel_l1
    desc = abi.make_desc(**CASES[case])
    dmix = abi.make_desc(**CASES[case], precision="mixed")
    P0 = _state(desc, 73)
    nu = 0.01
    Po = P0.copy()
    with Block(dmix) as blk:
        blk.set_viscous_flux(*VISC, (nu,))
        blk.upload(P0)
        for _ in range(4):
            err, do = oracle.courant_dt(desc, Po)
            assert err == 0
            blk.advance(0.3 * do)
            err, _ = oracle.advance(desc, Po, 0.0, 0.3 * do, None, ("constant", [nu]))
            assert err == 0
        Pg = blk.download()
    err = rel_l1(prim_to_cons(desc, interior(desc, Pg)), prim_to_cons(desc, interior(desc, Po)))
    assert err <= 1e-5, err
    # the hook took effect: without it the state differs by more than that
    Pn = P0.copy()
    for _ in range(4):
        err, do = oracle.courant_dt(desc, Pn)
        oracle.advance(desc, Pn, 0.0, 0.3 * do)
    assert rel_l1(prim_to_cons(desc, interior(desc, Pn)), prim_to_cons(desc, interior(desc, Po))) > 1e-4


def test_refused_configurations():
    with Block(abi.make_desc(law="hydro", cells=(16, 8, 1), lower=(1, 0.4, 0), upper=(2, 2.7, 1),
                             geometry="spherical")) as blk:
        with pytest.raises(runtime.MaraError):
            blk.set_viscous_flux(*VISC, (0.01,))
    desc = abi.make_desc(**CASES["hydro1d_pencil_to_flux"])
    with Block(desc) as blk:
        with pytest.raises(runtime.MaraError):
            blk.set_viscous_flux(os.path.join(SRC, "missing.hsaco"), "k")
        with pytest.raises(runtime.MaraError):
            blk.set_viscous_flux(VISC[0], "no_such_kernel")
        # the failed loads leave no HIP error behind for the next call
        blk.upload(_state(desc, 75))
        blk.advance(0.3 * blk.courant_dt())
