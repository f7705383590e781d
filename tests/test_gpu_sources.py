"""Source terms (mara_block_set_source): the per-zone function in a gfx950
code object (mara2_amd/sources/*.hsaco, written with include/mara_source.h)
against the same function in the oracle (oracle_advance_src), bit for bit in
FP64 on every stage path a source can take -- the flux-array path (a fused
hydro block switches to it), the 1-D pencil tiles and the fused 2-D MHD
tiles -- plus a closed form and the removal of a source."""
import os

import numpy as np
import pytest

from mara2_amd import abi, problems, runtime
from mara2_amd.runtime import Block

pytestmark = pytest.mark.gpu

SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mara2_amd", "sources")
GRAVITY = (os.path.join(SRC, "uniform_gravity.hsaco"), "mara_source_uniform_gravity")
POINT = (os.path.join(SRC, "point_mass.hsaco"), "mara_source_point_mass")


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


def _pair(desc, P0, steps, cfl, code, params, taux=(), oracle_name=None):
    """GPU block with the code object against the oracle with its restatement."""
    import oracle
    Po = P0.copy()
    with Block(desc) as blk:
        blk.set_source(code[0], code[1], params)
        if taux:
            blk.set_source_aux(taux)
        blk.upload(P0)
        for _ in range(steps):
            dg = cfl * blk.courant_dt()
            err, do = oracle.courant_dt(desc, Po)
            assert err == 0 and dg == cfl * do
            blk.advance(dg)
            err, _ = oracle.advance(desc, Po, 0.0, dg, (oracle_name, list(taux), list(params)))
            assert err == 0
        return blk.download(), Po


CASES = {
    "hydro2d_fused_to_flux": dict(law="hydro", recon="plm", riemann="hllc", bc="periodic", cells=(70, 37, 1)),
    "hydro2d_rk3": dict(law="hydro", recon="plm", riemann="hlle", bc="outflow", cells=(40, 24, 1), rk_order=3),
    "hydro1d_pencil": dict(law="hydro", recon="plm", riemann="hllc", bc="outflow", cells=(300, 1, 1)),
    "mhd2d_fused": dict(law="mhd", recon="plm", riemann="hlle", bc="periodic", cells=(70, 37, 1), field_ct=1),
    "hydro3d_flux": dict(law="hydro", recon="weno5", riemann="hlle", bc="periodic", cells=(20, 12, 10)),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_uniform_gravity_matches_oracle(case):
    desc = abi.make_desc(**CASES[case])
    P0 = _state(desc, 61)
    Pg, Po = _pair(desc, P0, 4, 0.3, GRAVITY, (0.3, -0.2, 0.1), oracle_name="uniform_gravity")
    diff = np.argwhere(Pg != Po)
    assert diff.size == 0, f"{len(diff)} differ, first {diff[0]}"


@pytest.mark.parametrize("case", ["hydro2d_fused_to_flux", "mhd2d_fused", "hydro1d_pencil"])
def test_point_mass_with_auxiliary_position_matches_oracle(case):
    desc = abi.make_desc(**CASES[case])
    P0 = _state(desc, 62)
    Pg, Po = _pair(desc, P0, 4, 0.3, POINT, (0.1, 0.01), taux=(0.4, 0.35, 0.5), oracle_name="point_mass")
    assert np.array_equal(Pg, Po)


def test_device_loop_with_source_matches_oracle_loop():
    """mara_block_run (graph replays) with a source against the oracle's loop."""
    import oracle
    desc = abi.make_desc(**CASES["mhd2d_fused"])
    P0 = _state(desc, 63)
    Po = P0.copy()
    err, to, so = oracle.run(desc, Po, 0.3, 0.05, max_steps=40, source=("point_mass", [0.4, 0.35, 0.5], [0.1, 0.01]))
    assert err == 0
    with Block(desc) as blk:
        blk.set_source(*POINT, (0.1, 0.01))
        blk.set_source_aux((0.4, 0.35, 0.5))
        blk.upload(P0)
        tg, sg, rep = blk.run(0.3, 0.05, 40)
        Pg = blk.download()
    assert (sg, tg) == (so, to)
    assert np.array_equal(Pg, Po)


def test_uniform_gravity_on_a_uniform_state():
    """No gradients: the fluxes cancel exactly, rho v = rho g t, and the energy
    source rho v . g adds exactly the kinetic energy, so p stays put."""
    desc = abi.make_desc(law="hydro", recon="plm", riemann="hllc", bc="periodic", cells=(32, 16, 1))
    P0 = problems.empty(desc)
    P0[..., 0] = 2.0
    P0[..., 4] = 1.5
    g = np.array([0.25, -0.5, 0.0])
    with Block(desc) as blk:
        blk.set_source(*GRAVITY, g)
        blk.upload(P0)
        t, steps, rep = blk.run(0.4, 0.3, 1000)
        P = blk.download()
    assert steps > 1 and t == 0.3
    np.testing.assert_allclose(P[..., 1:4], np.broadcast_to(g * t, P[..., 1:4].shape), rtol=0, atol=1e-14)
    np.testing.assert_allclose(P[..., 4], 1.5, rtol=0, atol=1e-13)
    assert np.all(P[..., 0] == 2.0)


def test_removing_the_source_restores_the_plain_step():
    desc = abi.make_desc(**CASES["hydro2d_fused_to_flux"])
    P0 = _state(desc, 64)
    with Block(desc) as plain:
        plain.upload(P0)
        plain.run(0.3, 1e300, 5)
        ref = plain.download()
    with Block(desc) as blk:
        blk.set_source(*GRAVITY, (1.0, 0.0, 0.0))
        blk.set_source(None)
        blk.upload(P0)
        blk.run(0.3, 1e300, 5)
        assert np.array_equal(blk.download(), ref)


def test_bad_code_object_or_kernel_is_refused():
    """Refused, and the block (with its HIP error state) stays usable."""
    desc = abi.make_desc(**CASES["hydro1d_pencil"])
    with Block(desc) as blk:
        with pytest.raises(runtime.MaraError):
            blk.set_source(os.path.join(SRC, "missing.hsaco"), "k")
        with pytest.raises(runtime.MaraError):
            blk.set_source(GRAVITY[0], "no_such_kernel")
        with pytest.raises(runtime.MaraError):
            blk.set_source(GRAVITY[0], GRAVITY[1], np.zeros(17))
        blk.upload(_state(desc, 65))
        blk.advance(0.3 * blk.courant_dt())


# ---------------------------------------------------------------------------
# Source terms in batches (mara_batch_set_source): the code object's
# <kernel>_batch kernel evaluates the function for every solve of the batch in
# one launch per stage.

BATCH_MIX = {
    "hydro1d": dict(law="hydro", recon="plm", riemann="hllc", bc="outflow", cells=(300, 1, 1)),
    "hydro2d": dict(law="hydro", recon="plm", riemann="hllc", bc="periodic", cells=(70, 37, 1)),
    "mhd2d_ct": dict(law="mhd", recon="plm", riemann="hlle", bc="periodic", cells=(48, 30, 1), field_ct=1),
    "mhd1d": dict(law="mhd", recon="plm", riemann="hlle", bc="outflow", cells=(120, 1, 1)),
    "hydro3d_weno": dict(law="hydro", recon="weno5", riemann="hlle", bc="periodic", cells=(20, 12, 10)),
}


def _batch(items, code, params, taux=(), precision="fp64", cfl=0.3, steps=5):
    """Admit every (name, desc, P0) at once under one source; step to the end."""
    from mara2_amd.runtime import Batch
    out = {}
    with Batch(1 << 28, rk_order=2, precision=precision, max_solves=8) as b:
        if code is not None:
            b.set_source(code[0], code[1], params)
        if taux:
            b.set_source_aux(taux)
        slots = {b.admit(d, P, cfl, 1e300, steps): name for name, d, P in items}
        while slots:
            b.step(3)
            for s in b.finished():
                done, t, n, st, rep = b.query(s)
                assert done and st == 0, rep.as_dict()
                out[slots.pop(s)] = (b.retire(s), t, n)
    return out


def _mix(seed, precision="fp64"):
    items = []
    for i, (name, kw) in enumerate(sorted(BATCH_MIX.items())):
        d = abi.make_desc(precision=precision, **kw)
        items.append((name, d, _state(d, seed + i)))
    return items


@pytest.mark.parametrize("which", ["gravity", "point_mass"])
def test_batched_sources_match_oracle(which):
    """Every solve of a 1-D/2-D/3-D hydro and MHD mix under one source,
    bit for bit against the oracle's loop with the same source."""
    import oracle
    if which == "gravity":
        code, params, taux, name = GRAVITY, (0.3, -0.2, 0.1), (), "uniform_gravity"
    else:
        code, params, taux, name = POINT, (0.1, 0.01), (0.4, 0.35, 0.5), "point_mass"
    items = _mix(71)
    got = _batch(items, code, params, taux)
    for nm, d, P0 in items:
        Po = P0.copy()
        err, to, so = oracle.run(d, Po, 0.3, 1e300, max_steps=5, source=(name, list(taux), list(params)))
        assert err == 0
        Pg, t, n = got[nm]
        assert (t, n) == (to, so), nm
        diff = np.argwhere(Pg != Po)
        assert diff.size == 0, f"{nm}: {len(diff)} differ, first {diff[0]}"


def test_batched_source_equals_single_block():
    items = _mix(72)
    got = _batch(items, GRAVITY, (0.5, 0.25, -0.125))
    for nm, d, P0 in items:
        with Block(d) as blk:
            blk.set_source(*GRAVITY, (0.5, 0.25, -0.125))
            blk.upload(P0)
            t, n, _ = blk.run(0.3, 1e300, 5)
            Pb = blk.download()
        assert (t, n) == got[nm][1:], nm
        assert np.array_equal(Pb, got[nm][0]), nm


def test_mixed_batch_with_source_close_to_oracle():
    """The mixed form's batch (binary32 faces) with the source within the
    smooth-flow bound 1e-5 of the FP64 oracle after 5 steps."""
    import oracle
    from helpers import interior, prim_to_cons, rel_l1
    items = _mix(73, precision="mixed")
    got = _batch(items, GRAVITY, (0.3, -0.2, 0.1), precision="mixed")
    for nm, d, P0 in items:
        d64 = abi.make_desc(**BATCH_MIX[nm])
        Po = P0.copy()
        err, to, so = oracle.run(d64, Po, 0.3, 1e300, max_steps=5, source=("uniform_gravity", [], [0.3, -0.2, 0.1]))
        assert err == 0 and so == got[nm][2]
        e = rel_l1(prim_to_cons(d64, interior(d64, got[nm][0])), prim_to_cons(d64, interior(d64, Po)))
        assert e <= 1e-5, (nm, e)


def test_removed_batch_source_restores_the_plain_kernels():
    """Setting a source and removing it again before admission gives the
    plain batch (fused tiles and pencils included) bit for bit."""
    items = _mix(74)
    plain = _batch(items, None, ())
    from mara2_amd.runtime import Batch
    out = {}
    with Batch(1 << 28, rk_order=2, precision="fp64", max_solves=8) as b:
        b.set_source(*GRAVITY, (1.0, 0.0, 0.0))
        b.set_source(None)
        slots = {b.admit(d, P, 0.3, 1e300, 5): nm for nm, d, P in items}
        while slots:
            b.step(3)
            for s in b.finished():
                out[slots.pop(s)] = b.retire(s)
    for nm in plain:
        assert np.array_equal(out[nm], plain[nm][0]), nm


def test_batch_source_refusals():
    from mara2_amd.runtime import Batch
    d = abi.make_desc(**BATCH_MIX["hydro2d"])
    dc = abi.make_desc(law="mhd", riemann="hlle", cooling_rate=0.5, cells=(40, 24, 1), bc="periodic", field_ct=1)
    with Batch(1 << 28, rk_order=2, precision="fp64", max_solves=4) as b:
        with pytest.raises(runtime.MaraError):
            b.set_source(GRAVITY[0], "no_such_kernel")
        with pytest.raises(runtime.MaraError):
            b.set_source(os.path.join(SRC, "missing.hsaco"), "k")
        b.set_source(*GRAVITY, (0.1, 0.0, 0.0))
        P = problems.empty(dc)
        P[..., 0] = 1.0
        P[..., 4] = 1.0
        with pytest.raises(runtime.MaraError) as e:
            b.admit(dc, problems.apply_bc(dc, P), 0.3, 1e300, 5)
        assert e.value.status == abi.ERR_UNSUPPORTED
        s = b.admit(d, _state(d, 75), 0.3, 1e300, 2)
        with pytest.raises(runtime.MaraError) as e:
            b.set_source(None)
        assert e.value.status == abi.ERR_INVALID_ARGUMENT
        while not b.finished():
            b.step(2)
        b.retire(s)
        b.set_source(None)  # allowed again once no solve is resident
