"""Launch shapes and the autotuner.  A shape (pencil length of 1-D stages,
z planes per work-group of the fused 3-D kernels, work-group rows of the
flux-array kernels and tile height of the fused 2-D MHD kernel) only redistributes zones among work-groups, so every
shape must give the default shape's result bit for bit; the autotuner must
leave the state and clock untouched and record its choice for blocks
created later with the same descriptor."""
import numpy as np
import pytest

from mara2_amd import abi, problems, runtime
from mara2_amd.runtime import Batch, Block

pytestmark = pytest.mark.gpu


def _random_state(desc, seed, amp=0.3):
    P = problems.empty(desc)
    rng = np.random.default_rng(seed)
    shp = P.shape[:3]
    P[..., 0] = 1.0 + 0.5 * rng.random(shp)
    P[..., 4] = 1.0 + 0.5 * rng.random(shp)
    P[..., 1:4] = amp * rng.standard_normal(shp + (3,))
    if desc.law != 0:
        P[..., 5:8] = 0.5 + 0.2 * rng.standard_normal(shp + (3,))
    return problems.apply_bc(desc, P)


def _case(name):
    if name == "sod_pencil":
        d = problems.sod_desc(700)
        return d, problems.sod_1d(d), "pencil", abi.PENCIL_SHAPES
    if name == "sod_pencil_mixed_rk3":
        d = abi.make_desc(law="hydro", gamma=1.4, recon="weno5", riemann="hllc", bc="outflow", cells=(600, 1, 1),
                          rk_order=3, precision="mixed")
        return d, problems.sod_1d(d), "pencil", abi.PENCIL_SHAPES
    if name == "mhd1d_pencil_ct":
        d = abi.make_desc(law="mhd", recon="plm", riemann="hlle", bc="periodic", cells=(530, 1, 1), field_ct=1)
        return d, _random_state(d, 41), "pencil", abi.PENCIL_SHAPES
    if name == "turbulence3d_zplanes":
        d = problems.turbulence_desc(40)
        return d, problems.turbulence_3d(d), "zplanes", abi.ZPLANE_SHAPES + (13, 40)
    if name == "turbulence3d_zplanes_mixed":
        d = problems.turbulence_desc(40, "mixed")
        return d, problems.turbulence_3d(d), "zplanes", abi.ZPLANE_SHAPES + (13, 40)
    if name == "weno3d_zplanes":
        d = abi.make_desc(law="hydro", recon="weno5", riemann="hlle", bc="periodic", cells=(70, 12, 37))
        return d, _random_state(d, 42), "zplanes", abi.ZPLANE_SHAPES + (1, 13, 37)
    if name == "mhd2d_rows":
        d = problems.mhd_desc(48)
        return d, problems.mhd_2d(d), "rows", abi.ROW_SHAPES
    if name == "mhd2d_rows_ragged_mixed":
        d = abi.make_desc(law="mhd", recon="plm", riemann="hlle", bc="outflow", cells=(70, 37, 1), field_ct=1,
                          precision="mixed")
        return d, _random_state(d, 43), "rows", abi.ROW_SHAPES
    if name == "mhd2d_rows_weno_noct":
        d = abi.make_desc(law="mhd", recon="weno5", riemann="hlle", bc="periodic", cells=(130, 21, 1), field_ct=0,
                          rk_order=3)
        return d, _random_state(d, 44), "rows", abi.ROW_SHAPES
    raise KeyError(name)


def _run(desc, P0, shape, steps=6, cfl=0.3):
    with Block(desc) as blk:
        if shape is not None:
            blk.set_shape(shape)
            assert blk.get_shape().as_dict() == shape.as_dict()
        blk.upload(P0)
        t, n, rep = blk.run(cfl, 1e300, steps)
        return t, n, blk.download()


@pytest.mark.parametrize("name", ["sod_pencil", "sod_pencil_mixed_rk3", "mhd1d_pencil_ct", "turbulence3d_zplanes",
                                  "turbulence3d_zplanes_mixed", "weno3d_zplanes", "mhd2d_rows",
                                  "mhd2d_rows_ragged_mixed", "mhd2d_rows_weno_noct"])
def test_every_shape_gives_the_default_result(name):
    desc, P0, knob, values = _case(name)
    t0, n0, ref = _run(desc, P0, None)
    for v in values:
        shape = abi.make_shape(**{knob: v})
        t, n, P = _run(desc, P0, shape)
        assert (t, n) == (t0, n0), (knob, v)
        assert np.array_equal(P, ref), f"{knob}={v}: results differ"


@pytest.mark.parametrize("name", ["sod_pencil", "turbulence3d_zplanes", "mhd2d_rows"])
def test_autotune_keeps_state_and_records_choice(name):
    desc, P0, knob, values = _case(name)
    t_ref, n_ref, P_ref = _run(desc, P0, None, steps=6)
    with Block(desc) as blk:
        blk.upload(P0)
        t, n, _ = blk.run(0.3, 1e300, 3)
        before = blk.download()
        shape, ms = blk.autotune(repeats=3)
        v = getattr(shape, knob)
        # z planes: the fixed candidates or a count balanced to device rounds
        assert (v in values or (knob == "zplanes" and 4 <= v <= abi.ZPLANES_MAX)) and ms > 0
        assert blk.get_shape().as_dict() == shape.as_dict()
        assert np.array_equal(blk.download(), before)
        t, n2, _ = blk.run(0.3, 1e300, 3, t=t)
        assert (t, n + n2) == (t_ref, n_ref)
        assert np.array_equal(blk.download(), P_ref)
    assert runtime.tune_lookup(desc).as_dict() == shape.as_dict()
    with Block(desc) as again:
        assert again.get_shape().as_dict() == shape.as_dict()


def test_invalid_shapes_are_refused():
    desc, P0, _, _ = _case("sod_pencil")
    with Block(desc) as blk:
        # a new block starts from the tuning table's entry, if an earlier
        # autotune recorded one for this descriptor
        start = blk.get_shape().as_dict()
        tuned = runtime.tune_lookup(desc)
        assert start == (tuned.as_dict() if tuned else {"pencil": 256, "zplanes": 16, "rows": 4})
        for bad in (abi.make_shape(pencil=100), abi.make_shape(zplanes=0), abi.make_shape(zplanes=abi.ZPLANES_MAX + 1),
                    abi.make_shape(rows=3)):
            with pytest.raises(runtime.MaraError):
                blk.set_shape(bad)
        assert blk.get_shape().as_dict() == start
    with Batch(1 << 24, rk_order=2, precision="fp64", max_solves=4) as b:
        with pytest.raises(runtime.MaraError):
            b.set_shape(abi.make_shape(pencil=64))
