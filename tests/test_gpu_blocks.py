"""Block groups (the single-device half of BlockDecomposition,
src/BlockDecomposition.cpp) against the undivided block: the same mesh split
into blocks that exchange guard layers after every stage must step zone for
zone like one block, so every comparison here is bitwise, in each precision
form, on each stage path (fused 3-D tiles, fast tiles, 1-D pencils, 2-D MHD
tiles, flux arrays), with periodic, outflow and reflecting meshes and with
partitions that leave blocks of unequal size."""
import numpy as np
import pytest

from mara2_amd import abi, problems
from mara2_amd.runtime import Block, Group

pytestmark = pytest.mark.gpu


def _random_state(desc, seed, mhd=False):
    rng = np.random.default_rng(seed)
    P = problems.empty(desc)
    shp = P.shape[:3]
    P[..., 0] = 1.0 + 0.5 * rng.random(shp)
    P[..., 4] = 1.0 + 0.5 * rng.random(shp)
    P[..., 1:4] = 0.4 * rng.standard_normal(shp + (3,))
    if mhd:
        P[..., 5:8] = 0.5 * rng.standard_normal(shp + (3,))
    return problems.apply_bc(desc, P)


def _step_both(desc, P0, dims, steps, cfl=0.3):
    """Advance one block and a group of dims blocks from P0 with the block's
    CFL time steps; the group's CFL must equal the block's bit for bit."""
    with Block(desc) as blk, Group(desc, dims) as grp:
        blk.upload(P0)
        grp.upload(P0)
        blk.apply_bc()
        grp.apply_bc()
        for n in range(steps):
            dtb = blk.courant_dt()
            dtg = grp.courant_dt()
            assert dtb == dtg, (n, dtb, dtg)
            blk.advance(cfl * dtb)
            grp.advance(cfl * dtb)
        return blk.download(), grp.download()


def _assert_same(Pb, Pg, what):
    assert np.all(np.isfinite(Pb)), what
    diff = np.argwhere(Pb != Pg)
    assert diff.size == 0, f"{what}: {len(diff)} values differ, first at {tuple(diff[0])}"


@pytest.mark.parametrize("precision", ["fp64", "mixed", "fp64_fast"])
@pytest.mark.parametrize("dims", [(2, 2, 2), (3, 1, 1), (1, 2, 3)])
def test_hydro_3d_periodic(precision, dims):
    desc = abi.make_desc(law="hydro", riemann="hllc", bc="periodic", cells=(26, 20, 23), rk_order=2,
                         precision=precision)
    P0 = _random_state(desc, 11)
    Pb, Pg = _step_both(desc, P0, dims, 4)
    _assert_same(Pb, Pg, f"hydro 3-D {precision} {dims}")


@pytest.mark.parametrize("bc", ["outflow", "reflecting"])
@pytest.mark.parametrize("rk_order", [1, 2, 3])
def test_hydro_2d_walls(bc, rk_order):
    # RK1 runs the flux arrays, RK2 the fused tiles, RK3 the three-buffer schedule
    desc = abi.make_desc(law="hydro", riemann="hlle", bc=bc, cells=(30, 22, 1), rk_order=rk_order)
    P0 = _random_state(desc, 12)
    Pb, Pg = _step_both(desc, P0, (3, 2, 1), 4)
    _assert_same(Pb, Pg, f"hydro 2-D {bc} RK{rk_order}")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sod_pencils(precision):
    desc = problems.sod_desc(200, precision)
    P0 = problems.sod_1d(desc)
    Pb, Pg = _step_both(desc, P0, (5, 1, 1), 6)
    _assert_same(Pb, Pg, f"Sod pencils {precision}")


def test_hydro_3d_weno5():
    desc = abi.make_desc(law="hydro", recon="weno5", riemann="hllc", bc="periodic", cells=(20, 18, 16),
                         rk_order=3)
    P0 = _random_state(desc, 13)
    Pb, Pg = _step_both(desc, P0, (2, 2, 1), 3)
    _assert_same(Pb, Pg, "hydro 3-D WENO5")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_mhd_2d_tiles_without_ct(precision):
    desc = abi.make_desc(law="mhd", riemann="hlle", bc="periodic", cells=(24, 20, 1), rk_order=2, field_ct=0,
                         precision=precision)
    P0 = _random_state(desc, 14, mhd=True)
    Pb, Pg = _step_both(desc, P0, (2, 2, 1), 4)
    _assert_same(Pb, Pg, f"MHD 2-D {precision}")


def test_mhd_3d_flux_arrays_outflow():
    desc = abi.make_desc(law="mhd", riemann="hlle", bc="outflow", cells=(16, 14, 12), rk_order=2, field_ct=0)
    P0 = _random_state(desc, 15, mhd=True)
    Pb, Pg = _step_both(desc, P0, (2, 1, 2), 3)
    _assert_same(Pb, Pg, "MHD 3-D outflow")


def test_rmhd_3d():
    desc = abi.make_desc(law="rmhd", riemann="hlle", bc="periodic", cells=(12, 12, 12), rk_order=2, field_ct=0)
    rng = np.random.default_rng(16)
    P0 = problems.empty(desc)
    shp = P0.shape[:3]
    P0[..., 0] = 1.0 + 0.2 * rng.random(shp)
    P0[..., 4] = 1.0 + 0.2 * rng.random(shp)
    P0[..., 1:4] = 0.1 * rng.standard_normal(shp + (3,))
    P0[..., 5:8] = 0.3 * rng.standard_normal(shp + (3,))
    P0 = problems.apply_bc(desc, P0)
    Pb, Pg = _step_both(desc, P0, (2, 1, 2), 3)
    _assert_same(Pb, Pg, "RMHD 3-D")


def test_regions_tile_the_mesh():
    desc = abi.make_desc(law="hydro", bc="periodic", cells=(26, 20, 23))
    with Group(desc, (3, 2, 4)) as grp:
        assert grp.num_blocks == 24
        hit = np.zeros(tuple(desc.cells[a] for a in range(3)), dtype=int)
        for m in range(grp.num_blocks):
            s, c = grp.block_region(m)
            hit[s[0]:s[0] + c[0], s[1]:s[1] + c[1], s[2]:s[2] + c[2]] += 1
        assert np.all(hit == 1)
        assert grp.device_bytes > 0

