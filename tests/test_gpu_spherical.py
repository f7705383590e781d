"""Spherical meshes (SphericalMeshGeometry, src/SphericalMeshGeometry.cpp) on
the GPU against the oracle.  The device forms the face areas, volumes and
cell lengths from host-evaluated edge tables with the reference's
expressions, so the binary64 form equals the oracle bit for bit on 1-D, 2-D
and 3-D meshes with linear and logarithmic radial edges; the mixed form is
held to the smooth-flow bound 1e-5.  With the ThermalConvection source
terms (gravity, heating and the geometric terms, which use tan, pow and exp
on the device) the step is held to relative L1 1e-12, and so is the
ThermalConvection program run through the C++ host layer."""
import math
import os
import subprocess

import numpy as np
import pytest

import oracle
from mara2_amd import abi, problems
from mara2_amd.runtime import Block, MaraError
from helpers import interior, prim_to_cons, rel_l1, run_pair

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "mara2_amd", "sources", "thermal_convection.hsaco")

MESHES = {
    "1d": dict(cells=(64, 1, 1), lower=(1.0, 0.0, 0.0), upper=(3.0, math.pi, 2 * math.pi)),
    "2d": dict(cells=(40, 24, 1), lower=(1.0, 0.3, 0.0), upper=(3.0, math.pi - 0.3, 2 * math.pi)),
    "3d": dict(cells=(16, 12, 10), lower=(1.0, 0.4, 0.0), upper=(2.5, 2.7, 1.2)),
}


def _desc(mesh, log, precision="fp64", bc="outflow", **kw):
    m = MESHES[mesh]
    return abi.make_desc(law="hydro", recon="plm", riemann="hllc", bc=bc, rk_order=2, precision=precision,
                         geometry="spherical", log_radial=log, **m, **kw)


def _blast(desc, seed=3):
    """A radial blast (high pressure inside r = 1.6) with a seeded velocity
    perturbation on the fleshed axes."""
    rng = np.random.default_rng(seed)
    P = problems.empty(desc)
    g = desc.guards
    n = desc.cells
    r = np.array([_centre(desc, 0, i) for i in range(-g[0], n[0] + g[0])])
    P[..., 0] = 1.0
    P[..., 4] = np.where(r < 1.6, 4.0, 1.0)[:, None, None]
    for a in range(3):
        if n[a] > 1:
            P[..., 1 + a] = 0.1 * rng.standard_normal(P.shape[:3])
    return problems.apply_bc(desc, P)


def _centre(desc, axis, i):
    lo, hi, n = desc.lower[axis], desc.upper[axis], desc.cells[axis]
    if axis == 0 and desc.log_radial:
        return lo * (hi / lo) ** ((i + 0.5) / n)
    return lo + (hi - lo) * (i + 0.5) / n


@pytest.mark.parametrize("mesh", list(MESHES))
@pytest.mark.parametrize("log", [False, True])
def test_fp64_equals_oracle(mesh, log):
    desc = _desc(mesh, log)
    P0 = _blast(desc)
    Pg, Po, dg, do = run_pair(desc, P0, 8, 0.4, Block)
    assert dg == do
    assert np.array_equal(Pg, Po)


@pytest.mark.parametrize("mesh", ["2d", "3d"])
def test_reflecting_and_periodic_walls(mesh):
    for bc in ("reflecting", "periodic"):
        desc = _desc(mesh, True, bc=bc)
        Pg, Po, dg, do = run_pair(desc, _blast(desc, 5), 6, 0.4, Block)
        assert dg == do
        assert np.array_equal(Pg, Po), bc


@pytest.mark.parametrize("mesh", ["1d", "2d"])
def test_mixed_close_to_oracle(mesh):
    desc = _desc(mesh, True)
    dmix = _desc(mesh, True, precision="mixed")
    P0 = _blast(desc)
    Po = P0.copy()
    with Block(dmix) as blk:
        blk.upload(P0)
        for _ in range(8):
            err, dto = oracle.courant_dt(desc, Po)
            assert err == 0
            blk.advance(0.4 * dto)
            err, _ = oracle.advance(desc, Po, 0.0, 0.4 * dto)
            assert err == 0
        Pg = blk.download()
    err = rel_l1(prim_to_cons(desc, interior(desc, Pg)), prim_to_cons(desc, interior(desc, Po)))
    print(f"mixed spherical {mesh}: rel L1 {err:.2e}")
    assert err <= 1e-5


def _polytrope(desc, g0=0.2):
    K = g0 * (5 / 3 - 1) / (5 / 3)
    P = problems.empty(desc)
    g, n = desc.guards, desc.cells
    r = np.array([_centre(desc, 0, i) for i in range(-g[0], n[0] + g[0])])
    d = r ** -1.5
    P[..., 0] = d[:, None, None]
    P[..., 4] = (K * d ** (5 / 3))[:, None, None]
    return problems.apply_bc(desc, P)


def test_thermal_convection_source():
    desc = abi.make_desc(law="hydro", recon="plm", plm_theta=2.0, riemann="hllc", bc="outflow", rk_order=2,
                         cells=(48, 32, 1), lower=(1.0, math.pi / 2 - math.pi / 12, 0.0),
                         upper=(100.0, math.pi / 2 + math.pi / 12, 0.1), geometry="spherical", log_radial=True)
    P0 = _polytrope(desc)
    params = [0.2, 1.0]
    Po = P0.copy()
    with Block(desc) as blk:
        blk.set_source(SRC, "mara_source_thermal_convection", params)
        blk.upload(P0)
        for _ in range(10):
            dtg = 0.5 * blk.courant_dt()
            err, dto = oracle.courant_dt(desc, Po)
            assert err == 0
            assert abs(dtg - 0.5 * dto) <= 1e-12 * dtg
            blk.advance(dtg)
            err, _ = oracle.advance(desc, Po, 0.0, dtg, ("thermal_convection", [], params))
            assert err == 0
        Pg = blk.download()
    err = rel_l1(prim_to_cons(desc, interior(desc, Pg)), prim_to_cons(desc, interior(desc, Po)))
    print(f"thermal convection source: rel L1 {err:.2e}")
    assert err <= 1e-12
    # the source moved the gas
    assert np.max(np.abs(interior(desc, Pg)[..., 1])) > 1e-6


def test_thermal_convection_program_through_host_layer(tmp_path):
    out = tmp_path / "tc.bin"
    exe = os.path.join(ROOT, "mara2_amd", "mara_regress")
    r = subprocess.run([exe, "thermal-convection", "ThermalConvection", "plm2", str(out), "1000000", SRC,
                        "mara_source_thermal_convection", "0.2", "1.0"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    with open(out, "rb") as f:
        h = np.frombuffer(f.read(32), dtype=np.int32)
        t = np.frombuffer(f.read(8), dtype=np.float64)[0]
        n = int(np.prod(h[:4]))
        P0 = np.frombuffer(f.read(8 * n), dtype=np.float64).reshape(h[:4]).copy()
        Pg = np.frombuffer(f.read(8 * n), dtype=np.float64).reshape(h[:4]).copy()
    steps = int(h[4])
    desc = abi.make_desc(law="hydro", recon="plm", plm_theta=2.0, riemann="hllc", bc="outflow", rk_order=2,
                         cells=(64, 64, 1), guards=(2, 2, 0), lower=(1.0, math.pi / 2 - math.pi / 12, 0.0),
                         upper=(100.0, math.pi / 2 + math.pi / 12, 0.1), geometry="spherical", log_radial=True)
    Po = P0.copy()
    err, to, so = oracle.run(desc, Po, 0.5, 0.5, source=("thermal_convection", [], [0.2, 1.0]))
    assert err == 0
    assert so == steps and abs(to - t) <= 1e-12
    assert rel_l1(prim_to_cons(desc, interior(desc, Pg)), prim_to_cons(desc, interior(desc, Po))) <= 1e-12


def test_batches_refuse_spherical_meshes():
    from mara2_amd.runtime import Batch
    desc = _desc("2d", False)
    with Batch(1 << 28) as b:
        with pytest.raises(MaraError) as e:
            b.admit(desc, _blast(desc), 0.4, 0.1, 10)
        assert e.value.status == abi.ERR_UNSUPPORTED
