"""The C++ host layer (mara2_amd/host, Mara2's plugin interfaces over the C
ABI) driving the reference's regression programs on the GPU, replayed on the
oracle from the same initial data with the same maraMainLoop contract."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from mara2_amd import abi

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "mara2_amd", "mara_regress")


def _read(path):
    with open(path, "rb") as f:
        h = np.frombuffer(f.read(32), dtype=np.int32)
        t = np.frombuffer(f.read(8), dtype=np.float64)[0]
        n = int(np.prod(h[:4]))
        P0 = np.frombuffer(f.read(8 * n), dtype=np.float64).reshape(h[:4])
        P = np.frombuffer(f.read(8 * n), dtype=np.float64).reshape(h[:4])
    return int(h[4]), t, P0.copy(), P.copy()


# (program, problem, scheme, descriptor of the same run, cfl, t_final)
CASES = [
    ("regress-1d", "Shocktube1", "plm2", dict(law="hydro", recon="plm", riemann="hllc", bc="outflow", cells=(128, 1, 1)), 0.5, 0.1),
    ("regress-1d", "Shocktube3", "pcm3", dict(law="hydro", recon="plm", riemann="hllc", bc="outflow", cells=(128, 1, 1)), 0.5, 0.005),
    ("regress-1d", "DensityWave", "plm1", dict(law="hydro", recon="plm", riemann="hllc", bc="periodic", cells=(128, 1, 1)), 0.5, 1.0),
    ("regress-rel", "RelativisticShocktube1", "plm2", dict(law="rmhd", recon="plm", riemann="hlle", bc="outflow", cells=(128, 1, 1)), 0.25, 0.15),
    ("regress-2d", "Shocktube2D", "plm2", dict(law="hydro", recon="plm", riemann="hlle", bc="outflow", cells=(128, 128, 1)), 0.5, 0.1),
    ("regress-2d", "ContactWave2D", "pcm2", dict(law="hydro", recon="pcm", riemann="hlle", bc="periodic", cells=(128, 128, 1)), 0.5, 0.1),
    ("regress-mhd", "CylindricalBlast", "plm2", dict(law="mhd", recon="plm", riemann="hlle", bc="periodic", cells=(64, 64, 1)), 0.5, 0.1),
    ("sod", "Sod", "plm2", dict(law="hydro", gamma=1.4, recon="plm", riemann="hlle", bc="outflow", cells=(1024, 1, 1)), 0.4, 0.2),
]


@pytest.mark.parametrize("program,problem,scheme,kw,cfl,tfinal", CASES, ids=[f"{c[0]}-{c[1]}-{c[2]}" for c in CASES])
def test_regression_program_matches_oracle(tmp_path, program, problem, scheme, kw, cfl, tfinal):
    out = tmp_path / "out.bin"
    r = subprocess.run([EXE, program, problem, scheme, str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    steps, t, P0, Pg = _read(out)
    kw = dict(kw)
    kw.setdefault("gamma", 5 / 3)
    desc = abi.make_desc(rk_order=int(scheme[3]), guards=(2, 2 if kw["cells"][1] > 1 else 0, 0), **kw)
    Po = P0.copy()
    err, to, so = oracle.run(desc, Po, cfl, tfinal)
    assert err == 0
    assert so == steps and to == t
    if kw["law"] == "rmhd":
        assert np.sum(np.abs(Pg - Po)) / np.sum(np.abs(Po)) < 1e-12
    else:
        assert np.array_equal(Pg, Po)


def test_regression_program_with_source_terms_matches_oracle(tmp_path):
    """GpuMethodOfLinesTVD::setSourceTermsCodeObject through the C++ layer:
    the 2-D shock tube under uniform gravity (the uniform_gravity code
    object) against the oracle's source path, bit for bit."""
    out = tmp_path / "out.bin"
    src = os.path.join(ROOT, "mara2_amd", "sources", "uniform_gravity.hsaco")
    g = (0.5, -1.0, 0.0)
    r = subprocess.run([EXE, "regress-2d", "Shocktube2D", "plm2", str(out), "1000000", src,
                        "mara_source_uniform_gravity"] + [repr(x) for x in g],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    steps, t, P0, Pg = _read(out)
    desc = abi.make_desc(law="hydro", recon="plm", riemann="hlle", bc="outflow", cells=(128, 128, 1),
                         guards=(2, 2, 0), rk_order=2)
    Po = P0.copy()
    err, to, so = oracle.run(desc, Po, 0.5, 0.1, source=("uniform_gravity", [], list(g)))
    assert err == 0
    assert so == steps and to == t
    assert np.array_equal(Pg, Po)


def test_regression_program_with_viscous_flux_matches_oracle(tmp_path):
    """GpuMethodOfLinesTVD::setViscousFluxCodeObject through the C++ layer:
    the 2-D shock tube with constant viscosity (the viscous_constant code
    object) against the oracle's viscous-flux path, bit for bit."""
    out = tmp_path / "out.bin"
    obj = os.path.join(ROOT, "mara2_amd", "sources", "viscous_constant.hsaco")
    r = subprocess.run([EXE, "regress-2d", "Shocktube2D", "plm2", str(out), "1000000", "viscous=" + obj,
                        "mara_viscous_constant", "0.002"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    steps, t, P0, Pg = _read(out)
    desc = abi.make_desc(law="hydro", recon="plm", riemann="hlle", bc="outflow", cells=(128, 128, 1),
                         guards=(2, 2, 0), rk_order=2)
    Po = P0.copy()
    err, to, so = oracle.run(desc, Po, 0.5, 0.1, viscous=("constant", [0.002]))
    assert err == 0
    assert so == steps and to == t
    assert np.array_equal(Pg, Po)
