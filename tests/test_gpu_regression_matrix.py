"""Every run of the reference's four regression programs (src/Problems.cpp:35-671)
through the C++ host layer on the GPU, each replayed on the oracle from the
same initial data with the same maraMainLoop contract:

    regress-1d   7 problems x 6 schemes (pcm1 plm1 pcm2 plm2 pcm3 plm3)   42 runs
    regress-rel  1 problem  x 2 schemes (pcm2 plm2)                         2 runs
    regress-2d   3 problems x 2 schemes (pcm2 plm2)                         6 runs
    regress-mhd  3 problems x 1 scheme  (plm2, flux-CT)                     3 runs

Hydro and MHD runs equal the oracle bit for bit; the relativistic runs within
relative L1 1e-12 (device pow/acos/cos/sin against the host libm).  Each run
also yields the norms the reference's Vis scripts plot
(Vis/RegressionHydro1d.py, Vis/RegressionMHD2d.py): relative L1 of density,
pressure and velocity1 (and magnetic1-3 for MHD, with max |div B|) against
the exact solution where one exists: Toro's exact Riemann solution
(tests/exact_riemann.py) for the shock tubes, along x + y for the diagonal
2-D tube; the initial state for the stationary contacts, for the density
waves and the field loop (advected exactly one period) and for the
force-free ABC field.  With MARA_REGRESSION_TABLE set to a path, the last
test writes the table there as markdown.

regress-1d replaces every scheme's flux scheme with PLM + HLLC
(src/Problems.cpp:162-174), so its pcm and plm runs of one RK order are the
same computation, as in the reference."""
import math
import os
import subprocess

import numpy as np
import pytest

import oracle
from mara2_amd import abi
from exact_riemann import sample

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "mara2_amd", "mara_regress")

G53 = 5.0 / 3.0
# problem: (final time, x0-left state, right state) of the 1-D tubes as
# (rho, vx, p), from src/Problems.cpp:35-89
TUBES_1D = {
    "Shocktube1": (0.100, (1.000, 0.000, 1.000), (0.125, 0.000, 0.100)),
    "Shocktube2": (0.100, (1.000, -2.000, 0.400), (1.000, 2.000, 0.400)),
    "Shocktube3": (0.005, (1.0, 0.0, 1e+3), (1.0, 0.0, 1e-2)),
    "Shocktube4": (0.01, (1.0, 0.0, 1e-2), (1.0, 0.0, 1e+2)),
    "Shocktube5": (0.01, (5.99924, 19.59750, 460.894), (5.99924, -6.19633, 46.095)),
}
FINAL = {"ContactWave": 0.1, "DensityWave": 1.0, "RelativisticShocktube1": 0.15, "Shocktube2D": 0.1,
         "ContactWave2D": 0.1, "DensityWave2D": 1.0, "CylindricalBlast": 0.1, "FieldLoop": 1.0, "AbcField": 0.1}
FINAL.update({k: v[0] for k, v in TUBES_1D.items()})

RUNS = []
for p in list(TUBES_1D) + ["ContactWave", "DensityWave"]:
    for s in ("pcm1", "plm1", "pcm2", "plm2", "pcm3", "plm3"):
        RUNS.append(("regress-1d", p, s))
for s in ("pcm2", "plm2"):
    RUNS.append(("regress-rel", "RelativisticShocktube1", s))
for p in ("Shocktube2D", "ContactWave2D", "DensityWave2D"):
    for s in ("pcm2", "plm2"):
        RUNS.append(("regress-2d", p, s))
for p in ("CylindricalBlast", "FieldLoop", "AbcField"):
    RUNS.append(("regress-mhd", p, "plm2"))
assert len(RUNS) == 53

ROWS = {}


def _read(path):
    with open(path, "rb") as f:
        h = np.frombuffer(f.read(32), dtype=np.int32)
        t = np.frombuffer(f.read(8), dtype=np.float64)[0]
        n = int(np.prod(h[:4]))
        P0 = np.frombuffer(f.read(8 * n), dtype=np.float64).reshape(h[:4])
        P = np.frombuffer(f.read(8 * n), dtype=np.float64).reshape(h[:4])
    return int(h[4]), t, P0.copy(), P.copy()


def _desc(program, scheme):
    """The descriptor of one run (src/Problems.cpp:141-174, 281-300, 421-440, 597-620)."""
    rk = int(scheme[3])
    recon = "plm" if scheme.startswith("plm") else "pcm"
    if program == "regress-1d":
        return abi.make_desc(law="hydro", gamma=G53, recon="plm", riemann="hllc", cells=(128, 1, 1),
                             guards=(2, 0, 0), rk_order=rk), 0.5
    if program == "regress-rel":
        return abi.make_desc(law="rmhd", gamma=G53, recon=recon, riemann="hlle", cells=(128, 1, 1),
                             guards=(2, 0, 0), rk_order=rk), 0.25
    if program == "regress-2d":
        return abi.make_desc(law="hydro", gamma=G53, recon=recon, riemann="hlle", cells=(128, 128, 1),
                             guards=(2, 2, 0), rk_order=rk), 0.5
    return abi.make_desc(law="mhd", gamma=G53, recon="plm", riemann="hlle", cells=(64, 64, 1),
                         guards=(2, 2, 0), rk_order=rk, field_ct=1), 0.5


def _bc(problem):
    periodic = problem.startswith("DensityWave") or problem in ("ContactWave2D", "CylindricalBlast", "FieldLoop",
                                                                  "AbcField")
    return abi.BCS["periodic" if periodic else "outflow"]


def _centres(n):
    return (np.arange(n) + 0.5) / n


def _exact(program, problem, desc, P0, t):
    """Exact interior state (rho, vx, p) -- plus B for MHD -- or None."""
    g = desc.guards
    n0, n1 = desc.cells[0], desc.cells[1]
    I0 = P0[g[0]:g[0] + n0, g[1]:g[1] + n1, 0, :]
    if problem in TUBES_1D:
        _, L, R = TUBES_1D[problem]
        ex = sample(L, R, G53, (_centres(n0) - 0.5) / t)
        out = I0.copy()
        out[:, 0, 0], out[:, 0, 1], out[:, 0, 4] = ex[:, 0], ex[:, 1], ex[:, 2]
        return out
    if problem == "Shocktube2D":
        # the Sod tube along the diagonal: s = (x + y - 1) / sqrt 2, normal velocity split evenly
        x, y = np.meshgrid(_centres(n0), _centres(n1), indexing="ij")
        s = (x + y - 1.0) / math.sqrt(2.0)
        ex = sample((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), G53, s.ravel() / t).reshape(n0, n1, 3)
        out = I0.copy()
        out[..., 0], out[..., 4] = ex[..., 0], ex[..., 2]
        out[..., 1] = out[..., 2] = ex[..., 1] / math.sqrt(2.0)
        return out
    if problem in ("ContactWave", "DensityWave", "ContactWave2D", "DensityWave2D", "FieldLoop", "AbcField"):
        return I0.copy()
    return None


def _l1(a, b):
    den = np.mean(np.abs(b))
    num = np.mean(np.abs(a - b))
    return num / den if den > 0 else num


@pytest.mark.parametrize("program,problem,scheme", RUNS, ids=[f"{r[0]}-{r[1]}-{r[2]}" for r in RUNS])
def test_run_matches_oracle(tmp_path, program, problem, scheme):
    out = tmp_path / "out.bin"
    r = subprocess.run([EXE, program, problem, scheme, str(out)], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    steps, t, P0, Pg = _read(out)
    desc, cfl = _desc(program, scheme)
    desc.bc = _bc(problem)
    Po = P0.copy()
    err, to, so = oracle.run(desc, Po, cfl, FINAL[problem])
    assert err == 0
    assert so == steps and to == t
    assert np.all(np.isfinite(Pg))
    if desc.law == abi.LAW_RMHD:
        vs_oracle = np.sum(np.abs(Pg - Po)) / np.sum(np.abs(Po))
        assert vs_oracle <= 1e-12
    else:
        assert np.array_equal(Pg, Po)
        vs_oracle = 0.0
    g = desc.guards
    I = Pg[g[0]:g[0] + desc.cells[0], g[1]:g[1] + desc.cells[1], 0, :]
    ex = _exact(program, problem, desc, P0, t)
    row = {"steps": steps, "t": t, "vs_oracle": vs_oracle}
    fields = [("density", 0), ("pressure", 4), ("velocity1", 1)]
    if desc.law == abi.LAW_MHD:
        fields += [("magnetic1", 5), ("magnetic2", 6), ("magnetic3", 7)]
        err, row["max_divB"] = oracle.monopole_max(desc, Pg)
        assert err == 0
    for name, q in fields:
        row[name] = None if ex is None else _l1(I[..., q], ex[..., q])
    ROWS[(program, problem, scheme)] = row


def test_write_norms_table():
    path = os.environ.get("MARA_REGRESSION_TABLE")
    if not path:
        pytest.skip("MARA_REGRESSION_TABLE not set")
    if len(ROWS) != len(RUNS):
        pytest.skip("the table needs every run of this module")
    cols = ["density", "pressure", "velocity1", "magnetic1", "magnetic2", "magnetic3"]

    def f(v):
        return "—" if v is None else f"{v:.3e}"

    lines = ["| program | problem | scheme | steps | t | GPU vs oracle | " + " | ".join(
        f"L1 {c} vs exact" for c in cols) + " | max \\|div B\\| |",
             "|" + "---|" * (6 + len(cols) + 1)]
    for key in RUNS:
        r = ROWS[key]
        lines.append(f"| {key[0]} | {key[1]} | {key[2]} | {r['steps']} | {r['t']:.6g} | "
                     + ("bit-equal" if r["vs_oracle"] == 0.0 else f"{r['vs_oracle']:.1e}") + " | "
                     + " | ".join(f(r.get(c)) for c in cols) + " | "
                     + (f"{r['max_divB']:.1e}" if "max_divB" in r else "—") + " |")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
