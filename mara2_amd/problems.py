"""Seeded initial data, in MeshData::P layout with guard zones filled.

Problems of the reference's regression programs (src/Problems.cpp:35-671)
and the synthetic workloads the benchmark and parity tests run:

  sod_1d          1-D Sod tube, 1024 zones, gamma 1.4, PLM + HLLE, RK2, CFL 0.4
  turbulence_3d   3-D decaying turbulence, periodic cube, gamma 5/3, PLM + HLLC, RK2
  mhd_2d          2-D ideal MHD, B = curl(A_z z) by the reference's CT initialisation
  rmhd_3d         3-D relativistic MHD, WENO5 + HLLE, CT, RK3
  riemann_problem one Riemann problem of the serving stream (1-D or 2-D, hydro or MHD)
  regress_mhd_2d  FieldLoop / AbcField of the reference's 2-D MHD regression program

Random fields are sums of Fourier modes with integer wave vectors in a band
kmin <= |k| <= kmax, complex Gaussian amplitudes with a power-law spectrum
and a numpy Generator seed.  All arithmetic here is host-side setup, done once.
"""
import math

import numpy as np

from . import abi

SEED_TURBULENCE = 2026120202
SEED_MHD = 2026120203
SEED_RMHD = 2026120204
SEED_SERVING = 2026120205


# ---------------------------------------------------------------------------
# guard fill (BoundaryCondition::applySimple, src/BoundaryConditions.cpp:8-176)
# ---------------------------------------------------------------------------
def apply_bc(desc, P):
    """Fill the guard zones of P in place, axis by axis, left then right."""
    nq = P.shape[3]
    for axis in range(3):
        g = desc.guards[axis]
        if g <= 0:
            continue
        N = P.shape[axis]
        for side in range(2):
            for n in range(g):
                if desc.bc == abi.BC_PERIODIC:
                    dst, src = (n, N - 2 * g + n) if side == 0 else (N - g + n, g + n)
                elif desc.bc == abi.BC_OUTFLOW:
                    dst, src = (n, g) if side == 0 else (N - 1 - n, N - g - 1)
                else:
                    dst, src = (n, 2 * g - n - 1) if side == 0 else (N - 1 - n, N - 2 * g + n)
                d = [slice(None)] * 4
                s = [slice(None)] * 4
                d[axis] = dst
                s[axis] = src
                P[tuple(d)] = P[tuple(s)]
                if desc.bc == abi.BC_REFLECTING:
                    d[3] = 1 + axis
                    P[tuple(d)] *= -1
    assert nq == P.shape[3]
    return P


def empty(desc):
    return np.zeros(abi.array_shape(desc))


def cell_centres(desc, axis):
    """Cell-centre coordinates along an axis, guards included
    (CartesianMeshGeometry::coordinateAtIndex, src/CartesianMeshGeometry.cpp:61-67)."""
    n, g = desc.cells[axis], desc.guards[axis]
    dx = (desc.upper[axis] - desc.lower[axis]) / n
    ic = np.arange(n + 2 * g) - g
    return desc.lower[axis] + dx * (ic + 0.5)


# ---------------------------------------------------------------------------
# Gaussian random fields on a periodic grid
# ---------------------------------------------------------------------------
def _spectrum(rng, n_axes, kmin, kmax, slope):
    """Random complex amplitudes on the integer modes of the band, stored on
    a small cube of side 2 kmax + 1 (index kmax is k = 0)."""
    m = 2 * kmax + 1
    ks = np.arange(m) - kmax
    grids = np.meshgrid(*([ks] * n_axes), indexing="ij")
    kk = np.sqrt(sum(g.astype(np.float64) ** 2 for g in grids))
    band = (kk >= kmin) & (kk <= kmax)
    amp = np.zeros(kk.shape)
    amp[band] = kk[band] ** (slope / 2.0)
    z = rng.standard_normal(kk.shape) + 1j * rng.standard_normal(kk.shape)
    return z * amp, ks


def _evaluate(spec, ks, n, shifts):
    """Real part of sum_k a_k exp(2 pi i k.(i + s)/n) on an n^d grid, with the
    shift s given in cells (s = 0.5 samples half a cell over)."""
    d = spec.ndim
    full = np.zeros((n,) * d, dtype=np.complex128)
    idx = np.ix_(*([ks % n] * d))
    phase = np.ones(spec.shape, dtype=np.complex128)
    for a in range(d):
        shape = [1] * d
        shape[a] = len(ks)
        phase = phase * np.exp(2j * np.pi * ks.reshape(shape) * shifts[a] / n)
    full[idx] += spec * phase
    field = np.fft.ifftn(full) * full.size
    return field.real.copy()


def random_field(rng, n_axes, n, kmin, kmax, slope=0.0, shifts=None):
    spec, ks = _spectrum(rng, n_axes, kmin, kmax, slope)
    return _evaluate(spec, ks, n, shifts or (0.0,) * n_axes)


# ---------------------------------------------------------------------------
# vector potential -> cell-centred B (src/Problems.cpp:660-667 with
# CellCenteredFieldCT::vectorPotentialToFluxes / correctGodunovFluxes and
# MeshOperator::divergence(G, +1))
# ---------------------------------------------------------------------------
def ct_correct(F):
    """CellCenteredFieldCT::correctGodunovFluxes on F of shape (s0,s1,s2,3,3)
    with magnetic index 0 (src/CellCenteredFieldCT.cpp:14-134)."""
    G = F.copy()
    s0, s1, s2 = F.shape[:3]
    G[..., 0, 0] = 0.0
    G[..., 1, 1] = 0.0
    G[..., 2, 2] = 0.0
    if s0 >= 3:
        G[1:, 1:-1, :, 1, 0] = 0.125 * (2 * F[1:, 1:-1, :, 1, 0] + F[1:, 2:, :, 1, 0] + F[1:, :-2, :, 1, 0]
                                        - F[1:, 2:, :, 0, 1] - F[:-1, 2:, :, 0, 1]
                                        - F[1:, 1:-1, :, 0, 1] - F[:-1, 1:-1, :, 0, 1])
        G[1:, :, 1:-1, 2, 0] = 0.125 * (2 * F[1:, :, 1:-1, 2, 0] + F[1:, :, 2:, 2, 0] + F[1:, :, :-2, 2, 0]
                                        - F[1:, :, 2:, 0, 2] - F[:-1, :, 2:, 0, 2]
                                        - F[1:, :, 1:-1, 0, 2] - F[:-1, :, 1:-1, 0, 2])
    if s1 >= 3:
        G[:, 1:, 1:-1, 2, 1] = 0.125 * (2 * F[:, 1:, 1:-1, 2, 1] + F[:, 1:, 2:, 2, 1] + F[:, 1:, :-2, 2, 1]
                                        - F[:, 1:, 2:, 1, 2] - F[:, :-1, 2:, 1, 2]
                                        - F[:, 1:, 1:-1, 1, 2] - F[:, :-1, 1:-1, 1, 2])
        G[1:-1, 1:, :, 0, 1] = 0.125 * (2 * F[1:-1, 1:, :, 0, 1] + F[2:, 1:, :, 0, 1] + F[:-2, 1:, :, 0, 1]
                                        - F[2:, 1:, :, 1, 0] - F[2:, :-1, :, 1, 0]
                                        - F[1:-1, 1:, :, 1, 0] - F[1:-1, :-1, :, 1, 0])
    if s2 >= 3:
        G[1:-1, :, 1:, 0, 2] = 0.125 * (2 * F[1:-1, :, 1:, 0, 2] + F[2:, :, 1:, 0, 2] + F[:-2, :, 1:, 0, 2]
                                        - F[2:, :, 1:, 2, 0] - F[2:, :, :-1, 2, 0]
                                        - F[1:-1, :, 1:, 2, 0] - F[1:-1, :, :-1, 2, 0])
        G[:, 1:-1, 1:, 1, 2] = 0.125 * (2 * F[:, 1:-1, 1:, 1, 2] + F[:, 2:, 1:, 1, 2] + F[:, :-2, 1:, 1, 2]
                                        - F[:, 2:, 1:, 2, 1] - F[:, 2:, :-1, 2, 1]
                                        - F[:, 1:-1, 1:, 2, 1] - F[:, 1:-1, :-1, 2, 1])
    return G


def divergence(F, factor, dA, dV):
    """MeshOperator::divergence (src/MeshOperator.cpp:323-381), uniform mesh."""
    f = factor / dV
    return f * (+F[1:, :-1, :-1, :, 0] * dA[0] - F[:-1, :-1, :-1, :, 0] * dA[0]
                + F[:-1, 1:, :-1, :, 1] * dA[1] - F[:-1, :-1, :-1, :, 1] * dA[1]
                + F[:-1, :-1, 1:, :, 2] * dA[2] - F[:-1, :-1, :-1, :, 2] * dA[2])


def spacing(desc):
    dx = [(desc.upper[a] - desc.lower[a]) / desc.cells[a] for a in range(3)]
    dA = (dx[1] * dx[2], dx[2] * dx[0], dx[0] * dx[1])
    return dx, dA, dx[0] * dx[1] * dx[2]


def vector_potential_to_B(desc, A):
    """A: (N1+1, N2+1, N3+1, 3, 3) sampled like MeshOperator::generate(face,
    scalars); returns cell B with guards, shape (N1, N2, N3, 3)."""
    F = np.zeros(A.shape)
    F[..., 1, 0] = -A[..., 2, 0]
    F[..., 2, 0] = +A[..., 1, 0]
    F[..., 2, 1] = -A[..., 0, 1]
    F[..., 0, 1] = +A[..., 2, 1]
    F[..., 0, 2] = -A[..., 1, 2]
    F[..., 1, 2] = +A[..., 0, 2]
    G = ct_correct(F)
    _, dA, dV = spacing(desc)
    return divergence(G, 1.0, dA, dV)


def _face_samples(desc, field_at):
    """Build A (N1+1, N2+1, N3+1, 3, 3): component q on the faces of axis a is
    field_at(q, a) -> periodic array on the interior grid sampled at the face
    centre; guard faces wrap periodically."""
    N = [desc.cells[a] + 2 * desc.guards[a] for a in range(3)]
    A = np.zeros((N[0] + 1, N[1] + 1, N[2] + 1, 3, 3))
    for a in range(3):
        for q in range(3):
            vals = field_at(q, a)
            if vals is None:
                continue
            vals = vals.reshape(desc.cells[0], desc.cells[1], desc.cells[2])
            idx = [(np.arange(N[e] + 1) - desc.guards[e]) % desc.cells[e] for e in range(3)]
            A[..., q, a] = vals[np.ix_(*idx)]
    return A


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------
def sod_desc(n=1024, precision="fp64"):
    return abi.make_desc(law="hydro", gamma=1.4, recon="plm", plm_theta=1.5, riemann="hlle",
                         bc="outflow", cells=(n, 1, 1), rk_order=2, precision=precision)


def sod_1d(desc):
    """Sod tube: x < 0.5 (rho, v, p) = (1, 0, 1), else (0.125, 0, 0.1)."""
    P = empty(desc)
    x = cell_centres(desc, 0)
    P[:, 0, 0, 0] = np.where(x < 0.5, 1.0, 0.125)
    P[:, 0, 0, 4] = np.where(x < 0.5, 1.0, 0.1)
    return apply_bc(desc, P)


def turbulence_desc(n=256, precision="fp64"):
    return abi.make_desc(law="hydro", gamma=5.0 / 3.0, recon="plm", plm_theta=1.5, riemann="hllc",
                         bc="periodic", cells=(n, n, n), rk_order=2, precision=precision)


def turbulence_3d(desc, seed=SEED_TURBULENCE, kmin=1, kmax=16, slope=-4.0, mach=1.0):
    """rho = 1, p = 1; each velocity component a Gaussian random field on
    kmin <= |k| <= kmax with |v_k|^2 ~ k^slope; rescaled to rms Mach `mach`."""
    n = desc.cells[0]
    rng = np.random.default_rng(seed)
    v = np.stack([random_field(rng, 3, n, kmin, kmax, slope) for _ in range(3)], axis=-1)
    cs = np.sqrt(desc.gamma * 1.0 / 1.0)
    v *= mach * cs / np.sqrt(np.mean(np.sum(v * v, axis=-1)))
    P = empty(desc)
    g = desc.guards[0]
    inner = (slice(g, g + n),) * 3
    P[..., 0] = 1.0
    P[..., 4] = 1.0
    P[inner + (slice(1, 4),)] = v
    return apply_bc(desc, P)


def mhd_desc(n=2048, precision="fp64"):
    return abi.make_desc(law="mhd", gamma=5.0 / 3.0, recon="plm", plm_theta=1.5, riemann="hlle",
                         bc="periodic", cells=(n, n, 1), rk_order=2, field_ct=1, precision=precision)


def mhd_2d(desc, seed=SEED_MHD, kmax=8, beta=1.0, vrms=0.5):
    """rho = 1, p = 1; B = curl(A_z z) via the reference's CT initialisation,
    A_z a Gaussian random field on 1 <= |k| <= kmax scaled to plasma beta
    2 p / <B^2> = beta; in-plane v random on the same band with rms |v| = vrms."""
    n = desc.cells[0]
    rng = np.random.default_rng(seed)
    spec, ks = _spectrum(rng, 2, 1, kmax, 0.0)

    def field_at(q, a):
        if q != 2 or a == 2:
            return None
        shifts = (0.0, 0.5) if a == 0 else (0.5, 0.0)  # x-faces / y-faces of a cell
        return _evaluate(spec, ks, n, shifts)

    A = _face_samples(desc, field_at)
    B = vector_potential_to_B(desc, A)
    g = desc.guards
    inner = (slice(g[0], g[0] + n), slice(g[1], g[1] + n), slice(0, 1))
    scale = np.sqrt(2.0 * 1.0 / beta / np.mean(np.sum(B[inner] ** 2, axis=-1)))
    B = vector_potential_to_B(desc, A * scale)
    v = np.stack([random_field(rng, 2, n, 1, kmax, 0.0) for _ in range(2)], axis=-1)
    v *= vrms / np.sqrt(np.mean(np.sum(v * v, axis=-1)))
    P = empty(desc)
    P[..., 0] = 1.0
    P[..., 4] = 1.0
    P[..., 5:8] = B
    P[inner[0], inner[1], 0, 1:3] = v
    return apply_bc(desc, P)


def rmhd_desc(n=128, precision="fp64"):
    return abi.make_desc(law="rmhd", gamma=4.0 / 3.0, recon="weno5", riemann="hlle", bc="periodic",
                         cells=(n, n, n), rk_order=3, field_ct=1, precision=precision)


def _solenoidal(specs, ks):
    """Project the complex amplitudes of a vector field onto the plane normal
    to each wave vector (k . a_k = 0), so its real part is divergence-free."""
    kx, ky, kz = np.meshgrid(ks, ks, ks, indexing="ij")
    kv = np.stack([kx, ky, kz]).astype(np.float64)
    k2 = np.sum(kv * kv, axis=0)
    k2[k2 == 0] = 1.0
    a = np.stack(specs)
    proj = np.sum(kv * a, axis=0) / k2
    return [a[c] - kv[c] * proj for c in range(3)]


def rmhd_3d(desc, seed=SEED_RMHD, kmax=8, wmax=5.0, sigma=1.0, u_slope=0.0, a_slope=0.0, solenoidal=False):
    """rho = 1, p = 0.1; spatial four-velocity u random on 1 <= |k| <= kmax
    with power spectrum |u_k|^2 ~ k^u_slope (flat by default), optionally
    projected to be divergence-free, scaled so max W = sqrt(1 + u^2) = wmax,
    v = u / W; B = curl A with A a random vector potential on the same band
    (|A_k|^2 ~ k^a_slope), scaled to <B^2> / rho = sigma."""
    n = desc.cells[0]
    rng = np.random.default_rng(seed)
    uspecs = [_spectrum(rng, 3, 1, kmax, u_slope) for _ in range(3)]
    ks_u = uspecs[0][1]
    ua = [sp for sp, _ in uspecs]
    if solenoidal:
        ua = _solenoidal(ua, ks_u)
    u = np.stack([_evaluate(a, ks_u, n, (0.0, 0.0, 0.0)) for a in ua], axis=-1)
    u *= np.sqrt(wmax * wmax - 1.0) / np.sqrt(np.max(np.sum(u * u, axis=-1)))
    W = np.sqrt(1.0 + np.sum(u * u, axis=-1))
    v = u / W[..., None]
    specs = [_spectrum(rng, 3, 1, kmax, a_slope) for _ in range(3)]
    face_shift = {0: (0.0, 0.5, 0.5), 1: (0.5, 0.0, 0.5), 2: (0.5, 0.5, 0.0)}

    def field_at(q, a):
        spec, ks = specs[q]
        return _evaluate(spec, ks, n, face_shift[a])

    A = _face_samples(desc, field_at)
    B = vector_potential_to_B(desc, A)
    g = desc.guards
    inner = tuple(slice(g[a], g[a] + n) for a in range(3))
    scale = np.sqrt(sigma * 1.0 / np.mean(np.sum(B[inner] ** 2, axis=-1)))
    B = vector_potential_to_B(desc, A * scale)
    P = empty(desc)
    P[..., 0] = 1.0
    P[..., 4] = 0.1
    P[..., 5:8] = B
    P[inner + (slice(1, 4),)] = v
    return apply_bc(desc, P)


# ---------------------------------------------------------------------------
# serving stream: independent Riemann problems
# ---------------------------------------------------------------------------
def _log_uniform(rng, lo, hi, size=None):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size))


def riemann_params(rng):
    """One request of the serving stream: 70% 1-D (zones log-uniform in
    256..16384), 30% 2-D (side log-uniform in 128..1024, discontinuity on a
    line through the centre at a random angle); left/right rho and p
    log-uniform in [0.1, 10], velocities uniform in [-1, 1]; half MHD with a
    uniform B (plasma beta log-uniform in [0.1, 10]) plus a random jump of
    the components transverse to the discontinuity.  Every request draws the
    same number of variates, so the stream does not depend on its mix."""
    dim = 1 if rng.random() < 0.7 else 2
    u = rng.random()
    if dim == 1:
        n = int(round(np.exp(np.log(256) + u * (np.log(16384) - np.log(256)))))
    else:
        n = int(round(np.exp(np.log(128) + u * (np.log(1024) - np.log(128)))))
    law = "mhd" if rng.random() < 0.5 else "hydro"
    rho = _log_uniform(rng, 0.1, 10.0, 2)
    pre = _log_uniform(rng, 0.1, 10.0, 2)
    vel = rng.uniform(-1.0, 1.0, (2, 3))
    angle = rng.uniform(0.0, np.pi) if dim == 2 else 0.0
    if dim == 1:
        rng.uniform(0.0, np.pi)
    beta = _log_uniform(rng, 0.1, 10.0)
    bdir = rng.standard_normal(3)
    jump = rng.standard_normal(3)
    jump_size = rng.random()
    nhat = np.array([np.cos(angle), np.sin(angle), 0.0])
    bmag = np.sqrt(2.0 * 0.5 * (pre[0] + pre[1]) / beta)
    B0 = bmag * bdir / np.linalg.norm(bdir)
    jump = jump - np.dot(jump, nhat) * nhat
    jump *= jump_size * bmag / max(np.linalg.norm(jump), 1e-300)
    BL, BR = B0 - 0.5 * jump, B0 + 0.5 * jump
    return dict(dim=dim, n=n, law=law, rho=rho, p=pre, v=vel, angle=angle, beta=beta,
                B=np.stack([BL, BR]) if law == "mhd" else None)


# NewtonianMHD::setPressureFloor for the stream's MHD solves: p = 1e-6 rho
# where the recovered pressure goes negative (src/ConservationLaws.cpp:394-400).
# Without it the reference scheme stops with an inversion failure on some of
# the stream's strong rarefactions; NewtonianHydro ignores the floor.
RIEMANN_PRESSURE_FLOOR = 1e-6


def riemann_desc(params, precision="mixed", rk_order=2):
    """PLM (theta 1.5) + HLLC for hydro, HLLE (flux-CT in 2-D) for MHD, gamma
    1.4, outflow boundaries on the unit interval or square."""
    cells = (params["n"], 1, 1) if params["dim"] == 1 else (params["n"], params["n"], 1)
    mhd = params["law"] == "mhd"
    return abi.make_desc(law=params["law"], gamma=1.4, pressure_floor=RIEMANN_PRESSURE_FLOOR,
                         recon="plm", plm_theta=1.5,
                         riemann="hlle" if mhd else "hllc", bc="outflow", cells=cells,
                         rk_order=rk_order, field_ct=1 if (mhd and params["dim"] == 2) else 0,
                         precision=precision)


def _vector_potential_z_to_B(desc, Ax, Ay):
    """In-plane cell B of a 2-D mesh from A_z sampled on x faces (Ax, shape
    (N1+1, N2+1)) and y faces (Ay): the z-potential terms of
    vector_potential_to_B with the same arithmetic, on 2-D arrays."""
    Fyx = -Ax          # F[..., 1, 0]
    Fxy = +Ay          # F[..., 0, 1]
    Gyx = Fyx.copy()
    Gxy = Fxy.copy()
    Gyx[1:, 1:-1] = 0.125 * (2 * Fyx[1:, 1:-1] + Fyx[1:, 2:] + Fyx[1:, :-2]
                             - Fxy[1:, 2:] - Fxy[:-1, 2:]
                             - Fxy[1:, 1:-1] - Fxy[:-1, 1:-1])
    Gxy[1:-1, 1:] = 0.125 * (2 * Fxy[1:-1, 1:] + Fxy[2:, 1:] + Fxy[:-2, 1:]
                             - Fyx[2:, 1:] - Fyx[2:, :-1]
                             - Fyx[1:-1, 1:] - Fyx[1:-1, :-1])
    _, dA, dV = spacing(desc)
    f = 1.0 / dV
    # components (0, 1) of divergence(G, 1, dA, dV) with G[..., 0, 0] = G[..., 1, 1] = 0
    Bx = f * (+0.0 * dA[0] - 0.0 * dA[0] + Gxy[:-1, 1:] * dA[1] - Gxy[:-1, :-1] * dA[1] + 0.0 * dA[2] - 0.0 * dA[2])
    By = f * (+Gyx[1:, :-1] * dA[0] - Gyx[:-1, :-1] * dA[0] + 0.0 * dA[1] - 0.0 * dA[1] + 0.0 * dA[2] - 0.0 * dA[2])
    return Bx, By


def riemann_problem(params, desc):
    """Initial primitives (guards filled) of one serving request."""
    P = empty(desc)
    rho, pre, vel = params["rho"], params["p"], params["v"]
    if params["dim"] == 1:
        x = cell_centres(desc, 0)
        left = (x < 0.5)[:, None, None]
    else:
        th = params["angle"]
        x = cell_centres(desc, 0)[:, None]
        y = cell_centres(desc, 1)[None, :]
        left = ((x - 0.5) * np.cos(th) + (y - 0.5) * np.sin(th) < 0.0)[:, :, None]
    P[..., 0] = np.where(left, rho[0], rho[1])
    for c in range(3):
        P[..., 1 + c] = np.where(left, vel[0, c], vel[1, c])
    P[..., 4] = np.where(left, pre[0], pre[1])
    if params["law"] == "mhd":
        BL, BR = params["B"]
        if params["dim"] == 1:
            for c in range(3):
                P[..., 5 + c] = np.where(left, BL[c], BR[c])
        else:
            # A_z = Bn u - Bt(s) s with s, u the coordinates along the normal
            # and the tangent of the discontinuity: curl(A_z z) is B0 with a
            # jump of the tangential component, divergence-free on the CT stencil
            th = params["angle"]
            n_, t_ = np.array([np.cos(th), np.sin(th)]), np.array([-np.sin(th), np.cos(th)])
            Bn = BL[0] * n_[0] + BL[1] * n_[1]
            BtL, BtR = BL[0] * t_[0] + BL[1] * t_[1], BR[0] * t_[0] + BR[1] * t_[1]
            g = desc.guards
            dx, _, _ = spacing(desc)
            N0, N1 = desc.cells[0] + 2 * g[0], desc.cells[1] + 2 * g[1]
            xf = desc.lower[0] + dx[0] * (np.arange(N0 + 1) - g[0]) - 0.5
            yf = desc.lower[1] + dx[1] * (np.arange(N1 + 1) - g[1]) - 0.5
            xc = xf + 0.5 * dx[0]
            yc = yf + 0.5 * dx[1]

            def Az(X, Y):
                s = X * n_[0] + Y * n_[1]
                u = X * t_[0] + Y * t_[1]
                return Bn * u - np.where(s < 0.0, BtL, BtR) * s

            Ax = Az(xf[:, None], yc[None, :])   # x faces: x on the face, y at the centre
            Ay = Az(xc[:, None], yf[None, :])   # y faces
            Bx, By = _vector_potential_z_to_B(desc, Ax, Ay)
            P[..., 0, 5] = Bx
            P[..., 0, 6] = By
            P[..., 7] = np.where(left, BL[2], BR[2])
    return apply_bc(desc, P)


# ---------------------------------------------------------------------------
# 2-D MHD regression problems with a vector potential
# (NewtonianMHD2DTestProgram, src/Problems.cpp:480-672)
# ---------------------------------------------------------------------------
def _generate_cells(desc, f):
    """MeshOperator::generate(F, cell) (src/MeshOperator.cpp:168-189): f at
    every cell centre, guards included; f is evaluated with the host libm."""
    xs = [cell_centres(desc, a) for a in range(3)]
    out = np.empty(abi.array_shape(desc))
    for i, x in enumerate(xs[0]):
        for j, y in enumerate(xs[1]):
            for k, z in enumerate(xs[2]):
                out[i, j, k, :] = f(float(x), float(y), float(z))
    return out


def _generate_faces(desc, f):
    """MeshOperator::generate(F, face, scalars) (src/MeshOperator.cpp:119-166):
    the 3 components of f on the faces of each axis, at
    coordinateAtIndex(ic - 1/2, jc, kc) and its permutations, guards included.
    Returns A of shape (N1+1, N2+1, N3+1, 3, 3) [..., component, axis]."""
    N = [desc.cells[a] + 2 * desc.guards[a] for a in range(3)]
    dx = [(desc.upper[a] - desc.lower[a]) / desc.cells[a] for a in range(3)]
    A = np.empty((N[0] + 1, N[1] + 1, N[2] + 1, 3, 3))

    def coord(a, i):
        return desc.lower[a] + dx[a] * (i + 0.5)

    for i in range(N[0] + 1):
        ic = i - desc.guards[0]
        for j in range(N[1] + 1):
            jc = j - desc.guards[1]
            for k in range(N[2] + 1):
                kc = k - desc.guards[2]
                A[i, j, k, :, 0] = f(coord(0, ic - 0.5), coord(1, jc), coord(2, kc))
                A[i, j, k, :, 1] = f(coord(0, ic), coord(1, jc - 0.5), coord(2, kc))
                A[i, j, k, :, 2] = f(coord(0, ic), coord(1, jc), coord(2, kc - 0.5))
    return A


def _field_loop_p(x, y, z):
    x -= 0.5
    y -= 0.5
    B, k = 1e-5, 10.0
    R = math.sqrt(x * x + y * y)
    bf = B * 3 * k * math.pow(k * R, 2) * math.exp(-math.pow(k * R, 3))
    return (1.0, 1.0, 1.0, 0.0, 1.0, bf * (-y / R), bf * (x / R), 0.0)


def _field_loop_a(x, y, z):
    x -= 0.5
    y -= 0.5
    B, k = 1e-5, 10.0
    R = math.sqrt(x * x + y * y)
    return (0.0, 0.0, B * math.exp(-math.pow(k * R, 3)))


_ABC = (1.0, 1.0, 0.0, 2 * math.pi)  # A, B, C, k


def _abc_field_p(x, y, z):
    A, B, C, k = _ABC
    bx = C * math.cos(k * z) - B * math.sin(k * y)
    by = A * math.cos(k * x) - C * math.sin(k * z)
    bz = B * math.cos(k * y) - A * math.sin(k * x)
    return (1.0, 0.0, 0.0, 0.0, 1.0, bx, by, bz)


def _abc_field_a(x, y, z):
    A, B, C, k = _ABC
    bx = C * math.cos(k * z) - B * math.sin(k * y)
    by = A * math.cos(k * x) - C * math.sin(k * z)
    bz = B * math.cos(k * y) - A * math.sin(k * x)
    return (bx / k, by / k, bz / k)


REGRESS_MHD_2D = {
    # name: (final time, primitive function, vector potential)
    "FieldLoop": (1.0, _field_loop_p, _field_loop_a),
    "AbcField": (0.1, _abc_field_p, _abc_field_a),
}
REGRESS_MHD_2D_CFL = 0.5


def regress_mhd_desc(n=64, precision="fp64"):
    """The program's mesh and scheme: n x n cells on the unit square, 2 guards,
    periodic, NewtonianMHD (gamma 5/3), PLM + HLLE with flux-CT, RK2
    (src/Problems.cpp:569-611)."""
    return abi.make_desc(law="mhd", gamma=5.0 / 3.0, recon="plm", riemann="hlle", bc="periodic",
                         cells=(n, n, 1), rk_order=2, field_ct=1, precision=precision)


def regress_mhd_2d(name, desc):
    """Initial MeshData::P of a vector-potential problem as runProblem builds
    it (src/Problems.cpp:655-667): primitives at cell centres, then the cell
    B of the whole array from A on faces -> vectorPotentialToFluxes -> CT ->
    divergence (factor +1), then the boundary condition.  Returns
    (P, final time)."""
    t_final, fp, fa = REGRESS_MHD_2D[name]
    P = _generate_cells(desc, fp)
    A = _generate_faces(desc, fa)
    P[..., 5:8] = vector_potential_to_B(desc, A)
    return apply_bc(desc, P), t_final
