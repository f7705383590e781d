"""Writes tests/golden/ref_vectors.npz: inputs and outputs of the reference's
own Reconstruction, QuarticPolynomial, ConsToPrimRMHD and UnitVector, called
through oracle/_ref/libmara_ref.so (built by oracle/Makefile from the
reference sources in place).  Run in a container that has the reference
tree:  python tests/golden/make_golden.py
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from mara2_amd import abi  # noqa: E402

dp = ctypes.POINTER(ctypes.c_double)


def main():
    oracle.build()
    R = oracle.ref()
    if R is None:
        raise SystemExit("oracle/_ref/libmara_ref.so is not built (reference tree absent)")
    rng = np.random.default_rng(20261202)
    out = {}

    uv = np.zeros(9)
    R.ref_unit_vectors(uv.ctypes.data_as(dp))
    out["unit_vectors"] = uv.reshape(3, 3)

    # PLM: random stencils spanning magnitudes, three limiter parameters
    n = 4000
    v = rng.standard_normal((n, 3)) * np.exp(rng.uniform(-6, 6, (n, 1)))
    v[: n // 10, 1] = v[: n // 10, 0]                            # flat pieces
    v[n // 10: n // 5, 2] = 2 * v[n // 10: n // 5, 1] - v[n // 10: n // 5, 0]  # linear pieces
    theta = rng.choice([1.0, 1.5, 2.0], n)
    sgn = rng.choice([-1.0, 1.0], n)
    plm = np.array([R.ref_plm(np.ascontiguousarray(v[i]).ctypes.data_as(dp), theta[i], sgn[i]) for i in range(n)])
    out.update(plm_v=v, plm_theta=theta, plm_sgn=sgn, plm_out=plm)

    # WENO5: six operations x three smoothness indicators
    m = 1500
    w = rng.standard_normal((m, 5)) * np.exp(rng.uniform(-4, 4, (m, 1)))
    w[: m // 5] = np.cumsum(rng.uniform(0, 1, (m // 5, 5)), axis=1)  # smooth monotone
    res = np.zeros((m, 6, 3))
    for i in range(m):
        wi = np.ascontiguousarray(w[i])
        for op in range(6):
            for mode in range(3):
                res[i, op, mode] = R.ref_weno5(wi.ctypes.data_as(dp), op, mode, 50.0)
    out.update(weno_v=w, weno_out=res)

    # quartic: random coefficients, plus the RMHD characteristic quartics of random states
    k = 3000
    d = rng.standard_normal((k, 5)) * np.exp(rng.uniform(-3, 3, (k, 5)))
    roots = np.full((k, 4), np.nan)
    nr = np.zeros(k, dtype=np.int64)
    for i in range(k):
        r = np.zeros(4)
        di = np.ascontiguousarray(d[i])
        nr[i] = R.ref_quartic(di.ctypes.data_as(dp), r.ctypes.data_as(dp))
        roots[i, :nr[i]] = r[:nr[i]]
    out.update(quartic_d=d, quartic_nr=nr, quartic_roots=roots)

    # ConsToPrimRMHD: conserved states of random primitive states, three gammas
    c = 3000
    P = np.zeros((c, 8))
    P[:, 0] = np.exp(rng.uniform(-3, 3, c))
    vdir = rng.standard_normal((c, 3))
    vdir /= np.linalg.norm(vdir, axis=1, keepdims=True)
    W = np.exp(rng.uniform(0, np.log(10.0), c))
    P[:, 1:4] = vdir * np.sqrt(1 - 1 / W ** 2)[:, None]
    P[:, 4] = P[:, 0] * np.exp(rng.uniform(-6, 3, c))
    P[:, 5:8] = rng.standard_normal((c, 3)) * np.exp(rng.uniform(-3, 2, (c, 1)))
    gam = rng.choice([4 / 3, 5 / 3, 1.4], c)
    U = np.zeros((c, 8))
    Pr = np.full((c, 8), np.nan)
    err = np.zeros(c, dtype=np.int64)
    which = np.zeros(c, dtype=np.int64)
    for i in range(c):
        e, Ui, _, _ = oracle.from_primitive(abi.LAW_RMHD, gam[i], P[i], [1.0, 0.0, 0.0])
        U[i] = Ui
        Po = np.zeros(8)
        wv = ctypes.c_int(0)
        err[i] = R.ref_rmhd_c2p(np.ascontiguousarray(Ui).ctypes.data_as(dp), gam[i], -1.0,
                                Po.ctypes.data_as(dp), ctypes.byref(wv))
        which[i] = wv.value
        if err[i] == 0:
            Pr[i] = Po
    out.update(c2p_U=U, c2p_gamma=gam, c2p_err=err, c2p_which=which, c2p_P=Pr, c2p_Pin=P)
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_vectors.npz"))


if __name__ == "__main__":
    main()
