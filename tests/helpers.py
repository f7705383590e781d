"""Shared helpers for the parity tests: run the HIP path and the oracle on
identical inputs and compare."""
import numpy as np

from mara2_amd import abi


def interior(desc, P):
    g = desc.guards
    return P[g[0]:P.shape[0] - g[0], g[1]:P.shape[1] - g[1], g[2]:P.shape[2] - g[2], :]


def rel_l1(a, b):
    """max over fields of sum|a - b| / sum|b| (fields with sum|b| == 0 use the
    absolute sum)."""
    worst = 0.0
    for q in range(a.shape[-1]):
        num = np.sum(np.abs(a[..., q] - b[..., q]))
        den = np.sum(np.abs(b[..., q]))
        worst = max(worst, num / den if den > 0 else num)
    return worst


def prim_to_cons(desc, P):
    """Vectorised U(P) for the hydro and MHD laws (the U part of
    fromPrimitive, src/ConservationLaws.cpp:253-290, 416-487)."""
    g = desc.gamma
    U = np.empty_like(P)
    vv = P[..., 1] ** 2 + P[..., 2] ** 2 + P[..., 3] ** 2
    U[..., 0] = P[..., 0]
    U[..., 1:4] = P[..., 0:1] * P[..., 1:4]
    if desc.law == abi.LAW_HYDRO:
        U[..., 4] = P[..., 0] * 0.5 * vv + P[..., 4] / (g - 1)
    else:
        BB = P[..., 5] ** 2 + P[..., 6] ** 2 + P[..., 7] ** 2
        U[..., 4] = P[..., 0] * 0.5 * vv + 0.5 * BB + P[..., 4] / (g - 1)
        U[..., 5:8] = P[..., 5:8]
    return U


def run_pair(desc, P0, steps, cfl, block_cls):
    """Advance `steps` steps of dt = cfl * courant on both the GPU block and
    the oracle; return (P_gpu, P_oracle, dts_gpu, dts_oracle)."""
    import oracle
    Po = P0.copy()
    dts_g, dts_o = [], []
    with block_cls(desc) as blk:
        blk.upload(P0)
        for _ in range(steps):
            dg = cfl * blk.courant_dt()
            err, do = oracle.courant_dt(desc, Po)
            assert err == 0
            do = cfl * do
            dts_g.append(dg)
            dts_o.append(do)
            blk.advance(dg)
            err, rep = oracle.advance(desc, Po, 0.0, do)
            assert err == 0, rep.as_dict()
        Pg = blk.download()
    return Pg, Po, dts_g, dts_o
