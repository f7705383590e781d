"""mara_block_run replaying its steps from a captured HIP graph against the
same runs launched kernel by kernel from the host: every call of a sequence
(each replaying the graph several times, with the host's own transfers and
waits between calls) must give bit-identical primitives, clocks and step
counts.  A run that stops on the clock inside a queued chunk is checked
against the oracle's main loop."""
import numpy as np
import pytest

from mara2_amd import abi, problems
from mara2_amd.runtime import Block

pytestmark = pytest.mark.gpu


def _random_state(desc, seed, amp=0.3):
    P = problems.empty(desc)
    rng = np.random.default_rng(seed)
    shp = P.shape[:3]
    P[..., 0] = 1.0 + 0.5 * rng.random(shp)
    P[..., 4] = 1.0 + 0.5 * rng.random(shp)
    P[..., 1:4] = amp * rng.standard_normal(shp + (3,))
    return problems.apply_bc(desc, P)


def _case(name):
    if name == "turbulence3d":
        d = problems.turbulence_desc(32)
        return d, problems.turbulence_3d(d), 0.3
    if name == "turbulence3d_mixed":
        d = problems.turbulence_desc(32, "mixed")
        return d, problems.turbulence_3d(d), 0.3
    if name == "hydro2d_rk3":
        d = abi.make_desc(law="hydro", recon="plm", riemann="hllc", bc="outflow", cells=(100, 37, 1), rk_order=3)
        return d, _random_state(d, 31), 0.4
    if name == "mhd2d":
        d = problems.mhd_desc(48)
        return d, problems.mhd_2d(d), 0.4
    if name == "sod":
        d = problems.sod_desc(256)
        return d, problems.sod_1d(d), 0.4
    raise KeyError(name)


def _runs(desc, P0, cfl, graphs, calls, steps):
    out = []
    with Block(desc) as blk:
        blk.set_graphs(graphs)
        blk.upload(P0)
        t = 0.0
        for _ in range(calls):
            t, n, rep = blk.run(cfl, 1e300, steps, t=t)
            out.append((t, n, rep.status, blk.download()))
    return out


@pytest.mark.parametrize("name", ["turbulence3d", "turbulence3d_mixed", "hydro2d_rk3", "mhd2d", "sod"])
def test_graph_replays_match_host_launches(name):
    desc, P0, cfl = _case(name)
    # 40 steps per call: two 16-step graph replays and 8 host-launched steps
    eager = _runs(desc, P0, cfl, False, 3, 40)
    graph = _runs(desc, P0, cfl, True, 3, 40)
    for call, (e, g) in enumerate(zip(eager, graph)):
        assert e[:3] == g[:3], (call, e[:3], g[:3])
        assert np.array_equal(e[3], g[3]), f"call {call}: primitives differ"


@pytest.mark.parametrize("rk", [1, 2])
def test_run_stopping_inside_a_chunk_matches_oracle(rk):
    """The clock reaches t_final after a few of the queued steps; the rest are
    no-ops, and the state must be the oracle's after exactly those steps."""
    import oracle
    desc = abi.make_desc(law="hydro", recon="plm", riemann="hllc", bc="outflow", cells=(100, 37, 1), rk_order=rk)
    P0 = _random_state(desc, 32)
    Po = P0.copy()
    err, to, so = oracle.run(desc, Po, 0.4, 0.01, max_steps=40)
    assert err == 0 and so < 40
    for graphs in (False, True):
        with Block(desc) as blk:
            blk.set_graphs(graphs)
            blk.upload(P0)
            tg, sg, rep = blk.run(0.4, 0.01, 40)
            Pg = blk.download()
        assert (sg, tg) == (so, to)
        assert np.array_equal(Pg, Po)
