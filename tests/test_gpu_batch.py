"""Batched solves (mara_batch_*): every solve of a mix gives the same result
alone and in the mix, bit for bit, and the FP64 batch reproduces the single
block and the oracle bit for bit.  Requests are the serving stream's own
(seeded Riemann problems, 1-D and 2-D, hydro and MHD) with the zone counts
scaled down so the oracle finishes in seconds."""
import numpy as np
import pytest

from mara2_amd import abi, problems, serving
from mara2_amd.runtime import Batch, Block
from helpers import interior, rel_l1

pytestmark = pytest.mark.gpu


def SMALL(dim, n):
    """1-D 256..16384 -> 10..682 zones, 2-D 128..1024 -> 12..102 per side."""
    return max(10, n // 24) if dim == 1 else max(12, n // 10)


N_REQ = 10
BUDGET = 1 << 30


def _stream(precision):
    reqs = serving.make_stream(N_REQ, precision=precision, resize=SMALL)
    kinds = {(r.params["dim"], r.params["law"]) for r in reqs}
    assert kinds == {(1, "hydro"), (1, "mhd"), (2, "hydro"), (2, "mhd")}, kinds
    return reqs


def _run_batch(reqs, precision, chunk=16, max_solves=64, graphs=True, shape=None):
    """Admit every request at once and step until all are done."""
    out = {}
    with Batch(BUDGET, rk_order=2, precision=precision, max_solves=max_solves) as b:
        b.set_graphs(graphs)
        if shape is not None:
            b.set_shape(shape)
            assert b.get_shape().as_dict() == shape.as_dict()
        slots = {}
        for r in reqs:
            s = b.admit(r.desc, r.P, serving.CFL, serving.T_FINAL, serving.MAX_STEPS)
            assert s is not None
            slots[s] = r
        while slots:
            b.step(chunk)
            for s in b.finished():
                done, t, steps, st, rep = b.query(s)
                assert done and st == 0, rep.as_dict()
                out[slots.pop(s).index] = (b.retire(s), t, steps)
        assert b.free_bytes == BUDGET
    return out


@pytest.fixture(scope="module")
def fp64_stream():
    return _stream("fp64")


@pytest.fixture(scope="module")
def fp64_mix(fp64_stream):
    return _run_batch(fp64_stream, "fp64")


def test_fp64_mix_equals_alone(fp64_stream, fp64_mix):
    for r in fp64_stream:
        alone = _run_batch([r], "fp64")[r.index]
        P, t, steps = fp64_mix[r.index]
        assert (t, steps) == alone[1:], r.index
        assert np.array_equal(P, alone[0]), r.index


def test_fp64_batch_equals_block(fp64_stream, fp64_mix):
    for r in fp64_stream:
        with Block(r.desc) as blk:
            blk.upload(r.P)
            t, steps, _ = blk.run(serving.CFL, serving.T_FINAL, serving.MAX_STEPS)
            Pb = blk.download()
        P, tb, sb = fp64_mix[r.index]
        assert (tb, sb) == (t, steps), r.index
        assert np.array_equal(P, Pb), (r.index, r.params["dim"], r.params["law"])


def test_fp64_batch_equals_oracle(fp64_stream, fp64_mix):
    import oracle
    for r in fp64_stream:
        Po = r.P.copy()
        err, t, steps = oracle.run(r.desc, Po, serving.CFL, serving.T_FINAL)
        assert err == 0
        P, tb, sb = fp64_mix[r.index]
        assert (tb, sb) == (t, steps), r.index
        assert np.array_equal(P, Po), r.index


def test_mixed_mix_equals_alone_and_close_to_fp64(fp64_mix):
    """Mixed form: alone and in the mix bit for bit; against the FP64 result
    of the same request, relative L1 of the interior primitives <= 2e-4 (the
    stream's discontinuities move a few binary32 roundings into shock
    positions; the smooth-flow bound of the mixed form is 1e-5)."""
    reqs = _stream("mixed")
    mix = _run_batch(reqs, "mixed", chunk=7)
    worst = 0.0
    for r in reqs:
        alone = _run_batch([r], "mixed", chunk=16)[r.index]
        P, t, steps = mix[r.index]
        assert (t, steps) == alone[1:], r.index
        assert np.array_equal(P, alone[0]), r.index
        P64 = fp64_mix[r.index][0]
        worst = max(worst, rel_l1(interior(r.desc, P), interior(r.desc, P64)))
    print(f"mixed vs fp64 worst rel L1 {worst:.3e}")
    assert worst <= 2e-4


def test_chunking_does_not_change_results(fp64_stream, fp64_mix):
    res = _run_batch(fp64_stream[:4], "fp64", chunk=1)
    for i, (P, t, steps) in res.items():
        assert (t, steps) == fp64_mix[i][1:]
        assert np.array_equal(P, fp64_mix[i][0])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_rounds_match_host_launches(fp64_stream, fp64_mix, precision):
    """Rounds replayed from captured graphs (recaptured as solves finish and
    leave the table) against every kernel launched from the host."""
    if precision == "fp64":
        reqs, ref = fp64_stream, fp64_mix
    else:
        reqs = _stream("mixed")
        ref = _run_batch(reqs, "mixed", graphs=True)
    eager = _run_batch(reqs, precision, graphs=False)
    assert eager.keys() == ref.keys()
    for i, (P, t, steps) in eager.items():
        assert (t, steps) == ref[i][1:], i
        assert np.array_equal(P, ref[i][0]), i


@pytest.mark.parametrize("pencil,rows", [(128, 1), (512, 4)])
def test_launch_shape_does_not_change_results(fp64_stream, fp64_mix, pencil, rows):
    """Pencil length, fused 3-D z planes and 2-D MHD tile height of a batch."""
    res = _run_batch(fp64_stream, "fp64", shape=abi.make_shape(pencil=pencil, zplanes=8, rows=rows))
    for i, (P, t, steps) in res.items():
        assert (t, steps) == fp64_mix[i][1:], i
        assert np.array_equal(P, fp64_mix[i][0]), i


def test_admission_by_budget_and_slot_reuse(fp64_stream):
    """A pool that holds only some solves: admit returns None when full,
    retire gives memory and slots back, and reused slots start clean."""
    reqs = fp64_stream
    sizes = [Batch.item_bytes_of(r.desc, "fp64") for r in reqs]
    budget = max(sizes) + min(sizes) + 4096
    want = _run_batch(reqs, "fp64")
    with Batch(budget, rk_order=2, precision="fp64", max_solves=2) as b:
        queue = list(reqs)
        slots = {}
        refused = 0
        while queue or slots:
            keep = []
            for r in queue:
                s = b.admit(r.desc, r.P, serving.CFL, serving.T_FINAL, serving.MAX_STEPS)
                if s is None:
                    keep.append(r)
                    refused += 1
                else:
                    slots[s] = r
            queue = keep
            assert slots
            b.step(32)
            for s in b.finished():
                r = slots.pop(s)
                P = b.retire(s)
                assert np.array_equal(P, want[r.index][0]), r.index
        assert refused > 0
        assert b.free_bytes == budget


def test_admit_rejects_other_rk_or_precision():
    r = serving.make_stream(1, precision="fp64", resize=SMALL)[0]
    with Batch(BUDGET, rk_order=3, precision="fp64") as b:
        with pytest.raises(Exception) as e:
            b.admit(r.desc, r.P, serving.CFL, serving.T_FINAL, 10)
        assert e.value.status == abi.ERR_UNSUPPORTED


def test_serve_loop_continuous_batching():
    """The serving loop on a short stream with staggered arrivals: every
    request completes with the same result as alone."""
    reqs = serving.make_stream(8, precision="mixed", resize=SMALL)
    serving.set_arrivals(reqs, rate=200.0)
    wall, done = serving.serve(reqs, BUDGET, max_solves=16, chunk=8, keep_results=True)
    assert len(done) == len(reqs)
    serving.check_errors(done)
    ref = _run_batch(reqs, "mixed")
    for r in done:
        assert r.finished >= r.admitted >= r.arrival
        assert np.array_equal(r.result, ref[r.index][0]), r.index
    st = serving.latency_stats(done)
    assert st["p99_s"] >= st["p50_s"] > 0


def _alone_vs_block(desc, P, steps=40, cfl=0.4):
    with Block(desc) as blk:
        blk.upload(P)
        t, n, _ = blk.run(cfl, 1e300, steps)
        return blk.download(), t, n


def test_two_lanes_with_a_large_tile_solve():
    """A 256^2 hydro solve makes the tile lane's steps slow, so the pencil
    lane takes several steps per batch step: every 1-D solve still ends
    bitwise equal to its single-block run."""
    prm = serving.make_stream(12, precision="fp64", resize=SMALL)
    ones = [r for r in prm if r.params["dim"] == 1][:3]
    assert len(ones) == 3
    rng = np.random.default_rng(5)
    params = problems.riemann_params(rng)
    params.update(dim=2, n=256, law="hydro")
    big = problems.riemann_desc(params, precision="fp64")
    Pbig = problems.riemann_problem(params, big)
    with Batch(BUDGET, rk_order=2, precision="fp64", max_solves=8) as b:
        sb = b.admit(big, Pbig, serving.CFL, 1e300, 60)
        slots = {b.admit(r.desc, r.P, serving.CFL, 1e300, 200): r for r in ones}
        left = 4
        got = {}
        while left:
            b.step(4)
            for s in b.finished():
                done, t, steps, st, rep = b.query(s)
                got[s] = (b.retire(s), t, steps)
                left -= 1
    P, t, n = got[sb]
    Pw, tw, nw = _alone_vs_block(big, Pbig, 60)
    assert (t, n) == (tw, nw) and np.array_equal(P, Pw)
    for s, r in slots.items():
        P, t, n = got[s]
        Pw, tw, nw = _alone_vs_block(r.desc, r.P, 200)
        assert (t, n) == (tw, nw) and np.array_equal(P, Pw), r.index


@pytest.mark.parametrize("graphs", [True, False])
def test_long_chain_moves_to_the_express_lane(graphs):
    """A 2-D MHD solve with far more steps left (1500) than the large hydro
    solves beside it (60 each) is promoted to the express lane after its
    first call and, while their steps are slow, takes several of its own per
    round; every solve ends bitwise equal to its single-block run."""
    rng = np.random.default_rng(7)
    big_p = problems.riemann_params(rng)
    while not (big_p["law"] == "mhd" and big_p["dim"] == 2):
        big_p = problems.riemann_params(rng)
    big_p.update(n=64)
    big = problems.riemann_desc(big_p, precision="fp64")
    Pbig = problems.riemann_problem(big_p, big)
    smalls = []
    while len(smalls) < 6:
        prm = problems.riemann_params(rng)
        if not (prm["law"] == "hydro" and prm["dim"] == 2):
            continue
        prm.update(n=256)
        d = problems.riemann_desc(prm, precision="fp64")
        smalls.append((d, problems.riemann_problem(prm, d)))
    with Batch(BUDGET, rk_order=2, precision="fp64", max_solves=16) as b:
        b.set_graphs(graphs)
        b.set_express(8)
        sb = b.admit(big, Pbig, serving.CFL, 1e300, 1500)
        ss = [b.admit(d, P, serving.CFL, 1e300, 60) for d, P in smalls]
        got = {}
        lanes = set()
        big_steps_when_others_done = None
        while len(got) < 7:
            b.step(4)
            if sb not in got:
                lanes.add(b.lane(sb))
            for s in b.finished():
                done, t, steps, st, rep = b.query(s)
                got[s] = (b.retire(s), t, steps, st)
            if big_steps_when_others_done is None and all(s in got for s in ss):
                big_steps_when_others_done = got[sb][2] if sb in got else b.query(sb)[2]
    assert 2 in lanes, lanes
    assert big_steps_when_others_done > 120, big_steps_when_others_done
    for s, (d, P0, n_max) in [(sb, (big, Pbig, 1500))] + [(s, (d, P, 60)) for s, (d, P) in zip(ss, smalls)]:
        P, t, n, st = got[s]
        with Block(d) as blk:
            blk.upload(P0)
            tw, nw, rep = blk.run(serving.CFL, 1e300, n_max, check=False)
            Pw = blk.download()
        assert (t, n, st) == (tw, nw, rep.status) and np.array_equal(P, Pw), s


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("zplanes", [16, 4])
def test_fused_3d_solve_in_a_batch(precision, zplanes):
    """3-D hydro + PLM solves take the batch's fused tiles (z-marching, 16 or
    4 planes per work-group): equal to the single block bit for bit, beside a
    1-D solve."""
    desc = problems.turbulence_desc(24, precision)
    P0 = problems.turbulence_3d(problems.turbulence_desc(24, "fp64"), kmax=4)
    one = serving.make_stream(2, precision=precision, resize=SMALL)[1]
    with Batch(BUDGET, rk_order=2, precision=precision) as b:
        b.set_shape(abi.make_shape(zplanes=zplanes))
        s3 = b.admit(desc, P0, 0.3, 1e300, 6)
        b.admit(one.desc, one.P, serving.CFL, serving.T_FINAL, serving.MAX_STEPS)
        got = {}
        while len(got) < 2:
            b.step(5)
            for s in b.finished():
                got[s] = b.retire(s)
    with Block(desc) as blk:
        blk.upload(P0)
        blk.run(0.3, 1e300, 6)
        Pw = blk.download()
    assert np.array_equal(got[s3], Pw)


def test_rk3_batch_uses_flux_arrays_and_matches_block():
    """Outside RK2 every solve takes the flux-array kernels: 1-D and 2-D
    hydro and MHD at RK3 equal their single-block runs."""
    reqs = serving.make_stream(N_REQ, precision="fp64", resize=SMALL)[:6]
    descs = []
    for r in reqs:
        d = problems.riemann_desc(r.params, precision="fp64", rk_order=3)
        descs.append(d)
    with Batch(BUDGET, rk_order=3, precision="fp64") as b:
        slots = {b.admit(d, r.P, serving.CFL, 1e300, 30): (d, r) for d, r in zip(descs, reqs)}
        got = {}
        while len(got) < len(slots):
            b.step(7)
            for s in b.finished():
                got[s] = b.retire(s)
    for s, (d, r) in slots.items():
        Pw, _, _ = _alone_vs_block(d, r.P, 30)
        assert np.array_equal(got[s], Pw), r.index


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("wpc", [1, 3])
def test_work_queue_tiles_equal_one_tile_per_work_group(fp64_stream, fp64_mix, precision, wpc):
    """mara_batch_set_persistent: the 2-D and 3-D hydro tiles drawn from a
    work queue by a fixed grid give the results of one work-group per tile,
    bit for bit, in a mix with 1-D, MHD and a 3-D solve, with graphs and
    host launches."""
    reqs = fp64_stream if precision == "fp64" else _stream("mixed")
    desc3 = problems.turbulence_desc(24, precision)
    P3 = problems.turbulence_3d(problems.turbulence_desc(24, "fp64"), kmax=4)

    def run(persistent, graphs):
        out = {}
        with Batch(BUDGET, rk_order=2, precision=precision, max_solves=64) as b:
            b.set_graphs(graphs)
            b.set_persistent(persistent)
            slots = {b.admit(r.desc, r.P, serving.CFL, serving.T_FINAL, serving.MAX_STEPS): r.index for r in reqs}
            slots[b.admit(desc3, P3, 0.3, 1e300, 8)] = "3d"
            while slots:
                b.step(7)
                for s in b.finished():
                    out[slots.pop(s)] = b.retire(s)
        return out

    ref = run(0, True)
    for graphs in (True, False):
        got = run(wpc, graphs)
        assert set(got) == set(ref)
        for k in ref:
            assert np.array_equal(got[k], ref[k]), (k, graphs)


@pytest.mark.parametrize("rk_order", [2, 3])
def test_cooled_mhd_solves_in_a_batch(rk_order):
    """NewtonianMHD with a cooling rate (nine fields, the cooling terms added
    to L in every stage) in a batch beside ordinary solves: bit-equal to the
    single block in FP64, 1-D and 2-D, flux-CT on and off."""
    def cooled(cells, ct, seed):
        d = abi.make_desc(law="mhd", recon="plm", riemann="hlle", bc="periodic", cells=cells, field_ct=ct,
                          cooling_rate=0.4, rk_order=rk_order)
        rng = np.random.default_rng(seed)
        P = problems.empty(d)
        shp = P.shape[:3]
        P[..., 0] = 1.0 + 0.3 * rng.random(shp)
        P[..., 4] = 1.0 + 0.3 * rng.random(shp)
        P[..., 1:4] = 0.2 * rng.standard_normal(shp + (3,))
        P[..., 5:8] = 0.3 + 0.1 * rng.standard_normal(shp + (3,))
        P[..., 8] = 0.1 * rng.random(shp)
        return d, problems.apply_bc(d, P)

    solves = [cooled((40, 1, 1), 1, 1), cooled((24, 20, 1), 1, 2), cooled((20, 18, 1), 0, 3)]
    plain = serving.make_stream(3, precision="fp64", resize=SMALL)
    plain = [(problems.riemann_desc(r.params, precision="fp64", rk_order=rk_order), r.P) for r in plain]
    with Batch(BUDGET, rk_order=rk_order, precision="fp64") as b:
        slots = {}
        for n, (d, P) in enumerate(solves + plain):
            if n < len(solves):
                slots[b.admit(d, P, 0.3, 1e300, 6)] = n
            else:
                slots[b.admit(d, P, serving.CFL, serving.T_FINAL, serving.MAX_STEPS)] = n
        got = {}
        while len(got) < len(slots):
            b.step(4)
            for s in b.finished():
                done, t, steps, st, rep = b.query(s)
                assert st == 0, rep.as_dict()
                got[slots[s]] = (b.retire(s), t, steps)
    for n, (d, P) in enumerate(solves + plain):
        with Block(d) as blk:
            blk.upload(P)
            if n < len(solves):
                t, steps, _ = blk.run(0.3, 1e300, 6)
            else:
                t, steps, _ = blk.run(serving.CFL, serving.T_FINAL, serving.MAX_STEPS)
            Pb = blk.download()
        assert (t, steps) == got[n][1:], n
        assert np.array_equal(Pb, got[n][0]), n
