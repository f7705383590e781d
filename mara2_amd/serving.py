"""Serving loop over independent solves: a seeded request stream, the
one-at-a-time baseline and continuous batching on a device pool.

A request is one Riemann problem (problems.riemann_params) solved to
t_final under the maraMainLoop contract (src/Mara.cpp:42-88: dt = min(cfl *
courant, t_final - t) while t < t_final).  The reference solves such a
problem by running one sub-program process per problem; here many solves
share the device.

serve() runs the continuous-batching loop: requests that have arrived are
admitted while the pool (an HBM budget) and the slot table have room,
mara_batch_step advances every admitted solve by a chunk of its own steps,
finished solves are retired (their primitives downloaded) between chunks,
and newcomers join the next chunk.  Latency is arrival to download.
sequential() serves requests one at a time through the single-block entry
points (create, upload, mara_block_run, download, destroy).
"""
import dataclasses
import time
from collections import deque

import numpy as np

from . import abi, problems
from .runtime import Batch, Block, MaraError

T_FINAL = 0.2
CFL = 0.4
# Work bound per request.  The stream's regular solves need at most about
# 1.6e5 steps (1-D, 16384 zones) or 7e9 zone-updates (2-D, 1024^2); an MHD
# problem that rarefies towards vacuum drives the Alfven speed B / sqrt(rho)
# up and dt down without bound, and is stopped here with its clock short of
# t_final (reported as truncated), in the batch and one-at-a-time alike.
MAX_STEPS = 200000
MAX_ZONE_UPDATES = 10**10


def max_steps(req):
    """Step bound of one request: MAX_STEPS and MAX_ZONE_UPDATES / zones."""
    return int(min(MAX_STEPS, MAX_ZONE_UPDATES // req.zones))


@dataclasses.dataclass
class Request:
    index: int
    params: dict
    desc: object
    P: np.ndarray = None
    arrival: float = 0.0       # seconds after the stream starts
    admitted: float = None
    finished: float = None
    steps: int = 0
    t: float = 0.0
    status: int = 0
    floors: int = 0            # zones where the pressure floor was applied
    result: np.ndarray = None

    @property
    def zones(self):
        return int(np.prod([self.desc.cells[a] for a in range(3)]))

    @property
    def latency(self):
        return self.finished - self.arrival


def make_stream(n_requests, seed=problems.SEED_SERVING, precision="mixed", resize=None,
                with_data=True):
    """The first n_requests requests of the seeded stream, arrival times unset.
    resize (tests only) maps (dim, zones per axis) to a smaller zone count."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_requests):
        prm = problems.riemann_params(rng)
        if resize is not None:
            prm["n"] = resize(prm["dim"], prm["n"])
        desc = problems.riemann_desc(prm, precision=precision)
        r = Request(i, prm, desc)
        if with_data:
            r.P = problems.riemann_problem(prm, desc)
        out.append(r)
    return out


def set_arrivals(requests, rate, seed=problems.SEED_SERVING + 1):
    """Poisson arrivals of the given mean rate (requests per second), seeded."""
    gaps = np.random.default_rng(seed).exponential(1.0, len(requests)) / rate
    t = np.cumsum(gaps) - gaps[0]
    for r, a in zip(requests, t):
        r.arrival = float(a)


def solve_alone(req, keep_result=True, graphs=True):
    """Serve one request on its own block: create, upload, run, download."""
    t0 = time.perf_counter()
    with Block(req.desc) as b:
        b.set_graphs(graphs)
        b.upload(req.P)
        t, steps, rep = b.run(CFL, T_FINAL, max_steps(req), check=False)
        P = b.download()
    req.t, req.steps, req.status, req.floors = t, steps, rep.status, rep.floor_applied
    if keep_result:
        req.result = P
    return time.perf_counter() - t0


def sequential(requests, keep_result=False, graphs=True):
    """Service time of each request served one at a time (seconds)."""
    return [solve_alone(r, keep_result, graphs) for r in requests]


def serve(requests, hbm_budget, max_solves=1024, chunk=16, rk_order=2, precision="mixed",
          keep_results=False, log=None, clock=time.perf_counter, graphs=True, shape=None, express=0,
          batch_factory=Batch, persistent=0):
    """Continuous batching over a request stream with arrival times set;
    shape (optional) is the batch's launch shape.  Returns (wall seconds,
    list of requests with admitted/finished stamps).  A request the pool
    cannot take even when it is empty (larger than the HBM budget) is not
    waited on: it finishes at once with status MARA_ERR_CAPACITY.
    batch_factory(hbm_budget, rk_order=, precision=, max_solves=) makes the
    batch (the device Batch by default); persistent > 0 runs the batch's
    hydro tiles as a work queue of that many work-groups per CU."""
    pending = deque(sorted(requests, key=lambda r: r.arrival))
    waiting = []
    in_flight = {}
    done = []
    with batch_factory(hbm_budget, rk_order=rk_order, precision=precision, max_solves=max_solves) as batch:
        batch.set_graphs(graphs)
        batch.set_express(express)
        if persistent:
            batch.set_persistent(persistent)
        if shape is not None:
            batch.set_shape(shape)
        t0 = clock()
        last_log = t0
        while pending or waiting or in_flight:
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            # admit in arrival order, skipping requests that do not fit yet
            still = []
            for r in waiting:
                slot = batch.admit(r.desc, r.P, CFL, T_FINAL, max_steps(r)) if len(in_flight) < max_solves else None
                if slot is None:
                    still.append(r)
                else:
                    r.admitted = clock() - t0
                    in_flight[slot] = r
            waiting = still
            if not in_flight:
                # the pool is empty: whatever still waits can never be admitted
                for r in waiting:
                    r.admitted = r.finished = clock() - t0
                    r.status = abi.ERR_CAPACITY
                    done.append(r)
                waiting = []
                if pending:
                    time.sleep(max(0.0, min(pending[0].arrival - (clock() - t0), 0.05)))
                continue
            batch.step(chunk)
            for slot in batch.finished():
                r = in_flight.pop(slot)
                _, r.t, r.steps, r.status, rep = batch.query(slot)
                r.floors = rep.floor_applied
                P = batch.retire(slot, download=True)
                r.finished = clock() - t0
                if keep_results:
                    r.result = P
                done.append(r)
            if log is not None and clock() - last_log > 10.0:
                last_log = clock()
                log(f"serve: t={clock() - t0:.1f}s done={len(done)} in_flight={len(in_flight)} "
                    f"waiting={len(waiting)} pending={len(pending)} free={batch.free_bytes / 2**30:.1f}GiB")
        wall = clock() - t0
    return wall, done


def latency_stats(done):
    lat = np.array([r.latency for r in done])
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max())}


def truncated(done):
    """Requests stopped by the step bound before t_final."""
    return [r for r in done if r.status == 0 and r.t < T_FINAL]


def slowest(done, n=5):
    """The n longest-latency requests: index, mesh, law, steps, clock."""
    out = []
    for r in sorted(done, key=lambda r: -r.latency)[:n]:
        out.append({"index": r.index, "dim": r.params["dim"], "n": r.params["n"], "law": r.params["law"],
                    "steps": r.steps, "t": r.t, "latency_s": r.latency})
    return out


def failures(done):
    return [r for r in done if r.status != 0]


def check_errors(done):
    bad = [r for r in done if r.status != 0]
    if bad:
        raise MaraError(bad[0].status, f"request {bad[0].index} ({len(bad)} failed)")
