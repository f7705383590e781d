#!/usr/bin/env python3
"""Throughput of the MI355X Godunov step on the 3-D decaying-turbulence box.

Workload: 256^3 zones, periodic unit cube, gamma 5/3, rho = 1, p = 1, seeded
Gaussian random velocity field on 1 <= |k| <= 16 with |v_k|^2 ~ k^-4 at rms
Mach 1; PLM (theta 1.5) + HLLC, RK2, CFL 0.3 (mara2_amd.problems.turbulence_3d).

One step = courant reduction + two RK stages + ghost fills, driven on the
device by mara_block_run (dt = 0.3 * courant each step).  The state is
resident in HBM before the timed region starts; the call returns after the
stream has drained.  Prints one JSON line (value = plain FP64 form; the
mixed form, its speed-up and its L1 distance to FP64 ride along).

    python bench.py [--gpus 1] [--steps 100] [--warmup 5]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from mara2_amd import abi, problems  # noqa: E402
from mara2_amd.runtime import Batch, Block  # noqa: E402

METRIC = "zone-updates/s, FP32-mixed vs FP64 at stated L1 error; solves/s + p99 latency"
HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
CFL = 0.3


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def alg_bytes_per_zone_update(desc):
    """(3s - 1) * nq * 8: each stage reads the state and writes the new one,
    stages after the first also read the step's starting state."""
    s = desc.rk_order
    return (3 * s - 1) * abi.num_fields(desc.law) * 8


def traffic_from_profile(kernel, workload):
    """HBM bytes per launch measured by rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE,
    committed under profiles/ (null when no such summary exists)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def time_form(desc, P0, steps, warmup, transfers=None, tune=None):
    """Upload, warm up, let the autotuner pick the launch shape (tune, if a
    dict, receives it), then time `steps` device-driven steps.  transfers, if
    a dict, receives the host<->device times of the MeshData::P arrays."""
    with Block(desc) as blk:
        t0 = time.perf_counter()
        blk.upload(P0)
        if transfers is not None:
            transfers["upload_s"] = time.perf_counter() - t0
        if warmup:
            blk.run(CFL, 1e300, warmup)
            blk.upload(P0)
        if tune is not None:
            shape, ms = blk.autotune(repeats=5)
            tune.update(shape.as_dict(), stage_ms=ms)
        # stage-kernel durations from HIP event pairs recorded on the block's
        # stream around every stage launch of the timed region (no host wait
        # between them; read after the run)
        blk.set_timing(True)
        t0 = time.perf_counter()
        t, taken, rep = blk.run(CFL, 1e300, steps)
        t1 = time.perf_counter()
        if taken != steps:
            raise RuntimeError(f"run stopped after {taken} of {steps} steps: {rep.as_dict()}")
        avg_ms, launches, name = blk.stage_kernel_time()
        t2 = time.perf_counter()
        P = blk.download()
        if transfers is not None:
            transfers["download_s"] = time.perf_counter() - t2
    return t1 - t0, P, avg_ms, launches, name


def fast_vs_plain_after(d64, dfast, P0, steps):
    """Relative L1 of the conserved interior between the fast and the plain
    binary64 forms after `steps` steps of the plain form's dt."""
    res = []
    dts = []
    for d in (d64, dfast):
        with Block(d) as blk:
            blk.upload(P0)
            for s in range(steps):
                if d is d64:
                    dts.append(CFL * blk.courant_dt())
                blk.advance(dts[s])
            res.append(blk.download())
    g = d64.guards
    inner = tuple(slice(g[a], g[a] + d64.cells[a]) for a in range(3))

    def cons(P):
        P = P[inner]
        U = np.empty_like(P)
        U[..., 0] = P[..., 0]
        U[..., 1:4] = P[..., :1] * P[..., 1:4]
        U[..., 4] = 0.5 * P[..., 0] * np.sum(P[..., 1:4] ** 2, axis=-1) + P[..., 4] / (d64.gamma - 1)
        return U
    a, b = cons(res[1]), cons(res[0])
    return max(float(np.sum(np.abs(a[..., q] - b[..., q])) / np.sum(np.abs(b[..., q]))) for q in range(5))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_rate(n_sample, threads, target_s, max_steps=2000):
    import oracle
    oracle.set_threads(threads)
    desc = problems.turbulence_desc(n_sample)
    P = problems.turbulence_3d(desc)
    err, dtc = oracle.courant_dt(desc, P)
    oracle.advance(desc, P, 0.0, CFL * dtc)  # warm-up step
    steps, t0 = 0, time.perf_counter()
    while True:
        err, dtc = oracle.courant_dt(desc, P)
        err, _ = oracle.advance(desc, P, 0.0, CFL * dtc)
        steps += 1
        el = time.perf_counter() - t0
        if el > target_s or steps >= max_steps:
            break
    return n_sample ** 3 * steps / el, steps, el


def cpu_baseline(n_sample, target_s=15.0):
    """The oracle (CPU restatement of the reference step) on the same workload
    at n_sample^3, OpenMP threads = the host's CPU share, plus a shorter
    1-thread run."""
    # the host's CPU share: OMP_NUM_THREADS where the environment sets it
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(os.cpu_count() or 1, 16)
    rate, steps, el = _oracle_rate(n_sample, threads, target_s)
    rate1, steps1, el1 = _oracle_rate(32, 1, 5.0, 200)
    return {"value": rate, "unit": "zone-updates/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(), "one_thread_value": rate1,
            "sample": f"turbulence_3d at {n_sample}^3 (same generator and scheme), {steps} steps after 1 warm-up "
                      f"step, {el:.1f} s, CPU restatement of the reference step (oracle/), OpenMP over zones; "
                      f"one thread: 32^3, {steps1} steps, {el1:.1f} s"}


def other_workload(args):
    """Secondary configurations (not the headline line): 2-D MHD with flux-CT,
    3-D relativistic MHD with WENO5, and the 1-D Sod tube to t = 0.2."""
    from mara2_amd.runtime import MaraError
    w = args.workload
    if w == "sod":
        d = problems.sod_desc(1024)
        P0 = problems.sod_1d(d)
        with Block(d) as blk:
            blk.set_graphs(not args.no_graphs)
            blk.upload(P0)
            blk.run(0.4, 0.2, 1 << 30)  # warm-up: kernel loading and graph capture
            blk.upload(P0)
            shape = None
            if not args.no_autotune:
                shape, _ = blk.autotune(repeats=20)
                shape = shape.as_dict()
                blk.run(0.4, 0.2, 1 << 30)  # graph capture in the tuned shape
                blk.upload(P0)
            t0 = time.perf_counter()
            t, steps, rep = blk.run(0.4, 0.2, 1 << 30)
            el = time.perf_counter() - t0
        out = {"metric": METRIC, "value": 1024 * steps / el, "unit": "zone-updates/s", "n_gpus": 1,
               "steps": steps, "warmup": 0, "ms_per_step": el / steps * 1e3, "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic: Sod tube",
               "config": {"workload": "1-D Sod tube 1024 zones, gamma 1.4, PLM + HLLE, RK2, CFL 0.4, to t = 0.2",
                          "t_reached": t, "time_to_solution_s": el, "graphs": not args.no_graphs,
                          "launch_shape": shape}}
        print(json.dumps(out), flush=True)
        return
    if w == "mhd2d":
        n = args.n or 2048
        make_desc, make, cfl, label = problems.mhd_desc, problems.mhd_2d, 0.4, \
            f"2-D ideal MHD {n}^2, PLM + HLLE, flux-CT, RK2, CFL 0.4, periodic"
        gen = lambda d: make(d)
    else:
        n = args.n or 128
        make_desc, make, cfl, label = problems.rmhd_desc, problems.rmhd_3d, 0.3, \
            f"3-D relativistic MHD {n}^3, WENO5 + HLLE, flux-CT, RK3, CFL 0.3, periodic, max W {args.wmax}"
        gen = lambda d: make(d, wmax=args.wmax)
    d64 = make_desc(n, "fp64")
    log(f"generating {label}")
    P0 = gen(d64)
    res = {}
    for form in ("fp64", "mixed"):
        d = make_desc(n, form)
        with Block(d) as blk:
            blk.set_graphs(not args.no_graphs)
            # warm-up (kernel loading, graph capture) and the launch shape
            blk.upload(P0)
            blk.run(cfl, 1e300, 2, check=False)
            shape = None
            if not args.no_autotune:
                shape = blk.autotune(repeats=5)[0].as_dict()
            # the maraMainLoop contract: stops at the first failure, as the reference does
            blk.upload(P0)
            t0 = time.perf_counter()
            try:
                t, steps, rep = blk.run(cfl, 1e300, args.steps, check=False)
            except MaraError as e:
                t, steps, rep = 0.0, 0, None
            el = time.perf_counter() - t0
            P = blk.download()
            mono = blk.monopole_max()
            timed = None
            if steps < args.steps:
                # throughput when the loop stopped early: up to 10 steps from
                # the initial state at its first dt, failures ignored.  Zones
                # past an inversion failure run the Newton solver to its
                # iteration limit every step and their number grows, so a
                # longer fixed-dt run would time the failure, not the scheme
                # (100 steps: 9.8 ms/step against 5.7 over the first 10).
                n_fixed = min(args.steps, 10)
                blk.upload(P0)
                dt = cfl * blk.courant_dt()
                blk.advance(dt, check=False)
                t1 = time.perf_counter()
                for _ in range(n_fixed):
                    blk.advance(dt, check=False)
                timed = (time.perf_counter() - t1, n_fixed)
            blk.upload(P0)
            blk.set_timing(True)
            blk.run(cfl, 1e300, 2, check=False)
            ms, nl, name = blk.stage_kernel_time()
        res[form] = dict(el=el, steps=steps, P=P, rep=rep.as_dict() if rep else None, mono=mono, ms=ms, name=name,
                         timed=timed, shape=shape)
        log(f"{form}: {steps} steps in {el:.3f} s, report {res[form]['rep']}, max |monopole| {mono:.3e}"
            + (f"; fixed-dt timing {timed[1]} steps in {timed[0]:.3f} s" if timed else ""))
    zones = int(np.prod([d64.cells[a] for a in range(3)]))
    g = d64.guards
    inner = tuple(slice(g[a], g[a] + d64.cells[a]) for a in range(3))
    a, b = res["mixed"]["P"][inner], res["fp64"]["P"][inner]
    l1 = max(float(np.sum(np.abs(a[..., q] - b[..., q])) / max(np.sum(np.abs(b[..., q])), 1e-300)) for q in range(8))
    r64 = res["fp64"]
    b_alg = alg_bytes_per_zone_update(d64)
    per_launch = b_alg * zones / d64.rk_order

    def rate(r):
        el, steps = r["timed"] if r["timed"] else (r["el"], r["steps"])
        return zones * steps / el, el / max(steps, 1) * 1e3, steps

    v64, ms64, st64 = rate(r64)
    vmx, msmx, _ = rate(res["mixed"])
    out = {"metric": METRIC, "value": v64, "unit": "zone-updates/s", "n_gpus": 1,
           "steps": st64, "warmup": 0, "ms_per_step": ms64,
           "steps_before_failure": {"fp64": r64["steps"], "mixed": res["mixed"]["steps"]},
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic: seeded Gaussian random fields",
           "config": {"workload": label, "zones": zones, "launch_shape": r64["shape"],
                      "launch_shape_mixed": res["mixed"]["shape"]},
           "roofline": {"bound": "hbm", "achieved": per_launch / (r64["ms"] * 1e-3) / 1e9, "peak": HBM_PEAK / 1e9,
                        "unit": "GB/s", "frac": per_launch / (r64["ms"] * 1e-3) / HBM_PEAK, "traffic": None,
                        "kernel": r64["name"], "avg_launch_ms": r64["ms"]},
           "fp64": {"max_abs_monopole": r64["mono"], "report": r64["rep"]},
           "mixed": {"value": vmx, "speedup_vs_fp64": vmx / v64,
                     "rel_l1_vs_fp64": l1, "max_abs_monopole": res["mixed"]["mono"], "report": res["mixed"]["rep"]}}
    print(json.dumps(out), flush=True)


def copy_workload(args):
    """A plain device-to-device copy of known size: 1 GiB read and 1 GiB
    written per launch (torch's elementwise copy kernel).  Under rocprofv3
    --pmc its FETCH_SIZE / WRITE_SIZE readings calibrate those counters for
    the stage kernel's traffic; its own rate is the HBM bandwidth a
    streaming kernel reaches on this GPU."""
    import torch
    n = 1 << 27  # doubles: 1 GiB
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(max(args.warmup, 1)):
        y.copy_(x)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        y.copy_(x)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    nbytes = 2 * n * 8
    print(json.dumps({"workload": "device copy, 1 GiB read + 1 GiB written per launch", "launches": args.steps,
                      "read_bytes_per_launch": n * 8, "write_bytes_per_launch": n * 8, "ms_per_launch": ms,
                      "GB_per_s": nbytes / (ms * 1e-3) / 1e9}), flush=True)


def serving_workload(args):
    """Continuous batching over the seeded Riemann-problem stream: 70% 1-D
    (256..16384 zones), 30% 2-D (128^2..1024^2), half MHD, PLM + HLLC / HLLE,
    RK2, CFL 0.4, to t = 0.2, mixed form.  The one-at-a-time service rate is
    measured on an evenly spaced sample of the stream (create, upload, run,
    download, destroy per request); arrivals are Poisson at twice that rate."""
    from mara2_amd import serving
    n_req = args.requests
    log(f"building {n_req} requests")
    reqs = serving.make_stream(n_req, precision="mixed")
    work = [r.zones for r in reqs]
    log("warm-up")
    warm = serving.make_stream(4, precision="mixed", resize=lambda dim, n: 32 if dim == 1 else 16)
    graphs = not args.no_graphs
    serving.sequential(warm, graphs=graphs)
    serving.serve(warm, 1 << 28, max_solves=8, chunk=args.chunk, graphs=graphs)
    stride = max(1, n_req // args.seq_sample)
    sample = reqs[::stride][:args.seq_sample]
    log(f"one-at-a-time service of {len(sample)} requests (every {stride}th)")
    times = serving.sequential(sample, graphs=graphs)
    seq_failed = [r.index for r in serving.failures(sample)]
    seq_rate = len(sample) / sum(times)
    seq_zone_rate = sum(r.zones * r.steps for r in sample) / sum(times)
    log(f"sequential: {seq_rate:.3f} solves/s, mean {np.mean(times):.3f} s, max {max(times):.3f} s")
    budget = int(args.budget_gib * 2**30)
    # the batch keeps its own launch shape (pencils of 128 faces, 64 x 8 MHD
    # tiles): tuned on one request alone, a block prefers longer pencils and
    # taller tiles, which are slower once the batch fills the device
    shape = None
    with Batch(1 << 20, rk_order=2, precision="mixed", max_solves=1) as b0:
        batch_shape = b0.get_shape().as_dict()
    out = {"metric": METRIC, "value": None, "unit": "solves/s", "n_gpus": 1, "steps": None,
           "warmup": None, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32/f64 mixed",
           "data": f"synthetic: seeded Riemann-problem stream (seed {problems.SEED_SERVING})",
           "config": {"workload": f"serving stream of {n_req} Riemann problems, 70% 1-D 256..16384 zones, "
                                  "30% 2-D 128^2..1024^2, half MHD, PLM + HLLC/HLLE, RK2, CFL 0.4, to t = 0.2, "
                                  "mixed form, Poisson arrivals at 2x the one-at-a-time rate",
                      "zones_total": int(sum(work)), "hbm_pool_gib": args.budget_gib, "chunk_steps": args.chunk,
                      "graphs": graphs, "launch_shape": batch_shape,
                      "express_lane_max": args.express},
           "arrival_rate": 2.0 * seq_rate,
           "sequential": {"solves_per_s": seq_rate, "zone_updates_per_s": seq_zone_rate,
                          "sample": f"{len(sample)} requests, every {stride}th of the stream",
                          "mean_s": float(np.mean(times)), "max_s": float(max(times)),
                          "failed_requests": seq_failed,
                          "truncated_requests": [r.index for r in serving.truncated(sample)]},
           "work_bound_per_request": {"steps": serving.MAX_STEPS, "zone_updates": serving.MAX_ZONE_UPDATES}}
    if not args.no_open_loop:
        serving.set_arrivals(reqs, 2.0 * seq_rate)
        log(f"serving {n_req} requests, Poisson arrivals at {2 * seq_rate:.3f}/s "
            f"(span {reqs[-1].arrival:.0f} s), pool {args.budget_gib:g} GiB")
        wall, done = serving.serve(reqs, budget, chunk=args.chunk, log=log, graphs=graphs, shape=shape,
                                   express=args.express)
        failed = [r.index for r in serving.failures(done)]
        st = serving.latency_stats(done)
        zone_updates = sum(r.zones * r.steps for r in done)
        log(f"served in {wall:.1f} s: p50 {st['p50_s']:.3f} s p99 {st['p99_s']:.3f} s")
        out.update({"value": len(done) / wall, "latency_s": st, "wall_s": wall,
                    "zone_updates_per_s": zone_updates / wall, "failed_requests": failed,
                    "truncated_requests": [r.index for r in serving.truncated(done)],
                    "slowest": serving.slowest(done),
                    "floor_applied_zones": int(sum(r.floors for r in done))})
    if args.saturate:
        for r in reqs:
            r.arrival = 0.0
        log("saturated run: every request at t = 0")
        wall_s, done_s = serving.serve(reqs, budget, chunk=args.chunk, log=log, graphs=graphs, shape=shape,
                                       express=args.express)
        zu_s = sum(r.zones * r.steps for r in done_s)
        # solves/s against the sample's solves/s, and zone-updates/s against
        # the sample's zone-updates/s: the second does not depend on how the
        # sample's mean work per solve compares with the stream's
        out["saturated"] = {"solves_per_s": len(done_s) / wall_s, "wall_s": wall_s,
                            "speedup_vs_sequential": len(done_s) / wall_s / seq_rate,
                            "work_speedup_vs_sequential": zu_s / wall_s / seq_zone_rate,
                            "stream_vs_sample_work_per_solve":
                                (zu_s / len(done_s)) / (sum(r.zones * r.steps for r in sample) / len(sample)),
                            "latency_s": serving.latency_stats(done_s),
                            "zone_updates_per_s": zu_s / wall_s,
                            "failed_requests": [r.index for r in serving.failures(done_s)],
                            "truncated_requests": [r.index for r in serving.truncated(done_s)],
                            "slowest": serving.slowest(done_s)}
        if out["value"] is None:
            out["value"] = out["saturated"]["solves_per_s"]
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=0)
    ap.add_argument("--cpu-sample", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", default="turbulence",
                    choices=["turbulence", "mhd2d", "rmhd3d", "sod", "serving", "copy"])
    ap.add_argument("--wmax", type=float, default=5.0)
    ap.add_argument("--requests", type=int, default=4096, help="serving: stream length")
    ap.add_argument("--seq-sample", type=int, default=256, help="serving: requests timed one at a time")
    ap.add_argument("--budget-gib", type=float, default=64.0, help="serving: HBM pool")
    ap.add_argument("--chunk", type=int, default=16, help="serving: steps per batch launch sequence")
    ap.add_argument("--saturate", action="store_true", help="serving: also run with every request at t = 0")
    ap.add_argument("--no-open-loop", action="store_true", help="serving: skip the Poisson-arrival run")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--express", type=int, default=0, help="serving: most solves on the batch's express lane")
    ap.add_argument("--no-graphs", action="store_true",
                    help="sod, mhd2d, rmhd3d, serving: launch every kernel from the host instead of replaying HIP graphs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(a multi-GPU run would shard the mesh into one block per GPU with a halo exchange)",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "serving":
        return serving_workload(args)
    if args.workload == "copy":
        return copy_workload(args)
    if args.workload != "turbulence":
        return other_workload(args)

    n = args.n or 256
    d64 = problems.turbulence_desc(n, "fp64")
    dmx = problems.turbulence_desc(n, "mixed")
    log(f"generating {n}^3 initial data")
    P0 = problems.turbulence_3d(d64)
    zones = n ** 3

    log("timing FP64 form")
    transfers = {}
    tune64 = None if args.no_autotune else {}
    tunemx = None if args.no_autotune else {}
    el64, P64, ms64, nl64, kname = time_form(d64, P0, args.steps, args.warmup, transfers, tune64)
    log(f"FP64: {el64:.3f} s for {args.steps} steps; stage kernel {ms64:.3f} ms ({kname}); shape {tune64}")
    dfast = problems.turbulence_desc(n, "fp64_fast")
    tunefast = None if args.no_autotune else {}
    log("timing fast FP64 form")
    elfa, Pfa, msfa, nlfa, kfa = time_form(dfast, P0, args.steps, args.warmup, None, tunefast)
    log(f"fast FP64: {elfa:.3f} s for {args.steps} steps; stage kernel {msfa:.3f} ms ({kfa}); shape {tunefast}")
    log("timing mixed form")
    elmx, Pmx, msmx, nlmx, _ = time_form(dmx, P0, args.steps, args.warmup, None, tunemx)
    log(f"mixed: shape {tunemx}")
    log(f"mixed: {elmx:.3f} s for {args.steps} steps; stage kernel {msmx:.3f} ms")

    g = d64.guards[0]
    inner = (slice(g, g + n),) * 3
    U64 = P64[inner]
    Umx = Pmx[inner]
    l1 = max(float(np.sum(np.abs(Umx[..., q] - U64[..., q])) / np.sum(np.abs(U64[..., q]))) for q in range(5))

    value = zones * args.steps / el64
    b_alg = alg_bytes_per_zone_update(d64)
    per_launch = b_alg * zones / d64.rk_order
    achieved = per_launch / (ms64 * 1e-3)
    workload = f"turbulence3d_{n}"
    roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": achieved / HBM_PEAK, "traffic": traffic_from_profile("stage_fp64", workload),
                "kernel": kname, "avg_launch_ms": ms64, "launches": nl64,
                "alg_bytes_per_launch": per_launch}
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el64 / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": f"synthetic: seeded Gaussian random velocity field (seed {problems.SEED_TURBULENCE})",
        "config": {"workload": f"3-D decaying turbulence {n}^3, hydro gamma 5/3, PLM theta 1.5 + HLLC, "
                               f"RK2, CFL {CFL}, periodic", "zones": zones, "form": "fp64",
                   "launch_shape": tune64, "launch_shape_mixed": tunemx},
        "roofline": roofline,
        # the C ABI takes host arrays: PCIe-inclusive rate of one upload, the
        # timed steps and one download of MeshData::P (not the headline value)
        "host_transfer": {"upload_s": transfers["upload_s"], "download_s": transfers["download_s"],
                          "pcie_inclusive_value": zones * args.steps
                          / (el64 + transfers["upload_s"] + transfers["download_s"])},
        "mixed": {"value": zones * args.steps / elmx, "ms_per_step": elmx / args.steps * 1e3,
                  "speedup_vs_fp64": el64 / elmx, "rel_l1_vs_fp64": l1,
                  "avg_stage_launch_ms": msmx,
                  "roofline_frac": per_launch / (msmx * 1e-3) / HBM_PEAK},
    }
    # the fast form against the plain one (bit-identical to the reference
    # restatement) after 10 steps of the same dt: the FP64 bound of 1e-12
    l1_fast10 = fast_vs_plain_after(d64, dfast, P0, 10)
    out["fp64_fast"] = {"value": zones * args.steps / elfa, "ms_per_step": elfa / args.steps * 1e3,
                        "speedup_vs_fp64": el64 / elfa, "avg_stage_launch_ms": msfa, "kernel": kfa,
                        "roofline_frac": per_launch / (msfa * 1e-3) / HBM_PEAK,
                        "rel_l1_vs_fp64_after_10_steps": l1_fast10, "launch_shape": tunefast}
    if not args.no_cpu_baseline:
        log("CPU baseline")
        out["cpu_baseline"] = cpu_baseline(args.cpu_sample)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
