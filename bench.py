#!/usr/bin/env python3
"""Throughput of the MI355X Godunov step, every configuration in one line.

Headline workload: 3-D decaying turbulence, 256^3 zones, periodic unit cube,
gamma 5/3, rho = 1, p = 1, seeded Gaussian random velocity field on
1 <= |k| <= 16 with |v_k|^2 ~ k^-4 at rms Mach 1; PLM (theta 1.5) + HLLC,
RK2, CFL 0.3 (mara2_amd.problems.turbulence_3d).  One step = the CFL
reduction, two RK stages and the ghost fills, driven on the device by
mara_block_run; the state is resident in HBM before the timed region starts
and the call returns after the stream has drained.

`value` is the fast binary64 form (every operation binary64, held to
relative L1 <= 1e-12 of the reference's conserved fields after 10 steps; the
distance to the bit-exact binary64 form is measured in the same run and
reported beside it).  The bit-exact binary64 form and the mixed form (binary32
fluxes, binary64 accumulation) ride along.  The other configurations follow
in `configs`: the Sod tube's time to solution, 2-D MHD 2048^2 with flux-CT,
3-D relativistic MHD 128^3, and a 1024-request serving stream; then the CPU
baseline (the oracle's -O3 build on the host's CPU share).

    python bench.py [--gpus 1] [--steps 20] [--warmup 5] [--workload all]
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
from mara2_amd.runtime import Batch, Block, Group, MaraError  # noqa: E402

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


def with_form(desc, form):
    d = type(desc).from_buffer_copy(desc)
    d.precision = abi.PRECISIONS[form]
    return d


def rel_l1_cons(desc, Pa, Pb):
    """max over fields of sum |U(Pa) - U(Pb)| / sum |U(Pb)| on the interior
    (hydro and MHD laws; the primitives for the relativistic one)."""
    g = desc.guards
    inner = tuple(slice(g[a], g[a] + desc.cells[a]) for a in range(3))
    a, b = Pa[inner], Pb[inner]

    def cons(P):
        if desc.law == abi.LAW_RMHD:
            return P
        U = np.empty_like(P)
        U[..., 0] = P[..., 0]
        U[..., 1:4] = P[..., :1] * P[..., 1:4]
        e = 0.5 * P[..., 0] * np.sum(P[..., 1:4] ** 2, axis=-1) + P[..., 4] / (desc.gamma - 1)
        if desc.law == abi.LAW_MHD:
            e = e + 0.5 * np.sum(P[..., 5:8] ** 2, axis=-1)
            U[..., 5:8] = P[..., 5:8]
        U[..., 4] = e
        return U
    ua, ub = cons(a), cons(b)
    return max(float(np.sum(np.abs(ua[..., q] - ub[..., q])) / max(np.sum(np.abs(ub[..., q])), 1e-300))
               for q in range(ua.shape[-1]))


def time_form(desc, P0, steps, warmup, cfl=CFL, tune=True, transfers=None):
    """Upload, warm up, let the autotuner pick the launch shape, then time
    `steps` device-driven steps with stage-kernel events.  Returns a dict."""
    out = {}
    with Block(desc) as blk:
        t0 = time.perf_counter()
        blk.upload(P0)
        if transfers is not None:
            transfers["upload_s"] = time.perf_counter() - t0
        if warmup:
            blk.run(cfl, 1e300, warmup, check=False)
            blk.upload(P0)
        if tune:
            shape, ms = blk.autotune(repeats=5)
            out["launch_shape"] = dict(shape.as_dict(), stage_ms=ms)
        blk.set_timing(True)
        t0 = time.perf_counter()
        t, taken, rep = blk.run(cfl, 1e300, steps, check=False)
        t1 = time.perf_counter()
        avg_ms, launches, name = blk.stage_kernel_time()
        t2 = time.perf_counter()
        out["P"] = blk.download()
        if transfers is not None:
            transfers["download_s"] = time.perf_counter() - t2
    out.update(el=t1 - t0, steps=taken, report=rep.as_dict(), stage_ms=avg_ms, launches=launches, kernel=name)
    return out


def run_fixed(desc, P0, dts):
    """The state after steps of the given sizes (mara_block_advance)."""
    with Block(desc) as blk:
        blk.upload(P0)
        for dt in dts:
            blk.advance(dt, check=False)
        return blk.download()


def courant_dts(desc, P0, steps, cfl):
    """The dt sequence of the bit-exact binary64 form over `steps` steps."""
    d = with_form(desc, "fp64")
    dts = []
    with Block(d) as blk:
        blk.upload(P0)
        for _ in range(steps):
            dts.append(cfl * blk.courant_dt())
            blk.advance(dts[-1], check=False)
    return dts


def roofline(desc, stage_ms, kernel, traffic=None):
    zones = int(np.prod([desc.cells[a] for a in range(3)]))
    per_launch = alg_bytes_per_zone_update(desc) * zones / desc.rk_order
    achieved = per_launch / (stage_ms * 1e-3)
    return {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "traffic": traffic, "kernel": kernel, "avg_launch_ms": stage_ms,
            "alg_bytes_per_launch": per_launch}


# ---------------------------------------------------------------------------
# headline: 3-D turbulence 256^3
# ---------------------------------------------------------------------------
def turbulence_leg(args):
    n = args.n or 256
    d64 = problems.turbulence_desc(n, "fp64")
    log(f"generating {n}^3 turbulence")
    P0 = problems.turbulence_3d(d64)
    zones = n ** 3
    res = {}
    transfers = {}
    for form in ("fp64_fast", "fp64", "mixed"):
        log(f"turbulence: timing {form}")
        r = time_form(with_form(d64, form), P0, args.steps, args.warmup, tune=not args.no_autotune,
                      transfers=transfers if form == "fp64_fast" else None)
        if r["steps"] != args.steps:
            raise RuntimeError(f"{form}: run stopped after {r['steps']} of {args.steps} steps: {r['report']}")
        r["value"] = zones * args.steps / r["el"]
        r.pop("P")
        log(f"{form}: {r['el']:.3f} s for {args.steps} steps; stage kernel {r['stage_ms']:.3f} ms ({r['kernel']})")
        res[form] = r
    # accuracy legs, untimed: the fast form after 10 steps against the
    # bit-exact one (which equals the reference restatement bit for bit), and
    # the mixed form over the configuration's stated 100 steps
    log("turbulence: accuracy legs")
    dts10 = courant_dts(d64, P0, 10, CFL)
    exact10 = run_fixed(d64, P0, dts10)
    l1_fast10 = rel_l1_cons(d64, run_fixed(with_form(d64, "fp64_fast"), P0, dts10), exact10)
    n_mixed = 100
    dts100 = courant_dts(d64, P0, n_mixed, CFL)
    l1_mixed100 = rel_l1_cons(d64, run_fixed(with_form(d64, "mixed"), P0, dts100), run_fixed(d64, P0, dts100))
    return d64, res, transfers, l1_fast10, l1_mixed100, n_mixed, P0


# ---------------------------------------------------------------------------
# one GPU's share of a split mesh, and the mesh split into blocks on one GPU
# ---------------------------------------------------------------------------
def blocks_leg(args, d64, P0):
    """(a) The block one GPU would hold if the n^3 mesh were split over 8 GPUs
    along x: n/8 x n x n, periodic, timed like the headline.  (b) The n^3 mesh
    as a 2 x 2 x 2 group of blocks on this GPU (guard exchange after every
    stage) against the undivided block: host-timed steps of mara_group_advance
    and mara_block_advance with the same time step, and the bitwise
    comparison of the two states afterwards."""
    n = d64.cells[0]
    out = {}
    share = abi.make_desc(law="hydro", gamma=d64.gamma, recon="plm", plm_theta=d64.plm_theta, riemann="hllc",
                          bc="periodic", cells=(n // 8, n, n), rk_order=2, precision="fp64",
                          upper=(1.0 / 8, 1.0, 1.0))
    Ps = problems.empty(share)
    Ps[...] = P0[:Ps.shape[0]]
    Ps = problems.apply_bc(share, Ps)
    zs = (n // 8) * n * n
    for form in ("fp64_fast", "fp64"):
        log(f"share block {n // 8}x{n}x{n}: timing {form}")
        r = time_form(with_form(share, form), Ps, args.steps, args.warmup, tune=not args.no_autotune)
        out[f"share_{form}"] = {"cells": [n // 8, n, n], "value": zs * r["steps"] / r["el"],
                                "ms_per_step": r["el"] / max(r["steps"], 1) * 1e3, "stage_ms": r["stage_ms"],
                                "kernel": r["kernel"],
                                "roofline_frac": roofline(share, r["stage_ms"], r["kernel"])["frac"],
                                "launch_shape": r.get("launch_shape")}
    # the same count of zones cut along the slowest axis of the device layout
    # (x is the contiguous one, and 64-lane tiles span it), and the cube of a
    # 2 x 2 x 2 split
    for cells in ((n, n, n // 8), (n // 2, n // 2, n // 2)):
        dd = abi.make_desc(law="hydro", gamma=d64.gamma, recon="plm", plm_theta=d64.plm_theta, riemann="hllc",
                           bc="periodic", cells=cells, rk_order=2, precision="fp64_fast",
                           upper=tuple(c / n for c in cells))
        Pd = problems.empty(dd)
        Pd[...] = P0[:Pd.shape[0], :Pd.shape[1], :Pd.shape[2]]
        Pd = problems.apply_bc(dd, Pd)
        z = int(np.prod(cells))
        log(f"share block {cells}: timing fp64_fast")
        r = time_form(dd, Pd, args.steps, args.warmup, tune=not args.no_autotune)
        out["share_fp64_fast_" + "x".join(str(c) for c in cells)] = {
            "cells": list(cells), "value": z * r["steps"] / r["el"],
            "ms_per_step": r["el"] / max(r["steps"], 1) * 1e3, "stage_ms": r["stage_ms"], "kernel": r["kernel"],
            "roofline_frac": roofline(dd, r["stage_ms"], r["kernel"])["frac"], "launch_shape": r.get("launch_shape")}
    dims = (2, 2, 2)
    d = with_form(d64, "fp64_fast")
    steps = args.steps
    log(f"group {dims} of {n}^3 against one block: {steps} steps")
    with Block(d) as blk:
        blk.upload(P0)
        dt = CFL * blk.courant_dt()
        for _ in range(args.warmup):
            blk.advance(dt, check=False)
        blk.upload(P0)
        t0 = time.perf_counter()
        for _ in range(steps):
            blk.advance(dt, check=False)
        t_blk = time.perf_counter() - t0
        Pb = blk.download()
    with Group(d, dims) as grp:
        grp.upload(P0)
        for _ in range(args.warmup):
            grp.advance(dt, check=False)
        grp.upload(P0)
        t0 = time.perf_counter()
        for _ in range(steps):
            grp.advance(dt, check=False)
        t_grp = time.perf_counter() - t0
        Pg = grp.download()
    zones = n ** 3
    out["group_2x2x2"] = {"form": "fp64_fast", "blocks": 8, "steps": steps,
                          "value": zones * steps / t_grp, "ms_per_step": t_grp / steps * 1e3,
                          "single_block_value": zones * steps / t_blk,
                          "single_block_ms_per_step": t_blk / steps * 1e3,
                          "group_over_single_time": t_grp / t_blk,
                          "bitwise_equal_to_single_block": bool(np.array_equal(Pb, Pg)),
                          "timing": "host clock around mara_group_advance / mara_block_advance calls, "
                                    "one synchronisation per step on both sides"}
    return out


# ---------------------------------------------------------------------------
# the other configurations
# ---------------------------------------------------------------------------
def sod_leg(args):
    d = problems.sod_desc(1024)
    P0 = problems.sod_1d(d)
    with Block(d) as blk:
        blk.upload(P0)
        blk.run(0.4, 0.2, 1 << 30)  # warm-up: kernel loading and graph capture
        blk.upload(P0)
        shape = None
        if not args.no_autotune:
            shape = blk.autotune(repeats=20)[0].as_dict()
            blk.run(0.4, 0.2, 1 << 30)  # graph capture in the tuned shape
            blk.upload(P0)
        t0 = time.perf_counter()
        t, steps, rep = blk.run(0.4, 0.2, 1 << 30)
        el = time.perf_counter() - t0
    return {"workload": "1-D Sod tube 1024 zones, gamma 1.4, PLM + HLLE, RK2, CFL 0.4, to t = 0.2, fp64",
            "time_to_solution_s": el, "steps": steps, "t": t, "value": 1024 * steps / el,
            "unit": "zone-updates/s", "launch_shape": shape,
            "roofline_frac": 1024 * steps / el * alg_bytes_per_zone_update(d) / HBM_PEAK}


def mhd2d_leg(args):
    n = 2048
    d64 = problems.mhd_desc(n, "fp64")
    log(f"generating 2-D MHD {n}^2")
    P0 = problems.mhd_2d(d64)
    zones = n * n
    out = {"workload": f"2-D ideal MHD {n}^2, PLM + HLLE, flux-CT, RK2, CFL 0.4, periodic"}
    for form in ("fp64", "fp64_fast", "mixed"):
        r = time_form(with_form(d64, form), P0, args.steps, 2, cfl=0.4, tune=not args.no_autotune)
        v = zones * r["steps"] / r["el"]
        out[form] = {"value": v, "ms_per_step": r["el"] / max(r["steps"], 1) * 1e3, "steps": r["steps"],
                     "stage_ms": r["stage_ms"], "kernel": r["kernel"],
                     "roofline_frac": v * alg_bytes_per_zone_update(d64) / HBM_PEAK,
                     "launch_shape": r.get("launch_shape")}
        log(f"mhd2d {form}: {v:.3e} zone-updates/s")
    # the configuration's 200 steps, untimed, with the bit-exact form's dt
    # (the fast form held to its bound after 10 of them)
    steps = 200
    dts = courant_dts(d64, P0, steps, 0.4)
    res = {}
    with Block(with_form(d64, "fp64")) as blk:
        blk.upload(P0)
        for dt in dts[:10]:
            blk.advance(dt, check=False)
        exact10 = blk.download()
    fast10 = run_fixed(with_form(d64, "fp64_fast"), P0, dts[:10])
    out["fp64_fast"]["rel_l1_vs_fp64_after_10_steps"] = rel_l1_cons(d64, fast10, exact10)
    out["fp64_fast"]["speedup_vs_fp64"] = out["fp64_fast"]["value"] / out["fp64"]["value"]
    for form in ("fp64", "mixed"):
        with Block(with_form(d64, form)) as blk:
            blk.upload(P0)
            for dt in dts:
                blk.advance(dt, check=False)
            res[form] = (blk.download(), blk.monopole_max())
    out["fp64"]["max_abs_monopole_after_200_steps"] = res["fp64"][1]
    out["mixed"]["max_abs_monopole_after_200_steps"] = res["mixed"][1]
    out["mixed"]["rel_l1_vs_fp64_after_200_steps"] = rel_l1_cons(d64, res["mixed"][0], res["fp64"][0])
    out["mixed"]["speedup_vs_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    return out


def rmhd_leg(args):
    n = 128
    d64 = problems.rmhd_desc(n, "fp64")
    log(f"generating 3-D RMHD {n}^3")
    P0 = problems.rmhd_3d(d64)
    zones = n ** 3
    out = {"workload": f"3-D relativistic MHD {n}^3, WENO5 + HLLE, flux-CT, RK3, CFL 0.3, periodic, max W 5, "
                       "mean sigma 1, flat spectra"}
    for form in ("fp64", "mixed"):
        d = with_form(d64, form)
        with Block(d) as blk:
            blk.upload(P0)
            blk.run(0.3, 1e300, 1, check=False)  # warm-up
            # the maraMainLoop contract: stops at the first inversion failure
            blk.upload(P0)
            t, steps, rep = blk.run(0.3, 1e300, 100, check=False)
            mono = blk.monopole_max()
            # throughput at the first step's dt from the initial state over a
            # fixed number of steps, failures counted (the reference would stop)
            blk.upload(P0)
            dt = 0.3 * blk.courant_dt()
            blk.advance(dt, check=False)
            n_fixed = min(args.steps, 10)
            blk.set_timing(True)
            t1 = time.perf_counter()
            fails = 0
            for _ in range(n_fixed):
                st, rp = blk.advance(dt, check=False)
                fails += rp.c2p_failures
            el = time.perf_counter() - t1
            ms, nl, name = blk.stage_kernel_time()
        v = zones * n_fixed / el
        out[form] = {"steps_before_failure": steps, "report_at_stop": rep.as_dict(),
                     "max_abs_monopole_at_stop": mono, "fixed_dt_value": v, "fixed_dt_steps": n_fixed,
                     "fixed_dt_c2p_failures": fails, "ms_per_step": el / n_fixed * 1e3, "stage_ms": ms,
                     "kernel": name, "roofline_frac": v * alg_bytes_per_zone_update(d64) / HBM_PEAK}
        log(f"rmhd3d {form}: {steps} steps before failure, fixed-dt {v:.3e} zone-updates/s")
    out["mixed"]["speedup_vs_fp64"] = out["mixed"]["fixed_dt_value"] / out["fp64"]["fixed_dt_value"]
    return out


def serving_leg(args, n_req, seq_sample, saturate=True, open_loop=True):
    """Continuous batching over the seeded Riemann-problem stream: 70% 1-D
    (256..16384 zones), 30% 2-D (128^2..1024^2), half MHD, PLM + HLLC / HLLE,
    RK2, CFL 0.4, to t = 0.2, mixed form.  The one-at-a-time service rate is
    measured on an evenly spaced sample of the stream (create, upload, run,
    download, destroy per request); arrivals are Poisson at twice that rate."""
    from mara2_amd import serving
    log(f"serving: building {n_req} requests")
    reqs = serving.make_stream(n_req, precision="mixed")
    warm = serving.make_stream(4, precision="mixed", resize=lambda dim, n: 32 if dim == 1 else 16)
    graphs = not args.no_graphs
    serving.sequential(warm, graphs=graphs)
    serving.serve(warm, 1 << 28, max_solves=8, chunk=args.chunk, graphs=graphs)
    stride = max(1, n_req // seq_sample)
    sample = reqs[::stride][:seq_sample]
    log(f"serving: one-at-a-time service of {len(sample)} requests (every {stride}th)")
    times = serving.sequential(sample, graphs=graphs)
    seq_rate = len(sample) / sum(times)
    seq_zone_rate = sum(r.zones * r.steps for r in sample) / sum(times)
    budget = int(args.budget_gib * 2**30)
    with Batch(1 << 20, rk_order=2, precision="mixed", max_solves=1) as b0:
        batch_shape = b0.get_shape().as_dict()
    out = {"workload": f"serving stream of {n_req} Riemann problems, 70% 1-D 256..16384 zones, 30% 2-D "
                       "128^2..1024^2, half MHD, PLM + HLLC/HLLE, RK2, CFL 0.4, to t = 0.2, mixed form",
           "unit": "solves/s", "hbm_pool_gib": args.budget_gib, "chunk_steps": args.chunk,
           "launch_shape": batch_shape,
           "sequential": {"solves_per_s": seq_rate, "zone_updates_per_s": seq_zone_rate,
                          "sample": f"{len(sample)} requests, every {stride}th of the stream",
                          "mean_s": float(np.mean(times)), "max_s": float(max(times)),
                          "failed_requests": [r.index for r in serving.failures(sample)],
                          "truncated_requests": [r.index for r in serving.truncated(sample)]},
           "work_bound_per_request": {"steps": serving.MAX_STEPS, "zone_updates": serving.MAX_ZONE_UPDATES}}
    if open_loop:
        serving.set_arrivals(reqs, 2.0 * seq_rate)
        log(f"serving: Poisson arrivals at {2 * seq_rate:.3f}/s (span {reqs[-1].arrival:.0f} s)")
        wall, done = serving.serve(reqs, budget, chunk=args.chunk, log=log, graphs=graphs)
        st = serving.latency_stats(done)
        out["open_loop"] = {"arrival_rate": 2.0 * seq_rate, "solves_per_s": len(done) / wall, "wall_s": wall,
                            "latency_s": st, "zone_updates_per_s": sum(r.zones * r.steps for r in done) / wall,
                            "failed_requests": [r.index for r in serving.failures(done)],
                            "truncated_requests": [r.index for r in serving.truncated(done)]}
        log(f"serving open loop: {len(done) / wall:.2f} solves/s, p50 {st['p50_s']:.3f} s, p99 {st['p99_s']:.3f} s")
    if saturate:
        for r in reqs:
            r.arrival = 0.0
        log("serving: saturated run, every request at t = 0")
        wall_s, done_s = serving.serve(reqs, budget, chunk=args.chunk, log=log, graphs=graphs)
        zu_s = sum(r.zones * r.steps for r in done_s)
        out["saturated"] = {"solves_per_s": len(done_s) / wall_s, "wall_s": wall_s,
                            "speedup_vs_sequential": len(done_s) / wall_s / seq_rate,
                            "work_speedup_vs_sequential": zu_s / wall_s / seq_zone_rate,
                            "stream_vs_sample_work_per_solve":
                                (zu_s / len(done_s)) / (sum(r.zones * r.steps for r in sample) / len(sample)),
                            "latency_s": serving.latency_stats(done_s), "zone_updates_per_s": zu_s / wall_s,
                            "failed_requests": [r.index for r in serving.failures(done_s)],
                            "truncated_requests": [r.index for r in serving.truncated(done_s)]}
        log(f"serving saturated: {len(done_s) / wall_s:.2f} solves/s "
            f"({out['saturated']['speedup_vs_sequential']:.2f}x one at a time)")
    out["value"] = out.get("saturated", out.get("open_loop", {})).get("solves_per_s")
    return out


def persistent_leg(args):
    """The serving batch's 2-D hydro tiles as a work queue
    (mara_batch_set_persistent: a fixed grid of work-groups per CU claiming
    the tiles from per-XCD counters) against one work-group per tile.  Every
    request at t = 0; the 2-D hydro requests of the stream alone (the tiles
    the queue serves) and the whole stream."""
    from mara2_amd import serving
    reqs = serving.make_stream(args.requests, precision="mixed")
    hydro2d = [r for r in reqs if r.params["dim"] == 2 and r.params["law"] == "hydro"]
    warm = serving.make_stream(4, precision="mixed", resize=lambda dim, n: 32 if dim == 1 else 16)
    serving.serve(warm, 1 << 28, max_solves=8, chunk=args.chunk)
    budget = int(args.budget_gib * 2**30)
    out = {"workload": f"saturated batches of the serving stream's first {args.requests} requests (mixed form)",
           "unit": "solves/s"}
    for name, subset in ((f"hydro_2d_{len(hydro2d)}", hydro2d), (f"stream_{len(reqs)}", reqs)):
        out[name] = {}
        for wpc in (0, 1, 2, 4):
            for r in subset:
                r.arrival = 0.0
            wall, done = serving.serve(subset, budget, chunk=args.chunk, persistent=wpc)
            key = "one_work_group_per_tile" if wpc == 0 else f"queue_{wpc}_per_cu"
            out[name][key] = {"solves_per_s": len(done) / wall, "wall_s": wall}
            log(f"persistent {name} {key}: {len(done) / wall:.2f} solves/s ({wall:.2f} s)")
    return out


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


# ---------------------------------------------------------------------------
# CPU baseline
# ---------------------------------------------------------------------------
def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_rate(n_sample, threads, target_s, min_steps=3, max_steps=2000):
    import ctypes
    import oracle
    from mara2_amd.abi import MaraReport
    L = oracle.lib(timing=True)
    L.oracle_set_threads(threads)
    desc = problems.turbulence_desc(n_sample)
    P = problems.turbulence_3d(desc)
    ptr = P.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    dt = ctypes.c_double(0)
    rep = MaraReport()

    def step():
        L.oracle_courant_dt(ctypes.byref(desc), ptr, ctypes.byref(dt))
        L.oracle_advance(ctypes.byref(desc), ptr, 0.0, CFL * dt.value, ctypes.byref(rep))
    step()  # warm-up step
    steps, t0 = 0, time.perf_counter()
    while True:
        step()
        steps += 1
        el = time.perf_counter() - t0
        if (el > target_s and steps >= min_steps) or steps >= max_steps:
            break
    return n_sample ** 3 * steps / el, steps, el


def cpu_baseline(n_sample, target_s=12.0):
    """The oracle (CPU restatement of the reference step, its -O3 timing
    build) on the same workload, OpenMP threads = the host CPU share this
    process may use (the affinity mask, capped by OMP_NUM_THREADS where the
    environment sets one), plus a shorter one-thread run."""
    affinity = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(range(os.cpu_count()))
    env_cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(len(affinity), env_cap) if env_cap else len(affinity)
    rate, steps, el = _oracle_rate(n_sample, threads, target_s)
    rate1, steps1, el1 = _oracle_rate(32, 1, 4.0, min_steps=5, max_steps=200)
    return {"value": rate, "unit": "zone-updates/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(), "os_cpu_count": os.cpu_count(), "affinity_cpus": len(affinity),
            "omp_num_threads_env": env_cap or None, "one_thread_value": rate1,
            "sample": f"turbulence_3d at {n_sample}^3 (same generator and scheme), {steps} steps after 1 warm-up "
                      f"step, {el:.1f} s; the CPU restatement of the reference step (oracle/), built -O3 "
                      f"(x86-64-v3, contraction on), OpenMP over zones on {threads} threads; one thread: 32^3, "
                      f"{steps1} steps, {el1:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=0, help="turbulence: cells per axis (default 256)")
    ap.add_argument("--cpu-sample", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", default="all",
                    choices=["all", "turbulence", "mhd2d", "rmhd3d", "sod", "serving", "copy", "persistent"])
    ap.add_argument("--requests", type=int, default=1024, help="serving: stream length")
    ap.add_argument("--seq-sample", type=int, default=64, help="serving: requests timed one at a time")
    ap.add_argument("--budget-gib", type=float, default=64.0, help="serving: HBM pool")
    ap.add_argument("--chunk", type=int, default=16, help="serving: steps per batch launch sequence")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--no-graphs", action="store_true", help="launch every kernel from the host")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(a multi-GPU run would shard the mesh into one block per GPU with a halo exchange)",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "copy":
        return copy_workload(args)
    base = {"metric": METRIC, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None}
    if args.workload in ("sod", "mhd2d", "rmhd3d", "serving", "persistent"):
        leg = {"sod": sod_leg, "mhd2d": mhd2d_leg, "rmhd3d": rmhd_leg, "persistent": persistent_leg,
               "serving": lambda a: serving_leg(a, a.requests, a.seq_sample)}[args.workload](args)
        out = dict(base)
        out.update(leg)
        print(json.dumps(out), flush=True)
        return

    d64, res, transfers, l1_fast10, l1_mixed100, n_mixed, turbulence_state = turbulence_leg(args)
    zones = int(np.prod([d64.cells[a] for a in range(3)]))
    fast, exact, mixed = res["fp64_fast"], res["fp64"], res["mixed"]
    workload = f"turbulence3d_{d64.cells[0]}"
    out = dict(base)
    out.update({
        "value": fast["value"], "unit": "zone-updates/s", "ms_per_step": fast["el"] / args.steps * 1e3,
        "dtype": "f64",
        "data": f"synthetic: seeded Gaussian random velocity field (seed {problems.SEED_TURBULENCE})",
        "config": {"workload": f"3-D decaying turbulence {d64.cells[0]}^3, hydro gamma 5/3, PLM theta 1.5 + HLLC, "
                               f"RK2, CFL {CFL}, periodic", "zones": zones,
                   "form": "fp64_fast: every operation binary64 (FMA, refined reciprocals and square roots), "
                           "held to relative L1 <= 1e-12 of the reference after 10 steps",
                   "rel_l1_vs_bit_exact_fp64_after_10_steps": l1_fast10,
                   "launch_shape": fast.get("launch_shape")},
        "roofline": roofline(d64, fast["stage_ms"], fast["kernel"], traffic_from_profile("stage_fast", workload)),
        "fp64_exact": {"value": exact["value"], "ms_per_step": exact["el"] / args.steps * 1e3,
                       "stage_ms": exact["stage_ms"], "kernel": exact["kernel"],
                       "roofline_frac": roofline(d64, exact["stage_ms"], exact["kernel"])["frac"],
                       "traffic": traffic_from_profile("stage_fp64", workload),
                       "launch_shape": exact.get("launch_shape"),
                       "note": "bit-identical to the CPU restatement of the reference step"},
        "mixed": {"value": mixed["value"], "ms_per_step": mixed["el"] / args.steps * 1e3,
                  "stage_ms": mixed["stage_ms"], "kernel": mixed["kernel"],
                  "speedup_vs_fp64_exact": mixed["value"] / exact["value"],
                  "speedup_vs_fp64_fast": mixed["value"] / fast["value"],
                  f"rel_l1_vs_fp64_after_{n_mixed}_steps": l1_mixed100,
                  "roofline_frac": roofline(d64, mixed["stage_ms"], mixed["kernel"])["frac"],
                  "traffic": traffic_from_profile("stage_mixed", workload),
                  "launch_shape": mixed.get("launch_shape")},
        "host_transfer": {"upload_s": transfers["upload_s"], "download_s": transfers["download_s"],
                          "pcie_inclusive_value": zones * args.steps
                          / (fast["el"] + transfers["upload_s"] + transfers["download_s"])},
    })
    if args.workload in ("all", "turbulence"):
        out["blocks"] = blocks_leg(args, d64, turbulence_state)
    if args.workload == "all":
        configs = {}
        for name, fn in (("sod", sod_leg), ("mhd2d", mhd2d_leg), ("rmhd3d", rmhd_leg),
                         ("serving", lambda a: serving_leg(a, a.requests, a.seq_sample))):
            try:
                configs[name] = fn(args)
            except (MaraError, RuntimeError) as e:
                configs[name] = {"error": str(e)}
                log(f"{name}: {e}")
        out["configs"] = configs
    if not args.no_cpu_baseline:
        log(f"CPU baseline at {args.cpu_sample}^3")
        out["cpu_baseline"] = cpu_baseline(args.cpu_sample)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
