"""Benchmark of the MI355X EM-PIC time step (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 1]

A step is one Plasma::Step of the beam-plasma workload: H half step, fused
gather/Boris push/charge-conserving deposit/departures, arrivals, J fold,
H second half, E update, periodic E, field energy.  Inputs are resident in
HBM before the timed region.  Prints ONE JSON line:

  value          particle pushes/s of the whole step (particles x steps / s)
  roofline       dominant kernel (the fused push/deposit): algorithmic bytes
                 per launch / its mean duration from HIP events recorded on
                 the stream it runs on, against the 8 TB/s HBM peak
  cpu_baseline   the FP64 CPU restatement (oracle/) timed on this host on a
                 bounded slab of the same workload
"""
import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload_name(cfg, idx, ppc):
    n = cfg["n"]
    return (f"config{idx}: beam-plasma {n[0]}x{n[1]}x{n[2]} cells, {sum(ppc)} ppc "
            f"(ions {ppc[0]}, electrons {ppc[1]}, beam {ppc[2]}), h=0.1, tau=0.05, periodic")


def algorithmic_bytes_push(P, ncell, width):
    """Push+deposit algorithmic HBM bytes per launch: each particle's position
    and momentum read and written (2 x 6 x width), plus per cell the 6 field
    components read once and the 3 current components read-modify-written."""
    per_cell = (6 + 2 * 3) * width if width == 8 else 6 * 4 + 2 * 3 * 8
    return P * 2 * 6 * width + ncell * per_cell


def cpu_baseline(pic, cfg, slab, steps=2, threads=1):
    import oracle
    n = (slab, cfg["n"][1], cfg["n"][2])
    params, parts = pic.generate_workload(cfg, stream_mode=1, scale_n=n, nthreads=8)
    o = oracle.Oracle(n, (params.lx, params.ly, params.lz), params.tau, threads=threads)
    for (sort, x, y, z, px, py, pz, m, qm) in parts:
        o.load_particles(sort, x, y, z, px, py, pz, m, qm)
    P = o.particle_count()
    o.step(1)
    t0 = time.perf_counter()
    for nt in range(2, 2 + steps):
        o.step(nt)
    dt = time.perf_counter() - t0
    o.close()
    return {"value": P * steps / dt, "unit": "particle pushes/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle FP64 restatement of Plasma::Step, full steps on a "
                      f"{n[0]}x{n[1]}x{n[2]}-cell slab of the same workload "
                      f"({P} particles), {steps} timed steps after 1 warm-up, "
                      f"{dt:.1f} s, single thread"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-slab", type=int, default=8)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the step would shard as x-slabs; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch
    pic = importlib.import_module("0buf_amd")
    cfg = pic.beam_plasma_config(args.config)
    ppc = tuple(s.ppc for s in cfg["species"])
    log(f"generating {workload_name(cfg, args.config, ppc)}")
    t0 = time.perf_counter()
    params, parts = pic.generate_workload(cfg, stream_mode=1, nthreads=16)
    log(f"generated {sum(len(p[1]) for p in parts)} particles in {time.perf_counter() - t0:.1f} s")
    g = pic.Pic(params)
    for (sort, x, y, z, px, py, pz, m, qm) in parts:
        g.load_particles(sort, x, y, z, px, py, pz, m, qm)
    del parts
    if cfg["bx0"]:
        g.load_field(pic.HX, np.full(g.G, cfg["bx0"]))
    P = g.particle_count()
    ncell = g.ncell
    log(f"loaded {P} particles into {ncell} cells")
    torch.cuda.synchronize()
    for nt in range(1, args.warmup + 1):
        g.step(nt)
    log(f"warm-up done ({args.warmup} steps)")
    g.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e = 0.0
    for nt in range(args.warmup + 1, args.warmup + args.steps + 1):
        e = g.step(nt)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    prof = g.profile_read()
    st = g.status()
    P_end = g.particle_count()
    log(f"timed {args.steps} steps in {dt:.3f} s; energy {e:.6e}; particles {P_end}")

    ms_step = dt * 1e3 / args.steps
    pushes = P * args.steps / dt
    push_ms = prof["push_deposit"][0] / max(prof["push_deposit"][1], 1)
    alg = algorithmic_bytes_push(P, ncell, 8)
    achieved = alg / (push_ms * 1e-3) / 1e9
    yee_ms = sum(prof[k][0] for k in ("h_half", "h2", "e", "e_periodic", "fold_j")) / args.steps
    stage_ms = {k: round(v[0] / max(v[1], 1), 4) for k, v in prof.items()}

    traffic = None
    tf = os.path.join(ROOT, "profiles", "push_deposit_f64_traffic.json")
    if os.path.exists(tf):
        try:
            traffic = json.load(open(tf)).get("bytes_per_launch")
        except Exception:
            traffic = None

    cpu = None
    if not args.no_cpu_baseline:
        log("timing the CPU baseline (oracle) on a slab of the workload")
        cpu = cpu_baseline(pic, cfg, args.cpu_slab)
        log(f"cpu baseline {cpu['value']:.3e} pushes/s")

    out = {
        "metric": "particle pushes/s/GPU (FP64 plain form, config 1)",
        "value": pushes,
        "unit": "particle pushes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic seeded beam-plasma (Wichmann-Hill per-cell streams)",
        "config": {"workload": workload_name(cfg, args.config, ppc), "particles": P,
                   "cells": ncell, "precision": "fp64 plain form"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_push_deposit (gather+Boris+VB deposit+departures)",
                     "algorithmic_bytes_per_launch": alg, "mean_launch_ms": push_ms},
        "cpu_baseline": cpu,
        "stage_ms": stage_ms,
        "yee_cell_updates_per_s": ncell / (yee_ms * 1e-3) if yee_ms > 0 else None,
        "particles_end": P_end,
        "status": st,
    }
    g.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
