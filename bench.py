"""Benchmark of the MI355X EM-PIC time step (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 1]

A step is one Plasma::Step of the beam-plasma workload: H half step, fused
gather/Boris push/charge-conserving deposit/departures, arrivals, J fold,
H second half, E update, periodic E, field energy.  Inputs are resident in
HBM before the timed region.  Prints ONE JSON line:

  value          particle pushes/s of the whole step (particles x steps / s)
                 of the FP32 precision form, with error_vs_fp64 (its energy
                 history against the FP64 run of the same particles) and
                 speedup_vs_fp64
  roofline       its dominant kernel (the fused push/deposit): algorithmic
                 bytes per launch / its mean duration from HIP events recorded
                 on the stream it runs on, against the 8 TB/s HBM peak
  fp64           the FP64 plain form (the parity form) on the same particles:
                 pushes/s, roofline, stage times
  cpu_baseline   the FP64 CPU restatement (oracle/) timed on this host on a
                 bounded slab of the same workload
  adjoint        config 3 gradient (differentiable form): wall time of
                 forward-with-checkpoints + recompute + adjoint sweep over a
                 plain forward run (adjoint_forward_ratio)
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


def cpu_baseline(pic, cfg, slab, steps=10, threads=1, all_threads=None):
    """The FP64 CPU restatement (oracle/) on a slab of the workload: `steps`
    full steps after one warm-up, single-threaded (about 10 s of CPU work at
    the default 16x64x64 slab), then the same with the oracle's cell loops
    spread over `all_threads` host threads."""
    import oracle
    n = (slab, cfg["n"][1], cfg["n"][2])
    params, parts = pic.generate_workload(cfg, stream_mode=1, scale_n=n, nthreads=8)

    def timed(nthreads):
        o = oracle.Oracle(n, (params.lx, params.ly, params.lz), params.tau, threads=nthreads)
        for (sort, x, y, z, px, py, pz, m, qm) in parts:
            o.load_particles(sort, x, y, z, px, py, pz, m, qm)
        P = o.particle_count()
        o.step(1)
        t0 = time.perf_counter()
        for nt in range(2, 2 + steps):
            o.step(nt)
        dt = time.perf_counter() - t0
        o.close()
        return P, dt
    P, dt = timed(threads)
    out = {"value": P * steps / dt, "unit": "particle pushes/s", "cores": threads,
           "kind": "port",
           "sample": f"oracle FP64 restatement of Plasma::Step, full steps on a "
                     f"{n[0]}x{n[1]}x{n[2]}-cell slab of the same workload "
                     f"({P} particles), {steps} timed steps after 1 warm-up, "
                     f"{dt:.1f} s, {threads} thread(s)"}
    if all_threads and all_threads > threads:
        P, dt = timed(all_threads)
        out["all_cores"] = {"value": P * steps / dt, "cores": all_threads,
                            "sample": f"the same slab and steps, cell loops over "
                                      f"{all_threads} threads, {dt:.1f} s"}
    return out


def run_form(pic, params, cfg, parts, bx0, precision, steps, warmup, seed=12345):
    """Load the workload into a context of the given precision (drawn on the
    GPU by the device initializer, or the host-generated `parts`), run warmup
    untimed steps then `steps` timed steps (torch.cuda.synchronize on both
    sides); per-stage mean launch times from HIP events on the context's
    stream."""
    import torch
    params.precision = precision
    g = pic.Pic(params)
    if parts is None:
        g.generate_device(cfg["species"], seed=seed)
    else:
        for (sort, x, y, z, px, py, pz, m, qm) in parts:
            g.load_particles(sort, x, y, z, px, py, pz, m, qm)
    if bx0:
        g.load_field(pic.HX, np.full(g.G, bx0))
    P = g.particle_count()
    torch.cuda.synchronize()
    energies = []
    for nt in range(1, warmup + 1):
        energies.append(g.step(nt))
    g.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for nt in range(warmup + 1, warmup + steps + 1):
        energies.append(g.step(nt))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    res = {"P": P, "ncell": g.ncell, "dt": dt, "prof": g.profile_read(), "status": g.status(),
           "P_end": g.particle_count(), "energies": np.array(energies)}
    g.close()
    return res


def push_roofline(res, width):
    prof = res["prof"]
    push_ms = prof["push_deposit"][0] / max(prof["push_deposit"][1], 1)
    alg = algorithmic_bytes_push(res["P"], res["ncell"], width)
    achieved = alg / (push_ms * 1e-3) / 1e9
    return push_ms, alg, achieved


def yee_roofline(ncell, yee_ms, width):
    """Field stages of one step (H half step, H second half, E update, periodic
    ghosts, J fold) against the HBM roofline.  The fused minimum per cell reads
    E, H and J once and writes E and H once: 5 x 3 x 8 = 120 B in FP64; in the
    FP32 form (FP32 E and H, int64 J) 4 x 3 x 4 + 3 x 8 = 72 B."""
    per_cell = 120 if width == 8 else 72
    if yee_ms <= 0:
        return None
    achieved = ncell * per_cell / (yee_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes_per_cell": per_cell,
            "ms_per_step": yee_ms}


def adjoint_ratio(pic):
    """Config 3 (64x16x16 cells, 32 ppc, 200 steps, checkpoints every 20):
    wall time of a gradient (forward with checkpoints + recompute + adjoint
    sweep) against a plain forward run of the same form.  Each call builds
    its own context (host particles loaded, checkpoint memory allocated), so
    both are the best of three runs after a warm-up of each."""
    import torch
    ag = importlib.import_module("0buf_amd.autograd")
    bp = ag.BeamPlasma(3)
    theta = torch.tensor([0.3] + [1e-3] * 8, dtype=torch.float64, requires_grad=True)
    bp.loss_value(theta.detach())                    # warm-up
    t_fwd = float("inf")
    for _ in range(3):
        t0 = time.perf_counter()
        bp.loss_value(theta.detach())
        t_fwd = min(t_fwd, time.perf_counter() - t0)
    runs = []
    for k in range(4):                               # the first is a warm-up
        theta.grad = None
        t0 = time.perf_counter()
        L = bp.loss(theta)
        t1 = time.perf_counter()
        L.backward()
        t2 = time.perf_counter()
        if k:
            runs.append((t2 - t0, t1 - t0, t2 - t1, tuple(bp.run.last.adjoint_timing())))
    grad_s, rec_s, back_s, dev = min(runs)
    bp.run.last.close()
    return {"workload": "config3: beam-plasma 64x16x16 cells, 32 ppc (ions 8, electrons 16, "
                        "beam 8), 200 steps, FP64 differentiable form, checkpoint every 20 "
                        "steps, loss = sum of field energy over steps 150..200, parameters "
                        "= beam drift + 8 displacement amplitudes",
            "forward_s": t_fwd, "gradient_s": grad_s, "ratio": grad_s / t_fwd,
            "forward_record_s": rec_s, "backward_s": back_s,
            "device_ms": {"forward_with_checkpoints": dev[0], "recompute": dev[1],
                          "adjoint_steps": dev[2]},
            "grad_norm": float(theta.grad.norm())}


def adjoint_device(pic, idx, precision, first=150, last=200, ckpt_every=20, seed=12345):
    """Gradient cost of the differentiable form at full size with the particles
    drawn on the GPU: a plain forward of the workload's steps, then
    forward-with-checkpoints + the reverse sweep of the loss sum_{t=first..last}
    W_t (steps after `last` have a zero adjoint and are not recomputed), both
    from the same s_0 on one context whose gradient memory is allocated."""
    import torch
    cfg = pic.beam_plasma_config(idx)
    n = cfg["n"]
    T = cfg["steps"]
    params = pic.make_params(n, tuple(k * 0.1 for k in n), cfg["tau"], precision=precision,
                             differentiable=1, ghost_mode=1,
                             departure_cap=cfg.get("departure_cap", 50))

    def ctx():
        g = pic.Pic(params)
        g.generate_device(cfg["species"], seed=seed)
        if cfg["bx0"]:
            g.load_field(pic.HX, np.full(g.G, cfg["bx0"]))
        torch.cuda.synchronize()
        return g
    g = ctx()
    P = g.particle_count()
    w = np.zeros(T)
    w[first - 1:last] = 1.0
    # one untimed gradient allocates the checkpoint and record memory; the
    # context is then rewound to s_0 for each timed leg
    g.forward_record(T, ckpt_every)
    g.adjoint(weights=w)
    g.rewind()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    W = g.run(1, T)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    g.rewind()
    t0 = time.perf_counter()
    Wr = g.forward_record(T, ckpt_every)
    t1 = time.perf_counter()
    gr = g.adjoint(weights=w)
    t2 = time.perf_counter()
    dev = g.adjoint_timing()
    g.close()
    pg = np.sqrt(sum(float(np.dot(gr[k], gr[k])) for k in ("x", "y", "z", "px", "py", "pz")))
    return {"workload": f"{workload_name(cfg, idx, tuple(s.ppc for s in cfg['species']))}, "
                        f"{T} steps, differentiable form "
                        f"{'FP32 (FP64 gradient accumulation)' if precision else 'FP64'}, "
                        f"ghost_mode 1, checkpoint every {ckpt_every}, loss = sum of field "
                        f"energy over steps {first}..{last}",
            "particles": P, "forward_s": t_fwd, "forward_record_s": t1 - t0,
            "backward_s": t2 - t1, "gradient_s": t2 - t0, "ratio": (t2 - t0) / t_fwd,
            "reversed_steps": last,
            "backward_per_reversed_step_over_forward_step": (t2 - t1) / last / (t_fwd / T),
            "device_ms": {"forward_with_checkpoints": dev[0], "recompute": dev[1],
                          "adjoint_steps": dev[2]},
            "record_matches_forward": bool(np.array_equal(W, Wr)),
            "particle_grad_norm": pg}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint-config4", action="store_true",
                    help="also time the config-4 gradient (FP64 and FP32 differentiable "
                         "forms, particles drawn on the GPU)")
    ap.add_argument("--adjoint-config4-only", action="store_true",
                    help="only the config-4 gradient timing (for profiling its kernels)")
    ap.add_argument("--cpu-slab", type=int, default=16)
    ap.add_argument("--host-init", action="store_true",
                    help="generate the particles on the host and upload them instead of "
                         "drawing them on the GPU (same particles, bit for bit)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a plain elementwise kernel (dst = src + 1, FP32) that reads "
                         "and writes BYTES each, 5 times: the known-size reference for the "
                         "FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            torch.add(src, 1.0, out=dst)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": n * 4, "bytes_written": n * 4,
                          "launches": 5}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the step would shard as x-slabs with halo and leaver exchange)", file=sys.stderr)
        sys.exit(2)

    # torch carries its own HIP runtime beside the system one the library links;
    # it initialises first, before any library call touches the device
    import torch
    torch.cuda.init()
    torch.cuda.synchronize()
    pic = importlib.import_module("0buf_amd")
    if args.adjoint_config4_only:
        out = {}
        for prec, nm in ((pic.FP64, "fp64"), (pic.FP32, "fp32")):
            log(f"timing the config 4 gradient, {nm}")
            out[nm] = adjoint_device(pic, 4, prec)
        print(json.dumps({"adjoint_config4": out}))
        return
    cfg = pic.beam_plasma_config(args.config)
    ppc = tuple(s.ppc for s in cfg["species"])
    log(f"workload {workload_name(cfg, args.config, ppc)}")
    parts = None
    n = cfg["n"]
    params = pic.make_params(n, tuple(k * 0.1 for k in n), cfg["tau"],
                             departure_cap=cfg.get("departure_cap", 50))
    if args.host_init:
        t0 = time.perf_counter()
        params, parts = pic.generate_workload(cfg, stream_mode=1, nthreads=16)
        log(f"generated {sum(len(p[1]) for p in parts)} particles on the host in "
            f"{time.perf_counter() - t0:.1f} s")

    r64 = run_form(pic, params, cfg, parts, cfg["bx0"], pic.FP64, args.steps, args.warmup)
    log(f"FP64: {r64['P']} particles, {args.steps} steps in {r64['dt']:.3f} s")
    r32 = None
    if not args.no_fp32:
        r32 = run_form(pic, params, cfg, parts, cfg["bx0"], pic.FP32, args.steps, args.warmup)
        log(f"FP32: {args.steps} steps in {r32['dt']:.3f} s")
    del parts

    P, ncell = r64["P"], r64["ncell"]

    def measured_traffic(name, alg):
        """HBM bytes per launch of a push kernel from the committed counter
        summary (FETCH_SIZE + WRITE_SIZE, calibrated), or None when that
        summary was measured on another workload (other algorithmic bytes)"""
        tf = os.path.join(ROOT, "profiles", name)
        try:
            rec = json.load(open(tf))
        except (OSError, ValueError):
            return None
        if rec.get("algorithmic_bytes") != alg:
            return None
        return rec.get("bytes_per_launch")

    def form_result(res, width, kernel, traffic_file):
        push_ms, alg, achieved = push_roofline(res, width)
        prof = res["prof"]
        yee_ms = sum(prof[k][0] for k in ("h_half", "h2", "e", "e_periodic",
                                          "fold_j")) / args.steps
        return {"value": P * args.steps / res["dt"], "unit": "particle pushes/s",
                "ms_per_step": res["dt"] * 1e3 / args.steps,
                "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                             "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                             "traffic": measured_traffic(traffic_file, alg), "kernel": kernel,
                             "algorithmic_bytes_per_launch": alg, "mean_launch_ms": push_ms},
                "stage_ms": {k: round(v[0] / max(v[1], 1), 4) for k, v in prof.items()},
                "yee_cell_updates_per_s": ncell / (yee_ms * 1e-3) if yee_ms > 0 else None,
                "yee_roofline": yee_roofline(ncell, yee_ms, width),
                "particles_end": res["P_end"], "status": res["status"]}

    fp64 = form_result(r64, 8, "k_push_deposit (gather+Boris+VB deposit+departures, FP64)",
                       "push_deposit_f64_traffic.json")
    fp32 = None
    if r32 is not None:
        fp32 = form_result(r32, 4, "k_push_deposit32 (gather+Boris+VB deposit+departures, "
                                   "FP32 offsets, int64 current)", "push_deposit_f32_traffic.json")
        rel = np.abs(r32["energies"] - r64["energies"]) / np.abs(r64["energies"])
        fp32["speedup_vs_fp64"] = r64["dt"] / r32["dt"]
        fp32["error_vs_fp64"] = {"energy_history_max_rel": float(rel.max()),
                                 "steps_compared": int(len(rel))}

    adj = None
    if not args.no_adjoint:
        log("timing the differentiable form (config 3 gradient)")
        adj = adjoint_ratio(pic)
        log(f"adjoint/forward ratio {adj['ratio']:.2f}")

    adj4 = None
    if args.adjoint_config4:
        adj4 = {}
        for prec, nm in ((pic.FP64, "fp64"), (pic.FP32, "fp32")):
            log(f"timing the config 4 gradient, {nm}")
            adj4[nm] = adjoint_device(pic, 4, prec)
            log(f"config 4 {nm}: gradient/forward {adj4[nm]['ratio']:.2f}")

    cpu = None
    if not args.no_cpu_baseline:
        log("timing the CPU baseline (oracle) on a slab of the workload")
        cpu = cpu_baseline(pic, cfg, args.cpu_slab,
                           all_threads=min(16, len(os.sched_getaffinity(0))))
        log(f"cpu baseline {cpu['value']:.3e} pushes/s")

    # the headline is the FP32 precision form with its error against the FP64
    # form stated; the FP64 plain form (the parity form) stands beside it
    head, form = (fp32, "fp32") if fp32 is not None else (fp64, "fp64")
    out = {
        "metric": "particle pushes/s/GPU (FP32 vs FP64 at stated error) + adjoint/forward ratio",
        "value": head["value"],
        "unit": "particle pushes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if form == "fp32" else "f64",
        "data": "synthetic seeded beam-plasma (Wichmann-Hill per-cell streams, seed 12345, "
                "drawn on the GPU by the device initializer)",
        "config": {"workload": workload_name(cfg, args.config, ppc), "particles": P,
                   "cells": ncell,
                   "form": ("FP32 precision form: FP32 push, gather and Yee curls on cell-local "
                            "offsets, int64 fixed-point current, FP64 energy reduction")
                   if form == "fp32" else "FP64 plain form"},
        "roofline": head["roofline"],
        "error_vs_fp64": head.get("error_vs_fp64"),
        "speedup_vs_fp64": head.get("speedup_vs_fp64"),
        "adjoint_forward_ratio": adj["ratio"] if adj else None,
        "cpu_baseline": cpu,
        "fp64": fp64,
        "adjoint": adj,
        "adjoint_config4": adj4,
        "stage_ms": head["stage_ms"],
        "yee_cell_updates_per_s": head["yee_cell_updates_per_s"],
        "yee_roofline": head["yee_roofline"],
        "particles_end": head["particles_end"],
        "status": head["status"],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
