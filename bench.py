"""Benchmark of the MI355X EM-PIC time step (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 2]

A step is one Plasma::Step of the beam-plasma workload: H half step, fused
gather/Boris push/charge-conserving deposit/departures, arrivals, J fold,
H second half, E update, periodic E, field energy.  Inputs are resident in
HBM before the timed region, which times K calls of pic_step exactly as a
caller makes them.  The default workload is the FP32-vs-FP64 comparison:
512x128x128 cells, 24 particles per cell per species (603,979,776
particles), a uniform Bx0 = 0.5.  Prints ONE JSON line:

  value          particle pushes/s of the whole step (particles x steps / s)
                 of the FP32 precision form, with error_vs_fp64 (its
                 field-energy history over every step run and its E field at
                 the end, against the FP64 run of the same particles) and
                 speedup_vs_fp64
  roofline       its dominant kernel (the fused push/deposit): algorithmic
                 bytes per launch / its mean duration from HIP events recorded
                 on the stream it runs on (a few stage-timed steps after the
                 timed region), against the 8 TB/s HBM peak
  fp64           the FP64 plain form (the parity form) on the same particles:
                 pushes/s, roofline, stage times
  secondary      the same pair on a second workload (default: 256x64x64
                 cells, 64 particles per cell, 67,108,864 particles)
  cpu_baseline   the FP64 CPU restatement (oracle/) timed on this host on an
                 x-slab of the same workload (median step after 3 warm-ups)
  adjoint        gradient of the differentiable form on 64x16x16 cells (200
                 steps): wall time of forward-with-checkpoints + recompute +
                 adjoint sweep over a plain forward run (adjoint_forward_ratio)
  adjoint_config4  the same at 256x64x64 cells, 48 ppc, Gaussian beam, 400
                 steps, FP64 and FP32 differentiable forms
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


def cpu_baseline(pic, cfg, slab, steps=7, warm=3, threads=1, all_threads=None):
    """The FP64 CPU restatement (oracle/) on an x-slab of the workload (the
    full y-z cross-section, `slab` planes of cells, the workload's species
    and Bx0): `warm` untimed steps, then `steps` steps timed one by one;
    the value is particles / median step time.  Single-threaded (the
    reference's serial semantics), then the same with the oracle's cell
    loops on `all_threads` host threads."""
    import oracle
    n = (slab, cfg["n"][1], cfg["n"][2])
    params, parts = pic.generate_workload(cfg, stream_mode=1, scale_n=n, nthreads=8)

    def timed(nthreads):
        o = oracle.Oracle(n, (params.lx, params.ly, params.lz), params.tau, threads=nthreads)
        for (sort, x, y, z, px, py, pz, m, qm) in parts:
            o.load_particles(sort, x, y, z, px, py, pz, m, qm)
        if cfg["bx0"]:
            o.load_field(pic.HX, np.full((n[0] + 2) * (n[1] + 2) * (n[2] + 2), cfg["bx0"]))
        P = o.particle_count()
        for nt in range(1, warm + 1):
            o.step(nt)
        dts = []
        for nt in range(warm + 1, warm + steps + 1):
            t0 = time.perf_counter()
            o.step(nt)
            dts.append(time.perf_counter() - t0)
        o.close()
        return P, float(np.median(dts)), sum(dts)
    P, med, tot = timed(threads)
    out = {"value": P / med, "unit": "particle pushes/s", "cores": threads,
           "kind": "port",
           "sample": f"oracle FP64 restatement of Plasma::Step, full steps on a "
                     f"{n[0]}x{n[1]}x{n[2]}-cell x-slab of the same workload "
                     f"({P} particles{', Bx0 = %g' % cfg['bx0'] if cfg['bx0'] else ''}), "
                     f"median of {steps} timed steps after {warm} warm-up steps "
                     f"({med:.2f} s per step, {tot:.1f} s timed), {threads} thread(s)"}
    if all_threads and all_threads > threads:
        P, med, tot = timed(all_threads)
        out["all_cores"] = {"value": P / med, "cores": all_threads,
                            "sample": f"the same slab and steps, cell loops on {all_threads} "
                                      f"threads ({med:.3f} s per step)"}
    return out


def run_form(pic, params, cfg, parts, bx0, precision, steps, warmup, stage_steps=5,
             keep_e=False, seed=12345):
    """Load the workload into a context of the given precision (drawn on the
    GPU by the device initializer, or the host-generated `parts`); `warmup`
    untimed steps; `steps` timed steps of pic_step as a caller runs it (stage
    timing off, so the arrival runs on its second stream beside the field
    stages), bracketed by torch.cuda.synchronize; then `stage_steps` more
    steps with stage timing on (HIP events on the context's stream around
    each stage) for the per-stage and push-kernel times."""
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
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for nt in range(warmup + 1, warmup + steps + 1):
        energies.append(g.step(nt))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    g.profile(True)
    for nt in range(warmup + steps + 1, warmup + steps + stage_steps + 1):
        energies.append(g.step(nt))
    torch.cuda.synchronize()
    res = {"P": P, "ncell": g.ncell, "dt": dt, "prof": g.profile_read(), "status": g.status(),
           "P_end": g.particle_count(), "energies": np.array(energies)}
    if keep_e:
        res["E"] = [g.get_field(k) for k in (pic.EX, pic.EY, pic.EZ)]
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
    FP32 form (FP32 E and H, int64 J) 4 x 3 x 4 + 3 x 8 = 72 B.  The step's
    kernels also keep the reference's Q between the two H half steps and zero
    J before the push: k_h_half moves 120 / 72 B per node (FP64 / FP32) and
    k_h2_e 168 / 84 B, reported beside the minimum as the bytes moved."""
    per_cell = 120 if width == 8 else 72
    moved = 288 if width == 8 else 156
    if yee_ms <= 0:
        return None
    achieved = ncell * per_cell / (yee_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes_per_cell": per_cell,
            "bytes_moved_per_cell": moved,
            "achieved_moved": ncell * moved / (yee_ms * 1e-3) / 1e9,
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


def adjoint_device(pic, idx, precision):
    """Gradient cost of the differentiable form at full size, particles drawn
    on the GPU (0buf_amd.autograd.device_gradient)"""
    ag = importlib.import_module("0buf_amd.autograd")
    r = ag.device_gradient(idx, precision)
    cfg = pic.beam_plasma_config(idx)
    r["workload"] = (f"{workload_name(cfg, idx, tuple(s.ppc for s in cfg['species']))}, "
                     f"{cfg['steps']} steps, differentiable form "
                     f"{'FP32 (FP64 gradient accumulation)' if precision else 'FP64'}, "
                     f"ghost_mode 1, checkpoint every {r['ckpt_every']}, loss = sum of field "
                     f"energy over steps {r['first']}..{r['last']}")
    return r


def measured_traffic(kernel, alg):
    """HBM bytes per launch of a push kernel from a committed counter summary
    (profiles/*traffic*.json: FETCH_SIZE + WRITE_SIZE, calibrated on a copy
    kernel) measured on this workload (same kernel, same algorithmic bytes),
    else None"""
    import glob
    for tf in sorted(glob.glob(os.path.join(ROOT, "profiles", "*traffic*.json"))):
        try:
            rec = json.load(open(tf))
        except (OSError, ValueError):
            continue
        if rec.get("kernel") == kernel and rec.get("algorithmic_bytes") == alg:
            return rec.get("bytes_per_launch")
    return None


def form_result(res, width, kernel, kernel_desc, steps):
    P, ncell = res["P"], res["ncell"]
    push_ms, alg, achieved = push_roofline(res, width)
    prof = res["prof"]
    nprof = max(prof["push_deposit"][1], 1)
    yee_ms = sum(prof[k][0] for k in ("h_half", "h2", "e", "e_periodic", "fold_j")) / nprof
    return {"value": P * steps / res["dt"], "unit": "particle pushes/s",
            "ms_per_step": res["dt"] * 1e3 / steps,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "traffic": measured_traffic(kernel, alg),
                         "kernel": f"{kernel} ({kernel_desc})",
                         "algorithmic_bytes_per_launch": alg, "mean_launch_ms": push_ms,
                         "launches_timed": nprof},
            "stage_ms": {k: round(v[0] / max(v[1], 1), 4) for k, v in prof.items()},
            "yee_cell_updates_per_s": ncell / (yee_ms * 1e-3) if yee_ms > 0 else None,
            "yee_roofline": yee_roofline(ncell, yee_ms, width),
            "particles_end": res["P_end"], "status": res["status"]}


K64 = ("k_push_deposit", "gather+Boris+VB deposit+departures, FP64")
K32 = ("k_push_deposit32", "gather+Boris+VB deposit+departures, FP32 offsets, int64 current")


def compare_forms(r64, r32):
    """FP32 form against the FP64 form from the same particles: field-energy
    history over every step run (warm-up, timed and stage-timed), and the
    electric field at the end, max-norm over all nodes relative to the FP64
    field's max-norm"""
    rel = np.abs(r32["energies"] - r64["energies"]) / np.abs(r64["energies"])
    out = {"energy_history_max_rel": float(rel.max()), "steps_compared": int(len(rel))}
    if "E" in r64 and "E" in r32:
        out["E_maxnorm_rel_at_end"] = {
            nm: float(np.abs(a - b).max() / max(np.abs(a).max(), 1e-300))
            for nm, a, b in zip(("Ex", "Ey", "Ez"), r64["E"], r32["E"])}
    return out


def run_config(pic, idx, args, fp32=True, keep_e=False):
    cfg = pic.beam_plasma_config(idx)
    ppc = tuple(s.ppc for s in cfg["species"])
    log(f"workload {workload_name(cfg, idx, ppc)}")
    n = cfg["n"]
    params = pic.make_params(n, tuple(k * 0.1 for k in n), cfg["tau"],
                             departure_cap=cfg.get("departure_cap", 50))
    parts = None
    if args.host_init:
        t0 = time.perf_counter()
        params, parts = pic.generate_workload(cfg, stream_mode=1, nthreads=16)
        log(f"generated {sum(len(p[1]) for p in parts)} particles on the host in "
            f"{time.perf_counter() - t0:.1f} s")
    r64 = run_form(pic, params, cfg, parts, cfg["bx0"], pic.FP64, args.steps, args.warmup,
                   keep_e=keep_e)
    log(f"config {idx} FP64: {r64['P']} particles, {args.steps} steps in {r64['dt']:.3f} s")
    out = {"workload": workload_name(cfg, idx, ppc), "particles": r64["P"],
           "cells": r64["ncell"], "bx0": cfg["bx0"],
           "fp64": form_result(r64, 8, *K64, args.steps)}
    if fp32:
        r32 = run_form(pic, params, cfg, parts, cfg["bx0"], pic.FP32, args.steps, args.warmup,
                       keep_e=keep_e)
        log(f"config {idx} FP32: {args.steps} steps in {r32['dt']:.3f} s")
        f32 = form_result(r32, 4, *K32, args.steps)
        f32["speedup_vs_fp64"] = r64["dt"] / r32["dt"]
        f32["error_vs_fp64"] = compare_forms(r64, r32)
        out["fp32"] = f32
    return cfg, ppc, out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", type=int, default=2,
                    help="headline workload (default 2: 512x128x128 cells, 604M particles, "
                         "Bx0 = 0.5, the FP32-vs-FP64 comparison)")
    ap.add_argument("--secondary-config", type=int, default=1,
                    help="second workload timed in both forms (default 1; -1 for none)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip both gradient legs (config 3 and config 4)")
    ap.add_argument("--no-adjoint-config4", action="store_true",
                    help="skip the config-4 gradient (FP64 and FP32 differentiable forms)")
    ap.add_argument("--adjoint-config4-only", action="store_true",
                    help="only the config-4 gradient timing (for profiling its kernels)")
    ap.add_argument("--cpu-slab", type=int, default=0,
                    help="x-planes of the CPU baseline's slab (0: about 2.4M particles, at "
                         "least 3 planes)")
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

    cfg, ppc, head_cfg = run_config(pic, args.config, args, fp32=not args.no_fp32, keep_e=True)
    second = None
    if args.secondary_config >= 0 and args.secondary_config != args.config:
        _, _, second = run_config(pic, args.secondary_config, args, fp32=not args.no_fp32)

    adj = None
    adj4 = None
    if not args.no_adjoint:
        log("timing the differentiable form (config 3 gradient)")
        adj = adjoint_ratio(pic)
        log(f"config 3 adjoint/forward ratio {adj['ratio']:.2f}")
        if not args.no_adjoint_config4:
            adj4 = {}
            for prec, nm in ((pic.FP64, "fp64"), (pic.FP32, "fp32")):
                log(f"timing the config 4 gradient, {nm}")
                adj4[nm] = adjoint_device(pic, 4, prec)
                log(f"config 4 {nm}: gradient/forward {adj4[nm]['ratio']:.2f}")

    cpu = None
    if not args.no_cpu_baseline:
        # at least 3 planes: the 27-neighbour arrival scan needs 3 distinct
        # x-neighbours (as the step itself needs nx >= 3)
        slab = args.cpu_slab or max(3, round(2.4e6 / (cfg["n"][1] * cfg["n"][2] * sum(ppc))))
        log(f"timing the CPU baseline (oracle) on a {slab}-plane slab of the workload")
        cpu = cpu_baseline(pic, cfg, slab, all_threads=min(16, len(os.sched_getaffinity(0))))
        log(f"cpu baseline {cpu['value']:.3e} pushes/s")

    # the headline is the FP32 precision form with its error against the FP64
    # form of the same particles stated; the FP64 plain form (the parity form)
    # stands beside it
    fp64 = head_cfg["fp64"]
    head, form = (head_cfg["fp32"], "fp32") if "fp32" in head_cfg else (fp64, "fp64")
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
        "config": {"workload": head_cfg["workload"], "particles": head_cfg["particles"],
                   "cells": head_cfg["cells"], "bx0": head_cfg["bx0"],
                   "form": ("FP32 precision form: FP32 push, gather and Yee curls on cell-local "
                            "offsets, int64 fixed-point current, FP64 energy reduction")
                   if form == "fp32" else "FP64 plain form"},
        "roofline": head["roofline"],
        "error_vs_fp64": head.get("error_vs_fp64"),
        "speedup_vs_fp64": head.get("speedup_vs_fp64"),
        "adjoint_forward_ratio": adj["ratio"] if adj else None,
        "cpu_baseline": cpu,
        "fp64": fp64,
        "secondary": second,
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
