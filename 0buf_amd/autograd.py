"""torch.autograd.Function over the C ABI of the differentiable form.

EnergyHistory.apply(run, x, y, z, px, py, pz, E, H) runs `run.nsteps` steps of
the step on the GPU (FP64, or the FP32 form with FP64 gradient accumulation,
per `run.params.precision`) from the given initial state and returns the energy
history W_1..W_T (W_t = sum over all ghosted nodes of Ex^2+Ey^2+Ez^2, the
reference's Diagnose, gpu_plasma.h:621-686).  Its backward calls
pic_adjoint: the checkpointed recompute and the hand-written adjoint
kernels.  Particle tensors are in load order (concatenated species, the ids
of the library); E and H are (3, G) ghosted arrays.

The initial-state maps of the differentiable workloads (beam drift, the 8
displacement amplitudes a_k sin(2 pi k x / Lx), the uniform Bx0) are plain
torch code on top, so torch's own autograd carries the adjoint of s_0 to the
parameters.  Tensors live on the CPU here: the GPU work happens inside the
library, whose boundary takes host arrays.
"""
import importlib
import math

import numpy as np
import torch

pic = importlib.import_module("0buf_amd")

PART_KEYS = ("x", "y", "z", "px", "py", "pz")
FIELD_KEYS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


class Run:
    """Grid, species (sort, m, q_m, count in load order), steps and checkpoint
    interval of a differentiable run; `last` keeps the context of the latest
    forward for its backward and timing."""

    def __init__(self, params, species, nsteps, ckpt_every):
        self.params = params
        self.params.differentiable = 1
        self.species = species
        self.nsteps = nsteps
        self.ckpt_every = ckpt_every
        self.last = None
        self.record_tape = False   # the next forward records its decision tape
        self.tape = None           # (words, cells) of the last recorded forward

    def make(self, x, y, z, px, py, pz, E, H):
        g = pic.Pic(self.params)
        off = 0
        for (sort, m, q_m, n) in self.species:
            sl = slice(off, off + n)
            g.load_particles(sort, x[sl], y[sl], z[sl], px[sl], py[sl], pz[sl], m, q_m)
            off += n
        for i in range(3):
            g.load_field(pic.EX + i, E[i])
            g.load_field(pic.HX + i, H[i])
        return g


def _np(t):
    return t.detach().cpu().numpy().astype(np.float64, copy=False)


class EnergyHistory(torch.autograd.Function):
    @staticmethod
    def forward(ctx, run, x, y, z, px, py, pz, E, H):
        if run.last is not None:
            run.last.close()
        g = run.make(*[_np(t) for t in (x, y, z, px, py, pz)], _np(E), _np(H))
        if run.record_tape:
            g.tape(g.TAPE_RECORD, run.nsteps)
        W = g.forward_record(run.nsteps, run.ckpt_every)
        if run.record_tape:
            run.tape = g.tape_get(run.nsteps)
        run.last = g
        ctx.run = run
        ctx.g = g
        return torch.from_numpy(W)

    @staticmethod
    def backward(ctx, gW):
        res = ctx.g.adjoint(weights=_np(gW))
        parts = [torch.from_numpy(res[k].copy()) for k in PART_KEYS]
        E = torch.from_numpy(np.stack([res[k] for k in FIELD_KEYS[:3]]))
        H = torch.from_numpy(np.stack([res[k] for k in FIELD_KEYS[3:]]))
        return (None, *parts, E, H)


def energy_loss_window(W, first, last):
    """sum of W_t for steps t = first..last (1-based, inclusive)"""
    return W[first - 1:last].sum()


# ---------------------------------------------------------------------------
# Differentiable beam-plasma workloads
# ---------------------------------------------------------------------------

class BeamPlasma:
    """Config 3/4 style run.  theta = (drift, a_1..a_8, bx0): the beam's drift
    velocity along x, the amplitudes of the beam displacement
    x -> x + sum_k a_k sin(2 pi k x / Lx), and a uniform external Hx = bx0
    on every node.  Base state: the seeded generator's particles with the
    beam drawn around drift0 and no displacement.

    ghost_mode 1 (every E ghost node periodic) is the default here: with the
    reference's ghosts (0) the field a particle sees jumps when it crosses a
    y face, so the discrete loss is discontinuous in theta and central
    differences at steps above ~1e-7 pick up the jumps."""

    def __init__(self, cfg, drift0=0.3, scale_n=None, nsteps=None, ckpt_every=20,
                 stream_mode=None, ghost_mode=1, precision=pic.FP64, workload=None):
        spec_cfg = pic.beam_plasma_config(cfg) if isinstance(cfg, int) else cfg
        if workload is None:
            workload = pic.generate_workload(spec_cfg, scale_n=scale_n, stream_mode=stream_mode)
        params, parts = workload
        params = pic.PicParams.from_buffer_copy(params)
        params.ghost_mode = ghost_mode
        params.precision = precision
        self.params = params
        self.lx = params.lx
        self.species = []
        cols = [[] for _ in range(6)]
        beam = []
        for (sort, x, y, z, px, py, pz, m, q_m) in parts:
            self.species.append((sort, m, q_m, len(x)))
            for c, a in zip(cols, (x, y, z, px, py, pz)):
                c.append(a)
            beam.append(np.full(len(x), sort == pic.BEAM_ELECTRON))
        base = [torch.from_numpy(np.concatenate(c)) for c in cols]
        self.beam = torch.from_numpy(np.concatenate(beam))
        self.beam_idx = torch.nonzero(self.beam).flatten()
        self.x0, self.y0, self.z0 = base[0], base[1], base[2]
        p = torch.stack(base[3:6])
        u = p / torch.sqrt(1.0 + (p * p).sum(0))
        u[0] = torch.where(self.beam, u[0] - drift0, u[0])
        self.u_th = u          # thermal part of the beam velocity, full velocity otherwise
        self.G = (params.nx + 2) * (params.ny + 2) * (params.nz + 2)
        nsteps = nsteps or spec_cfg["steps"]
        self.run = Run(params, self.species, nsteps, ckpt_every)

    def initial_state(self, theta):
        drift, amps = theta[0], theta[1:9]
        bx0 = theta[9] if theta.numel() > 9 else theta.new_zeros(())
        k = torch.arange(1, 9, dtype=theta.dtype)
        xb = self.x0[self.beam_idx]
        disp = (amps[:, None] * torch.sin(2.0 * math.pi * k[:, None] * xb[None, :] /
                                          self.lx)).sum(0)
        x = self.x0.index_put((self.beam_idx,), xb + disp)
        ux = self.u_th[0].index_put((self.beam_idx,), self.u_th[0][self.beam_idx] + drift)
        u = torch.stack([ux, self.u_th[1], self.u_th[2]])
        p = u / torch.sqrt(1.0 - (u * u).sum(0))
        E = torch.zeros(3, self.G, dtype=theta.dtype)
        H = torch.stack([bx0 * torch.ones(self.G, dtype=theta.dtype),
                         torch.zeros(self.G, dtype=theta.dtype),
                         torch.zeros(self.G, dtype=theta.dtype)])
        return x, self.y0, self.z0, p[0], p[1], p[2], E, H

    def energy_history(self, theta):
        return EnergyHistory.apply(self.run, *self.initial_state(theta))

    def loss(self, theta, first=150, last=200):
        return energy_loss_window(self.energy_history(theta), first, last)

    def loss_value(self, theta, first=150, last=200, tape=None):
        """forward only (no checkpoints), for finite differences; with tape
        = (words, cells) of a recorded run, the run replays its decisions"""
        with torch.no_grad():
            st = self.initial_state(theta)
        g = self.run.make(*[_np(t) for t in st[:6]], _np(st[6]), _np(st[7]))
        try:
            if tape is not None:
                g.tape_set(*tape)
                g.tape(g.TAPE_REPLAY, self.run.nsteps)
            W = g.run(1, self.run.nsteps)
        finally:
            g.close()
        return float(W[first - 1:last].sum())


def fd_check(cfg=3, deltas=(1e-5,), ghost_mode=1, theta=None, first=150, last=200):
    """Adjoint gradient of the windowed energy loss against central finite
    differences of the same forward, for each step in `deltas`.  Returns a
    dict: theta, adjoint gradient, {delta: finite differences}, loss."""
    bp = BeamPlasma(cfg, ghost_mode=ghost_mode)
    if theta is None:
        theta = [0.3] + [1e-3] * 8
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    L = bp.loss(th, first, last)
    L.backward()
    grad = th.grad.numpy().copy()
    bp.run.last.close()
    base = th.detach().numpy()
    fds = {}
    for d in deltas:
        fd = np.zeros(len(base))
        for i in range(len(base)):
            e = np.zeros(len(base))
            e[i] = d
            fd[i] = (bp.loss_value(torch.tensor(base + e), first, last) -
                     bp.loss_value(torch.tensor(base - e), first, last)) / (2 * d)
        fds[d] = fd
    return {"theta": base, "grad": grad, "fd": fds, "loss": float(L.detach())}


def fd_check_frozen(cfg=3, deltas=(1e-5,), ghost_mode=1, theta=None, first=150, last=200):
    """The adjoint gradient against central differences of the map it
    differentiates: the base forward records its decision tape (load cells,
    gather indices and weight floors, relay cells, wraps, receivers) and each
    perturbed forward replays it, so every run stays on the base run's
    smooth piece.  FP64 differentiable form.  Returns the fd_check dict plus
    the free-run loss of the base state replayed (equal to the recorded
    run's, bit for bit, since replaying a run's own decisions changes
    nothing)."""
    bp = BeamPlasma(cfg, ghost_mode=ghost_mode)
    if theta is None:
        theta = [0.3] + [1e-3] * 8
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    bp.run.record_tape = True
    L = bp.loss(th, first, last)
    L.backward()
    bp.run.record_tape = False
    tape = bp.run.tape
    grad = th.grad.numpy().copy()
    bp.run.last.close()
    bp.run.last = None
    base = th.detach().numpy()
    L_replay = bp.loss_value(torch.tensor(base), first, last, tape=tape)
    fds = {}
    for d in deltas:
        fd = np.zeros(len(base))
        for i in range(len(base)):
            e = np.zeros(len(base))
            e[i] = d
            fd[i] = (bp.loss_value(torch.tensor(base + e), first, last, tape=tape) -
                     bp.loss_value(torch.tensor(base - e), first, last, tape=tape)) / (2 * d)
        fds[d] = fd
    return {"theta": base, "grad": grad, "fd": fds, "loss": float(L.detach()), "loss_replay": L_replay}


def default_theta(cfg):
    """drift 0.3, a_k = 1e-3, and Bx0 at the workload's value"""
    spec_cfg = pic.beam_plasma_config(cfg) if isinstance(cfg, int) else cfg
    return [0.3] + [1e-3] * 8 + [spec_cfg["bx0"]]


def precision_compare(cfg=4, theta=None, first=150, last=200, ckpt_every=20, scale_n=None,
                      nsteps=None, log=None):
    """Gradient of the windowed energy loss by the FP32 form (FP32 forward and
    particle adjoint, FP64 field adjoints and accumulation) against the FP64
    adjoint on the same shape and seeded particles.  Returns, per precision,
    the loss, the gradient with respect to theta = (drift, a_1..a_8, Bx0) and
    the device times (ms) of the forward with checkpoints, the recompute and
    the adjoint steps.  log (optional) receives a progress line per leg."""
    spec_cfg = pic.beam_plasma_config(cfg) if isinstance(cfg, int) else cfg
    theta = default_theta(spec_cfg) if theta is None else theta
    workload = pic.generate_workload(spec_cfg, scale_n=scale_n)
    if log:
        log(f"generated {sum(len(p[1]) for p in workload[1])} particles")
    out = {}
    for prec in (pic.FP64, pic.FP32):
        bp = BeamPlasma(spec_cfg, precision=prec, nsteps=nsteps, ckpt_every=ckpt_every,
                        workload=workload)
        th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
        L = bp.loss(th, first, last)
        L.backward()
        out[prec] = {"loss": float(L.detach()), "grad": th.grad.numpy().copy(),
                     "ms": bp.run.last.adjoint_timing(), "n": int(bp.beam.numel())}
        bp.run.last.close()
        bp.run.last = None
        if log:
            log(f"{'FP32' if prec == pic.FP32 else 'FP64'} gradient done: loss "
                f"{out[prec]['loss']:.12e}")
    return out


def device_gradient(cfg=4, precision=pic.FP64, first=150, last=200, ckpt_every=20,
                    seed=12345):
    """Cost of a gradient of the differentiable form at full size, particles
    drawn on the GPU by the device initializer (no displacement: the drawn
    state is s_0): a plain forward of the workload's steps against
    forward-with-checkpoints + the reverse sweep of the loss
    sum_{t=first..last} W_t (steps after `last` have a zero adjoint and are
    not recomputed), both from the same s_0 on one context whose gradient
    memory a first, untimed gradient allocated.  Also reports whether the
    recorded forward reproduced the plain forward's energy history bit for
    bit, the step status (dropped particles, flags) after the sweep, and the
    norm of the particle adjoint."""
    import time
    spec_cfg = pic.beam_plasma_config(cfg) if isinstance(cfg, int) else cfg
    n = spec_cfg["n"]
    T = spec_cfg["steps"]
    params = pic.make_params(n, tuple(k * 0.1 for k in n), spec_cfg["tau"], precision=precision,
                             differentiable=1, ghost_mode=1,
                             departure_cap=spec_cfg.get("departure_cap", 50))
    g = pic.Pic(params)
    try:
        g.generate_device(spec_cfg["species"], seed=seed)
        if spec_cfg["bx0"]:
            g.load_field(pic.HX, np.full(g.G, spec_cfg["bx0"]))
        torch.cuda.synchronize()
        P = g.particle_count()
        w = np.zeros(T)
        w[first - 1:last] = 1.0
        # the gradient lands in pinned host arrays, allocated once
        pinned = {k: torch.empty(max(g.id_count(), 1), dtype=torch.float64,
                                 pin_memory=True).numpy() for k in PART_KEYS}
        pinned.update({k: torch.empty(g.G, dtype=torch.float64, pin_memory=True).numpy()
                       for k in FIELD_KEYS[:6]})
        g.forward_record(T, ckpt_every)
        g.adjoint(weights=w, out=pinned)
        g.rewind()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        W = g.run(1, T)
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t0
        st_fwd = g.status()
        g.rewind()
        t0 = time.perf_counter()
        Wr = g.forward_record(T, ckpt_every)
        t1 = time.perf_counter()
        gr = g.adjoint(weights=w, out=pinned)
        t2 = time.perf_counter()
        dev = g.adjoint_timing()
        st = g.status()
        P_end = g.particle_count()
    finally:
        g.close()
    pg = math.sqrt(sum(float(np.dot(gr[k], gr[k])) for k in PART_KEYS))
    return {"particles": P, "particles_end": P_end, "steps": T, "first": first, "last": last,
            "ckpt_every": ckpt_every, "forward_s": t_fwd, "forward_record_s": t1 - t0,
            "backward_s": t2 - t1, "gradient_s": t2 - t0, "ratio": (t2 - t0) / t_fwd,
            "reversed_steps": last,
            "backward_per_reversed_step_over_forward_step": (t2 - t1) / last / (t_fwd / T),
            "device_ms": {"forward_with_checkpoints": dev[0], "recompute": dev[1],
                          "adjoint_steps": dev[2]},
            "record_matches_forward": bool(np.array_equal(W, Wr)),
            "status_forward": st_fwd, "status": st, "particle_grad_norm": pg}


def _main():
    import argparse
    ap = argparse.ArgumentParser(description="adjoint vs central finite differences")
    ap.add_argument("--config", type=int, default=3)
    ap.add_argument("--ghost-mode", type=int, default=1)
    ap.add_argument("--deltas", default="1e-5")
    ap.add_argument("--compare-precision", action="store_true",
                    help="FP32 adjoint against the FP64 adjoint instead of differences")
    ap.add_argument("--frozen", action="store_true",
                    help="differences of forwards that replay the base run's decision tape")
    args = ap.parse_args()
    if args.compare_precision:
        r = precision_compare(args.config)
        g64, g32 = r[pic.FP64]["grad"], r[pic.FP32]["grad"]
        names = ["drift"] + [f"a_{k}" for k in range(1, 9)] + ["Bx0"]
        print(f"config {args.config}, {r[pic.FP64]['n']} particles, loss FP64 "
              f"{r[pic.FP64]['loss']:.12e} FP32 {r[pic.FP32]['loss']:.12e}")
        print("| parameter | FP64 adjoint | FP32 adjoint | abs diff / max abs FP64 |")
        print("|---|---|---|---|")
        gmax = np.max(np.abs(g64))
        for i, nm in enumerate(names):
            print(f"| {nm} | {g64[i]:.9e} | {g32[i]:.9e} | {abs(g32[i] - g64[i]) / gmax:.2e} |")
        print(f"| normwise | | | {np.linalg.norm(g32 - g64) / np.linalg.norm(g64):.2e} |")
        for prec, nm in ((pic.FP64, "FP64"), (pic.FP32, "FP32")):
            f, rc, rv = r[prec]["ms"]
            print(f"{nm}: forward+checkpoints {f:.1f} ms, recompute {rc:.1f} ms, adjoint steps "
                  f"{rv:.1f} ms, (recompute+adjoint)/forward {(rc + rv) / f:.2f}")
        return
    deltas = [float(v) for v in args.deltas.split(",")]
    r = (fd_check_frozen if args.frozen else fd_check)(args.config, deltas, args.ghost_mode)
    names = ["drift"] + [f"a_{k}" for k in range(1, 9)]
    print(f"config {args.config}, ghost_mode {args.ghost_mode}, "
          f"{'frozen decisions, ' if args.frozen else ''}loss {r['loss']:.15e}")
    print("| parameter | adjoint | " + " | ".join(f"FD {d:.0e} (rel err)" for d in deltas) + " |")
    print("|---|---|" + "---|" * len(deltas))
    for i, nm in enumerate(names):
        cells = [f"{r['fd'][d][i]:.9e} ({abs(r['grad'][i] - r['fd'][d][i]) / abs(r['fd'][d][i]):.1e})"
                 for d in deltas]
        print(f"| {nm} | {r['grad'][i]:.9e} | " + " | ".join(cells) + " |")
    row = [f"{np.linalg.norm(r['grad'] - r['fd'][d]) / np.linalg.norm(r['fd'][d]):.2e}"
           for d in deltas]
    print("| normwise | | " + " | ".join(row) + " |")


if __name__ == "__main__":
    _main()
