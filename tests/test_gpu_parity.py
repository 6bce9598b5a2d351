"""Parity of the HIP plain FP64 form against the oracle (MI355X only).

Tolerances (FP64):
  * order-free stages (H half step, H second half, migration decisions,
    per-cell counts) are compared bit for bit;
  * stages fed by atomic sums (J, then E and the energy) differ only in
    summation order: max|a-b| <= 1e-13 * max|b| per step;
  * over the 100 steps of config 0: fields, energy history and particle
    moments within 1e-10 relative (max-norm relative to the field's scale,
    moments relative to their magnitude sums), counts exact at every step.
The push uses a correctly rounded pow(v,-0.5); glibc's pow may round the
other way in rare near-halfway cases, so particle coordinates are compared
to 1e-14 relative with the bit-exact fraction reported.
"""
import math

import numpy as np
import pytest

from helpers import (div_staggered, load_state, load_workload, rel_max, rho_cic,
                     sorted_particles)

pytestmark = pytest.mark.gpu

NAMES = ["Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "Jx", "Jy", "Jz", "Qx", "Qy", "Qz"]


def geometry(params):
    n = (params.nx, params.ny, params.nz)
    L = (params.lx, params.ly, params.lz)
    return n, L, tuple(a / b for a, b in zip(L, n))


@pytest.fixture(scope="module")
def cfg0(picmod):
    return picmod.generate_workload(0)


def pair(picmod, orc, params, parts, **kw):
    n, L, _ = geometry(params)
    o = orc.Oracle(n, L, params.tau, departure_cap=kw.get("departure_cap", 50))
    p = picmod.make_params(n, L, params.tau, **kw)
    g = picmod.Pic(p)
    load_workload(o, parts)
    load_workload(g, parts)
    return o, g


def oracle_state(o):
    return {nm: o.get_field(i) for i, nm in enumerate(NAMES)}, o.particles()


def assert_particles_match(pg, po, rtol=1e-14, min_exact=0.99):
    assert len(pg["x"]) == len(po["x"])
    a, b = sorted_particles(pg), sorted_particles(po)
    assert np.array_equal(a["cell"], b["cell"])
    assert np.array_equal(a["sort"], b["sort"])
    exact = 1.0
    for k in ("x", "y", "z", "px", "py", "pz"):
        d = np.abs(a[k] - b[k])
        scale = np.maximum(np.abs(b[k]), 1e-300)
        assert (d / scale).max() <= rtol or d.max() <= rtol, k
        exact = min(exact, float(np.mean(a[k] == b[k])))
    assert exact >= min_exact, exact
    return exact


def test_stage_parity_from_identical_state(picmod, orc, cfg0):
    params, parts = cfg0
    n, L, h = geometry(params)
    o = orc.Oracle(n, L, params.tau)
    load_workload(o, parts)
    for nt in range(1, 6):
        o.step(nt)
    fields, prt = oracle_state(o)
    g = picmod.Pic(picmod.make_params(n, L, params.tau))
    load_state(g, fields, prt)

    o.stage("h_half")
    g.stage("yee_h_half")
    for nm in ("Hx", "Hy", "Hz", "Qx", "Qy", "Qz", "Ex", "Ey", "Ez"):
        i = NAMES.index(nm)
        assert np.array_equal(g.get_field(i), o.get_field(i)), nm
    for i in (6, 7, 8):
        assert not g.get_field(i).any()

    o.stage("push")
    o.stage("deposit")
    g.stage("push_deposit")
    for i in (6, 7, 8):
        assert rel_max(g.get_field(i), o.get_field(i)) < 1e-13, NAMES[i]

    assert o.stage("migrate") == 0
    g.stage("migrate")
    g.sync()
    assert np.array_equal(g.cell_counts(), o.cell_counts())
    assert_particles_match(g.particles(), o.particles())

    o.stage("fold_j")
    g.stage("fold_j")
    for i in (6, 7, 8):
        assert rel_max(g.get_field(i), o.get_field(i)) < 1e-13, NAMES[i]
    o.stage("h2")
    g.stage("yee_h2")
    for i in (3, 4, 5):
        assert np.array_equal(g.get_field(i), o.get_field(i)), NAMES[i]
    o.stage("e")
    g.stage("yee_e")
    o.stage("e_periodic")
    g.stage("e_periodic")
    for i in (0, 1, 2):
        assert rel_max(g.get_field(i), o.get_field(i)) < 1e-13, NAMES[i]
    assert g.energy() == pytest.approx(o.energy(), rel=1e-13)
    g.close()


@pytest.mark.parametrize("ghost_mode", [0, 1])
def test_fused_field_pass_equals_stages(picmod, cfg0, ghost_mode):
    """The step runs H second half, the E update and the periodic E ghosts
    as one pass (k_h2_e); the stage entry points run k_h2, k_e and the
    k_e_periodic_* copies.  With the differentiable form's order-free
    fixed-point current both give the same fields and energies over 10 steps
    of config 0, bit for bit (a different rounding order in the E update or
    a ghost written from the wrong node shows here)."""
    params, parts = cfg0
    n, L, _ = geometry(params)
    names = ("yee_h_half", "push_deposit", "migrate", "fold_j", "yee_h2", "yee_e", "e_periodic")
    runs = []
    for mode in ("step", "stages"):
        g = picmod.Pic(picmod.make_params(n, L, params.tau, differentiable=1,
                                          ghost_mode=ghost_mode))
        load_workload(g, parts)
        e = []
        for nt in range(1, 11):
            if mode == "stages":
                for nm in names:
                    g.stage(nm)
                g.sync()
                e.append(g.energy())
            else:
                e.append(g.step(nt))
        runs.append((np.array(e), [g.get_field(i) for i in range(12)]))
        g.close()
    assert np.array_equal(runs[0][0], runs[1][0])
    for i, (a, b) in enumerate(zip(runs[0][1], runs[1][1])):
        assert np.array_equal(a, b), NAMES[i]


@pytest.mark.parametrize("n", [(3, 3, 3), (3, 5, 4), (5, 4, 3)])
@pytest.mark.parametrize("form", ["fp32", "fp64-diff"])
def test_fused_field_pass_small_grids(picmod, n, form):
    """The step's fused field passes against the stage calls on grids of 3 to
    5 cells per axis (every node on a fold or ghost layer, face classes with
    Nx - 2 = 1), random particles, both ghost modes, 8 steps: energies and all
    12 fields bit for bit (the FP32 form's and the differentiable form's
    currents are order-free)."""
    L = tuple(0.1 * k for k in n)
    rng = np.random.default_rng(11)
    N = 600
    parts = []
    for sort, qm, m in ((1, -1.0, -0.02), (0, 1 / 1836.0, 0.02)):
        x = [rng.uniform(0, L[d], N) for d in range(3)]
        p = rng.normal(0, 0.5, (3, N)) if sort == 1 else np.zeros((3, N))
        parts.append((sort, x[0], x[1], x[2], p[0], p[1], p[2], m, qm))
    names = ("yee_h_half", "push_deposit", "migrate", "fold_j", "yee_h2", "yee_e", "e_periodic")
    kw = dict(precision=picmod.FP32) if form == "fp32" else dict(differentiable=1)
    for gm in (0, 1):
        runs = []
        for mode in ("step", "stages"):
            g = picmod.Pic(picmod.make_params(n, L, 0.05, departure_cap=0, ghost_mode=gm, **kw))
            load_workload(g, parts)
            e = []
            for nt in range(1, 9):
                if mode == "stages":
                    for nm in names:
                        g.stage(nm)
                    g.sync()
                    e.append(g.energy())
                else:
                    e.append(g.step(nt))
            runs.append((np.array(e), [g.get_field(i) for i in range(12)]))
            g.close()
        assert np.array_equal(runs[0][0], runs[1][0]), gm
        for i, (a, b) in enumerate(zip(runs[0][1], runs[1][1])):
            assert np.array_equal(a, b), (gm, NAMES[i])


def test_full_run_config0(picmod, orc, cfg0):
    """Config 0: 32x8x8 cells, 3 x 16 ppc, 100 steps."""
    params, parts = cfg0
    o, g = pair(picmod, orc, params, parts)
    eo, eg = [], []
    for nt in range(1, 101):
        e1, st = o.step(nt)
        assert st == 0
        eo.append(e1)
        eg.append(g.step(nt))
        assert np.array_equal(g.cell_counts(), o.cell_counts()), nt
    eo, eg = np.array(eo), np.array(eg)
    assert np.abs(eg - eo).max() <= 1e-10 * np.abs(eo).max()
    assert np.all(np.abs(eg - eo) <= 1e-10 * np.abs(eo))
    for i in range(9):
        assert rel_max(g.get_field(i), o.get_field(i)) < 1e-10, NAMES[i]
    mg, mo = g.moments(), o.moments()
    assert np.array_equal(mg[:, 0], mo[:, 0])
    for s in range(3):
        assert np.allclose(mg[s, 1:4], mo[s, 1:4], rtol=1e-10, atol=0)
        assert np.abs(mg[s, 4:7] - mo[s, 4:7]).max() <= 1e-10 * mo[s, 8]
        assert mg[s, 7] == pytest.approx(mo[s, 7], rel=1e-10, abs=1e-300)
    st = g.status()
    assert st["dropped"] == 0 and st["flags"] == 0
    g.close()


@pytest.mark.parametrize("start", [10, 60])
def test_step_from_oracle_dump_bitexact_migration(picmod, orc, cfg0, start):
    """Given identical state (a dump of the oracle at step `start`), one step
    moves every particle to the same cell."""
    params, parts = cfg0
    n, L, h = geometry(params)
    o = orc.Oracle(n, L, params.tau)
    load_workload(o, parts)
    for nt in range(1, start + 1):
        o.step(nt)
    fields, prt = oracle_state(o)
    g = picmod.Pic(picmod.make_params(n, L, params.tau))
    load_state(g, fields, prt)
    e_o, _ = o.step(start + 1)
    e_g = g.step(start + 1)
    assert np.array_equal(g.cell_counts(), o.cell_counts())
    frac = assert_particles_match(g.particles(), o.particles())
    assert e_g == pytest.approx(e_o, rel=1e-13)
    print(f"bit-exact particle coordinate fraction after step {start + 1}: {frac:.6f}")
    g.close()


def test_minimum_grid_corner_crossings(picmod, orc):
    """3x3x3 cells (smallest supported) with fast particles: every neighbour
    direction including corners, periodic wrap on all faces."""
    n = (3, 3, 3)
    L = (0.3, 0.3, 0.3)
    rng = np.random.default_rng(4)
    N = 3000
    parts = []
    for sort, qm, m in ((1, -1.0, -0.02), (0, 1 / 1836.0, 0.02)):
        x = rng.uniform(0, 0.3, (3, N))
        p = rng.normal(0, 0.6, (3, N)) if sort == 1 else np.zeros((3, N))
        parts.append((sort, x[0], x[1], x[2], p[0], p[1], p[2], m, qm))
    params = picmod.make_params(n, L, 0.05)
    o, g = pair(picmod, orc, params, parts, departure_cap=0)
    for nt in range(1, 31):
        e1, _ = o.step(nt)
        e2 = g.step(nt)
        assert np.array_equal(g.cell_counts(), o.cell_counts()), nt
        assert e2 == pytest.approx(e1, rel=1e-10, abs=1e-300)
    assert_particles_match(g.particles(), o.particles(), rtol=1e-10, min_exact=0.0)
    g.close()


@pytest.mark.parametrize("case", ["empty", "single"])
@pytest.mark.parametrize("prec", [0, 1])
def test_empty_and_single_particle_domains(picmod, orc, case, prec):
    """Random E, H on a grid whose z extent is not a whole number of tiles:
    no particles at all (every tile empty), or one electron crossing faces
    (one lane of one chunk, every other tile empty); 40 steps of either form
    against the oracle. FP64: energies to 1e-10, counts exact, position to
    1e-12 h. FP32: energies to 1e-4, position to 1e-5 h (FP32 fields, push
    and offsets)."""
    n = (10, 8, 12)
    L = tuple(0.1 * k for k in n)
    G = (n[0] + 2) * (n[1] + 2) * (n[2] + 2)
    rng = np.random.default_rng(77)
    fields = [0.01 * rng.standard_normal(G) for _ in range(6)]
    parts = []
    if case == "single":
        parts = [(1, np.array([0.41]), np.array([0.37]), np.array([0.93]), np.array([0.3]),
                  np.array([-0.2]), np.array([0.25]), -0.01, -1.0)]
    params = picmod.make_params(n, L, 0.05)
    o, g = pair(picmod, orc, params, parts, precision=prec)
    for i, f in enumerate(fields):
        o.load_field(i, f)
        g.load_field(i, f)
    for nt in range(1, 41):
        e1, _ = o.step(nt)
        e2 = g.step(nt)
        assert e2 == pytest.approx(e1, rel=1e-10 if prec == 0 else 1e-4), nt
        if prec == 0:
            assert np.array_equal(g.cell_counts(), o.cell_counts()), nt
    assert g.particle_count() == len(parts)
    assert g.status()["flags"] == 0
    if case == "single":
        pg, po = g.particles(), o.particles()
        # in units of the cell size h = 0.1: FP32 offsets carry about 1e-7 h
        # of rounding per step
        for k in ("x", "y", "z"):
            assert abs(pg[k][0] - po[k][0]) <= (1e-12 if prec == 0 else 1e-5) * 0.1, k
    g.close()


def test_bucket_regrow_keeps_parity(picmod, orc, cfg0):
    """Buckets sized to the initial maximum overflow on the first arrivals;
    the library regrows them and replaces the overflowed arrivals."""
    params, parts = cfg0
    o, g = pair(picmod, orc, params, parts, bucket_capacity=48)
    for nt in range(1, 21):
        o.step(nt)
        g.step(nt)
        assert np.array_equal(g.cell_counts(), o.cell_counts()), nt
    st = g.status()
    assert st["bucket_regrows"] >= 1
    assert st["bucket_capacity"] > 48
    assert_particles_match(g.particles(), o.particles(), rtol=1e-12, min_exact=0.0)
    g.close()


def test_departure_cap_error(picmod):
    n, L = (8, 6, 5), (0.8, 0.6, 0.5)
    for count, fails in ((50, False), (51, True)):
        g = picmod.Pic(picmod.make_params(n, L, 0.05, departure_cap=50))
        x = np.full(count, 0.295)
        y = np.linspace(0.31, 0.39, count)
        z = np.full(count, 0.35)
        g.load_particles(1, x, y, z, np.full(count, 0.9), np.zeros(count), np.zeros(count),
                         -0.01, -1.0)
        if fails:
            with pytest.raises(picmod.PicError) as ei:
                g.step(1)
            assert ei.value.code == 4
        else:
            g.step(1)
        g.close()


def find_overlap(h_len=0.1, nmax=200):
    """A grid where fl(fl(i*h) + h) > fl((i+1)*h): cells i and i+1 overlap
    by one ulp under isPointInCell (cell.h:717)."""
    for N in range(5, nmax):
        L = N * h_len
        h = L / N
        for i in range(N - 2):
            if (i * h) + h > (i + 1) * h:
                return N, L, h, i
    return None


def test_overlap_duplicates_like_reference(picmod, orc):
    """A leaver lying in the one-ulp overlap of two cells is inserted by both
    receivers in GPU_ArrangeFlights (kernels.cu:316-370): two copies."""
    found = find_overlap()
    assert found is not None
    N, Lx, h, i = found
    n, L = (N, 4, 4), (Lx, 0.4, 0.4)
    x = np.array([(i + 1) * h])
    o = orc.Oracle(n, L, 0.05)
    g = picmod.Pic(picmod.make_params(n, L, 0.05))
    for t in (o, g):
        t.load_particles(1, x, np.array([0.195]), np.array([0.15]), np.array([0.0]),
                         np.array([0.3]), np.array([0.0]), -0.01, -1.0)
    o.step(1)
    g.step(1)
    assert o.particle_count() == 2
    assert g.particle_count() == 2
    assert np.array_equal(g.cell_counts(), o.cell_counts())
    assert g.status()["duplicated"] == 1
    g.close()


def test_continuity_and_gauss_on_gpu(picmod, cfg0):
    params, parts = cfg0
    n, L, h = geometry(params)
    g = picmod.Pic(picmod.make_params(n, L, params.tau))
    load_workload(g, parts)
    g0 = None
    for nt in range(1, 8):
        r0 = rho_cic(g.particles(), n, h)
        g.step(nt)
        p1 = g.particles()
        r1 = rho_cic(p1, n, h)
        J = [g.get_field(6 + k) for k in range(3)]
        res = r1 - r0 + params.tau * div_staggered(*J, n, h)
        assert np.abs(res).max() < 1e-13 * np.abs(r1).max()
        E = [g.get_field(k) for k in range(3)]
        gauss = div_staggered(*E, n, h) - r1
        if g0 is None:
            g0 = gauss
        assert np.abs(gauss - g0).max() < 1e-13
    g.close()


def test_yee_mode_recurrence_on_gpu(picmod):
    n = (12, 10, 14)
    L = tuple(0.1 * k for k in n)
    tau, h, mode = 0.05, 0.1, 2
    g = picmod.Pic(picmod.make_params(n, L, tau))
    gi = np.arange(n[0] + 2, dtype=np.float64)
    prof = np.sin(2 * np.pi * mode * (gi - 0.5) / n[0])
    arr = np.broadcast_to(prof.reshape(-1, 1, 1), tuple(k + 2 for k in n)).copy()
    g.load_field(1, arr)
    lam = 4 * math.sin(np.pi * mode / n[0]) ** 2 / h ** 2
    idx = np.ravel_multi_index((2, 4, 5), tuple(k + 2 for k in n))
    s = [g.get_field(1)[idx]]
    for nt in range(1, 40):
        g.step(nt)
        s.append(g.get_field(1)[idx])
    s = np.array(s)
    res = s[2:] + s[:-2] - (2 - tau ** 2 * lam) * s[1:-1]
    assert np.abs(res).max() < 1e-12
    g.close()


def test_archapi_memory_roundtrip(picmod):
    lib = picmod.load_library()
    import ctypes as C
    a = np.arange(1000, dtype=np.float64)
    b = np.zeros_like(a)
    d = C.c_void_p()
    assert lib.MemoryAllocate(C.byref(d), a.nbytes) == 0
    assert lib.MemoryCopy(d, a.ctypes.data_as(C.c_void_p), a.nbytes, picmod.HOST_TO_DEVICE) == 0
    assert lib.MemoryCopy(b.ctypes.data_as(C.c_void_p), d, a.nbytes, picmod.DEVICE_TO_HOST) == 0
    assert np.array_equal(a, b)
    assert lib.MemorySet(d, 0, a.nbytes) == 0
    assert lib.MemoryCopy(b.ctypes.data_as(C.c_void_p), d, a.nbytes, picmod.DEVICE_TO_HOST) == 0
    assert not b.any()
    assert lib.MemoryCopy(d, d, 8, 12345) != 0
    assert lib.MemoryFree(d) == 0
