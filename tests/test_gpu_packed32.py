"""Packed layout of the FP32 form: the push writes the stayers of each tile
into the other particle set, packed; arrivals go to the cell's inbox records
and, past those, to the destination tile's spill list; the host unpacks into
buckets for every other reader and repairs a spill list that ran full.

Workload: 16x8x8 cells, 8 background electrons and ions per cell, plus jets
of 200 fast particles (px = 3, so |du| = 0.47 cells per step) placed near the
+x face of a cell: every jet particle enters the next cell in the same step,
200 arrivals into one cell, 168 of them past its 32 inbox records.  One jet
fills a spill list part way; three jets into cells of one tile overflow it
(504 records for 256 places), which the host repairs from the departure
queues.

* against the FP64 form on the same particles, matched by particle id: no
  particle lost or duplicated, positions and momenta within FP32 round-off,
  the same field-energy history;
* unpacking between steps (a reader in the middle of a run) leaves the run
  bit for bit unchanged;
* the differentiable form's recompute after a repaired spill list matches
  the first forward bit for bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = (16, 8, 8)
H = 0.1
SINGLE = ((4, 3, 3),)
TRIPLE = ((4, 3, 3), (4, 2, 3), (4, 1, 3))


def jet_parts(jets, seed=5, njet=200, ppc=8):
    rng = np.random.default_rng(seed)
    L = [H * v for v in N]
    parts = []
    nb = ppc * N[0] * N[1] * N[2]
    for sort, qm, m in ((1, -1.0, -0.02), (0, 1 / 1836., 0.02)):
        x = [rng.uniform(0.0005, L[d] - 0.0005, nb) for d in range(3)]
        p = [rng.normal(0, 0.01, nb) for _ in range(3)]
        parts.append([sort] + x + p + [m, qm])
    for (i, j, k) in jets:
        x = (i + rng.uniform(0.6, 0.95, njet)) * H
        y = (j + rng.uniform(0.2, 0.8, njet)) * H
        z = (k + rng.uniform(0.2, 0.8, njet)) * H
        px = np.full(njet, 3.0)
        py, pz = rng.normal(0, 0.01, njet), rng.normal(0, 0.01, njet)
        parts.append([2, x, y, z, px, py, pz, -0.001, -1.0])
    return parts


def make(picmod, parts, prec, diff=0):
    prm = picmod.make_params(N, tuple(H * v for v in N), 0.05, precision=prec,
                             differentiable=diff, departure_cap=0)
    g = picmod.Pic(prm)
    for (sort, x, y, z, px, py, pz, m, qm) in parts:
        g.load_particles(sort, x, y, z, px, py, pz, m, qm)
    return g


def by_id(g):
    p, ids = g.particles(), g.particle_ids()
    order = np.argsort(ids)
    return ids[order], {k: v[order] for k, v in p.items()}


@pytest.mark.parametrize("jets,overflow", [(SINGLE, False), (TRIPLE, True)],
                         ids=["spill", "spill-overflow"])
def test_spill_lists_against_fp64(picmod, jets, overflow):
    parts = jet_parts(jets)
    n0 = sum(len(p[1]) for p in parts)
    a = make(picmod, parts, picmod.FP32, diff=1)
    b = make(picmod, parts, picmod.FP64, diff=1)
    Wa, Wb = a.run(1, 4), b.run(1, 4)
    st = a.status()
    print(st, flush=True)
    assert st["dropped"] == 0 and st["duplicated"] == 0 and st["flags"] & ~4 == 0
    assert (st["spill_repairs"] > 0) == overflow
    ia, pa = by_id(a)
    ib, pb = by_id(b)
    assert len(ia) == n0 and np.array_equal(ia, ib) and np.array_equal(ia, np.arange(n0))
    for k in ("x", "y", "z"):
        assert np.abs(pa[k] - pb[k]).max() <= 1e-5, k
    for k in ("px", "py", "pz"):
        assert np.abs(pa[k] - pb[k]).max() <= 1e-5 * max(1.0, np.abs(pb[k]).max()), k
    assert np.abs(Wa / Wb - 1).max() <= 1e-4
    a.close()
    b.close()


def test_unpack_between_steps_is_bitwise(picmod):
    parts = jet_parts(SINGLE)
    a = make(picmod, parts, picmod.FP32)
    b = make(picmod, parts, picmod.FP32)
    Wa = a.run(1, 6)
    Wb = [b.run(1, 1)]
    pb = b.particles()                  # the spill list of step 1 unpacked
    assert len(pb["x"]) == sum(len(p[1]) for p in parts)
    Wb.append(b.run(2, 2))
    b.moments()
    Wb.append(b.run(4, 3))
    assert np.array_equal(Wa, np.concatenate(Wb))
    for i in range(6):
        assert np.array_equal(a.get_field(i), b.get_field(i))
    assert np.array_equal(a.cell_counts(), b.cell_counts())
    a.close()
    b.close()


def test_recompute_after_spill_repair_is_bitwise(picmod):
    parts = jet_parts(TRIPLE)
    a = make(picmod, parts, picmod.FP32, diff=1)
    Wa = a.run(1, 8)
    b = make(picmod, parts, picmod.FP32, diff=1)
    Wb = b.forward_record(8, 4)
    assert b.status()["spill_repairs"] > 0
    assert np.array_equal(Wa, Wb)
    b.adjoint(weights=np.ones(8))       # raises if a recomputed segment differs
    for i in range(6):
        assert np.array_equal(a.get_field(i), b.get_field(i))
    assert np.array_equal(a.cell_counts(), b.cell_counts())
    a.close()
    b.close()
