"""Device-side seeded initializer against the host generator (MI355X only).

pic_generate_device runs the stream_mode-1 draws of pic_generate_species on
the GPU and writes each particle into its cell's bucket; the result must be
the host path's (generate, then pic_load_particles in species order) bit for
bit: bucket contents in order, counts, ids, and therefore every later step.
"""
import numpy as np
import pytest

from helpers import load_workload

pytestmark = pytest.mark.gpu


def pair(picmod, cfg, n, precision=0, differentiable=0, seed=12345):
    params, parts = picmod.generate_workload(cfg, scale_n=n, stream_mode=1, seed=seed,
                                             nthreads=8)
    params.precision = precision
    params.differentiable = differentiable
    a = picmod.Pic(params)
    load_workload(a, parts)
    b = picmod.Pic(params)
    nb = b.generate_device(cfg["species"], seed=seed)
    assert nb == sum(len(p[1]) for p in parts)
    return a, b


def same_particles(a, b):
    pa, pb = a.particles(), b.particles()
    for k in pa:
        assert np.array_equal(pa[k], pb[k]), k


@pytest.mark.parametrize("precision", [0, 1])
def test_device_init_matches_host(picmod, precision):
    cfg = picmod.beam_plasma_config(1)
    a, b = pair(picmod, cfg, (32, 16, 16), precision)
    assert np.array_equal(a.cell_counts(), b.cell_counts())
    same_particles(a, b)
    for nt in range(1, 6):
        ea, eb = a.step(nt), b.step(nt)
        if precision == picmod.FP32:
            assert ea == eb           # integer current: order-free, bitwise
        else:
            # FP64 atomics sum J in a run-dependent order: last-bit differences
            assert ea == pytest.approx(eb, rel=1e-12)
    a.close()
    b.close()


def test_device_init_gaussian_beam_and_ids(picmod):
    cfg = picmod.beam_plasma_config(4)
    a, b = pair(picmod, cfg, (16, 32, 32), 0, differentiable=1)
    assert np.array_equal(a.cell_counts(), b.cell_counts())
    same_particles(a, b)
    assert np.array_equal(a.particle_ids(), b.particle_ids())
    assert a.id_count() == b.id_count()
    for nt in range(1, 4):
        assert a.step(nt) == b.step(nt)
    a.close()
    b.close()


def test_device_init_refuses_displacement(picmod):
    cfg = picmod.beam_plasma_config(3)
    cfg["species"][2].perturb[0] = 1e-3
    params, _ = picmod.generate_workload(cfg, scale_n=(8, 8, 8))
    g = picmod.Pic(params)
    with pytest.raises(picmod.PicError):
        g.generate_device(cfg["species"])
    g.close()
