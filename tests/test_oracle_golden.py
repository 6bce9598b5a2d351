"""Pin the oracle (and the host-side initializer) to the reference's own code.

Golden vectors come from oracle/_ref binaries built from the reference's
particle.h (Particle::Move) and maxwell.cxx (g05cae_, g05dde_), see
tests/golden/gen_golden.py.  Bit-exact comparisons.
"""
import ctypes as C
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_boris_oracle_matches_reference_particle_h(orc):
    g = np.load(os.path.join(GOLDEN, "boris_ref.npz"))
    n = len(g["q_m"])
    for j in range(n):
        p, x1 = orc.boris(g["E"][j], g["H"][j], g["tau"][j], g["q_m"][j], g["x"][j], g["p"][j])
        assert np.array_equal(p, g["p_out"][j]), j
        assert np.array_equal(x1, g["x1_out"][j]), j


def test_rng_uniform_matches_reference_g05cae(picmod):
    g = np.load(os.path.join(GOLDEN, "rng_ref.npz"))
    lib = picmod.load_library()
    r = picmod.PicRng()
    lib.pic_rng_seed(C.byref(r), 1, 255, 25555)
    u = np.array([lib.pic_rng_uniform(C.byref(r)) for _ in range(len(g["uniform"]))])
    assert np.array_equal(u, g["uniform"])
    # the normal deviates continue the same generator state (ref_rng draws
    # all uniforms first, then all normals)
    v = np.array([lib.pic_rng_normal(C.byref(r), float(g["mean"]), float(g["sd"]))
                  for _ in range(len(g["normal"]))])
    assert np.array_equal(v, g["normal"])


def test_rng_normal_moments(picmod):
    """g05dde_ is a standard normal: mean, variance and a tail quantile."""
    lib = picmod.load_library()
    r = picmod.PicRng()
    lib.pic_rng_seed(C.byref(r), 1, 255, 25555)
    v = np.array([lib.pic_rng_normal(C.byref(r), 0.0, 1.0) for _ in range(200000)])
    assert abs(v.mean()) < 0.01
    assert abs(v.var() - 1.0) < 0.01
    assert abs(np.mean(np.abs(v) > 1.959964) - 0.05) < 0.003


def test_compare_metric_closed_forms(orc):
    a = np.array([0.0, 1.0, 2.0, 3.0])
    b = np.array([0.0, 1.0 + 1e-16, 2.5, 3.0])
    assert orc.compare(a, b, 1e-15) == pytest.approx(0.75)
    assert orc.rms(a, b) == pytest.approx(np.sqrt(0.25 / 4), rel=1e-12)
