"""The differentiable workloads' initial-state map (torch, CPU) against the
seeded C generator.

BeamPlasma.initial_state(theta) rebuilds the beam from a base draw at drift
0.3 with no displacement; for any (drift, a_1..a_8) it must give the
particles that pic_generate_species draws with that drift and those
displacement amplitudes (same Wichmann-Hill stream): positions to 1e-15 of
the box, momenta to 1e-13 relative (the velocity is shifted by the drift in
floating point rather than drawn around it).
"""
import importlib

import numpy as np
import pytest

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ag():
    return importlib.import_module("0buf_amd.autograd")


@pytest.fixture(scope="module")
def beam3(ag):
    return ag.BeamPlasma(3)


def c_particles(picmod, drift, amps):
    cfg = picmod.beam_plasma_config(3)
    beam = cfg["species"][2]
    beam.drift[0] = drift
    for k, a in enumerate(amps):
        beam.perturb[k] = a
    _, parts = picmod.generate_workload(cfg)
    return [np.concatenate([p[i] for p in parts]) for i in range(1, 7)]


@pytest.mark.parametrize("drift,amps", [(0.3, [0.0] * 8),
                                        (0.27, [1e-3 * (k + 1) for k in range(8)]),
                                        (0.31, [1e-3, -2e-3, 0, 5e-4, 0, 0, 1e-3, -1e-3])])
def test_initial_state_matches_generator(picmod, beam3, drift, amps):
    theta = torch.tensor([drift] + amps, dtype=torch.float64)
    st = beam3.initial_state(theta)
    ref = c_particles(picmod, drift, amps)
    for k in range(3):
        assert np.abs(st[k].numpy() - ref[k]).max() <= 1e-15 * beam3.lx * 10
    for k in range(3, 6):
        a, b = st[k].numpy(), ref[k]
        assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()
    assert not st[6].any() and not st[7].any()


def test_bx0_enters_as_uniform_hx(beam3):
    theta = torch.tensor([0.3] + [0.0] * 8 + [0.5], dtype=torch.float64, requires_grad=True)
    st = beam3.initial_state(theta)
    H = st[7]
    assert torch.all(H[0] == 0.5) and not H[1:].any()
    H[0].sum().backward()
    assert theta.grad[9].item() == beam3.G


def test_parameter_chain_rule(beam3):
    """d(sum of beam x)/d a_k = sum over beam particles of sin(2 pi k x0 / Lx)"""
    theta = torch.tensor([0.3] + [1e-3] * 8, dtype=torch.float64, requires_grad=True)
    x = beam3.initial_state(theta)[0]
    x.sum().backward()
    x0 = beam3.x0.numpy()[beam3.beam.numpy()]
    for k in range(1, 9):
        want = np.sin(2 * np.pi * k * x0 / beam3.lx).sum()
        assert theta.grad[k].item() == pytest.approx(want, rel=1e-12, abs=1e-9)
