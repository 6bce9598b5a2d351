"""Per-cell particle moments (DENSITY, KINETIC_ENERGY, VELOCITY, CURRENT of
particle_types.h:15-16) on the GPU against the same sums over the oracle's
particles after 10 config-0 steps (the two states agree to 1e-14), reduced
with math.fsum per cell."""
import math

import numpy as np
import pytest

from helpers import load_workload

pytestmark = pytest.mark.gpu


def expected(parts, ncell, sort, op):
    sel = parts["sort"] == sort
    cell = parts["cell"][sel]
    m, qm = parts["m"][sel], parts["q_m"][sel]
    p = np.stack([parts[k][sel] for k in ("px", "py", "pz")])
    gm = np.sqrt(1.0 + (p * p).sum(0))
    u = p / gm
    if op == "density":
        vals = [m]
    elif op == "kinetic_energy":
        vals = [(m / qm) * ((p * p).sum(0) / (gm + 1.0))]
    elif op == "velocity":
        vals = list(u)
    else:
        vals = list(m * u)
    out = np.zeros((len(vals), ncell))
    groups = {}
    for j, c in enumerate(cell):
        groups.setdefault(int(c), []).append(j)
    for c, idx in groups.items():
        for r, v in enumerate(vals):
            s = math.fsum(v[idx])
            out[r, c] = s / len(idx) if op == "velocity" else s
    return out


@pytest.mark.parametrize("precision", [0, 1])
def test_cell_moments_against_oracle_state(picmod, orc, precision):
    params, parts = picmod.generate_workload(0)
    n = (params.nx, params.ny, params.nz)
    o = orc.Oracle(n, (params.lx, params.ly, params.lz), params.tau)
    load_workload(o, parts)
    params.precision = precision
    g = picmod.Pic(params)
    load_workload(g, parts)
    for nt in range(1, 11):
        o.step(nt)
        g.step(nt)
    po = o.particles()
    tol = 1e-12 if precision == 0 else 1e-5     # FP32 momenta: ~1e-7 each, summed
    for sort in range(3):
        for op in ("density", "kinetic_energy", "velocity", "current"):
            want = expected(po, g.ncell, sort, op)
            got = g.cell_moments(sort, op).reshape(want.shape)
            scale = max(np.abs(want).max(), 1e-300)
            err = np.abs(got - want).max() / scale
            assert err <= tol, (sort, op, err)
    g.close()
