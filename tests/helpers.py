"""Test helpers: workload setup and the discrete charge/current geometry.

Geometry of the reference's deposit (cell.h:1511-1693): the charge that the
current deposit conserves sits on the cell-centre lattice, global index g at
position (g - 0.5) h (Fortran numbering, cell.h:680-692 getNodeX), with
linear (cloud-in-cell) weights; Jx sits at x = g h (the face between centres
g and g+1) and at centres in y and z, and likewise for Jy, Jz.  The
discrete continuity equation is then

    rho^{n+1}_g - rho^n_g + tau * sum_d (J_d[g] - J_d[g - e_d]) / h_d = 0

for every centre g of the periodic box.  Ex, Ey, Ez share the J staggering,
so div E is taken with the same stencil.
"""
import numpy as np


def rho_cic(parts, n, h):
    """Charge density on the periodic centre lattice, bulk indices 0..N-1
    (bulk b <-> global g = b + 1)."""
    rho = np.zeros(n, dtype=np.float64)
    pos = [parts["x"], parts["y"], parts["z"]]
    a = parts["m"]
    idx, w = [], []
    for d in range(3):
        u = pos[d] / h[d] + 0.5
        g = np.floor(u).astype(np.int64)
        f = u - g
        b0 = (g - 1) % n[d]
        b1 = g % n[d]
        idx.append((b0, b1))
        w.append((1.0 - f, f))
    for ix in range(2):
        for iy in range(2):
            for iz in range(2):
                np.add.at(rho, (idx[0][ix], idx[1][iy], idx[2][iz]),
                          a * w[0][ix] * w[1][iy] * w[2][iz])
    return rho


def div_staggered(Fx, Fy, Fz, n, h):
    """Discrete divergence at the bulk centres of arrays laid out like J / E
    (ghosted, k fastest): (F_d[g] - F_d[g - e_d]) / h_d for g = 1..N."""
    sh = (n[0] + 2, n[1] + 2, n[2] + 2)
    Fx, Fy, Fz = Fx.reshape(sh), Fy.reshape(sh), Fz.reshape(sh)
    s = np.s_[1:n[0] + 1, 1:n[1] + 1, 1:n[2] + 1]
    dx = (Fx[1:n[0] + 1, 1:n[1] + 1, 1:n[2] + 1] - Fx[0:n[0], 1:n[1] + 1, 1:n[2] + 1]) / h[0]
    dy = (Fy[1:n[0] + 1, 1:n[1] + 1, 1:n[2] + 1] - Fy[1:n[0] + 1, 0:n[1], 1:n[2] + 1]) / h[1]
    dz = (Fz[1:n[0] + 1, 1:n[1] + 1, 1:n[2] + 1] - Fz[1:n[0] + 1, 1:n[1] + 1, 0:n[2]]) / h[2]
    del s
    return dx + dy + dz


def load_workload(target, parts):
    """Load a generate_workload() particle list into an Oracle or a Pic."""
    dropped = 0
    for (sort, x, y, z, px, py, pz, m, qm) in parts:
        dropped += target.load_particles(sort, x, y, z, px, py, pz, m, qm)
    return dropped


def load_state(target, fields, parts):
    """Load fields (dict name->array) and a particles() dict."""
    names = ["Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "Jx", "Jy", "Jz", "Qx", "Qy", "Qz"]
    for i, nm in enumerate(names):
        if nm in fields:
            target.load_field(i, fields[nm])
    keys = sorted(set(zip(parts["sort"].tolist(), parts["m"].tolist(), parts["q_m"].tolist())))
    for (s, m, qm) in keys:
        sel = (parts["sort"] == s) & (parts["m"] == m) & (parts["q_m"] == qm)
        target.load_particles(s, parts["x"][sel], parts["y"][sel], parts["z"][sel],
                              parts["px"][sel], parts["py"][sel], parts["pz"][sel], m, qm)


def sorted_particles(p):
    """Particles ordered by (cell, sort, x, y, z, px) for order-free comparison."""
    order = np.lexsort((p["px"], p["z"], p["y"], p["x"], p["sort"], p["cell"]))
    return {k: v[order] for k, v in p.items()}


def rel_max(a, b):
    """max|a-b| / max|b| (0 when both vanish)."""
    s = np.abs(b).max()
    d = np.abs(a - b).max()
    return 0.0 if s == 0 and d == 0 else d / (s if s > 0 else 1.0)
