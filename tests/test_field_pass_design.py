"""The step's field update after the push, restated on the CPU with the same
decomposition as its two kernels (k_fold_j_classes, then k_h2_e, in
0buf_amd/csrc/k_fields.hip), against the stage sequence it replaces: the
three current-fold launches (k_fold_j: x, y, z), k_h2, k_e and the periodic
E ghost copies (k_e_periodic_x, _z, and _y in ghost mode 1), as the reference
runs them (gpu_plasma.h:926-993, 540-551, 455-505; kernels.cu:1024-1181).

Random FP64 fields on small grids (N >= 3, as pic_create requires), both
ghost modes.  Each class of the fold and each node of the pass reads the
input state only (snapshots), so one that needed another's result would
differ here.  Checked: the results equal the sequence's bit for bit; every
node's H^{n+1} is written once and every E node at most once; the face
enumeration of the fold covers each class of more than one node exactly
once.  The GPU tests compare the kernels themselves with the stage entry
points (test_fused_field_pass_equals_stages, test_fp32_step_paths_bit_identical).
"""
import numpy as np
import pytest

C1X, C1Y, C1Z, TAU = 0.3, 0.7, 1.1, 0.05


def stage_sequence(N, mode, E, H, Q, J):
    """k_fold_j x, y, z; k_h2; k_e; k_e_periodic_x, _z (, _y)."""
    Nx, Ny, Nz = N
    E = [a.copy() for a in E]
    H = [a.copy() for a in H]
    J = [a.copy() for a in J]
    for d in range(3):
        Nd = N[d]
        for cc in range(3):
            A = np.moveaxis(J[cc], d, 0)
            if d != cc:
                A[1] = A[1] + A[Nd + 1]
            if d != 1 or cc != 1:
                A[Nd + 1] = A[1]
            A[0] = A[0] + A[Nd]
            A[Nd] = A[0]
    H[0][:, :Ny + 1, :Nz + 1] += Q[0][:, :Ny + 1, :Nz + 1]
    H[1][:Nx + 1, :, :Nz + 1] += Q[1][:Nx + 1, :, :Nz + 1]
    H[2][:Nx + 1, :Ny + 1, :] += Q[2][:Nx + 1, :Ny + 1, :]
    Hx, Hy, Hz = H
    E0 = [a.copy() for a in E]
    for i in range(Nx + 1):
        for l in range(Ny + 1):
            for k in range(Nz + 1):
                if l >= 1 and k >= 1:
                    E[0][i, l, k] = E0[0][i, l, k] + (
                        C1Y * (Hz[i, l, k] - Hz[i, l - 1, k]) -
                        C1Z * (Hy[i, l, k] - Hy[i, l, k - 1]) - TAU * J[0][i, l, k])
                if i >= 1 and k >= 1:
                    E[1][i, l, k] = E0[1][i, l, k] + (
                        C1Z * (Hx[i, l, k] - Hx[i, l, k - 1]) -
                        C1X * (Hz[i, l, k] - Hz[i - 1, l, k]) - TAU * J[1][i, l, k])
                if i >= 1 and l >= 1:
                    E[2][i, l, k] = E0[2][i, l, k] + (
                        C1X * (Hy[i, l, k] - Hy[i - 1, l, k]) -
                        C1Y * (Hx[i, l, k] - Hx[i, l - 1, k]) - TAU * J[2][i, l, k])
    for l in range(Ny + 2):
        for k in range(Nz + 2):
            if l <= Ny and 1 <= k <= Nz:
                E[1][0, l, k] = E[1][Nx, l, k]
                E[1][Nx + 1, l, k] = E[1][1, l, k]
            if 1 <= l <= Ny and k <= Nz:
                E[2][0, l, k] = E[2][Nx, l, k]
                E[2][Nx + 1, l, k] = E[2][1, l, k]
    for i in range(Nx + 2):
        for l in range(Ny + 2):
            if i <= Nx:
                E[0][i, l, 0] = E[0][i, l, Nz]
                E[0][i, l, Nz + 1] = E[0][i, l, 1]
            if l <= Ny:
                E[1][i, l, 0] = E[1][i, l, Nz]
                E[1][i, l, Nz + 1] = E[1][i, l, 1]
    if mode == 1:
        for c in (0, 2):
            E[c][:, 0, :] = E[c][:, Ny, :]
            E[c][:, Ny + 1, :] = E[c][:, 1, :]
    return E, H, J


def ghost_src3(N, mode, comp, i, l, k):
    Nx, Ny, Nz = N
    if mode == 1 and comp != 1:
        l = Ny if l == 0 else 1 if l == Ny + 1 else l
    if (comp == 0 and i <= Nx) or (comp == 1 and l <= Ny):
        k = Nz if k == 0 else 1 if k == Nz + 1 else k
    if (comp == 1 and l <= Ny and 1 <= k <= Nz) or (comp == 2 and 1 <= l <= Ny and k <= Nz):
        i = Nx if i == 0 else 1 if i == Nx + 1 else i
    return i, l, k


def face_representatives(N):
    """k_fold_j_classes' thread -> representative map"""
    Nx, Ny, Nz = N
    fx, fy = 2 * Ny * Nz, 2 * (Nx - 2) * Nz
    for t in range(fx + fy + 2 * (Nx - 2) * (Ny - 2)):
        if t < fx:
            a, r = divmod(t, Ny * Nz)
            yield (Nx if a else 1), r // Nz + 1, r % Nz + 1
        elif t - fx < fy:
            a, r = divmod(t - fx, (Nx - 2) * Nz)
            yield r // Nz + 2, (Ny if a else 1), r % Nz + 1
        else:
            a, r = divmod(t - fx - fy, (Nx - 2) * (Ny - 2))
            yield r // (Ny - 2) + 2, r % (Ny - 2) + 2, (Nz if a else 1)


def fold_pair(v, a, b, d, cc, at_n):
    r, p = v[a], v[b]
    if at_n:
        v[a] = v[b] = p + r
    elif d != cc:
        v[a] = v[b] = r + p
    elif d != 1:
        v[b] = r


def fused(N, mode, E, H, Q, J, fold_axes=(0, 1, 2)):
    Nx, Ny, Nz = N
    sh = (Nx + 2, Ny + 2, Nz + 2)
    # k_fold_j_classes: reads the deposited current only
    Jf = [a.copy() for a in J]
    reps = list(face_representatives(N))
    assert len(set(reps)) == len(reps)
    assert set(reps) == {(i, l, k) for i in range(1, Nx + 1) for l in range(1, Ny + 1)
                         for k in range(1, Nz + 1)
                         if i in (1, Nx) or l in (1, Ny) or k in (1, Nz)}
    for ri, rl, rk in reps:
        px = 0 if ri == Nx else Nx + 1 if ri == 1 else -1
        py = 0 if rl == Ny else Ny + 1 if rl == 1 else -1
        pz = 0 if rk == Nz else Nz + 1 if rk == 1 else -1
        X, Y, Z = (ri, px), (rl, py), (rk, pz)

        def member(s):
            return ((s & 1) == 0 or px >= 0) and ((s & 2) == 0 or py >= 0) and \
                   ((s & 4) == 0 or pz >= 0)

        def at(s):
            return X[s & 1], Y[(s >> 1) & 1], Z[s >> 2]

        for cc in range(3):
            v = [J[cc][at(s)] if member(s) else 0.0 for s in range(8)]
            for d in fold_axes:
                if (px, py, pz)[d] < 0:
                    continue
                bit = 1 << d
                for s in range(8):
                    if not s & bit and member(s):
                        fold_pair(v, s, s | bit, d, cc, (ri, rl, rk)[d] == N[d])
            for s in range(8):
                if member(s):
                    Jf[cc][at(s)] = v[s]
    # k_h2_e: reads E, H, Q and the folded current; one thread per node

    def hn1(d, i, l, k):
        inr = (l <= Ny and k <= Nz) if d == 0 else (i <= Nx and k <= Nz) if d == 1 \
            else (i <= Nx and l <= Ny)
        return H[d][i, l, k] + Q[d][i, l, k] if inr else H[d][i, l, k]

    Eo = [a.copy() for a in E]
    Ho = [np.full(sh, np.nan) for _ in range(3)]
    wE = [np.zeros(sh, int) for _ in range(3)]
    wH = np.zeros(sh, int)
    for i in range(Nx + 2):
        for l in range(Ny + 2):
            for k in range(Nz + 2):
                hx0, hy0, hz0 = hn1(0, i, l, k), hn1(1, i, l, k), hn1(2, i, l, k)
                Ho[0][i, l, k], Ho[1][i, l, k], Ho[2][i, l, k] = hx0, hy0, hz0
                wH[i, l, k] += 1
                e = [None] * 3
                if i <= Nx and l <= Ny and k <= Nz:
                    if l >= 1 and k >= 1:
                        e[0] = E[0][i, l, k] + (C1Y * (hz0 - hn1(2, i, l - 1, k)) -
                                                C1Z * (hy0 - hn1(1, i, l, k - 1)) -
                                                TAU * Jf[0][i, l, k])
                    if i >= 1 and k >= 1:
                        e[1] = E[1][i, l, k] + (C1Z * (hx0 - hn1(0, i, l, k - 1)) -
                                                C1X * (hz0 - hn1(2, i - 1, l, k)) -
                                                TAU * Jf[1][i, l, k])
                    if i >= 1 and l >= 1:
                        e[2] = E[2][i, l, k] + (C1X * (hy0 - hn1(1, i - 1, l, k)) -
                                                C1Y * (hx0 - hn1(0, i, l - 1, k)) -
                                                TAU * Jf[2][i, l, k])
                    for cc in range(3):
                        if e[cc] is not None:
                            Eo[cc][i, l, k] = e[cc]
                            wE[cc][i, l, k] += 1
                gi = 0 if i == Nx else Nx + 1 if i == 1 else -1
                gl = 0 if l == Ny else Ny + 1 if l == 1 else -1
                gk = 0 if k == Nz else Nz + 1 if k == 1 else -1
                for s in range(1, 8):
                    ci, cl, ck = (gi if s & 1 else i), (gl if s & 2 else l), (gk if s & 4 else k)
                    if min(ci, cl, ck) < 0:
                        continue
                    for cc in range(3):
                        if ghost_src3(N, mode, cc, ci, cl, ck) != (i, l, k):
                            continue
                        Eo[cc][ci, cl, ck] = E[cc][i, l, k] if e[cc] is None else e[cc]
                        wE[cc][ci, cl, ck] += 1
    assert (wH == 1).all()
    for cc in range(3):
        assert (wE[cc] <= 1).all()
    return Eo, Ho, Jf


@pytest.mark.parametrize("ghost_mode", [0, 1])
@pytest.mark.parametrize("N", [(3, 3, 3), (3, 4, 5), (5, 3, 4), (4, 6, 3), (6, 5, 7)])
def test_fused_field_pass_matches_stage_sequence(N, ghost_mode):
    rng = np.random.default_rng(1)
    sh = tuple(n + 2 for n in N)
    E, H, Q, J = ([rng.standard_normal(sh) for _ in range(3)] for _ in range(4))
    Es, Hs, Js = stage_sequence(N, ghost_mode, E, H, Q, J)
    Ef, Hf, Jf = fused(N, ghost_mode, E, H, Q, J)
    for cc in range(3):
        assert np.array_equal(Ef[cc], Es[cc]), ("E", cc)
        assert np.array_equal(Hf[cc], Hs[cc]), ("H", cc)
        assert np.array_equal(Jf[cc], Js[cc]), ("J", cc)


def test_fused_field_pass_design_catches_a_wrong_fold_order():
    """The same restatement with the class fold applied z, y, x instead of
    x, y, z rounds the FP64 edge and corner sums differently, and the
    comparison sees it (the check above is not vacuous)."""
    rng = np.random.default_rng(2)
    N = (4, 5, 3)
    sh = tuple(n + 2 for n in N)
    E, H, Q, J = ([rng.standard_normal(sh) for _ in range(3)] for _ in range(4))
    _, _, Js = stage_sequence(N, 0, E, H, Q, J)
    _, _, Jw = fused(N, 0, E, H, Q, J, fold_axes=(2, 1, 0))
    assert any(not np.array_equal(a, b) for a, b in zip(Jw, Js))
