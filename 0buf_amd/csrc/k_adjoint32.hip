/*
 * k_adjoint32.hip -- particle adjoint of the FP32 precision form
 * (differentiable form at FP32): the reverse of k_push_deposit32 for one
 * step, in FP32 arithmetic, with the field adjoints accumulated in FP64.
 *
 * Positions are cell-local offsets, so the particle adjoints kept by id are
 * with respect to the offsets (cell units); the host converts them to
 * absolute coordinates (d/dx = d/do / h) at the ends of the sweep.  The
 * forward's dual cells, crossing code and relay point are recomputed with
 * the push's own expressions (k_push_math32.h, same contraction setting).
 *
 * Per particle, the chain reversed (cell units, J = fixed point x 2^-40):
 *   o1 = o + du, du = (tau/h) g1 p1            (Boris32)
 *   J  = seg(rd; o -> e1) + [m != 0] seg(r1; x2 -> o1),   e1 = m ? x2 : o1
 *   x2 on crossed axes 1/2, elsewhere o + du sr, sr = num / den
 * seg(rd; a -> b): q = kj (b - a) per axis, midpoint weights as vb_seg.
 */
#include "k_push_math32.h"

namespace {

constexpr int TXC = 2, TYC = 4, TZC = 8;

/* copies of the LDS adjoint box the gather adjoint scatters into: lane l adds
 * into copy l % ACOPIES, dividing the same-address conflicts of the FP64 LDS
 * atomics (the particles of one cell scatter to the same nodes); the copies
 * are summed before the global add */
#ifndef PIC_ACOPIES32
#define PIC_ACOPIES32 4
#endif
constexpr int ACOPIES = PIC_ACOPIES32;

struct AdjArgs32 {
    Grid g;
    TileGeom t;
    AdjFwd32 fwd;
    PartPtrs32 p;          /* compact record of s_{t-1} */
    const long *snap_off;
    float qm[PIC_MAX_SLOTS];
    float kj[PIC_MAX_SLOTS][3];   /* m h_d / tau: J per unit cell displacement along d */
    float tdh[3], hh[3], tau;
    AdjLam lam;
    LamParts lp;           /* FP64, by id, offset units */
    DevStatus *st;
};

/* adjoint of one vb_seg (k_particles_f32.hip) */
template <int BY, int BZ, int JY0, int JZ0>
__device__ __forceinline__ void seg_adj(const float *sL, int ca, int cb, int cc, int rdx, int rdy,
                                        int rdz, float ax, float ay, float az, float bx, float by,
                                        float bz, const float *kj, float gA[3], float gB[3]) {
    const float dx = bx - ax, dy = by - ay, dz = bz - az;
    const float mx = 0.5f * (ax + bx) - ((float)rdx - 0.5f);
    const float my = 0.5f * (ay + by) - ((float)rdy - 0.5f);
    const float mz = 0.5f * (az + bz) - ((float)rdz - 0.5f);
    const float lx = 1.0f - mx, ly = 1.0f - my, lz = 1.0f - mz;
    const float syz = dy * dz * (1.0f / 12.0f), sxz = dx * dz * (1.0f / 12.0f),
                sxy = dx * dy * (1.0f / 12.0f);
    const float qx = kj[0] * dx, qy = kj[1] * dy, qz = kj[2] * dz;
    const int n = ((ca + rdx) * BY + (cb + rdy)) * BZ + (cc + rdz);
    const float *Lx = sL, *Ly = sL + JY0, *Lz = sL + JZ0;
    const float a0 = Lx[n], a1 = Lx[n + 1], a2 = Lx[n + BZ], a3 = Lx[n + BZ + 1];
    const float b0 = Ly[n], b1 = Ly[n + 1], b2 = Ly[n + BY * BZ], b3 = Ly[n + BY * BZ + 1];
    const float c0 = Lz[n], c1 = Lz[n + BZ], c2 = Lz[n + BY * BZ], c3 = Lz[n + BY * BZ + BZ];
    const float Bx = a0 * (ly * lz + syz) + a1 * (ly * mz - syz) + a2 * (my * lz - syz) +
                     a3 * (my * mz + syz);
    const float By = b0 * (lx * lz + sxz) + b1 * (lx * mz - sxz) + b2 * (mx * lz - sxz) +
                     b3 * (mx * mz + sxz);
    const float Bz = c0 * (lx * ly + sxy) + c1 * (lx * my - sxy) + c2 * (mx * ly - sxy) +
                     c3 * (mx * my + sxy);
    const float Bx_my = -a0 * lz - a1 * mz + a2 * lz + a3 * mz;
    const float Bx_mz = -a0 * ly + a1 * ly - a2 * my + a3 * my;
    const float Bx_s = a0 - a1 - a2 + a3;
    const float By_mx = -b0 * lz - b1 * mz + b2 * lz + b3 * mz;
    const float By_mz = -b0 * lx + b1 * lx - b2 * mx + b3 * mx;
    const float By_s = b0 - b1 - b2 + b3;
    const float Bz_mx = -c0 * ly - c1 * my + c2 * ly + c3 * my;
    const float Bz_my = -c0 * lx + c1 * lx - c2 * mx + c3 * mx;
    const float Bz_s = c0 - c1 - c2 + c3;
    const float gmx = qy * By_mx + qz * Bz_mx;
    const float gmy = qx * Bx_my + qz * Bz_my;
    const float gmz = qx * Bx_mz + qy * By_mz;
    const float gdx = kj[0] * Bx + (qy * By_s * dz + qz * Bz_s * dy) * (1.0f / 12.0f);
    const float gdy = kj[1] * By + (qx * Bx_s * dz + qz * Bz_s * dx) * (1.0f / 12.0f);
    const float gdz = kj[2] * Bz + (qx * Bx_s * dy + qy * By_s * dx) * (1.0f / 12.0f);
    gA[0] = 0.5f * gmx - gdx;
    gA[1] = 0.5f * gmy - gdy;
    gA[2] = 0.5f * gmz - gdz;
    gB[0] = 0.5f * gmx + gdx;
    gB[1] = 0.5f * gmy + gdy;
    gB[2] = 0.5f * gmz + gdz;
}

/* adjoint of tri(): scatter lam x weight into the FP64 adjoint box, return
 * d(lam * value) / d(near weights) */
template <int BY, int BZ>
__device__ __forceinline__ void tri_adj(const float *F, double *A, int u, int v, int w, float x0,
                                        float y0, float z0, float lam, float &dx, float &dy,
                                        float &dz) {
    const float x1 = 1.0f - x0, y1 = 1.0f - y0, z1 = 1.0f - z0;
    const int o = (u * BY + v) * BZ + w;
    const float f000 = F[o], f001 = F[o + 1], f010 = F[o + BZ], f011 = F[o + BZ + 1];
    const float f100 = F[o + BY * BZ], f101 = F[o + BY * BZ + 1];
    const float f110 = F[o + BY * BZ + BZ], f111 = F[o + BY * BZ + BZ + 1];
    const float yz00 = y0 * z0, yz01 = y0 * z1, yz10 = y1 * z0, yz11 = y1 * z1;
    const float p0 = yz00 * f000 + yz01 * f001 + yz10 * f010 + yz11 * f011;
    const float p1 = yz00 * f100 + yz01 * f101 + yz10 * f110 + yz11 * f111;
    dx += lam * (p0 - p1);
    const float ly0 = x0 * (z0 * f000 + z1 * f001) + x1 * (z0 * f100 + z1 * f101);
    const float ly1 = x0 * (z0 * f010 + z1 * f011) + x1 * (z0 * f110 + z1 * f111);
    dy += lam * (ly0 - ly1);
    const float lz0 = x0 * (y0 * f000 + y1 * f010) + x1 * (y0 * f100 + y1 * f110);
    const float lz1 = x0 * (y0 * f001 + y1 * f011) + x1 * (y0 * f101 + y1 * f111);
    dz += lam * (lz0 - lz1);
    const double l0 = (double)(lam * x0), l1 = (double)(lam * x1);
    atomicAdd(&A[o], l0 * yz00);
    atomicAdd(&A[o + 1], l0 * yz01);
    atomicAdd(&A[o + BZ], l0 * yz10);
    atomicAdd(&A[o + BZ + 1], l0 * yz11);
    atomicAdd(&A[o + BY * BZ], l1 * yz00);
    atomicAdd(&A[o + BY * BZ + 1], l1 * yz01);
    atomicAdd(&A[o + BY * BZ + BZ], l1 * yz10);
    atomicAdd(&A[o + BY * BZ + BZ + 1], l1 * yz11);
}

__device__ __forceinline__ void cross3(const float a[3], const float b[3], float r[3]) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ float dot3f(const float a[3], const float b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

template <int TX, int TY, int TZ>
__global__ __launch_bounds__(PIC_BLOCK) void k_adj_particles32(AdjArgs32 a) {
    constexpr int C = TX * TY * TZ;
    constexpr int BX = TX + 2, BY = TY + 2, BZ = TZ + 2, NB = BX * BY * BZ;
    /* the push's box layout: Ex, Hy, Hz (node-type along x) take x planes
     * 0..TX only (NBX nodes), the others all BX planes; the adjoint current
     * box has the E boxes' layout (Jx takes x planes 0..TX) */
    constexpr int NBX = (TX + 1) * BY * BZ;
    constexpr int FB[6] = {0, NBX, NBX + NB, NBX + 2 * NB, NBX + 3 * NB, 2 * NBX + 3 * NB};
    constexpr int NF = 3 * NBX + 3 * NB, NJ = NBX + 2 * NB;
    constexpr int JY0 = FB[1], JZ0 = FB[2];
    __shared__ float sF[NF];
    __shared__ float sL[NJ];
    __shared__ double sA[ACOPIES * NF];
    double *const myA = sA + (threadIdx.x % ACOPIES) * NF;
    __shared__ int sOff[C + 1];
    __shared__ int sCnt[C];

    const Grid &g = a.g;
    const int tid = threadIdx.x;
    long tile;
    {
        long nt = a.t.ntiles, b = blockIdx.x;
        long q = nt / 8, r = nt % 8, x = b % 8;
        tile = x * q + (x < r ? x : r) + b / 8;
    }
    const int tz = (int)(tile % a.t.ntz);
    const int ty = (int)((tile / a.t.ntz) % a.t.nty);
    const int tx = (int)(tile / ((long)a.t.ntz * a.t.nty));
    const int I0 = tx * TX, L0 = ty * TY, K0 = tz * TZ;
    const int cap = a.p.cap;

    /* thread n takes box node n in every component box (one decode, one
     * 32-bit offset: the grid has fewer than 2^31 nodes) and cell n of the
     * tile; every global load is issued before the first LDS write */
    static_assert(NB <= PIC_BLOCK && C <= PIC_BLOCK, "one box node and one cell per thread");
    constexpr int FN[6] = {NBX, NB, NB, NB, NBX, NBX};   /* nodes in each box */
    const float *const Fsrc[6] = {a.fwd.E[0], a.fwd.E[1], a.fwd.E[2],
                                  a.fwd.H[0], a.fwd.H[1], a.fwd.H[2]};
    const int nu = tid / (BY * BZ), nv = (tid / BZ) % BY, nw = tid % BZ;
    const bool node_in = tid < NB && I0 + nu <= g.Nx + 1 && L0 + nv <= g.Ny + 1 &&
                         K0 + nw <= g.Nz + 1;
    const unsigned noff = (unsigned)(I0 + nu) * g.sx + (unsigned)(L0 + nv) * g.sy + (K0 + nw);
    {
        float fv[6], lj[3];
#pragma unroll
        for (int q = 0; q < 6; q++) fv[q] = 0.0f;
#pragma unroll
        for (int q = 0; q < 3; q++) lj[q] = 0.0f;
        if (node_in) {
#pragma unroll
            for (int q = 0; q < 6; q++) fv[q] = Fsrc[q][noff];
#pragma unroll
            for (int q = 0; q < 3; q++) lj[q] = (float)a.lam.J[q][noff];
        }
        const int gi = I0 + tid / (TY * TZ), gl = L0 + (tid / TZ) % TY, gk = K0 + tid % TZ;
        int cnt = 0;
        if (tid < C && gi < g.Nx && gl < g.Ny && gk < g.Nz) {
            cnt = a.p.count[((long)gi * g.Ny + gl) * g.Nz + gk];
            cnt = cnt < cap ? cnt : cap;
        }
        if (tid < NB) {
#pragma unroll
            for (int q = 0; q < 6; q++) {
                if (tid >= FN[q]) continue;
                sF[FB[q] + tid] = fv[q];
                if (q < 3) sL[FB[q] + tid] = lj[q];
            }
        }
        if (tid < C) sCnt[tid] = cnt;
    }
    for (int e = tid; e < ACOPIES * NF; e += PIC_BLOCK) sA[e] = 0.0;
    __syncthreads();
    if (tid == 0) {
        int s = 0;
        for (int c = 0; c < C; c++) {
            sOff[c] = s;
            s += sCnt[c];
        }
        sOff[C] = s;
    }
    __syncthreads();
    const int T = sOff[C];

    for (int fidx = tid; fidx < T; fidx += PIC_BLOCK) {
        int blo = 0, bhi = C;
        while (bhi - blo > 1) {
            int mid = (blo + bhi) >> 1;
            if (sOff[mid] <= fidx) blo = mid; else bhi = mid;
        }
        const int c = blo;
        const int ca = c / (TY * TZ), cb = (c / TZ) % TY, cc = c % TZ;
        const long pidx = a.snap_off[tile] + fidx;    /* compact record */
        const float o[3] = {a.p.p[0][pidx], a.p.p[1][pidx], a.p.p[2][pidx]};
        const float pu = a.p.p[3][pidx], pv = a.p.p[4][pidx], pw = a.p.p[5][pidx];
        const unsigned char s = a.p.sp[pidx];
        const int pid = a.p.id[pidx];
        const float tau1 = a.qm[s] * a.tau * 0.5f;

        /* ---- forward recompute ---- */
        const Weights32 W = weights32(o[0], o[1], o[2]);
        const int rdx = W.rd[0], rdy = W.rd[1], rdz = W.rd[2];
        const float cx0 = W.c0[0], cx1 = 1.0f - cx0, cy0 = W.c0[1], cy1 = 1.0f - cy0,
                    cz0 = W.c0[2], cz1 = 1.0f - cz0;
        const float nx0 = W.n0[0], ny0 = W.n0[1], nz0 = W.n0[2];
        const float ox = o[0], oy = o[1], oz = o[2];
        float Ef[3], Hf[3];
        Ef[0] = tri<BY, BZ>(sF + FB[0], ca, cb + rdy, cc + rdz, ox, cy1, cz1);
        Ef[1] = tri<BY, BZ>(sF + FB[1], ca + rdx, cb, cc + rdz, cx1, oy, cz1);
        Ef[2] = tri<BY, BZ>(sF + FB[2], ca + rdx, cb + rdy, cc, cx1, cy1, oz);
        Hf[0] = tri<BY, BZ>(sF + FB[3], ca + rdx, cb, cc, cx1, oy, oz);
        Hf[1] = tri<BY, BZ>(sF + FB[4], ca, cb + rdy, cc, ox, cy1, oz);
        Hf[2] = tri<BY, BZ>(sF + FB[5], ca, cb, cc + rdz, ox, oy, cz1);
        const Boris32 B = boris32(pu, pv, pw, Ef[0], Ef[1], Ef[2], Hf[0], Hf[1], Hf[2], tau1, a.tdh);
        const float o1[3] = {end_offset(ox, B.du[0]), end_offset(oy, B.du[1]),
                             end_offset(oz, B.du[2])};
        const Relay32 R = relay32(ox, oy, oz, o1[0], o1[1], o1[2], B.du, W.rd, a.hh);
        const int m = R.m;

        /* ---- reverse ---- */
        double2 *const lq = (double2 *)(a.lp.v + (long)pid * PIC_LSTRIDE);
        const double2 q0 = lq[0], q1 = lq[1], q2 = lq[2];
        float lo1[3] = {(float)q0.x, (float)q0.y, (float)q1.x};
        float lp1[3] = {(float)q1.y, (float)q2.x, (float)q2.y};
        float lo[3], ldu[3] = {0.0f, 0.0f, 0.0f};
        const float *kj = a.kj[s];
        {
            float gA[3], gB[3];
            const float e1[3] = {m ? R.x2[0] : o1[0], m ? R.x2[1] : o1[1], m ? R.x2[2] : o1[2]};
            seg_adj<BY, BZ, JY0, JZ0>(sL, ca, cb, cc, rdx, rdy, rdz, ox, oy, oz, e1[0], e1[1], e1[2],
                                kj, gA, gB);
            for (int d = 0; d < 3; d++) lo[d] = gA[d];
            if (m == 0) {
                for (int d = 0; d < 3; d++) lo1[d] += gB[d];
            } else {
                float hA[3], hB[3];
                seg_adj<BY, BZ, JY0, JZ0>(sL, ca, cb, cc, R.r1[0], R.r1[1], R.r1[2], R.x2[0], R.x2[1],
                                    R.x2[2], o1[0], o1[1], o1[2], kj, hA, hB);
                for (int d = 0; d < 3; d++) lo1[d] += hB[d];
                const float lx2[3] = {gB[0] + hA[0], gB[1] + hA[1], gB[2] + hA[2]};
                const bool cr[3] = {(m & 4) != 0, (m & 2) != 0, (m & 1) != 0};
                float lsr = 0.0f;
                for (int d = 0; d < 3; d++) {
                    if (cr[d]) continue;
                    lo[d] += lx2[d];
                    ldu[d] += lx2[d] * R.sr;
                    lsr += lx2[d] * B.du[d];
                }
                const int nc = (int)cr[0] + (int)cr[1] + (int)cr[2];
                /* (the crossed axes by an unrolled loop: a computed index
                 * would put lo, ldu and du in LDS, a slice per thread) */
                if (nc == 1) {
                    for (int d = 0; d < 3; d++) {
                        if (!cr[d]) continue;
                        lo[d] += lsr * (-1.0f / B.du[d]);
                        ldu[d] += lsr * (-R.sr / B.du[d]);
                    }
                } else if (nc == 2) {
                    for (int d = 0; d < 3; d++) {
                        if (!cr[d]) continue;
                        lo[d] += lsr * (-B.du[d] * a.hh[d]) / R.den;
                        ldu[d] += lsr * ((0.5f - o[d]) * a.hh[d] - R.sr * 2.0f * B.du[d] * a.hh[d]) /
                                  R.den;
                    }
                }
            }
        }
        /* o1 = o + du;  du = (tau/h) g1 p1 */
        for (int d = 0; d < 3; d++) {
            lo[d] += lo1[d];
            ldu[d] += lo1[d];
        }
        {
            const float w[3] = {a.tdh[0] * ldu[0], a.tdh[1] * ldu[1], a.tdh[2] * ldu[2]};
            const float g2 = B.gi2, g23 = g2 * g2 * g2;
            const float pl = dot3f(B.p1, w);
            for (int d = 0; d < 3; d++) lp1[d] += g2 * w[d] - g23 * pl * B.p1[d];
        }
        /* Boris reverse (Cayley rotation, as the FP64 adjoint) */
        float lE[3], lH[3], lpm[3];
        {
            for (int d = 0; d < 3; d++) lE[d] = tau1 * lp1[d];
            const float *b = B.b, *sv = B.s;
            const float rD = B.rden;
            const float lN[3] = {lp1[0] * rD, lp1[1] * rD, lp1[2] * rD};
            float sxb[3];
            cross3(sv, b, sxb);
            const float bs = dot3f(b, sv);
            float pp[3];
            for (int d = 0; d < 3; d++) pp[d] = (sv[d] + bs * b[d] + sxb[d]) * rD;
            const float bl = dot3f(b, lN), lpp = dot3f(lp1, pp);
            float bxl[3], lxs[3];
            cross3(b, lN, bxl);
            cross3(lN, sv, lxs);
            float ls[3], lb[3];
            for (int d = 0; d < 3; d++) {
                ls[d] = lN[d] + b[d] * bl + bxl[d];
                lb[d] = sv[d] * bl + bs * lN[d] + lxs[d] - 2.0f * lpp * rD * b[d];
            }
            float bxls[3], lsxpm[3];
            cross3(b, ls, bxls);
            cross3(ls, B.pm, lsxpm);
            for (int d = 0; d < 3; d++) {
                lpm[d] = ls[d] + bxls[d];
                lb[d] += lsxpm[d];
            }
            const float lps = dot3f(lb, Hf);
            for (int d = 0; d < 3; d++) lH[d] = B.ps * lb[d];
            if (tau1 != 0.0f) {
                const float g1 = B.ps / tau1;
                const float k = -tau1 * g1 * g1 * g1 * lps;
                for (int d = 0; d < 3; d++) lpm[d] += k * B.pm[d];
            }
            for (int d = 0; d < 3; d++) lE[d] += tau1 * lpm[d];
        }
        /* gather adjoint: every near weight has d/do = -1 */
        {
            float dx = 0.0f, dy = 0.0f, dz = 0.0f;
            tri_adj<BY, BZ>(sF + FB[0], myA + FB[0], ca, cb + rdy, cc + rdz, nx0, cy0, cz0, lE[0], dx, dy, dz);
            tri_adj<BY, BZ>(sF + FB[1], myA + FB[1], ca + rdx, cb, cc + rdz, cx0, ny0, cz0, lE[1], dx, dy, dz);
            tri_adj<BY, BZ>(sF + FB[2], myA + FB[2], ca + rdx, cb + rdy, cc, cx0, cy0, nz0, lE[2], dx, dy, dz);
            tri_adj<BY, BZ>(sF + FB[3], myA + FB[3], ca + rdx, cb, cc, cx0, ny0, nz0, lH[0], dx, dy, dz);
            tri_adj<BY, BZ>(sF + FB[4], myA + FB[4], ca, cb + rdy, cc, nx0, cy0, nz0, lH[1], dx, dy, dz);
            tri_adj<BY, BZ>(sF + FB[5], myA + FB[5], ca, cb, cc + rdz, nx0, ny0, cz0, lH[2], dx, dy, dz);
            lo[0] -= dx;
            lo[1] -= dy;
            lo[2] -= dz;
        }
        lq[0] = make_double2(lo[0], lo[1]);
        lq[1] = make_double2(lo[2], lpm[0]);
        lq[2] = make_double2(lpm[1], lpm[2]);
    }
    __syncthreads();
    /* thread n adds box node n of the six adjoint boxes */
    if (tid < NB) {
#pragma unroll
        for (int q = 0; q < 6; q++) {
            if (tid >= FN[q]) continue;
            const int e = FB[q] + tid;
            double val = sA[e];
            for (int r = 1; r < ACOPIES; r++) val += sA[r * NF + e];
            if (val != 0.0 && node_in)
                atomicAdd(q < 3 ? &a.lam.E[q][noff] : &a.lam.H[q - 3][noff], val);
        }
    }
}

/* particle adjoints between absolute (d/dx) and offset units (d/do = h d/dx) */
__global__ __launch_bounds__(PIC_BLOCK) void k_scale_lam(LamParts lp, double sx, double sy,
                                                         double sz) {
    const long e = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (e >= lp.n) return;
    double *const v = lp.v + e * PIC_LSTRIDE;
    v[0] *= sx;
    v[1] *= sy;
    v[2] *= sz;
}

__global__ __launch_bounds__(PIC_BLOCK) void k_lam_copy(LamParts lp, double *flat, int k,
                                                        int to_lines) {
    const long e = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (e >= lp.n) return;
    if (to_lines) lp.v[e * PIC_LSTRIDE + k] = flat[e];
    else flat[e] = lp.v[e * PIC_LSTRIDE + k];
}

__global__ __launch_bounds__(PIC_BLOCK) void k_adj_seed_energy32(Grid g, const float *Ex,
                                                                 const float *Ey, const float *Ez,
                                                                 double w2, AdjLam l) {
    const long n = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (n >= g.G) return;
    l.E[0][n] += w2 * (double)Ex[n];
    l.E[1][n] += w2 * (double)Ey[n];
    l.E[2][n] += w2 * (double)Ez[n];
}

inline unsigned blocks_for(long n) { return (unsigned)((n + PIC_BLOCK - 1) / PIC_BLOCK); }

}  // namespace

void launch_adj_particles_f32(const Grid &g, const TileGeom &t, AdjFwd32 fwd, PartPtrs32 snap,
                              const long *snap_off, SpeciesTable sp, AdjLam l, LamParts lp,
                              DevStatus *st, hipStream_t s) {
    AdjArgs32 a;
    a.g = g;
    a.t = t;
    a.fwd = fwd;
    a.p = snap;
    a.snap_off = snap_off;
    const double h[3] = {g.hx, g.hy, g.hz};
    for (int k = 0; k < PIC_MAX_SLOTS; k++) {
        a.qm[k] = (float)sp.q_m[k];
        for (int d = 0; d < 3; d++) a.kj[k][d] = (float)(sp.m[k] * h[d] / g.tau);
    }
    for (int d = 0; d < 3; d++) {
        a.tdh[d] = (float)(g.tau / h[d]);
        a.hh[d] = (float)(h[d] * h[d]);
    }
    a.tau = (float)g.tau;
    a.lam = l;
    a.lp = lp;
    a.st = st;
    hipLaunchKernelGGL((k_adj_particles32<TXC, TYC, TZC>), dim3((unsigned)t.ntiles),
                       dim3(PIC_BLOCK), 0, s, a);
}

void launch_lam_copy(LamParts lp, double *flat, int k, bool to_lines, hipStream_t s) {
    hipLaunchKernelGGL(k_lam_copy, dim3((unsigned)((lp.n + PIC_BLOCK - 1) / PIC_BLOCK)),
                       dim3(PIC_BLOCK), 0, s, lp, flat, k, (int)to_lines);
}

void launch_scale_lam(LamParts lp, double sx, double sy, double sz, hipStream_t s) {
    if (lp.n <= 0) return;
    hipLaunchKernelGGL(k_scale_lam, dim3(blocks_for(lp.n)), dim3(PIC_BLOCK), 0, s, lp, sx, sy, sz);
}

void launch_adj_seed_energy32(const Grid &g, const float *const E[3], double w, AdjLam l,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_adj_seed_energy32, dim3(blocks_for(g.G)), dim3(PIC_BLOCK), 0, s, g, E[0],
                       E[1], E[2], 2.0 * w, l);
}
