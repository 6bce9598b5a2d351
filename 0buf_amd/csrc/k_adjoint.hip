/*
 * k_adjoint.hip -- adjoint kernels of the FP64 step (differentiable form).
 *
 * Forward step t (s_{t-1} -> s_t, Plasma::Step, gpu_plasma.h:605-620):
 *     Q   = Ch E                     (MagneticStageOne, on the Q ranges)
 *     H'  = H + Q                    gather reads E and H'
 *     p1  = Boris(p, E(x), H'(x)),   x1 = x + tau u(p1)
 *     J   = VB deposit of x -> x1    (per segment pqr)
 *     x   = Reflect(x1)              (a shift: derivative 1)
 *     Jf  = Fold(J)
 *     H'' = H' + Q                   (same Q)
 *     E'  = P(E + Ce H'' - tau Jf)   (E ranges, then the periodic ghost copies)
 *     W_t = sum E'^2
 * The reverse step takes (lamE, lamH, lamP) at s_t to s_{t-1}:
 *   1. lamE += 2 w_t E_t                                (k_adj_seed_energy)
 *   2. lamE <- P^T lamE                                 (k_adj_e_periodic_*)
 *   3. lamH += Ce^T lamE|R ; lamQ := lamH|RQ ;
 *      lamJ := -tau lamE|R                              (k_adj_curl_e)
 *   4. lamJ <- Fold^T lamJ                              (k_adj_fold_j)
 *   5. particle adjoint: reverse deposit (a gather of lamJ), reverse push,
 *      reverse Boris, reverse gather (an atomic scatter into lamE, lamH)
 *                                                       (k_adj_particles)
 *   6. lamE += Ch^T (lamQ + lamH|RQ)                    (k_adj_curl_h)
 * after which lamH is the adjoint of H_{t-1} and lamE that of E_{t-1}.
 * Field transposes are gathers (one node per lane, deterministic); the
 * particle stage accumulates lamE, lamH in an LDS box per tile with FP64
 * atomics and adds the box to the global arrays with FP64 atomics.
 */
#include "k_push_math.h"

#pragma clang fp contract(off)

namespace {

constexpr int TXC = 2, TYC = 4, TZC = 8;

/* copies of the LDS adjoint box the gather adjoint scatters into: lane l adds
 * into copy l % ACOPIES, dividing the same-address conflicts of the FP64 LDS
 * atomics (the particles of one cell scatter to the same nodes); the copies
 * are summed before the global add */
#ifndef PIC_ACOPIES64
#define PIC_ACOPIES64 3
#endif
constexpr int ACOPIES = PIC_ACOPIES64;

inline unsigned blocks_for(long n) { return (unsigned)((n + PIC_BLOCK - 1) / PIC_BLOCK); }

__device__ __forceinline__ void decode(const Grid &g, long n, int &i, int &l, int &k) {
    k = (int)(n % g.sy);
    long t = n / g.sy;
    l = (int)(t % (g.Ny + 2));
    i = (int)(t / (g.Ny + 2));
}

__global__ __launch_bounds__(PIC_BLOCK) void k_adj_seed_energy(Grid g, const double *Ex,
                                                               const double *Ey, const double *Ez,
                                                               double w2, AdjLam l) {
    const long n = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (n >= g.G) return;
    l.E[0][n] += w2 * Ex[n];
    l.E[1][n] += w2 * Ey[n];
    l.E[2][n] += w2 * Ez[n];
}

/* transpose of E[0] = E[N]; E[N+1] = E[1] along one line */
__device__ __forceinline__ void ghost_copy_T(double *L, long n0, long n1, long nN, long nN1) {
    L[n1] += L[nN1];
    L[nN1] = 0.0;
    L[nN] += L[n0];
    L[n0] = 0.0;
}

/* transpose of k_e_periodic_z (Ex over i in [0,Nx], l in [0,Ny+1]; Ey over
 * i in [0,Nx+1], l in [0,Ny]) */
__global__ __launch_bounds__(PIC_BLOCK) void k_adj_e_periodic_z(Grid g, AdjLam lam) {
    const long t = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (t >= (long)(g.Nx + 2) * (g.Ny + 2)) return;
    const int i = (int)(t / (g.Ny + 2)), l = (int)(t % (g.Ny + 2));
    const long base = (long)i * g.sx + (long)l * g.sy;
    if (l <= g.Ny) ghost_copy_T(lam.E[1], base, base + 1, base + g.Nz, base + g.Nz + 1);
    if (i <= g.Nx) ghost_copy_T(lam.E[0], base, base + 1, base + g.Nz, base + g.Nz + 1);
}

/* transpose of k_e_periodic_x (Ey over l in [0,Ny], k in [1,Nz]; Ez over l
 * in [1,Ny], k in [0,Nz]) */
__global__ __launch_bounds__(PIC_BLOCK) void k_adj_e_periodic_x(Grid g, AdjLam lam) {
    const long t = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (t >= (long)(g.Ny + 2) * g.sy) return;
    const int l = (int)(t / g.sy), k = (int)(t % g.sy);
    const long base = (long)l * g.sy + k;
    const long n0 = base, n1 = (long)g.sx + base, nN = (long)g.Nx * g.sx + base,
               nN1 = (long)(g.Nx + 1) * g.sx + base;
    if (l >= 1 && l <= g.Ny && k <= g.Nz) ghost_copy_T(lam.E[2], n0, n1, nN, nN1);
    if (l <= g.Ny && k >= 1 && k <= g.Nz) ghost_copy_T(lam.E[1], n0, n1, nN, nN1);
}

/* transpose of k_e_periodic_y (ghost_mode 1: Ex and Ez along y) */
__global__ __launch_bounds__(PIC_BLOCK) void k_adj_e_periodic_y(Grid g, AdjLam lam) {
    const long t = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (t >= (long)(g.Nx + 2) * g.sy) return;
    const int i = (int)(t / g.sy), k = (int)(t % g.sy);
    const long base = (long)i * g.sx + k;
    const long n0 = base, n1 = base + g.sy, nN = base + (long)g.Ny * g.sy,
               nN1 = base + (long)(g.Ny + 1) * g.sy;
    ghost_copy_T(lam.E[2], n0, n1, nN, nN1);
    ghost_copy_T(lam.E[0], n0, n1, nN, nN1);
}

/* E-update ranges (emeElement, kernels.cu:1162-1181) */
__device__ __forceinline__ bool in_ex(const Grid &g, int i, int l, int k) {
    return i >= 0 && i <= g.Nx && l >= 1 && l <= g.Ny && k >= 1 && k <= g.Nz;
}
__device__ __forceinline__ bool in_ey(const Grid &g, int i, int l, int k) {
    return i >= 1 && i <= g.Nx && l >= 0 && l <= g.Ny && k >= 1 && k <= g.Nz;
}
__device__ __forceinline__ bool in_ez(const Grid &g, int i, int l, int k) {
    return i >= 1 && i <= g.Nx && l >= 1 && l <= g.Ny && k >= 0 && k <= g.Nz;
}
/* Q ranges (emh1_Element, kernels.cu:1052-1089) */
__device__ __forceinline__ bool in_qx(const Grid &g, int i, int l, int k) {
    return i >= 0 && i <= g.Nx + 1 && l >= 0 && l <= g.Ny && k >= 0 && k <= g.Nz;
}
__device__ __forceinline__ bool in_qy(const Grid &g, int i, int l, int k) {
    return i >= 0 && i <= g.Nx && l >= 0 && l <= g.Ny + 1 && k >= 0 && k <= g.Nz;
}
__device__ __forceinline__ bool in_qz(const Grid &g, int i, int l, int k) {
    return i >= 0 && i <= g.Nx && l >= 0 && l <= g.Ny && k >= 0 && k <= g.Nz + 1;
}

/* transpose of the E update  E += c1y(Hz - Hz[l-1]) - c1z(Hy - Hy[k-1]) - tau Jx  etc.
 * per H node; also lamJ = -tau lamE on the E ranges and lamQ = lamH'' on the
 * Q ranges */
__global__ __launch_bounds__(PIC_BLOCK) void k_adj_curl_e(Grid g, AdjLam lam) {
    const long n = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (n >= g.G) return;
    int i, l, k;
    decode(g, n, i, l, k);
    const double c1x = g.tau / g.hx, c1y = g.tau / g.hy, c1z = g.tau / g.hz, tau = g.tau;
    const double *LEx = lam.E[0], *LEy = lam.E[1], *LEz = lam.E[2];
    const long sy = g.sy, sx = g.sx;
    const double ex = in_ex(g, i, l, k) ? LEx[n] : 0.0;
    const double ey = in_ey(g, i, l, k) ? LEy[n] : 0.0;
    const double ez = in_ez(g, i, l, k) ? LEz[n] : 0.0;
    const double ex_l = in_ex(g, i, l + 1, k) ? LEx[n + sy] : 0.0;
    const double ex_k = in_ex(g, i, l, k + 1) ? LEx[n + 1] : 0.0;
    const double ey_i = in_ey(g, i + 1, l, k) ? LEy[n + sx] : 0.0;
    const double ey_k = in_ey(g, i, l, k + 1) ? LEy[n + 1] : 0.0;
    const double ez_i = in_ez(g, i + 1, l, k) ? LEz[n + sx] : 0.0;
    const double ez_l = in_ez(g, i, l + 1, k) ? LEz[n + sy] : 0.0;
    const double hx = lam.H[0][n] + (c1z * (ey - ey_k) - c1y * (ez - ez_l));
    const double hy = lam.H[1][n] + (-c1z * (ex - ex_k) + c1x * (ez - ez_i));
    const double hz = lam.H[2][n] + (c1y * (ex - ex_l) - c1x * (ey - ey_i));
    lam.H[0][n] = hx;
    lam.H[1][n] = hy;
    lam.H[2][n] = hz;
    lam.Q[0][n] = in_qx(g, i, l, k) ? hx : 0.0;
    lam.Q[1][n] = in_qy(g, i, l, k) ? hy : 0.0;
    lam.Q[2][n] = in_qz(g, i, l, k) ? hz : 0.0;
    lam.J[0][n] = -tau * ex;
    lam.J[1][n] = -tau * ey;
    lam.J[2][n] = -tau * ez;
}

/* transpose of the J fold for one face direction (k_fold_j: per component
 * c: if dir != c: J[1] += J[N-1]; if !(dir == y && c == y): J[N-1] = J[1];
 * J[0] += J[N-2]; J[N-2] = J[0]), operations reversed */
__global__ __launch_bounds__(PIC_BLOCK) void k_adj_fold_j(Grid g, AdjLam lam, int dir) {
    const long t = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    int na, nb, N;
    if (dir == 0) { na = g.Ny + 2; nb = g.Nz + 2; N = g.Nx + 2; }
    else if (dir == 1) { na = g.Nx + 2; nb = g.Nz + 2; N = g.Ny + 2; }
    else { na = g.Nx + 2; nb = g.Ny + 2; N = g.Nz + 2; }
    if (t >= (long)na * nb) return;
    const int a = (int)(t / nb), b = (int)(t % nb);
    long base, stride;
    if (dir == 0) { base = (long)a * g.sy + b; stride = g.sx; }
    else if (dir == 1) { base = (long)a * g.sx + b; stride = g.sy; }
    else { base = (long)a * g.sx + (long)b * g.sy; stride = 1; }
    const long n1 = base + stride, nNm1 = base + (long)(N - 1) * stride;
    const long nNm2 = base + (long)(N - 2) * stride, n0 = base;
    for (int c = 2; c >= 0; c--) {
        double *L = lam.J[c];
        L[n0] += L[nNm2];
        L[nNm2] = 0.0;
        L[nNm2] += L[n0];
        if (dir != 1 || c != 1) {
            L[n1] += L[nNm1];
            L[nNm1] = 0.0;
        }
        if (dir != c) L[nNm1] += L[n1];
    }
}

/* transpose of Q = Ch E:  Qx = 1/2(c1z(Ey[k+1]-Ey) - c1y(Ez[l+1]-Ez)) etc.,
 * with the Q adjoint mu = lamQ + lamH on the Q ranges */
__global__ __launch_bounds__(PIC_BLOCK) void k_adj_curl_h(Grid g, AdjLam lam) {
    const long n = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (n >= g.G) return;
    int i, l, k;
    decode(g, n, i, l, k);
    const double c1x = g.tau / g.hx, c1y = g.tau / g.hy, c1z = g.tau / g.hz;
    const long sy = g.sy, sx = g.sx;
    auto mux = [&](int ii, int ll, int kk, long m) {
        return in_qx(g, ii, ll, kk) ? lam.Q[0][m] + lam.H[0][m] : 0.0;
    };
    auto muy = [&](int ii, int ll, int kk, long m) {
        return in_qy(g, ii, ll, kk) ? lam.Q[1][m] + lam.H[1][m] : 0.0;
    };
    auto muz = [&](int ii, int ll, int kk, long m) {
        return in_qz(g, ii, ll, kk) ? lam.Q[2][m] + lam.H[2][m] : 0.0;
    };
    const double mx = mux(i, l, k, n), my = muy(i, l, k, n), mz = muz(i, l, k, n);
    const double mx_k = k >= 1 ? mux(i, l, k - 1, n - 1) : 0.0;
    const double mx_l = l >= 1 ? mux(i, l - 1, k, n - sy) : 0.0;
    const double my_k = k >= 1 ? muy(i, l, k - 1, n - 1) : 0.0;
    const double my_i = i >= 1 ? muy(i - 1, l, k, n - sx) : 0.0;
    const double mz_l = l >= 1 ? muz(i, l - 1, k, n - sy) : 0.0;
    const double mz_i = i >= 1 ? muz(i - 1, l, k, n - sx) : 0.0;
    lam.E[0][n] += 0.5 * (c1z * (my - my_k) + c1y * (mz_l - mz));
    lam.E[1][n] += 0.5 * (c1z * (mx_k - mx) + c1x * (mz - mz_i));
    lam.E[2][n] += 0.5 * (c1y * (mx - mx_l) + c1x * (my_i - my));
}

/* ---------------------------------------------------------------- particles */

struct AdjArgs {
    Grid g;
    DivConst dc;
    TileGeom t;
    AdjFwd fwd;
    PartPtrs p;          /* compact record of s_{t-1} */
    const long *snap_off;
    SpeciesTable sp;
    AdjLam lam;
    LamParts lp;
    DevStatus *st;
};

/* adjoint of one pqr segment (cell.h:1511-1693): gradient of
 * sum_nodes lamJ[node] * contribution(A -> B) with respect to the segment
 * ends A and B; lamJ is read from the LDS box at the forward's nodes */
template <int BY, int BZ, int NB>
__device__ __forceinline__ void pqr_adj(const double *sL, const Geo &c, const DivConst &dc, int ix,
                                        int iy, int iz, double ax, double ay, double az, double bx,
                                        double by, double bz, double q, double gA[3],
                                        double gB[3]) {
    const double dx = (0.50 * (ax + bx) - c.hx * ((double)ix - 2.50) - c.x0);
    const double dy = (0.50 * (ay + by) - c.hy * ((double)iy - 2.50) - c.y0);
    const double dz = (0.50 * (az + bz) - c.hz * ((double)iz - 2.50) - c.z0);
    const double dx1 = c.hx - dx, dy1 = c.hy - dy, dz1 = c.hz - dz;
    const double Dx = bx - ax, Dy = by - ay, Dz = bz - az;
    const double s1 = Dy * Dz * dc.r12, s2 = Dx * Dz * dc.r12, s3 = Dx * Dy * dc.r12;
    const double kx = q * dc.rDx, ky = q * dc.rDy, kz = q * dc.rDz;
    const double SU = Dx * kx, SV = Dy * ky, SW = Dz * kz;
    const int u = c.ca + ix - 2, v = c.cb + iy - 2, w = c.cc + iz - 2;
    const int n = (u * BY + v) * BZ + w;
    /* the forward deposits nothing outside the tile stencil (and flags it) */
    if (ix < 2 || ix > 3 || iy < 2 || iy > 3 || iz < 2 || iz > 3) {
        for (int d = 0; d < 3; d++) gA[d] = gB[d] = 0.0;
        return;
    }
    const double *Lx = sL, *Ly = sL + NB, *Lz = sL + 2 * NB;
    const double a0 = Lx[n], a1 = Lx[n + 1], a2 = Lx[n + BZ], a3 = Lx[n + BZ + 1];
    const double b0 = Ly[n], b1 = Ly[n + 1], b2 = Ly[n + BY * BZ], b3 = Ly[n + BY * BZ + 1];
    const double c0 = Lz[n], c1 = Lz[n + BZ], c2 = Lz[n + BY * BZ], c3 = Lz[n + BY * BZ + BZ];
    const double Bx = a0 * (dy1 * dz1 + s1) + a1 * (dy1 * dz - s1) + a2 * (dy * dz1 - s1) +
                      a3 * (dy * dz + s1);
    const double By = b0 * (dx1 * dz1 + s2) + b1 * (dx1 * dz - s2) + b2 * (dx * dz1 - s2) +
                      b3 * (dx * dz + s2);
    const double Bz = c0 * (dx1 * dy1 + s3) + c1 * (dx1 * dy - s3) + c2 * (dx * dy1 - s3) +
                      c3 * (dx * dy + s3);
    const double Bx_dy = -a0 * dz1 - a1 * dz + a2 * dz1 + a3 * dz;
    const double Bx_dz = -a0 * dy1 + a1 * dy1 - a2 * dy + a3 * dy;
    const double Bx_s = a0 - a1 - a2 + a3;
    const double By_dx = -b0 * dz1 - b1 * dz + b2 * dz1 + b3 * dz;
    const double By_dz = -b0 * dx1 + b1 * dx1 - b2 * dx + b3 * dx;
    const double By_s = b0 - b1 - b2 + b3;
    const double Bz_dx = -c0 * dy1 - c1 * dy + c2 * dy1 + c3 * dy;
    const double Bz_dy = -c0 * dx1 + c1 * dx1 - c2 * dx + c3 * dx;
    const double Bz_s = c0 - c1 - c2 + c3;
    /* d/d(midpoint offsets) and d/d(segment deltas) */
    const double gdx = SV * By_dx + SW * Bz_dx;
    const double gdy = SU * Bx_dy + SW * Bz_dy;
    const double gdz = SU * Bx_dz + SV * By_dz;
    const double gDx = kx * Bx + (SV * By_s * Dz + SW * Bz_s * Dy) * dc.r12;
    const double gDy = ky * By + (SU * Bx_s * Dz + SW * Bz_s * Dx) * dc.r12;
    const double gDz = kz * Bz + (SU * Bx_s * Dy + SV * By_s * Dx) * dc.r12;
    gA[0] = 0.5 * gdx - gDx;
    gA[1] = 0.5 * gdy - gDy;
    gA[2] = 0.5 * gdz - gDz;
    gB[0] = 0.5 * gdx + gDx;
    gB[1] = 0.5 * gdy + gDy;
    gB[2] = 0.5 * gdz + gDz;
}

/* adjoint of Interpolate3D: scatter lam * weight into the LDS adjoint box and
 * return d(lam * value)/d(sx, sy, sz), the weights of the near planes
 * (the far-plane weights are 1 - s) */
template <int BY, int BZ>
__device__ __forceinline__ void interp_adj(const double *F, double *A, int u, int v, int w,
                                           double sx, double sy, double sz, double lam,
                                           double &dsx, double &dsy, double &dsz) {
    const double sx1 = 1. - sx, sy1 = 1. - sy, sz1 = 1. - sz;
    const int o = (u * BY + v) * BZ + w;
    const double f000 = F[o], f001 = F[o + 1], f010 = F[o + BZ], f011 = F[o + BZ + 1];
    const double f100 = F[o + BY * BZ], f101 = F[o + BY * BZ + 1];
    const double f110 = F[o + BY * BZ + BZ], f111 = F[o + BY * BZ + BZ + 1];
    const double yz00 = sy * sz, yz01 = sy * sz1, yz10 = sy1 * sz, yz11 = sy1 * sz1;
    const double p0 = yz00 * f000 + yz01 * f001 + yz10 * f010 + yz11 * f011;
    const double p1 = yz00 * f100 + yz01 * f101 + yz10 * f110 + yz11 * f111;
    dsx += lam * (p0 - p1);
    const double ly0 = sx * (sz * f000 + sz1 * f001) + sx1 * (sz * f100 + sz1 * f101);
    const double ly1 = sx * (sz * f010 + sz1 * f011) + sx1 * (sz * f110 + sz1 * f111);
    dsy += lam * (ly0 - ly1);
    const double lz0 = sx * (sy * f000 + sy1 * f010) + sx1 * (sy * f100 + sy1 * f110);
    const double lz1 = sx * (sy * f001 + sy1 * f011) + sx1 * (sy * f101 + sy1 * f111);
    dsz += lam * (lz0 - lz1);
    const double lx0 = lam * sx, lx1 = lam * sx1;
    atomicAdd(&A[o], lx0 * yz00);
    atomicAdd(&A[o + 1], lx0 * yz01);
    atomicAdd(&A[o + BZ], lx0 * yz10);
    atomicAdd(&A[o + BZ + 1], lx0 * yz11);
    atomicAdd(&A[o + BY * BZ], lx1 * yz00);
    atomicAdd(&A[o + BY * BZ + 1], lx1 * yz01);
    atomicAdd(&A[o + BY * BZ + BZ], lx1 * yz10);
    atomicAdd(&A[o + BY * BZ + BZ + 1], lx1 * yz11);
}

__device__ __forceinline__ void cross(const double a[3], const double b[3], double r[3]) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ double dot3(const double a[3], const double b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

template <int TX, int TY, int TZ>
__global__ __launch_bounds__(PIC_BLOCK) void k_adj_particles(AdjArgs a) {
    constexpr int C = TX * TY * TZ;
    constexpr int BX = TX + 2, BY = TY + 2, BZ = TZ + 2, NB = BX * BY * BZ;
    __shared__ double sF[6 * NB];    /* E, H' of the forward step */
    __shared__ double sL[3 * NB];    /* lamJ */
    __shared__ double sA[ACOPIES * 6 * NB];   /* lamE, lamH' accumulated by the gather adjoint */
    double *const myA = sA + (threadIdx.x % ACOPIES) * (6 * NB);
    __shared__ int sOff[C + 1];
    __shared__ int sCnt[C];
    __shared__ unsigned sFlags;

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

    /* thread n takes box node n of every box (one decode, one 32-bit
     * offset: the grid has fewer than 2^31 nodes) and cell n of the tile;
     * every global load is issued before the first LDS write */
    static_assert(NB <= PIC_BLOCK && C <= PIC_BLOCK, "one box node and one cell per thread");
    const double *const Fsrc[6] = {a.fwd.E[0], a.fwd.E[1], a.fwd.E[2],
                                   a.fwd.H[0], a.fwd.H[1], a.fwd.H[2]};
    const int nu = tid / (BY * BZ), nv = (tid / BZ) % BY, nw = tid % BZ;
    const bool node_in = tid < NB && I0 + nu <= g.Nx + 1 && L0 + nv <= g.Ny + 1 &&
                         K0 + nw <= g.Nz + 1;
    const unsigned noff = (unsigned)(I0 + nu) * g.sx + (unsigned)(L0 + nv) * g.sy + (K0 + nw);
    {
        double fv[6], lj[3];
#pragma unroll
        for (int q = 0; q < 6; q++) fv[q] = 0.0;
#pragma unroll
        for (int q = 0; q < 3; q++) lj[q] = 0.0;
        if (node_in) {
#pragma unroll
            for (int q = 0; q < 6; q++) fv[q] = Fsrc[q][noff];
#pragma unroll
            for (int q = 0; q < 3; q++) lj[q] = a.lam.J[q][noff];
        }
        const int gi = I0 + tid / (TY * TZ), gl = L0 + (tid / TZ) % TY, gk = K0 + tid % TZ;
        int cnt = 0;
        if (tid < C && gi < g.Nx && gl < g.Ny && gk < g.Nz) {
            cnt = a.p.count[((long)gi * g.Ny + gl) * g.Nz + gk];
            cnt = cnt < cap ? cnt : cap;
        }
        if (tid < NB) {
#pragma unroll
            for (int q = 0; q < 6; q++) sF[q * NB + tid] = fv[q];
#pragma unroll
            for (int q = 0; q < 3; q++) sL[q * NB + tid] = lj[q];
        }
        if (tid < C) sCnt[tid] = cnt;
    }
    for (int e = tid; e < ACOPIES * 6 * NB; e += PIC_BLOCK) sA[e] = 0.0;
    if (tid == 0) sFlags = 0;
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

    const double hx = g.hx, hy = g.hy, hz = g.hz, tau = g.tau;
    unsigned bad = 0;
    for (int fidx = tid; fidx < T; fidx += PIC_BLOCK) {
        int lo = 0, hi = C;
        while (hi - lo > 1) {
            int mid = (lo + hi) >> 1;
            if (sOff[mid] <= fidx) lo = mid; else hi = mid;
        }
        const int c = lo;
        Geo G;
        G.ca = c / (TY * TZ);
        G.cb = (c / TZ) % TY;
        G.cc = c % TZ;
        const int gi = I0 + G.ca, gl = L0 + G.cb, gk = K0 + G.cc;
        const long pidx = a.snap_off[tile] + fidx;    /* compact record */
        G.hx = hx; G.hy = hy; G.hz = hz; G.tau = tau;
        G.x0 = (double)gi * hx;
        G.y0 = (double)gl * hy;
        G.z0 = (double)gk * hz;

        const double x = a.p.p[0][pidx], y = a.p.p[1][pidx], z = a.p.p[2][pidx];
        const double pu = a.p.p[3][pidx], pv = a.p.p[4][pidx], pw = a.p.p[5][pidx];
        const unsigned char s = a.p.sp[pidx];
        const int pid = a.p.id[pidx];
        const double mass = a.sp.m[s], q_m = a.sp.q_m[s];
        const double tau1 = q_m * tau * 0.5;

        /* ---- forward recompute, the push kernel's arithmetic ---- */
        const double frx = divc(x - G.x0, hx, a.dc.rhx);
        const double fry = divc(y - G.y0, hy, a.dc.rhy);
        const double frz = divc(z - G.z0, hz, a.dc.rhz);
        double Ef[3] = {0, 0, 0}, Hf[3] = {0, 0, 0};
        const bool live = !(x < 0 || y < 0 || z < 0);
        GatherIdx W{};
        if (live) {
            W = gather_index(x, y, z, frx, fry, frz, g, a.dc);
            if (!W.ok) {
                bad = 1;
            } else {
                const double s1 = W.s1, s2 = W.s2, s3 = W.s3, s4 = W.s4, s5 = W.s5, s6 = W.s6;
                const double s11 = 1. - s1, s21 = 1. - s2, s31 = 1. - s3, s41 = 1. - s4,
                             s51 = 1. - s5, s61 = 1. - s6;
                const int ua = G.ca - 2, va = G.cb - 2, wa = G.cc - 2;
                Ef[0] = interp<BY, BZ>(sF + 0 * NB, ua + W.ix, va + W.i1y, wa + W.i1z, s1, s11, s4, s41, s6, s61);
                Ef[1] = interp<BY, BZ>(sF + 1 * NB, ua + W.i1x, va + W.iy, wa + W.i1z, s2, s21, s3, s31, s6, s61);
                Ef[2] = interp<BY, BZ>(sF + 2 * NB, ua + W.i1x, va + W.i1y, wa + W.iz, s2, s21, s4, s41, s5, s51);
                Hf[0] = interp<BY, BZ>(sF + 3 * NB, ua + W.i1x, va + W.iy, wa + W.iz, s2, s21, s3, s31, s5, s51);
                Hf[1] = interp<BY, BZ>(sF + 4 * NB, ua + W.ix, va + W.i1y, wa + W.iz, s1, s11, s4, s41, s5, s51);
                Hf[2] = interp<BY, BZ>(sF + 5 * NB, ua + W.ix, va + W.iy, wa + W.i1z, s1, s11, s3, s31, s6, s61);
            }
        }
        const Boris B = boris_move(pu, pv, pw, Ef[0], Ef[1], Ef[2], Hf[0], Hf[1], Hf[2], tau1);
        const double x1 = x + tau * (B.g2 * B.p1[0]);
        const double y1 = y + tau * (B.g2 * B.p1[1]);
        const double z1 = z + tau * (B.g2 * B.p1[2]);
        const Relay R = relay_point(x, y, z, x1, y1, z1, frx, fry, frz, G, a.dc);
        const int m = R.m;

        /* ---- reverse ---- */
        double2 *const lq = (double2 *)(a.lp.v + (long)pid * PIC_LSTRIDE);
        const double2 q0 = lq[0], q1 = lq[1], q2 = lq[2];
        double lx1[3] = {q0.x, q0.y, q1.x};
        double lp1[3] = {q1.y, q2.x, q2.y};
        double lx[3];
        const double X0[3] = {x, y, z}, X1[3] = {x1, y1, z1};
        /* deposit: J = pqr(i1; x -> e1) + [m != 0] pqr(i2; x2 -> x1) */
        {
            double gA[3], gB[3];
            const double e1x = m ? R.x2x : x1, e1y = m ? R.x2y : y1, e1z = m ? R.x2z : z1;
            pqr_adj<BY, BZ, NB>(sL, G, a.dc, R.i1x, R.i1y, R.i1z, x, y, z, e1x, e1y, e1z, mass,
                                gA, gB);
            lx[0] = gA[0];
            lx[1] = gA[1];
            lx[2] = gA[2];
            if (m == 0) {
                lx1[0] += gB[0];
                lx1[1] += gB[1];
                lx1[2] += gB[2];
            } else {
                double hA[3], hB[3];
                pqr_adj<BY, BZ, NB>(sL, G, a.dc, R.i2x, R.i2y, R.i2z, R.x2x, R.x2y, R.x2z, x1, y1,
                                    z1, mass, hA, hB);
                lx1[0] += hB[0];
                lx1[1] += hB[1];
                lx1[2] += hB[2];
                /* relay point x2(x, x1): crossed axes sit on a face (constant),
                 * the others on the straight path at sr = num/den */
                const double lx2[3] = {gB[0] + hA[0], gB[1] + hA[1], gB[2] + hA[2]};
                const bool cr[3] = {(m & 4) != 0, (m & 2) != 0, (m & 1) != 0};
                const double F[3] = {R.fx, R.fy, R.fz};
                double lsr = 0.0;
                for (int d = 0; d < 3; d++) {
                    if (cr[d]) continue;
                    lsr += lx2[d] * (X1[d] - X0[d]);
                    lx[d] += lx2[d] * (1.0 - R.sr);
                    lx1[d] += lx2[d] * R.sr;
                }
                const int nc = (int)cr[0] + (int)cr[1] + (int)cr[2];
                if (nc == 1) {
                    const int d = cr[0] ? 0 : (cr[1] ? 1 : 2);
                    const double D = X1[d] - X0[d];
                    lx[d] += lsr * (R.sr - 1.0) / D;
                    lx1[d] += lsr * (-R.sr / D);
                } else if (nc == 2) {
                    for (int d = 0; d < 3; d++) {
                        if (!cr[d]) continue;
                        const double D = X1[d] - X0[d], f = F[d] - X0[d];
                        /* num = sum (f_d - x_d) D_d, den = sum D_d^2 */
                        const double dnum_x = -D - f, dnum_x1 = f;
                        const double dden_x = -2.0 * D, dden_x1 = 2.0 * D;
                        lx[d] += lsr * (dnum_x - R.sr * dden_x) / R.den;
                        lx1[d] += lsr * (dnum_x1 - R.sr * dden_x1) / R.den;
                    }
                }
            }
        }
        /* x1 = x + tau g2 p1 */
        for (int d = 0; d < 3; d++) lx[d] += lx1[d];
        {
            const double lu[3] = {tau * lx1[0], tau * lx1[1], tau * lx1[2]};
            const double g2 = B.g2, g23 = g2 * g2 * g2;
            const double pl = dot3(B.p1, lu);
            for (int d = 0; d < 3; d++) lp1[d] += g2 * lu[d] - g23 * pl * B.p1[d];
        }
        /* p1 = pp + tau1 E;  pp = (s + (b.s) b + s x b) / (1 + b.b) */
        double lE[3], lH[3], lpm[3];
        {
            for (int d = 0; d < 3; d++) lE[d] = tau1 * lp1[d];
            const double *b = B.b, *sv = B.s;
            const double rD = 1.0 / B.sden;
            double lN[3] = {lp1[0] * rD, lp1[1] * rD, lp1[2] * rD};
            /* pp recomputed from s and b for the denominator term */
            double sxb[3];
            cross(sv, b, sxb);
            const double bs = dot3(b, sv);
            double pp[3];
            for (int d = 0; d < 3; d++) pp[d] = (sv[d] + bs * b[d] + sxb[d]) * rD;
            const double bl = dot3(b, lN), lpp = dot3(lp1, pp);
            double bxl[3], lxs[3];
            cross(b, lN, bxl);
            cross(lN, sv, lxs);
            double ls[3], lb[3];
            for (int d = 0; d < 3; d++) {
                ls[d] = lN[d] + b[d] * bl + bxl[d];
                lb[d] = sv[d] * bl + bs * lN[d] + lxs[d] - 2.0 * lpp * rD * b[d];
            }
            /* s = pm + pm x b */
            double bxls[3], lsxpm[3];
            cross(b, ls, bxls);
            cross(ls, B.pm, lsxpm);
            for (int d = 0; d < 3; d++) {
                lpm[d] = ls[d] + bxls[d];
                lb[d] += lsxpm[d];
            }
            /* b = ps H, ps = tau1 g1, g1 = (pm.pm + 1)^-1/2 */
            const double lps = dot3(lb, Hf);
            for (int d = 0; d < 3; d++) lH[d] = B.ps * lb[d];
            if (tau1 != 0.0) {
                const double g1 = B.ps / tau1;
                const double k = -tau1 * g1 * g1 * g1 * lps;
                for (int d = 0; d < 3; d++) lpm[d] += k * B.pm[d];
            }
            /* pm = p + tau1 E */
            for (int d = 0; d < 3; d++) lE[d] += tau1 * lpm[d];
        }
        /* gather: scatter into the adjoint box, position derivative through
         * the weights (every s has ds/dx = -1/h) */
        if (live && W.ok) {
            const int ua = G.ca - 2, va = G.cb - 2, wa = G.cc - 2;
            double dsx = 0, dsy = 0, dsz = 0;
            interp_adj<BY, BZ>(sF + 0 * NB, myA + 0 * NB, ua + W.ix, va + W.i1y, wa + W.i1z, W.s1, W.s4, W.s6, lE[0], dsx, dsy, dsz);
            interp_adj<BY, BZ>(sF + 1 * NB, myA + 1 * NB, ua + W.i1x, va + W.iy, wa + W.i1z, W.s2, W.s3, W.s6, lE[1], dsx, dsy, dsz);
            interp_adj<BY, BZ>(sF + 2 * NB, myA + 2 * NB, ua + W.i1x, va + W.i1y, wa + W.iz, W.s2, W.s4, W.s5, lE[2], dsx, dsy, dsz);
            interp_adj<BY, BZ>(sF + 3 * NB, myA + 3 * NB, ua + W.i1x, va + W.iy, wa + W.iz, W.s2, W.s3, W.s5, lH[0], dsx, dsy, dsz);
            interp_adj<BY, BZ>(sF + 4 * NB, myA + 4 * NB, ua + W.ix, va + W.i1y, wa + W.iz, W.s1, W.s4, W.s5, lH[1], dsx, dsy, dsz);
            interp_adj<BY, BZ>(sF + 5 * NB, myA + 5 * NB, ua + W.ix, va + W.iy, wa + W.i1z, W.s1, W.s3, W.s6, lH[2], dsx, dsy, dsz);
            lx[0] -= dsx * a.dc.rhx;
            lx[1] -= dsy * a.dc.rhy;
            lx[2] -= dsz * a.dc.rhz;
        }
        lq[0] = make_double2(lx[0], lx[1]);
        lq[1] = make_double2(lx[2], lpm[0]);
        lq[2] = make_double2(lpm[1], lpm[2]);
    }
    if (bad) atomicOr(&sFlags, (unsigned)PIC_F_RANGE);
    __syncthreads();
    if (tid < NB) {   /* thread n adds box node n of the six adjoint boxes */
#pragma unroll
        for (int q = 0; q < 6; q++) {
            const int e = q * NB + tid;
            double val = sA[e];
            for (int r = 1; r < ACOPIES; r++) val += sA[r * 6 * NB + e];
            if (val != 0.0 && node_in)
                atomicAdd(q < 3 ? &a.lam.E[q][noff] : &a.lam.H[q - 3][noff], val);
        }
    }
    if (tid == 0 && sFlags) atomicOr(&a.st->flags, sFlags);
}

/* bitwise comparison of a recomputed array with its checkpoint */
template <typename T>
__global__ __launch_bounds__(PIC_BLOCK) void k_mismatch(const T *a, const T *b, long n,
                                                        DevStatus *st) {
    const long e = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (e < n && a[e] != b[e]) atomicOr(&st->flags, (unsigned)PIC_F_RECOMPUTE);
}

}  // namespace

void launch_mismatch(const void *a, const void *b, long n, int elem_bytes, DevStatus *st,
                     hipStream_t s) {
    if (elem_bytes == 8)
        hipLaunchKernelGGL(k_mismatch<unsigned long long>, dim3(blocks_for(n)), dim3(PIC_BLOCK),
                           0, s, (const unsigned long long *)a, (const unsigned long long *)b, n,
                           st);
    else
        hipLaunchKernelGGL(k_mismatch<unsigned int>, dim3(blocks_for(n)), dim3(PIC_BLOCK), 0, s,
                           (const unsigned int *)a, (const unsigned int *)b, n, st);
}

void launch_adj_seed_energy(const Grid &g, const double *const E[3], double w, AdjLam l,
                            hipStream_t s) {
    hipLaunchKernelGGL(k_adj_seed_energy, dim3(blocks_for(g.G)), dim3(PIC_BLOCK), 0, s, g, E[0],
                       E[1], E[2], 2.0 * w, l);
}

void launch_adj_e_periodic(const Grid &g, AdjLam l, hipStream_t s) {
    if (g.ghost_mode == 1)
        hipLaunchKernelGGL(k_adj_e_periodic_y, dim3(blocks_for((long)(g.Nx + 2) * g.sy)),
                           dim3(PIC_BLOCK), 0, s, g, l);
    hipLaunchKernelGGL(k_adj_e_periodic_z, dim3(blocks_for((long)(g.Nx + 2) * (g.Ny + 2))),
                       dim3(PIC_BLOCK), 0, s, g, l);
    hipLaunchKernelGGL(k_adj_e_periodic_x, dim3(blocks_for((long)(g.Ny + 2) * g.sy)),
                       dim3(PIC_BLOCK), 0, s, g, l);
}

void launch_adj_curl_e(const Grid &g, AdjLam l, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_curl_e, dim3(blocks_for(g.G)), dim3(PIC_BLOCK), 0, s, g, l);
}

void launch_adj_fold_j(const Grid &g, AdjLam l, hipStream_t s) {
    const long faces[3] = {(long)(g.Ny + 2) * (g.Nz + 2), (long)(g.Nx + 2) * (g.Nz + 2),
                           (long)(g.Nx + 2) * (g.Ny + 2)};
    for (int dir = 2; dir >= 0; dir--)
        hipLaunchKernelGGL(k_adj_fold_j, dim3(blocks_for(faces[dir])), dim3(PIC_BLOCK), 0, s, g,
                           l, dir);
}

void launch_adj_particles_f64(const Grid &g, const TileGeom &t, AdjFwd fwd, PartPtrs snap,
                              const long *snap_off, SpeciesTable sp, AdjLam l, LamParts lp,
                              DevStatus *st, hipStream_t s) {
    AdjArgs a;
    a.g = g;
    a.dc = make_divconst(g);
    a.t = t;
    a.fwd = fwd;
    a.p = snap;
    a.snap_off = snap_off;
    a.sp = sp;
    a.lam = l;
    a.lp = lp;
    a.st = st;
    hipLaunchKernelGGL((k_adj_particles<TXC, TYC, TZC>), dim3((unsigned)t.ntiles), dim3(PIC_BLOCK),
                       0, s, a);
}

void launch_adj_curl_h(const Grid &g, AdjLam l, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_curl_h, dim3(blocks_for(g.G)), dim3(PIC_BLOCK), 0, s, g, l);
}
