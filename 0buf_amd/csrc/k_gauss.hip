/*
 * k_gauss.hip -- Gauss-law residual div E - rho in FP64 (both forms).
 *
 * The charge that the current deposit conserves lives on the cell-centre
 * lattice (centre g at (g - 1/2) h, the reference's getNodeX numbering,
 * cell.h:680-692) with cloud-in-cell weights; E shares the J staggering, so
 * div E at centre g is sum_d (E_d[g] - E_d[g - e_d]) / h_d.  The Yee update
 * changes div E by -tau div J and the deposit changes rho by the same amount
 * (discrete continuity), so the residual is constant in time up to round-off.
 * rho is accumulated in int64 fixed point (2^-40): exact and order-free.
 */
#include "pic_internal.h"

namespace {

__device__ __forceinline__ int wrap1(int g, int N) {   /* centre index into 1..N */
    int b = (g - 1) % N;
    if (b < 0) b += N;
    return b + 1;
}

__device__ __forceinline__ long long fixd(double v) {
    return (long long)llrint(v * PIC_J_SCALE);
}

__device__ __forceinline__ void cic(const Grid &g, long long *rho, double ux, double uy,
                                    double uz, double a) {
    const double fx = floor(ux + 0.5), fy = floor(uy + 0.5), fz = floor(uz + 0.5);
    const double wx1 = ux + 0.5 - fx, wy1 = uy + 0.5 - fy, wz1 = uz + 0.5 - fz;
    const double wx[2] = {1.0 - wx1, wx1}, wy[2] = {1.0 - wy1, wy1}, wz[2] = {1.0 - wz1, wz1};
    const int gx = (int)fx, gy = (int)fy, gz = (int)fz;
    for (int a0 = 0; a0 < 2; a0++)
        for (int b0 = 0; b0 < 2; b0++)
            for (int c0 = 0; c0 < 2; c0++) {
                const long n = (long)wrap1(gx + a0, g.Nx) * g.sx +
                               (long)wrap1(gy + b0, g.Ny) * g.sy + wrap1(gz + c0, g.Nz);
                atomicAdd((unsigned long long *)&rho[n],
                          (unsigned long long)fixd(a * wx[a0] * wy[b0] * wz[c0]));
            }
}

__global__ __launch_bounds__(PIC_BLOCK) void k_rho64(Grid g, long ncell, PartPtrs p,
                                                    SpeciesTable sp, long long *rho) {
    const long e = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (e >= ncell * p.cap) return;
    const long c = e / p.cap;
    if ((int)(e % p.cap) >= p.count[c]) return;
    cic(g, rho, p.p[0][e] / g.hx, p.p[1][e] / g.hy, p.p[2][e] / g.hz, sp.m[p.sp[e]]);
}

__global__ __launch_bounds__(PIC_BLOCK) void k_rho32(Grid g, long ncell, PartPtrs32 p,
                                                    SpeciesTable sp, long long *rho) {
    const long e = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (e >= ncell * p.cap) return;
    const long c = e / p.cap;
    if ((int)(e % p.cap) >= p.count[c]) return;
    const int i = (int)(c / ((long)g.Ny * g.Nz)), l = (int)((c / g.Nz) % g.Ny),
              k = (int)(c % g.Nz);
    cic(g, rho, (double)i + (double)p.p[0][e], (double)l + (double)p.p[1][e],
        (double)k + (double)p.p[2][e], sp.m[p.sp[e]]);
}

template <typename T>
__global__ __launch_bounds__(PIC_BLOCK) void k_gauss(Grid g, const T *Ex, const T *Ey,
                                                    const T *Ez, const long long *rho,
                                                    double *res) {
    const long b = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (b >= (long)g.Nx * g.Ny * g.Nz) return;
    const int i = (int)(b / ((long)g.Ny * g.Nz)) + 1, l = (int)((b / g.Nz) % g.Ny) + 1,
              k = (int)(b % g.Nz) + 1;
    const long n = (long)i * g.sx + (long)l * g.sy + k;
    const double div = ((double)Ex[n] - (double)Ex[n - g.sx]) / g.hx +
                       ((double)Ey[n] - (double)Ey[n - g.sy]) / g.hy +
                       ((double)Ez[n] - (double)Ez[n - 1]) / g.hz;
    res[b] = div - (double)rho[n] * PIC_J_INV;
}

inline unsigned blocks_for(long n) { return (unsigned)((n + PIC_BLOCK - 1) / PIC_BLOCK); }

}  // namespace

void launch_rho64(const Grid &g, long ncell, PartPtrs p, SpeciesTable sp, long long *rho,
                  hipStream_t s) {
    hipLaunchKernelGGL(k_rho64, dim3(blocks_for(ncell * p.cap)), dim3(PIC_BLOCK), 0, s, g, ncell,
                       p, sp, rho);
}
void launch_rho32(const Grid &g, long ncell, PartPtrs32 p, SpeciesTable sp, long long *rho,
                  hipStream_t s) {
    hipLaunchKernelGGL(k_rho32, dim3(blocks_for(ncell * p.cap)), dim3(PIC_BLOCK), 0, s, g, ncell,
                       p, sp, rho);
}
void launch_gauss64(const Grid &g, FieldPtrs f, const long long *rho, double *res,
                    hipStream_t s) {
    hipLaunchKernelGGL(k_gauss<double>, dim3(blocks_for((long)g.Nx * g.Ny * g.Nz)),
                       dim3(PIC_BLOCK), 0, s, g, (const double *)f.f[PIC_EX],
                       (const double *)f.f[PIC_EY], (const double *)f.f[PIC_EZ], rho, res);
}
void launch_gauss32(const Grid &g, FieldPtrs32 f, const long long *rho, double *res,
                    hipStream_t s) {
    hipLaunchKernelGGL(k_gauss<float>, dim3(blocks_for((long)g.Nx * g.Ny * g.Nz)),
                       dim3(PIC_BLOCK), 0, s, g, (const float *)f.E[0], (const float *)f.E[1],
                       (const float *)f.E[2], rho, res);
}

/* ------------------------------------------------------------------------
 * Per-cell particle moments (the reference's diagnostics_operations
 * DENSITY, KINETIC_ENERGY, VELOCITY, CURRENT of particle_types.h:15-16,
 * declared there and never implemented): for one sort, per bulk cell,
 *   DENSITY         sum of the charge weights m (= q n of the cell)
 *   KINETIC_ENERGY  sum of (m / q_m)(gamma - 1), gamma - 1 = p.p / (gamma + 1)
 *   VELOCITY        mean of u = p / gamma (3 components)
 *   CURRENT         sum of m u (3 components)
 * accumulated in FP64 over the cell's bucket, one lane per cell.
 */
namespace {

template <class P>
__global__ __launch_bounds__(PIC_BLOCK) void k_cell_moments(long ncell, P p, SpeciesTable sp,
                                                            const int *slot_sort, int sort,
                                                            int op, double *out) {
    const long c = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (c >= ncell) return;
    int n = p.count[c];
    if (n > p.cap) n = p.cap;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    long cnt = 0;
    for (int j = 0; j < n; j++) {
        const long e = c * p.cap + j;
        const int s = p.sp[e];
        if (slot_sort[s] != sort) continue;
        const double px = p.p[3][e], py = p.p[4][e], pz = p.p[5][e];
        const double gm = sqrt(1.0 + px * px + py * py + pz * pz);
        const double m = sp.m[s];
        cnt++;
        switch (op) {
        case 0: a0 += m; break;
        /* gamma - 1 as p.p / (gamma + 1): no cancellation for slow particles */
        case 1: a0 += (m / sp.q_m[s]) * ((px * px + py * py + pz * pz) / (gm + 1.0)); break;
        case 2: a0 += px / gm; a1 += py / gm; a2 += pz / gm; break;
        default: a0 += m * (px / gm); a1 += m * (py / gm); a2 += m * (pz / gm); break;
        }
    }
    if (op <= 1) {
        out[c] = a0;
    } else {
        const double d = (op == 2 && cnt > 0) ? (double)cnt : 1.0;
        out[c] = a0 / d;
        out[ncell + c] = a1 / d;
        out[2 * ncell + c] = a2 / d;
    }
}

}  // namespace

void launch_cell_moments64(long ncell, PartPtrs p, SpeciesTable sp, const int *slot_sort, int sort,
                           int op, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_cell_moments<PartPtrs>, dim3((unsigned)((ncell + PIC_BLOCK - 1) / PIC_BLOCK)),
                       dim3(PIC_BLOCK), 0, s, ncell, p, sp, slot_sort, sort, op, out);
}

void launch_cell_moments32(long ncell, PartPtrs32 p, SpeciesTable sp, const int *slot_sort,
                           int sort, int op, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_cell_moments<PartPtrs32>, dim3((unsigned)((ncell + PIC_BLOCK - 1) / PIC_BLOCK)),
                       dim3(PIC_BLOCK), 0, s, ncell, p, sp, slot_sort, sort, op, out);
}
