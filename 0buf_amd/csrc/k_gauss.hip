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
