/*
 * k_fields.hip -- Yee field stages of the FP64 plain form as coalesced 3-D
 * stencils over the ghosted arrays (k fastest), one node per lane.
 *
 * Built with -ffp-contract=off: each update keeps the reference's rounding
 * sequence, so these kernels are bit-identical to the CPU restatement.
 * HBM roofline: every kernel is a streaming stencil; the neighbour reads at
 * +-1 in k, +-(nz+2) and +-(ny+2)(nz+2) hit L2.
 */
#include "pic_internal.h"

#pragma clang fp contract(off)

namespace {

__device__ __forceinline__ void decode(const Grid &g, long n, int &i, int &l, int &k) {
    k = (int)(n % g.sy);
    long t = n / g.sy;
    l = (int)(t % (g.Ny + 2));
    i = (int)(t / (g.Ny + 2));
}

/* MagneticStageOne (gpu_plasma.h:524-538) -> GPU_emh1 / emh1_Element
 * (kernels.cu:1052-1089) for the three components, ranges from
 * MagneticFieldTrace (gpu_plasma.h:735-767), shifts from
 * getMagneticFieldTraceShifts (722-733); fused with the J := 0 of
 * SetCurrentsToZero (gpu_plasma.h:2058-2069). */
__global__ __launch_bounds__(PIC_BLOCK) void k_h_half(Grid g, FieldPtrs f) {
    long n = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (n >= g.G) return;
    int i, l, k;
    decode(g, n, i, l, k);
    const double c1x = g.tau / g.hx, c1y = g.tau / g.hy, c1z = g.tau / g.hz;
    const double *Ex = f.f[PIC_EX], *Ey = f.f[PIC_EY], *Ez = f.f[PIC_EZ];
    if (l <= g.Ny && k <= g.Nz) {   /* Qx over [0,Nx+1]x[0,Ny]x[0,Nz] */
        double e1_n1 = Ey[n + 1], e1_n = Ey[n], e2_n2 = Ez[n + g.sy], e2_n = Ez[n];
        double t = 0.5 * (c1z * (e1_n1 - e1_n) - c1y * (e2_n2 - e2_n));
        f.f[PIC_QX][n] = t;
        f.f[PIC_HX][n] += t;
    }
    if (i <= g.Nx && k <= g.Nz) {   /* Qy over [0,Nx]x[0,Ny+1]x[0,Nz] */
        double e1_n1 = Ez[n + g.sx], e1_n = Ez[n], e2_n2 = Ex[n + 1], e2_n = Ex[n];
        double t = 0.5 * (c1x * (e1_n1 - e1_n) - c1z * (e2_n2 - e2_n));
        f.f[PIC_QY][n] = t;
        f.f[PIC_HY][n] += t;
    }
    if (i <= g.Nx && l <= g.Ny) {   /* Qz over [0,Nx]x[0,Ny]x[0,Nz+1] */
        double e1_n1 = Ex[n + g.sy], e1_n = Ex[n], e2_n2 = Ey[n + g.sx], e2_n = Ey[n];
        double t = 0.5 * (c1y * (e1_n1 - e1_n) - c1x * (e2_n2 - e2_n));
        f.f[PIC_QZ][n] = t;
        f.f[PIC_HZ][n] += t;
    }
    f.f[PIC_JX][n] = 0.0;
    f.f[PIC_JY][n] = 0.0;
    f.f[PIC_JZ][n] = 0.0;
}

/* MagneticFieldStageTwo (gpu_plasma.h:540-551) -> GPU_emh2 (kernels.cu:1024-1049) */
__global__ __launch_bounds__(PIC_BLOCK) void k_h2(Grid g, FieldPtrs f) {
    long n = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (n >= g.G) return;
    int i, l, k;
    decode(g, n, i, l, k);
    if (l <= g.Ny && k <= g.Nz) f.f[PIC_HX][n] += f.f[PIC_QX][n];
    if (i <= g.Nx && k <= g.Nz) f.f[PIC_HY][n] += f.f[PIC_QY][n];
    if (i <= g.Nx && l <= g.Ny) f.f[PIC_HZ][n] += f.f[PIC_QZ][n];
}

/* ElectricFieldEvaluate (gpu_plasma.h:455-481) -> GPU_eme / emeElement
 * (kernels.cu:1091-1181); starts and shifts from GetElectricFieldStartsDirs
 * (gpu_plasma.h:286-304), finish (Nx,Ny,Nz). */
__global__ __launch_bounds__(PIC_BLOCK) void k_e(Grid g, FieldPtrs f) {
    long n = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (n >= g.G) return;
    int i, l, k;
    decode(g, n, i, l, k);
    if (i > g.Nx || l > g.Ny || k > g.Nz) return;
    const double c1x = g.tau / g.hx, c1y = g.tau / g.hy, c1z = g.tau / g.hz, tau = g.tau;
    const double *Hx = f.f[PIC_HX], *Hy = f.f[PIC_HY], *Hz = f.f[PIC_HZ];
    if (l >= 1 && k >= 1) {   /* Ex: [0,Nx]x[1,Ny]x[1,Nz], H1=Hz(l-1), H2=Hy(k-1) */
        double *E = f.f[PIC_EX];
        E[n] += c1y * (Hz[n] - Hz[n - g.sy]) - c1z * (Hy[n] - Hy[n - 1]) - tau * f.f[PIC_JX][n];
    }
    if (i >= 1 && k >= 1) {   /* Ey: [1,Nx]x[0,Ny]x[1,Nz], H1=Hx(k-1), H2=Hz(i-1) */
        double *E = f.f[PIC_EY];
        E[n] += c1z * (Hx[n] - Hx[n - 1]) - c1x * (Hz[n] - Hz[n - g.sx]) - tau * f.f[PIC_JY][n];
    }
    if (i >= 1 && l >= 1) {   /* Ez: [1,Nx]x[1,Ny]x[0,Nz], H1=Hy(i-1), H2=Hx(l-1) */
        double *E = f.f[PIC_EZ];
        E[n] += c1x * (Hy[n] - Hy[n - g.sx]) - c1y * (Hx[n] - Hx[n - g.sy]) - tau * f.f[PIC_JZ][n];
    }
}

/* PeriodicBoundaries along x (gpu_plasma.h:802-870 with the arguments at
 * 486-505): Ey over l in [0,Ny], k in [1,Nz]; Ez over l in [1,Ny], k in
 * [0,Nz]; E[0]=E[Nx] and E[Nx+1]=E[1] (periodicElement, kernels.cu:1104). */
__global__ __launch_bounds__(PIC_BLOCK) void k_e_periodic_x(Grid g, FieldPtrs f) {
    long t = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (t >= (long)(g.Ny + 2) * g.sy) return;
    int l = (int)(t / g.sy), k = (int)(t % g.sy);
    long base = (long)l * g.sy + k;
    long n0 = base, n1 = (long)g.sx + base, nN = (long)g.Nx * g.sx + base,
         nN1 = (long)(g.Nx + 1) * g.sx + base;
    if (l <= g.Ny && k >= 1 && k <= g.Nz) {
        double *E = f.f[PIC_EY];
        E[n0] = E[nN];
        E[nN1] = E[n1];
    }
    if (l >= 1 && l <= g.Ny && k <= g.Nz) {
        double *E = f.f[PIC_EZ];
        E[n0] = E[nN];
        E[nN1] = E[n1];
    }
}

/* PeriodicBoundaries along z: Ex over i in [0,Nx], l in [0,Ny+1]; Ey over
 * i in [0,Nx+1], l in [0,Ny] (after the x copies). */
__global__ __launch_bounds__(PIC_BLOCK) void k_e_periodic_z(Grid g, FieldPtrs f) {
    long t = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (t >= (long)(g.Nx + 2) * (g.Ny + 2)) return;
    int i = (int)(t / (g.Ny + 2)), l = (int)(t % (g.Ny + 2));
    long base = (long)i * g.sx + (long)l * g.sy;
    long n0 = base, n1 = base + 1, nN = base + g.Nz, nN1 = base + g.Nz + 1;
    if (i <= g.Nx) {
        double *E = f.f[PIC_EX];
        E[n0] = E[nN];
        E[nN1] = E[n1];
    }
    if (l <= g.Ny) {
        double *E = f.f[PIC_EY];
        E[n0] = E[nN];
        E[nN1] = E[n1];
    }
}

/* SetPeriodicCurrentComponent (gpu_plasma.h:926-978) -> GPU_CurrentPeriodic /
 * periodicCurrentElement (kernels.cu:1124-1159), all three components for one
 * face direction `dir`; the three directions run as three launches in the
 * reference's order. */
__global__ __launch_bounds__(PIC_BLOCK) void k_fold_j(Grid g, FieldPtrs f, int dir) {
    long t = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    int na, nb, N;
    if (dir == 0) { na = g.Ny + 2; nb = g.Nz + 2; N = g.Nx + 2; }
    else if (dir == 1) { na = g.Nx + 2; nb = g.Nz + 2; N = g.Ny + 2; }
    else { na = g.Nx + 2; nb = g.Ny + 2; N = g.Nz + 2; }
    if (t >= (long)na * nb) return;
    int a = (int)(t / nb), b = (int)(t % nb);
    long base, stride;
    if (dir == 0) { base = (long)a * g.sy + b; stride = g.sx; }
    else if (dir == 1) { base = (long)a * g.sx + b; stride = g.sy; }
    else { base = (long)a * g.sx + (long)b * g.sy; stride = 1; }
    long n1 = base + stride, nNm1 = base + (long)(N - 1) * stride;
    long nNm2 = base + (long)(N - 2) * stride, n0 = base;
    for (int dirE = 0; dirE < 3; dirE++) {
        double *E = f.f[PIC_JX + dirE];
        if (dir != dirE) E[n1] += E[nNm1];
        if (dir != 1 || dirE != 1) E[nNm1] = E[n1];
        E[n0] += E[nNm2];
        E[nNm2] = E[n0];
    }
}

/* getElectricEnergy (gpu_plasma.h:621-657) -> GPU_getCellEnergy
 * (wrap_kernel.cu:14-36): sum over every ghosted node of Ex^2+Ey^2+Ez^2.
 * Deterministic two-level tree instead of one-address atomics. */
__global__ __launch_bounds__(PIC_BLOCK) void k_energy(Grid g, FieldPtrs f, double *partials) {
    __shared__ double red[PIC_BLOCK];
    const double *ex = f.f[PIC_EX], *ey = f.f[PIC_EY], *ez = f.f[PIC_EZ];
    double s = 0.0;
    for (long n = (long)blockIdx.x * PIC_BLOCK + threadIdx.x; n < g.G;
         n += (long)gridDim.x * PIC_BLOCK) {
        double a = ex[n], b = ey[n], c = ez[n];
        s += a * a + b * b + c * c;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = PIC_BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(PIC_BLOCK) void k_reduce_final(const double *partials, int n,
                                                            double *out) {
    __shared__ double red[PIC_BLOCK];
    double s = 0.0;
    for (int j = threadIdx.x; j < n; j += PIC_BLOCK) s += partials[j];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = PIC_BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

__global__ void k_reset_status(DevStatus *st) {
    st->flags = 0;
    st->max_dep = 0;
    st->departures = 0;
    st->dropped = 0;
    st->duplicated = 0;
}

inline unsigned blocks_for(long n) { return (unsigned)((n + PIC_BLOCK - 1) / PIC_BLOCK); }

}  // namespace

void launch_h_half(const Grid &g, FieldPtrs f, hipStream_t s) {
    hipLaunchKernelGGL(k_h_half, dim3(blocks_for(g.G)), dim3(PIC_BLOCK), 0, s, g, f);
}
void launch_h2(const Grid &g, FieldPtrs f, hipStream_t s) {
    hipLaunchKernelGGL(k_h2, dim3(blocks_for(g.G)), dim3(PIC_BLOCK), 0, s, g, f);
}
void launch_e(const Grid &g, FieldPtrs f, hipStream_t s) {
    hipLaunchKernelGGL(k_e, dim3(blocks_for(g.G)), dim3(PIC_BLOCK), 0, s, g, f);
}
void launch_e_periodic(const Grid &g, FieldPtrs f, hipStream_t s) {
    hipLaunchKernelGGL(k_e_periodic_x, dim3(blocks_for((long)(g.Ny + 2) * g.sy)), dim3(PIC_BLOCK),
                       0, s, g, f);
    hipLaunchKernelGGL(k_e_periodic_z, dim3(blocks_for((long)(g.Nx + 2) * (g.Ny + 2))),
                       dim3(PIC_BLOCK), 0, s, g, f);
}
void launch_fold_j(const Grid &g, FieldPtrs f, hipStream_t s) {
    long faces[3] = {(long)(g.Ny + 2) * (g.Nz + 2), (long)(g.Nx + 2) * (g.Nz + 2),
                     (long)(g.Nx + 2) * (g.Ny + 2)};
    for (int dir = 0; dir < 3; dir++)
        hipLaunchKernelGGL(k_fold_j, dim3(blocks_for(faces[dir])), dim3(PIC_BLOCK), 0, s, g, f,
                           dir);
}
void launch_energy(const Grid &g, FieldPtrs f, double *partials, int nblocks, double *out,
                   hipStream_t s) {
    hipLaunchKernelGGL(k_energy, dim3(nblocks), dim3(PIC_BLOCK), 0, s, g, f, partials);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(PIC_BLOCK), 0, s, partials, nblocks, out);
}
void launch_reset_status(DevStatus *st, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(1), 0, s, st);
}
