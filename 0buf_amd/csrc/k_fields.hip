/*
 * k_fields.hip -- Yee field stages as coalesced 3-D stencils over the
 * ghosted arrays (k fastest), one node per lane.
 *
 * Each kernel is a template over the field type T and the current type JT:
 *   FP64 plain form   T = double, JT = double  (built with -ffp-contract=off:
 *                     every update keeps the reference's rounding sequence,
 *                     bit-identical to the CPU restatement);
 *   FP32 form         T = float,  JT = int64 fixed point (2^-40), so that the
 *                     fold and the E update read an exact, order-free J.
 * HBM roofline: streaming stencils; neighbour reads at +-1 in k, +-(nz+2) and
 * +-(ny+2)(nz+2) hit L2.
 */
#include "pic_internal.h"

#pragma clang fp contract(off)

namespace {

template <typename T, typename JT>
struct FV {
    T *E[3], *H[3], *Q[3];
    JT *J[3];
};

inline FV<double, double> view(FieldPtrs f) {
    FV<double, double> v;
    for (int c = 0; c < 3; c++) {
        v.E[c] = f.f[PIC_EX + c];
        v.H[c] = f.f[PIC_HX + c];
        v.Q[c] = f.f[PIC_QX + c];
        v.J[c] = f.f[PIC_JX + c];
    }
    return v;
}

inline FV<float, long long> view(FieldPtrs32 f) {
    FV<float, long long> v;
    for (int c = 0; c < 3; c++) {
        v.E[c] = f.E[c];
        v.H[c] = f.H[c];
        v.Q[c] = f.Q[c];
        v.J[c] = f.J[c];
    }
    return v;
}

__device__ __forceinline__ double jval(double j) { return j; }
__device__ __forceinline__ float jval(long long j) { return (float)((double)j * PIC_J_INV); }

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
template <typename T, typename JT>
__global__ __launch_bounds__(PIC_BLOCK) void k_h_half(Grid g, FV<T, JT> f) {
    long n = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (n >= g.G) return;
    int i, l, k;
    decode(g, n, i, l, k);
    const T c1x = (T)(g.tau / g.hx), c1y = (T)(g.tau / g.hy), c1z = (T)(g.tau / g.hz);
    const T *Ex = f.E[0], *Ey = f.E[1], *Ez = f.E[2];
    if (l <= g.Ny && k <= g.Nz) {   /* Qx over [0,Nx+1]x[0,Ny]x[0,Nz] */
        T e1_n1 = Ey[n + 1], e1_n = Ey[n], e2_n2 = Ez[n + g.sy], e2_n = Ez[n];
        T t = (T)0.5 * (c1z * (e1_n1 - e1_n) - c1y * (e2_n2 - e2_n));
        f.Q[0][n] = t;
        f.H[0][n] += t;
    }
    if (i <= g.Nx && k <= g.Nz) {   /* Qy over [0,Nx]x[0,Ny+1]x[0,Nz] */
        T e1_n1 = Ez[n + g.sx], e1_n = Ez[n], e2_n2 = Ex[n + 1], e2_n = Ex[n];
        T t = (T)0.5 * (c1x * (e1_n1 - e1_n) - c1z * (e2_n2 - e2_n));
        f.Q[1][n] = t;
        f.H[1][n] += t;
    }
    if (i <= g.Nx && l <= g.Ny) {   /* Qz over [0,Nx]x[0,Ny]x[0,Nz+1] */
        T e1_n1 = Ex[n + g.sy], e1_n = Ex[n], e2_n2 = Ey[n + g.sx], e2_n = Ey[n];
        T t = (T)0.5 * (c1y * (e1_n1 - e1_n) - c1x * (e2_n2 - e2_n));
        f.Q[2][n] = t;
        f.H[2][n] += t;
    }
    f.J[0][n] = (JT)0;
    f.J[1][n] = (JT)0;
    f.J[2][n] = (JT)0;
}

/* MagneticFieldStageTwo (gpu_plasma.h:540-551) -> GPU_emh2 (kernels.cu:1024-1049) */
template <typename T, typename JT>
__global__ __launch_bounds__(PIC_BLOCK) void k_h2(Grid g, FV<T, JT> f) {
    long n = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (n >= g.G) return;
    int i, l, k;
    decode(g, n, i, l, k);
    if (l <= g.Ny && k <= g.Nz) f.H[0][n] += f.Q[0][n];
    if (i <= g.Nx && k <= g.Nz) f.H[1][n] += f.Q[1][n];
    if (i <= g.Nx && l <= g.Ny) f.H[2][n] += f.Q[2][n];
}

/* ElectricFieldEvaluate (gpu_plasma.h:455-481) -> GPU_eme / emeElement
 * (kernels.cu:1091-1181); starts and shifts from GetElectricFieldStartsDirs
 * (gpu_plasma.h:286-304), finish (Nx,Ny,Nz). */
template <typename T, typename JT>
__global__ __launch_bounds__(PIC_BLOCK) void k_e(Grid g, FV<T, JT> f) {
    long n = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (n >= g.G) return;
    int i, l, k;
    decode(g, n, i, l, k);
    if (i > g.Nx || l > g.Ny || k > g.Nz) return;
    const T c1x = (T)(g.tau / g.hx), c1y = (T)(g.tau / g.hy), c1z = (T)(g.tau / g.hz),
            tau = (T)g.tau;
    const T *Hx = f.H[0], *Hy = f.H[1], *Hz = f.H[2];
    if (l >= 1 && k >= 1) {   /* Ex: [0,Nx]x[1,Ny]x[1,Nz], H1=Hz(l-1), H2=Hy(k-1) */
        T *E = f.E[0];
        E[n] += c1y * (Hz[n] - Hz[n - g.sy]) - c1z * (Hy[n] - Hy[n - 1]) - tau * jval(f.J[0][n]);
    }
    if (i >= 1 && k >= 1) {   /* Ey: [1,Nx]x[0,Ny]x[1,Nz], H1=Hx(k-1), H2=Hz(i-1) */
        T *E = f.E[1];
        E[n] += c1z * (Hx[n] - Hx[n - 1]) - c1x * (Hz[n] - Hz[n - g.sx]) - tau * jval(f.J[1][n]);
    }
    if (i >= 1 && l >= 1) {   /* Ez: [1,Nx]x[1,Ny]x[0,Nz], H1=Hy(i-1), H2=Hx(l-1) */
        T *E = f.E[2];
        E[n] += c1x * (Hy[n] - Hy[n - g.sx]) - c1y * (Hx[n] - Hx[n - g.sy]) - tau * jval(f.J[2][n]);
    }
}

/* PeriodicBoundaries along x (gpu_plasma.h:802-870 with the arguments at
 * 486-505): Ey over l in [0,Ny], k in [1,Nz]; Ez over l in [1,Ny], k in
 * [0,Nz]; E[0]=E[Nx] and E[Nx+1]=E[1] (periodicElement, kernels.cu:1104). */
template <typename T, typename JT>
__global__ __launch_bounds__(PIC_BLOCK) void k_e_periodic_x(Grid g, FV<T, JT> f) {
    long t = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (t >= (long)(g.Ny + 2) * g.sy) return;
    int l = (int)(t / g.sy), k = (int)(t % g.sy);
    long base = (long)l * g.sy + k;
    long n0 = base, n1 = (long)g.sx + base, nN = (long)g.Nx * g.sx + base,
         nN1 = (long)(g.Nx + 1) * g.sx + base;
    if (l <= g.Ny && k >= 1 && k <= g.Nz) {
        T *E = f.E[1];
        E[n0] = E[nN];
        E[nN1] = E[n1];
    }
    if (l >= 1 && l <= g.Ny && k <= g.Nz) {
        T *E = f.E[2];
        E[n0] = E[nN];
        E[nN1] = E[n1];
    }
}

/* PeriodicBoundaries along z: Ex over i in [0,Nx], l in [0,Ny+1]; Ey over
 * i in [0,Nx+1], l in [0,Ny] (after the x copies). */
template <typename T, typename JT>
__global__ __launch_bounds__(PIC_BLOCK) void k_e_periodic_z(Grid g, FV<T, JT> f) {
    long t = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (t >= (long)(g.Nx + 2) * (g.Ny + 2)) return;
    int i = (int)(t / (g.Ny + 2)), l = (int)(t % (g.Ny + 2));
    long base = (long)i * g.sx + (long)l * g.sy;
    long n0 = base, n1 = base + 1, nN = base + g.Nz, nN1 = base + g.Nz + 1;
    if (i <= g.Nx) {
        T *E = f.E[0];
        E[n0] = E[nN];
        E[nN1] = E[n1];
    }
    if (l <= g.Ny) {
        T *E = f.E[1];
        E[n0] = E[nN];
        E[nN1] = E[n1];
    }
}

/* ghost_mode 1 only (no reference counterpart): Ex and Ez along y over the
 * full ghosted extent of x and z, after the reference's copies */
template <typename T, typename JT>
__global__ __launch_bounds__(PIC_BLOCK) void k_e_periodic_y(Grid g, FV<T, JT> f) {
    long t = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (t >= (long)(g.Nx + 2) * g.sy) return;
    int i = (int)(t / g.sy), k = (int)(t % g.sy);
    long base = (long)i * g.sx + k;
    long n0 = base, n1 = base + g.sy, nN = base + (long)g.Ny * g.sy,
         nN1 = base + (long)(g.Ny + 1) * g.sy;
    for (int c = 0; c < 3; c += 2) {
        T *E = f.E[c];
        E[n0] = E[nN];
        E[nN1] = E[n1];
    }
}

/* SetPeriodicCurrentComponent (gpu_plasma.h:926-978) -> GPU_CurrentPeriodic /
 * periodicCurrentElement (kernels.cu:1124-1159), all three components for one
 * face direction `dir`; the three directions run as three launches in the
 * reference's order. */
template <typename T, typename JT>
__global__ __launch_bounds__(PIC_BLOCK) void k_fold_j(Grid g, FV<T, JT> f, int dir) {
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
        JT *E = f.J[dirE];
        if (dir != dirE) E[n1] += E[nNm1];
        if (dir != 1 || dirE != 1) E[nNm1] = E[n1];
        E[n0] += E[nNm2];
        E[nNm2] = E[n0];
    }
}

/* The periodic E ghost copies and the field energy in one pass (the step's
 * own sequence: x copies, z copies, in ghost_mode 1 the y copies, then the
 * energy over every ghosted node).  Each node's value after the copies is
 * read where the copy chain takes it from: ghost_source follows the copies
 * backwards (the last one first), each applied where its range covers the
 * current node, down to a node no copy writes -- so every read is of a node
 * this kernel does not write, and each copied node gets the value the
 * sequential copies give it.  The energy is summed in k_energy's order. */
__device__ __forceinline__ long ghost_source(const Grid &g, int comp, int i, int l, int k) {
    if (g.ghost_mode == 1 && comp != 1) {            /* y copies of Ex, Ez */
        if (l == 0) l = g.Ny; else if (l == g.Ny + 1) l = 1;
    }
    if ((comp == 0 && i <= g.Nx) || (comp == 1 && l <= g.Ny)) {        /* z copies */
        if (k == 0) k = g.Nz; else if (k == g.Nz + 1) k = 1;
    }
    if ((comp == 1 && l <= g.Ny && k >= 1 && k <= g.Nz) ||
        (comp == 2 && l >= 1 && l <= g.Ny && k <= g.Nz)) {               /* x copies */
        if (i == 0) i = g.Nx; else if (i == g.Nx + 1) i = 1;
    }
    return (long)i * g.sx + (long)l * g.sy + k;
}

template <typename T, typename JT>
__global__ __launch_bounds__(PIC_BLOCK) void k_e_ghosts_energy(Grid g, FV<T, JT> f,
                                                               double *partials) {
    __shared__ double red[PIC_BLOCK];
    double s = 0.0;
    for (long n = (long)blockIdx.x * PIC_BLOCK + threadIdx.x; n < g.G;
         n += (long)gridDim.x * PIC_BLOCK) {
        int i, l, k;
        decode(g, n, i, l, k);
        T v[3];
        for (int c = 0; c < 3; c++) {
            const long src = ghost_source(g, c, i, l, k);
            v[c] = f.E[c][src];
            if (src != n) f.E[c][n] = v[c];
        }
        double a = (double)v[0], b = (double)v[1], c = (double)v[2];
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

/* getElectricEnergy (gpu_plasma.h:621-657) -> GPU_getCellEnergy
 * (wrap_kernel.cu:14-36): sum over every ghosted node of Ex^2+Ey^2+Ez^2,
 * accumulated in FP64 by a deterministic two-level tree instead of
 * one-address atomics. */
template <typename T, typename JT>
__global__ __launch_bounds__(PIC_BLOCK) void k_energy(Grid g, FV<T, JT> f, double *partials) {
    __shared__ double red[PIC_BLOCK];
    const T *ex = f.E[0], *ey = f.E[1], *ez = f.E[2];
    double s = 0.0;
    for (long n = (long)blockIdx.x * PIC_BLOCK + threadIdx.x; n < g.G;
         n += (long)gridDim.x * PIC_BLOCK) {
        double a = (double)ex[n], b = (double)ey[n], c = (double)ez[n];
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

template <typename T, typename JT>
void h_half(const Grid &g, FV<T, JT> f, hipStream_t s) {
    hipLaunchKernelGGL((k_h_half<T, JT>), dim3(blocks_for(g.G)), dim3(PIC_BLOCK), 0, s, g, f);
}
template <typename T, typename JT>
void h2(const Grid &g, FV<T, JT> f, hipStream_t s) {
    hipLaunchKernelGGL((k_h2<T, JT>), dim3(blocks_for(g.G)), dim3(PIC_BLOCK), 0, s, g, f);
}
template <typename T, typename JT>
void e_update(const Grid &g, FV<T, JT> f, hipStream_t s) {
    hipLaunchKernelGGL((k_e<T, JT>), dim3(blocks_for(g.G)), dim3(PIC_BLOCK), 0, s, g, f);
}
template <typename T, typename JT>
void e_periodic(const Grid &g, FV<T, JT> f, hipStream_t s) {
    hipLaunchKernelGGL((k_e_periodic_x<T, JT>), dim3(blocks_for((long)(g.Ny + 2) * g.sy)),
                       dim3(PIC_BLOCK), 0, s, g, f);
    hipLaunchKernelGGL((k_e_periodic_z<T, JT>), dim3(blocks_for((long)(g.Nx + 2) * (g.Ny + 2))),
                       dim3(PIC_BLOCK), 0, s, g, f);
    if (g.ghost_mode == 1)
        hipLaunchKernelGGL((k_e_periodic_y<T, JT>), dim3(blocks_for((long)(g.Nx + 2) * g.sy)),
                           dim3(PIC_BLOCK), 0, s, g, f);
}
template <typename T, typename JT>
void fold_j(const Grid &g, FV<T, JT> f, hipStream_t s) {
    long faces[3] = {(long)(g.Ny + 2) * (g.Nz + 2), (long)(g.Nx + 2) * (g.Nz + 2),
                     (long)(g.Nx + 2) * (g.Ny + 2)};
    for (int dir = 0; dir < 3; dir++)
        hipLaunchKernelGGL((k_fold_j<T, JT>), dim3(blocks_for(faces[dir])), dim3(PIC_BLOCK), 0, s,
                           g, f, dir);
}
template <typename T, typename JT>
void ghosts_energy(const Grid &g, FV<T, JT> f, double *partials, int nblocks, hipStream_t s) {
    hipLaunchKernelGGL((k_e_ghosts_energy<T, JT>), dim3(nblocks), dim3(PIC_BLOCK), 0, s, g, f,
                       partials);
}
template <typename T, typename JT>
void energy(const Grid &g, FV<T, JT> f, double *partials, int nblocks, double *out,
            hipStream_t s) {
    hipLaunchKernelGGL((k_energy<T, JT>), dim3(nblocks), dim3(PIC_BLOCK), 0, s, g, f, partials);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(PIC_BLOCK), 0, s, partials, nblocks, out);
}

}  // namespace

void launch_h_half(const Grid &g, FieldPtrs f, hipStream_t s) { h_half(g, view(f), s); }
void launch_h2(const Grid &g, FieldPtrs f, hipStream_t s) { h2(g, view(f), s); }
void launch_e(const Grid &g, FieldPtrs f, hipStream_t s) { e_update(g, view(f), s); }
void launch_e_periodic(const Grid &g, FieldPtrs f, hipStream_t s) { e_periodic(g, view(f), s); }
void launch_fold_j(const Grid &g, FieldPtrs f, hipStream_t s) { fold_j(g, view(f), s); }
void launch_energy(const Grid &g, FieldPtrs f, double *partials, int nblocks, double *out,
                   hipStream_t s) {
    energy(g, view(f), partials, nblocks, out, s);
}

void launch_h_half32(const Grid &g, FieldPtrs32 f, hipStream_t s) { h_half(g, view(f), s); }
void launch_h2_32(const Grid &g, FieldPtrs32 f, hipStream_t s) { h2(g, view(f), s); }
void launch_e32(const Grid &g, FieldPtrs32 f, hipStream_t s) { e_update(g, view(f), s); }
void launch_e_periodic32(const Grid &g, FieldPtrs32 f, hipStream_t s) {
    e_periodic(g, view(f), s);
}
void launch_fold_j32(const Grid &g, FieldPtrs32 f, hipStream_t s) { fold_j(g, view(f), s); }
void launch_energy32(const Grid &g, FieldPtrs32 f, double *partials, int nblocks, double *out,
                     hipStream_t s) {
    energy(g, view(f), partials, nblocks, out, s);
}

void launch_ghosts_energy(const Grid &g, FieldPtrs f, double *partials, int nblocks,
                          hipStream_t s) {
    ghosts_energy(g, view(f), partials, nblocks, s);
}
void launch_ghosts_energy32(const Grid &g, FieldPtrs32 f, double *partials, int nblocks,
                            hipStream_t s) {
    ghosts_energy(g, view(f), partials, nblocks, s);
}
void launch_reduce_final(const double *partials, int nblocks, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(PIC_BLOCK), 0, s, partials, nblocks, out);
}

void launch_reset_status(DevStatus *st, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(1), 0, s, st);
}
