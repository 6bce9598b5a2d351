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

/* the same decode in 32-bit arithmetic (pic_create keeps G below 2^31) */
__device__ __forceinline__ void decode32(const Grid &g, unsigned n, int &i, int &l, int &k) {
    const unsigned sy = (unsigned)g.sy, ny2 = (unsigned)(g.Ny + 2);
    const unsigned t = n / sy;
    k = (int)(n - t * sy);
    const unsigned ii = t / ny2;
    l = (int)(t - ii * ny2);
    i = (int)ii;
}

/* Workgroups are dispatched round-robin over the 8 XCDs (workgroup b on XCD
 * b mod 8), and each XCD has its own L2.  Renumbered this way each XCD takes
 * one contiguous run of the nodes in order, so that the stencil's +-1, +-sy
 * and +-sx neighbours are read by workgroups of the same XCD close in time
 * (L2 hits) instead of by the other seven. */
__device__ __forceinline__ unsigned xcd_block() {
    const unsigned nb = gridDim.x, b = blockIdx.x, x = b & 7u, per = nb >> 3, rem = nb & 7u;
    return x * per + min(x, rem) + (b >> 3);
}

/* curl and current coefficients tau/h and tau in the field type, rounded on
 * the host exactly as (T)(tau / h) rounds on the device */
template <typename T>
struct YeeC {
    T c1x, c1y, c1z, tau;
};
template <typename T>
YeeC<T> yee_coef(const Grid &g) {
    return YeeC<T>{(T)(g.tau / g.hx), (T)(g.tau / g.hy), (T)(g.tau / g.hz), (T)g.tau};
}

/* the second H array the fused H / E pass writes H^{n+1} into */
template <typename T>
struct HOut {
    T *H[3];
};

/* MagneticStageOne (gpu_plasma.h:524-538) -> GPU_emh1 / emh1_Element
 * (kernels.cu:1052-1089) for the three components, ranges from
 * MagneticFieldTrace (gpu_plasma.h:735-767), shifts from
 * getMagneticFieldTraceShifts (722-733); fused with the J := 0 of
 * SetCurrentsToZero (gpu_plasma.h:2058-2069). */
template <typename T, typename JT>
__global__ __launch_bounds__(PIC_BLOCK) void k_h_half(Grid g, FV<T, JT> f, YeeC<T> c) {
    const unsigned n = xcd_block() * PIC_BLOCK + threadIdx.x;
    if (n >= (unsigned)g.G) return;
    int i, l, k;
    decode32(g, n, i, l, k);
    const T c1x = c.c1x, c1y = c.c1y, c1z = c.c1z;
    const unsigned sy = (unsigned)g.sy, sx = (unsigned)g.sx;
    const T *Ex = f.E[0], *Ey = f.E[1], *Ez = f.E[2];
    if (l <= g.Ny && k <= g.Nz) {   /* Qx over [0,Nx+1]x[0,Ny]x[0,Nz] */
        T e1_n1 = Ey[n + 1], e1_n = Ey[n], e2_n2 = Ez[n + sy], e2_n = Ez[n];
        T t = (T)0.5 * (c1z * (e1_n1 - e1_n) - c1y * (e2_n2 - e2_n));
        f.Q[0][n] = t;
        f.H[0][n] += t;
    }
    if (i <= g.Nx && k <= g.Nz) {   /* Qy over [0,Nx]x[0,Ny+1]x[0,Nz] */
        T e1_n1 = Ez[n + sx], e1_n = Ez[n], e2_n2 = Ex[n + 1], e2_n = Ex[n];
        T t = (T)0.5 * (c1x * (e1_n1 - e1_n) - c1z * (e2_n2 - e2_n));
        f.Q[1][n] = t;
        f.H[1][n] += t;
    }
    if (i <= g.Nx && l <= g.Ny) {   /* Qz over [0,Nx]x[0,Ny]x[0,Nz+1] */
        T e1_n1 = Ex[n + sy], e1_n = Ex[n], e2_n2 = Ey[n + sx], e2_n = Ey[n];
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

/* The periodic E ghost copies (the step's own sequence: x copies, z copies,
 * in ghost_mode 1 the y copies) as a map from each node to the node whose
 * value it holds after them: the copies followed backwards (the last one
 * first), each applied where its range covers the current node, down to a
 * node no copy writes.  A node no copy writes maps to itself.  Each step
 * moves one coordinate from 0 to N or from N+1 to 1. */
__device__ __forceinline__ void ghost_src3(const Grid &g, int comp, int &i, int &l, int &k) {
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
}

/* H^{n+1} = H^{n+1/2} + Q of component d at node m = (i, l, k), over k_h2's
 * ranges (H unchanged outside them) */
template <typename T, typename JT>
__device__ __forceinline__ T hn1(const Grid &g, const FV<T, JT> &f, int d, unsigned m, int i, int l,
                                 int k) {
    T h = f.H[d][m];
    const bool in = d == 0 ? (l <= g.Ny && k <= g.Nz)
                  : d == 1 ? (i <= g.Nx && k <= g.Nz) : (i <= g.Nx && l <= g.Ny);
    if (in) h += f.Q[d][m];
    return h;
}

/* k_fold_j's four operations on one line, restricted to one partner pair:
 * r at the representative (position N or 1), p at its partner (0 or N+1),
 * for fold direction dir and current component cc:
 *   J[1] += J[N+1] (dir != cc);  J[N+1] = J[1] (unless dir = cc = y);
 *   J[0] += J[N];  J[N] = J[0]. */
template <typename JT>
__device__ __forceinline__ void fold_pair(JT &r, JT &p, int dir, int cc, bool at_n) {
    if (at_n) {
        const JT s = p + r;
        r = s;
        p = s;
    } else if (dir != cc) {
        const JT s = r + p;
        r = s;
        p = s;
    } else if (dir != 1) {
        p = r;
    }
}

/* the classes of nodes the current fold combines: along each axis the fold
 * only ever combines position N with 0 and position 1 with N+1 (N >= 3, so
 * the pairs are distinct), so the nodes whose coordinates are a
 * representative's own or its partner's (representative in [1,N]^3) form a
 * class no other class's fold touches.  The classes of more than one node
 * have their representative on a face of [1,N]^3: 2 Ny Nz with ri in
 * {1, Nx}, 2 (Nx-2) Nz with ri inside and rl in {1, Ny}, 2 (Nx-2)(Ny-2) with
 * ri, rl inside and rk in {1, Nz}. */
__device__ __forceinline__ unsigned shell_count(const Grid &g) {
    return 2u * (unsigned)g.Ny * g.Nz + 2u * (unsigned)(g.Nx - 2) * g.Nz +
           2u * (unsigned)(g.Nx - 2) * (g.Ny - 2);
}

/* SetPeriodicCurrentComponent (gpu_plasma.h:926-978), the three k_fold_j
 * launches x, y, z in one: one thread per class of more than one node, the
 * class's current folded in registers (x pairs, then y, then z, as the
 * launches run, with the same additions) and written back. */
template <typename T, typename JT>
__global__ __launch_bounds__(PIC_BLOCK) void k_fold_j_classes(Grid g, FV<T, JT> f) {
    unsigned t = blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (t >= shell_count(g)) return;
    const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
    int ri, rl, rk;
    const unsigned fx = 2u * Ny * Nz, fy = 2u * (unsigned)(Nx - 2) * Nz;
    if (t < fx) {                                    /* ri in {1, Nx} */
        const unsigned a = t / ((unsigned)Ny * Nz), r = t - a * ((unsigned)Ny * Nz);
        ri = a ? Nx : 1;
        rl = (int)(r / Nz) + 1;
        rk = (int)(r % Nz) + 1;
    } else if ((t -= fx) < fy) {                     /* rl in {1, Ny} */
        const unsigned a = t / ((unsigned)(Nx - 2) * Nz), r = t - a * ((unsigned)(Nx - 2) * Nz);
        rl = a ? Ny : 1;
        ri = (int)(r / Nz) + 2;
        rk = (int)(r % Nz) + 1;
    } else {                                         /* rk in {1, Nz} */
        t -= fy;
        const unsigned a = t / ((unsigned)(Nx - 2) * (Ny - 2)),
                       r = t - a * ((unsigned)(Nx - 2) * (Ny - 2));
        rk = a ? Nz : 1;
        ri = (int)(r / (Ny - 2)) + 2;
        rl = (int)(r % (Ny - 2)) + 2;
    }
    const unsigned sy = (unsigned)g.sy, sx = (unsigned)g.sx;
    const int px = ri == Nx ? 0 : ri == 1 ? Nx + 1 : -1;
    const int py = rl == Ny ? 0 : rl == 1 ? Ny + 1 : -1;
    const int pz = rk == Nz ? 0 : rk == 1 ? Nz + 1 : -1;
    const int X[2] = {ri, px}, Y[2] = {rl, py}, Z[2] = {rk, pz};
    auto member = [&](int s) {
        return ((s & 1) == 0 || px >= 0) && ((s & 2) == 0 || py >= 0) && ((s & 4) == 0 || pz >= 0);
    };
    unsigned m[8];
#pragma unroll
    for (int s = 0; s < 8; s++)
        m[s] = (unsigned)X[s & 1] * sx + (unsigned)Y[(s >> 1) & 1] * sy + (unsigned)Z[s >> 2];
    for (int cc = 0; cc < 3; cc++) {
        JT *J = f.J[cc];
        JT v[8];
#pragma unroll
        for (int s = 0; s < 8; s++) v[s] = member(s) ? J[m[s]] : (JT)0;
        if (px >= 0) {
#pragma unroll
            for (int s = 0; s < 8; s += 2)
                if (member(s)) fold_pair(v[s], v[s | 1], 0, cc, ri == Nx);
        }
        if (py >= 0) {
#pragma unroll
            for (int s = 0; s < 8; s++)
                if (!(s & 2) && member(s)) fold_pair(v[s], v[s | 2], 1, cc, rl == Ny);
        }
        if (pz >= 0) {
#pragma unroll
            for (int s = 0; s < 4; s++)
                if (member(s)) fold_pair(v[s], v[s | 4], 2, cc, rk == Nz);
        }
#pragma unroll
        for (int s = 0; s < 8; s++)
            if (member(s)) J[m[s]] = v[s];
    }
}

/* The step's field update after the current fold in one pass:
 * MagneticFieldStageTwo (H^{n+1} = H^{n+1/2} + Q, as k_h2), the E update of
 * ElectricFieldEvaluate (as k_e) and the periodic E ghosts (as the
 * k_e_periodic_* sequence), with the same floating-point operations in the
 * same order, so the fields equal the separate stages' bit for bit.
 *   - H^{n+1} is formed where it is read: at the node and at the -1, -sy,
 *     -sx neighbours the curl takes, each as H + Q over k_h2's ranges, and
 *     written to the second H array `out` (every node), so no node reads
 *     another's H^{n+1}.
 *   - The E update reads only its own node's E and writes it in place.
 *   - Each periodic ghost of E is written by the thread of the node its
 *     value comes from (ghost_src3): the candidates are the nodes whose
 *     coordinate along each axis is the node's own or the ghost that copies
 *     from it (0 from N, N+1 from 1).  Ghosts lie outside the E update
 *     ranges and are never a source, so no node is both written and read. */
template <typename T, typename JT>
__global__ __launch_bounds__(PIC_BLOCK) void k_h2_e(Grid g, FV<T, JT> f, HOut<T> out, YeeC<T> c) {
    const unsigned n = xcd_block() * PIC_BLOCK + threadIdx.x;
    if (n >= (unsigned)g.G) return;
    int i, l, k;
    decode32(g, n, i, l, k);
    const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
    const T hx0 = hn1(g, f, 0, n, i, l, k), hy0 = hn1(g, f, 1, n, i, l, k),
            hz0 = hn1(g, f, 2, n, i, l, k);
    out.H[0][n] = hx0;
    out.H[1][n] = hy0;
    out.H[2][n] = hz0;
    T e[3];
    bool have[3] = {false, false, false};
    if (i <= Nx && l <= Ny && k <= Nz) {
        const unsigned sy = (unsigned)g.sy, sx = (unsigned)g.sx;
        if (l >= 1 && k >= 1) {   /* Ex: [0,Nx]x[1,Ny]x[1,Nz] */
            e[0] = f.E[0][n] + (c.c1y * (hz0 - hn1(g, f, 2, n - sy, i, l - 1, k)) -
                                c.c1z * (hy0 - hn1(g, f, 1, n - 1, i, l, k - 1)) -
                                c.tau * jval(f.J[0][n]));
            f.E[0][n] = e[0];
            have[0] = true;
        }
        if (i >= 1 && k >= 1) {   /* Ey: [1,Nx]x[0,Ny]x[1,Nz] */
            e[1] = f.E[1][n] + (c.c1z * (hx0 - hn1(g, f, 0, n - 1, i, l, k - 1)) -
                                c.c1x * (hz0 - hn1(g, f, 2, n - sx, i - 1, l, k)) -
                                c.tau * jval(f.J[1][n]));
            f.E[1][n] = e[1];
            have[1] = true;
        }
        if (i >= 1 && l >= 1) {   /* Ez: [1,Nx]x[1,Ny]x[0,Nz] */
            e[2] = f.E[2][n] + (c.c1x * (hy0 - hn1(g, f, 1, n - sx, i - 1, l, k)) -
                                c.c1y * (hx0 - hn1(g, f, 0, n - sy, i, l - 1, k)) -
                                c.tau * jval(f.J[2][n]));
            f.E[2][n] = e[2];
            have[2] = true;
        }
    }
    /* periodic ghosts taking this node's value (only nodes at 1 or N along
     * some axis have any) */
    const int gi = i == Nx ? 0 : i == 1 ? Nx + 1 : -1;
    const int gl = l == Ny ? 0 : l == 1 ? Ny + 1 : -1;
    const int gk = k == Nz ? 0 : k == 1 ? Nz + 1 : -1;
    if (gi < 0 && gl < 0 && gk < 0) return;
    for (int s = 1; s < 8; s++) {
        const int ci = (s & 1) ? gi : i, cl = (s & 2) ? gl : l, ck = (s & 4) ? gk : k;
        if (ci < 0 || cl < 0 || ck < 0) continue;
        const unsigned m = (unsigned)ci * g.sx + (unsigned)cl * g.sy + (unsigned)ck;
        for (int cc = 0; cc < 3; cc++) {
            int si = ci, sl = cl, sk = ck;
            ghost_src3(g, cc, si, sl, sk);
            if (si != i || sl != l || sk != k) continue;
            if (!have[cc]) {
                e[cc] = f.E[cc][n];
                have[cc] = true;
            }
            f.E[cc][m] = e[cc];
        }
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
    hipLaunchKernelGGL((k_h_half<T, JT>), dim3(blocks_for(g.G)), dim3(PIC_BLOCK), 0, s, g, f,
                       yee_coef<T>(g));
}
template <typename T, typename JT>
void fold_h2_e(const Grid &g, FV<T, JT> f, HOut<T> out, hipStream_t s) {
    const long shell = 2L * g.Ny * g.Nz + 2L * (g.Nx - 2) * g.Nz + 2L * (g.Nx - 2) * (g.Ny - 2);
    hipLaunchKernelGGL((k_fold_j_classes<T, JT>), dim3(blocks_for(shell)), dim3(PIC_BLOCK), 0, s,
                       g, f);
    hipLaunchKernelGGL((k_h2_e<T, JT>), dim3(blocks_for(g.G)), dim3(PIC_BLOCK), 0, s, g, f, out,
                       yee_coef<T>(g));
}template <typename T, typename JT>
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

void launch_fold_h2_e(const Grid &g, FieldPtrs f, double *const h_out[3], hipStream_t s) {
    fold_h2_e(g, view(f), HOut<double>{{h_out[0], h_out[1], h_out[2]}}, s);
}
void launch_fold_h2_e32(const Grid &g, FieldPtrs32 f, float *const h_out[3], hipStream_t s) {
    fold_h2_e(g, view(f), HOut<float>{{h_out[0], h_out[1], h_out[2]}}, s);
}
void launch_reduce_final(const double *partials, int nblocks, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(PIC_BLOCK), 0, s, partials, nblocks, out);
}

void launch_reset_status(DevStatus *st, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(1), 0, s, st);
}
