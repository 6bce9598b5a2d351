/*
 * pic_oracle.cpp -- FP64 CPU restatement of the reference Plasma::Step.
 *
 * TEST INFRASTRUCTURE ONLY (see pic_oracle.h).  It is the checker for the HIP
 * path, never the thing measured as the product.  Build: oracle/Makefile.
 *
 * Arithmetic is written in the reference's evaluation order, built with
 * -ffp-contract=off and without -ffast-math, so every rounding matches the
 * reference's double-precision code except where the reference itself is
 * order-free (atomic sums: GPU_WriteAllCurrents, GPU_getCellEnergy).
 */
#include "pic_oracle.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <memory>
#include <thread>

namespace {

/* Particle (particle.h:38-54).  `sort` and `direction` are kept as ints. */
struct OParticle {
    double x, y, z, pu, pv, pw, m, q_m, x1, y1, z1;
    int direction;
    int sort;
};

struct OCellTiles {           /* CellDouble M[5][5][5] (cell.h:45-47) x 6 */
    double E[6][5][5][5];
};

struct I3 { int x, y, z; };
struct D3 { double x, y, z; };

}  // namespace

struct orc_plasma {
    int Nx, Ny, Nz;
    double Lx, Ly, Lz, tau;
    double hx, hy, hz;
    int departure_cap;
    int nthreads;
    long G;
    std::vector<double> f[ORC_NFIELDS];
    /* one particle list per ghosted cell, index getGlobalCellNumber (cell.h:796) */
    std::vector<std::vector<OParticle>> cells;

    /* cell.h:785-801 getGlobalCellNumberTriplet / getGlobalCellNumber */
    inline long gcn(int i, int l, int k) const {
        if (i >= Nx + 2) i -= Nx + 2;
        if (l >= Ny + 2) l -= Ny + 2;
        if (k >= Nz + 2) k -= Nz + 2;
        return (long)i * (Ny + 2) * (Nz + 2) + (long)l * (Nz + 2) + k;
    }
    /* cell.h:806-853 getWrapCellNumber (periodic wrap into 0..N-1) */
    inline long wrap_cell(int i, int l, int k) const {
        if (i >= Nx) i -= Nx; else if (i < 0) i += Nx;
        if (l >= Ny) l -= Ny; else if (l < 0) l += Ny;
        if (k >= Nz) k -= Nz; else if (k < 0) k += Nz;
        return (long)i * (Ny + 2) * (Nz + 2) + (long)l * (Nz + 2) + k;
    }
};

namespace {

/* Cell geometry (cell.h:731-737): hx = Lx/Nx, x0 = (double)i*hx, xm = Lx */
struct CellGeom {
    int i, l, k, Nx, Ny, Nz;
    double hx, hy, hz, tau, x0, y0, z0, xm, ym, zm;
};

inline CellGeom geom(const orc_plasma *P, int i, int l, int k) {
    CellGeom c;
    c.i = i; c.l = l; c.k = k;
    c.Nx = P->Nx; c.Ny = P->Ny; c.Nz = P->Nz;
    c.hx = P->Lx / ((double)P->Nx);
    c.hy = P->Ly / ((double)P->Ny);
    c.hz = P->Lz / ((double)P->Nz);
    c.xm = P->Lx; c.ym = P->Ly; c.zm = P->Lz;
    c.x0 = (double)i * c.hx;
    c.y0 = (double)l * c.hy;
    c.z0 = (double)k * c.hz;
    c.tau = P->tau;
    return c;
}

/* cell.h:717 isPointInCell */
inline bool in_cell(const CellGeom &c, double x, double y, double z) {
    return (c.x0 <= x) && (x < c.x0 + c.hx) && (c.y0 <= y) && (y < c.y0 + c.hy) &&
           (c.z0 <= z) && (z < c.z0 + c.hz);
}

/* cell.h:527-530 getCellFraction, getCellNumber */
inline double cell_fraction(double x, double x0, double h) { return (x - x0) / h; }
inline int cell_number(double x, double x0, double h) {
    return ((int)(cell_fraction(x, x0, h) + 1.0) + 1);
}
/* cell.h:544-549 getCellNumberCenter */
inline int cell_number_center(double x, double x0, double h) {
    double t = ((cell_fraction(x, x0, h) + 1.0) + 1.5);
    return (int)(t);
}
/* cell.h:560-566 getCellNumberCenterCurrent */
inline int cell_number_center_current(double x, double x0, double h) {
    double t = (cell_fraction(x, x0, h) + 1.5);
    return (int)(t + 1);
}

/* cell.h:1193-1236 s1..s6_interpolate: weights from the global coordinate */
inline double s_node(double x, double h) { return (int)(x / h + 1.0) - x / h; }
inline double s_center(double x, double h) { return (int)(x / h + 1.5) - 0.5 - x / h; }

/* cell.h:1054-1153 Interpolate3D (returns t2 + t1) */
inline double interpolate3d(const double M[5][5][5], int i, int l, int k, double sx,
                            double sx1, double sy, double sy1, double sz, double sz1) {
    if (i < 0 || i > 5) return 0.0;
    if (l < 0 || l > 5) return 0.0;
    if (k < 0 || k > 5) return 0.0;
    double t_ilk = M[i][l][k];
    double t_ilk1 = M[i][l][k + 1];
    double t_il1k = M[i][l + 1][k];
    double t_il1k1 = M[i][l + 1][k + 1];
    double t_i1lk = M[i + 1][l][k];
    double t_i1lk1 = M[i + 1][l][k + 1];
    double t_i1l1k = M[i + 1][l + 1][k];
    double t_i1l1k1 = M[i + 1][l + 1][k + 1];
    double t1 = sx * (sy * (sz * t_ilk + sz1 * t_ilk1) + sy1 * (sz * t_il1k + sz1 * t_il1k1));
    double t2 = sx1 * (sy * (sz * t_i1lk + sz1 * t_i1lk1) + sy1 * (sz * t_i1l1k + sz1 * t_i1l1k1));
    return t2 + t1;
}

/* cell.h:1984-1993 getFortranArrayValue and cell.h:2000-2036 readField:
 * tile M[a][b][c] = F[global(i+a-2, l+b-2, k+c-2)], zero if flat n<0 or n>G */
void read_tile(const orc_plasma *P, const double *F, int i, int l, int k, double M[5][5][5]) {
    for (int a = 0; a < 5; a++)
        for (int b = 0; b < 5; b++)
            for (int c = 0; c < 5; c++) {
                long n = P->gcn(i + a - 2, l + b - 2, k + c - 2);
                M[a][b][c] = (n < 0 || n > P->G) ? 0.0 : F[n];
            }
}

/* particle.h:160-202 Particle::Move, with ElectricMove (65-75) and MagneticMove (77-99) */
inline void boris_move(OParticle &p, const D3 &E, const D3 &H, double tau) {
    double tau1 = p.q_m * tau * 0.5;
    p.pu += tau1 * E.x;
    p.pv += tau1 * E.y;
    p.pw += tau1 * E.z;
    double ps = tau1 * std::pow((p.pu * p.pu + p.pv * p.pv + p.pw * p.pw) * 1. + 1.0, -0.5);

    double bx = ps * H.x, by = ps * H.y, bz = ps * H.z;
    double su = p.pu + p.pv * bz - p.pw * by;
    double sv = p.pv + p.pw * bx - p.pu * bz;
    double sw = p.pw + p.pu * by - p.pv * bx;
    double s1 = bx * bx, s2 = by * by, s3 = bz * bz;
    double s4 = bx * by, s5 = by * bz, s6 = bz * bx;
    double s = s1 + 1. + s2 + s3;
    double pu1 = ((s1 + 1.) * su + (s4 + bz) * sv + (s6 - by) * sw) / s;
    double pv1 = ((s4 - bz) * su + (s2 + 1.) * sv + (s5 + bx) * sw) / s;
    double pw1 = ((s6 + by) * su + (s5 - bx) * sv + (s3 + 1.) * sw) / s;

    D3 sx3 = {E.x * tau1, E.y * tau1, E.z * tau1};   /* mult (particle.h:128-135) */
    p.pu = pu1 + sx3.x;
    p.pv = pv1 + sx3.y;
    p.pw = pw1 + sx3.z;
    ps = std::pow(((p.pu * p.pu + p.pv * p.pv + p.pw * p.pw) + 1.0), -0.5);
    double u = ps * p.pu, v = ps * p.pv, w = ps * p.pw;
    p.x1 = p.x + tau * u;
    p.y1 = p.y + tau * v;
    p.z1 = p.z + tau * w;
}

/* cell.h:1257-1290 GetField = InverseKernel (1010-1047) + GetElectricField
 * (746-772) + GetMagneticField (1161-1188) */
inline void get_field(const CellGeom &c, const OCellTiles &T, const OParticle &p, D3 &E, D3 &H) {
    if (p.x < 0 || p.y < 0 || p.z < 0) {
        E = {0, 0, 0};
        H = {0, 0, 0};
        return;
    }
    I3 i, i1;
    i.x = cell_number(p.x, c.x0, c.hx);
    i1.x = cell_number_center(p.x, c.x0, c.hx);
    double s1 = s_node(p.x, c.hx);
    double s2 = s_center(p.x, c.hx);
    i.y = cell_number(p.y, c.y0, c.hy);
    i1.y = cell_number_center(p.y, c.y0, c.hy);
    double s3 = s_node(p.y, c.hy);
    double s4 = s_center(p.y, c.hy);
    i.z = cell_number(p.z, c.z0, c.hz);
    i1.z = cell_number_center(p.z, c.z0, c.hz);
    double s5 = s_node(p.z, c.hz);
    double s6 = s_center(p.z, c.hz);
    double s11 = 1. - s1, s21 = 1. - s2, s31 = 1. - s3, s41 = 1. - s4, s51 = 1. - s5,
           s61 = 1. - s6;
    E.x = interpolate3d(T.E[0], i.x, i1.y, i1.z, s1, s11, s4, s41, s6, s61);
    E.y = interpolate3d(T.E[1], i1.x, i.y, i1.z, s2, s21, s3, s31, s6, s61);
    E.z = interpolate3d(T.E[2], i1.x, i1.y, i.z, s2, s21, s4, s41, s5, s51);
    H.x = interpolate3d(T.E[3], i1.x, i.y, i.z, s2, s21, s3, s31, s5, s51);
    H.y = interpolate3d(T.E[4], i.x, i1.y, i.z, s1, s11, s4, s41, s5, s51);
    H.z = interpolate3d(T.E[5], i.x, i.y, i1.z, s1, s11, s3, s31, s6, s61);
}

/* tensor.h:26-62 CurrentTensorComponent: 4 (node, value) pairs */
struct TComp {
    int n[4][3];
    double t[4];
};
struct TTensor {
    TComp J[3];
};

/* cell.h:1511-1693 pqr (the tensor-filling overload) */
inline void pqr(const CellGeom &c, const I3 &i, const D3 &x, const D3 &x1, double a1, double tau,
                TTensor &t) {
    /* cell.h:626-629 getCenterRelatedShift */
    double dx = (0.50 * (x.x + x1.x) - c.hx * ((double)i.x - 2.50) - c.x0);
    double dy = (0.50 * (x.y + x1.y) - c.hy * ((double)i.y - 2.50) - c.y0);
    double dz = (0.50 * (x.z + x1.z) - c.hz * ((double)i.z - 2.50) - c.z0);
    double a = a1;
    double dx1 = c.hx - dx, dy1 = c.hy - dy, dz1 = c.hz - dz;
    double su = x1.x - x.x, sv = x1.y - x.y, sw = x1.z - x.z;
    double s1 = sv * sw / 12.0, s2 = su * sw / 12.0, s3 = su * sv / 12.0;
    su = su * a / (tau * c.hy * c.hz);
    sv = sv * a / (tau * c.hx * c.hz);
    sw = sw * a / (tau * c.hx * c.hy);

    TComp &jx = t.J[0];
    jx.n[0][0] = i.x; jx.n[0][1] = i.y;     jx.n[0][2] = i.z;     jx.t[0] = su * (dy1 * dz1 + s1);
    jx.n[1][0] = i.x; jx.n[1][1] = i.y;     jx.n[1][2] = i.z + 1; jx.t[1] = su * (dy1 * dz - s1);
    jx.n[2][0] = i.x; jx.n[2][1] = i.y + 1; jx.n[2][2] = i.z;     jx.t[2] = su * (dy * dz1 - s1);
    jx.n[3][0] = i.x; jx.n[3][1] = i.y + 1; jx.n[3][2] = i.z + 1; jx.t[3] = su * (dy * dz + s1);

    TComp &jy = t.J[1];
    jy.n[0][0] = i.x;     jy.n[0][1] = i.y; jy.n[0][2] = i.z;     jy.t[0] = sv * (dx1 * dz1 + s2);
    jy.n[1][0] = i.x;     jy.n[1][1] = i.y; jy.n[1][2] = i.z + 1; jy.t[1] = sv * (dx1 * dz - s2);
    jy.n[2][0] = i.x + 1; jy.n[2][1] = i.y; jy.n[2][2] = i.z;     jy.t[2] = sv * (dx * dz1 - s2);
    jy.n[3][0] = i.x + 1; jy.n[3][1] = i.y; jy.n[3][2] = i.z + 1; jy.t[3] = sv * (dx * dz + s2);

    TComp &jz = t.J[2];
    jz.n[0][0] = i.x;     jz.n[0][1] = i.y;     jz.n[0][2] = i.z; jz.t[0] = sw * (dx1 * dy1 + s3);
    jz.n[1][0] = i.x;     jz.n[1][1] = i.y + 1; jz.n[1][2] = i.z; jz.t[1] = sw * (dx1 * dy - s3);
    jz.n[2][0] = i.x + 1; jz.n[2][1] = i.y;     jz.n[2][2] = i.z; jz.t[2] = sw * (dx * dy1 - s3);
    jz.n[3][0] = i.x + 1; jz.n[3][1] = i.y + 1; jz.n[3][2] = i.z; jz.t[3] = sw * (dx * dy + s3);
}

/* cell.h:600-620 getCellTransitAverage / Ratio / Product / getRatioBasedX */
inline double transit_average(double h, int i1, int i2, double x0) {
    return (h * (((double)i1 + (double)i2) * 0.5 - 2.0) + x0);
}
inline double transit_ratio(double z1, double z, double z2) { return (z2 - z) / (z1 - z); }
inline double transit_product(double z1, double z, double z2) { return (z2 - z) * (z1 - z); }
inline double ratio_based(double x1, double x, double s) { return (x + (x1 - x) * s); }

/* cell.h:1427-1478 Reflect (periodic wrap; exactly 0 and exactly xm map to 0) */
inline double reflect1(double x, double xm) {
    return (x > xm) * (x - xm) + (x < 0.0) * (xm + x) + (x > 0 && x < xm) * x;
}

/* cell.h:1306-1403 the zig-zag relay point of CurrentToMesh: dual cells i1
 * of x and i2 of x1 (getCellNumberCenterCurrent), crossing code
 * m = 4 di + 2 dl + dk, and the relay point x2 on the crossed dual faces;
 * for two crossed faces the third coordinate follows the path at the
 * parameter s of the least-squares projection onto the crossed edge
 * (cell.h:1374-1403). Returns m (0 when no face is crossed). */
inline int relay_point(const CellGeom &c, const D3 &x, const D3 &x1, I3 &i1, I3 &i2, D3 &x2) {
    i1.x = cell_number_center_current(x.x, c.x0, c.hx);
    i1.y = cell_number_center_current(x.y, c.y0, c.hy);
    i1.z = cell_number_center_current(x.z, c.z0, c.hz);
    i2.x = cell_number_center_current(x1.x, c.x0, c.hx);
    i2.y = cell_number_center_current(x1.y, c.y0, c.hy);
    i2.z = cell_number_center_current(x1.z, c.z0, c.hz);
    int di = std::abs(i2.x - i1.x), dl = std::abs(i2.y - i1.y), dk = std::abs(i2.z - i1.z);
    int m = 4 * di + 2 * dl + dk;
    double s;
    switch (m) {
    case 1:
        x2.z = transit_average(c.hz, i1.z, i2.z, c.z0);
        s = transit_ratio(x1.z, x.z, x2.z);
        x2.x = ratio_based(x1.x, x.x, s);
        x2.y = ratio_based(x1.y, x.y, s);
        break;
    case 2:
        x2.y = transit_average(c.hy, i1.y, i2.y, c.y0);
        s = transit_ratio(x1.y, x.y, x2.y);
        x2.x = ratio_based(x1.x, x.x, s);
        x2.z = ratio_based(x1.z, x.z, s);
        break;
    case 3:
        x2.y = transit_average(c.hy, i1.y, i2.y, c.y0);
        x2.z = transit_average(c.hz, i1.z, i2.z, c.z0);
        s = (transit_product(x1.z, x.z, x2.z) + transit_product(x1.y, x.y, x2.y)) /
            (std::pow(x1.z - x.z, 2.0) + std::pow(x1.y - x.y, 2.0));
        x2.x = ratio_based(x1.x, x.x, s);
        break;
    case 4:
        x2.x = transit_average(c.hx, i1.x, i2.x, c.x0);
        s = transit_ratio(x1.x, x.x, x2.x);
        x2.y = ratio_based(x1.y, x.y, s);
        x2.z = ratio_based(x1.z, x.z, s);
        break;
    case 5:
        x2.x = transit_average(c.hx, i1.x, i2.x, c.x0);
        x2.z = transit_average(c.hz, i1.z, i2.z, c.z0);
        s = (transit_product(x1.z, x.z, x2.z) + transit_product(x1.x, x.x, x2.x)) /
            (std::pow(x1.z - x.z, 2.0) + std::pow(x1.x - x.x, 2.0));
        x2.y = ratio_based(x1.y, x.y, s);
        break;
    case 6:
        x2.x = transit_average(c.hx, i1.x, i2.x, c.x0);
        x2.y = transit_average(c.hy, i1.y, i2.y, c.y0);
        s = (transit_product(x1.y, x.y, x2.y) + transit_product(x1.x, x.x, x2.x)) /
            (std::pow(x1.y - x.y, 2.0) + std::pow(x1.x - x.x, 2.0));
        x2.z = ratio_based(x1.z, x.z, s);
        break;
    case 7:
        x2.x = transit_average(c.hx, i1.x, i2.x, c.x0);
        x2.y = transit_average(c.hy, i1.y, i2.y, c.y0);
        x2.z = transit_average(c.hz, i1.z, i2.z, c.z0);
        break;
    default:
        m = 0;
        break;
    }
    return m;
}

/* cell.h:1306-1422 CurrentToMesh: zig-zag relay + pqr; then x=x1 and Reflect */
inline int current_to_mesh(const CellGeom &c, OParticle &p, TTensor &t1, TTensor &t2) {
    D3 x = {p.x, p.y, p.z};
    D3 x1 = {p.x1, p.y1, p.z1};
    double mass = p.m;
    int cells = 1;
    I3 i1, i2;
    D3 x2;
    int m = relay_point(c, x, x1, i1, i2, x2);
    if (m == 0) {
        pqr(c, i1, x, x1, mass, c.tau, t1);
    } else {
        pqr(c, i1, x, x2, mass, c.tau, t1);
        pqr(c, i2, x2, x1, mass, c.tau, t2);
        cells = 2;
    }
    p.x = p.x1;
    p.y = p.y1;
    p.z = p.z1;
    p.x = reflect1(p.x, c.xm);
    p.y = reflect1(p.y, c.ym);
    p.z = reflect1(p.z, c.zm);
    return cells;
}

/* cell.h:694-701 flyDirection */
inline int fly_direction(const CellGeom &c, const OParticle &p) {
    int dx = (((p.x > c.x0 + c.hx) && (p.x < c.x0 + 2 * c.hx)) || ((p.x < c.hx) && (c.i == c.Nx - 1)))
                 ? 2
                 : (((p.x < c.x0) || ((p.x > c.xm - c.hx) && (c.i == 0))) ? 0 : 1);
    int dy = (((p.y > c.y0 + c.hy) && (p.y < c.y0 + 2 * c.hy)) || ((p.y < c.hy) && (c.l == c.Ny - 1)))
                 ? 2
                 : (((p.y < c.y0) || ((p.y > c.ym - c.hy) && (c.l == 0))) ? 0 : 1);
    int dz = (((p.z > c.z0 + c.hz) && (p.z < c.z0 + 2 * c.hz)) || ((p.z < c.hz) && (c.k == c.Nz - 1)))
                 ? 2
                 : (((p.z < c.z0) || ((p.z > c.zm - c.hz) && (c.k == 0))) ? 0 : 1);
    return dx * 9 + dy * 3 + dz;
}

/* contiguous ranges [t*n/T, (t+1)*n/T) of 0..n-1 on T threads */
template <class F>
void parallel_ranges(int nthreads, long n, F fn) {
    if (nthreads <= 1) {
        fn(0, 0L, n);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++)
        th.emplace_back([&, t]() { fn(t, n * t / nthreads, n * (t + 1) / nthreads); });
    for (auto &x : th) x.join();
}

template <class F>
void parallel_cells(int nthreads, long ncells, F fn) {
    if (nthreads <= 1 || ncells < 64) {
        for (long c = 0; c < ncells; c++) fn(c);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++)
        th.emplace_back([&, t]() {
            for (long c = t; c < ncells; c += nthreads) fn(c);
        });
    for (auto &x : th) x.join();
}

}  // namespace

extern "C" {

orc_plasma *orc_create(int nx, int ny, int nz, double lx, double ly, double lz, double tau,
                       int departure_cap) {
    orc_plasma *P = new orc_plasma();
    P->Nx = nx; P->Ny = ny; P->Nz = nz;
    P->Lx = lx; P->Ly = ly; P->Lz = lz;
    P->tau = tau;
    P->hx = lx / (double)nx;
    P->hy = ly / (double)ny;
    P->hz = lz / (double)nz;
    P->departure_cap = departure_cap;
    P->nthreads = 1;
    P->G = (long)(nx + 2) * (ny + 2) * (nz + 2);
    for (int f = 0; f < ORC_NFIELDS; f++) P->f[f].assign(P->G, 0.0);
    P->cells.resize(P->G);
    return P;
}

void orc_destroy(orc_plasma *p) { delete p; }
void orc_set_threads(orc_plasma *p, int n) { p->nthreads = n < 1 ? 1 : n; }

long orc_load_particles(orc_plasma *P, int sort, const double *x, const double *y, const double *z,
                        const double *px, const double *py, const double *pz, long n, double m,
                        double q_m) {
    long dropped = 0;
    for (long j = 0; j < n; j++) {
        /* cell.h:593-598 getPointCell: (int)((x-0)/hx) per axis */
        int i = (int)cell_fraction(x[j], 0.0, P->hx);
        int l = (int)cell_fraction(y[j], 0.0, P->hy);
        int k = (int)cell_fraction(z[j], 0.0, P->hz);
        if (i < 0 || l < 0 || k < 0 || i >= P->Nx || l >= P->Ny || k >= P->Nz) {
            dropped++;   /* outside the bulk: the reference would index a ghost cell */
            continue;
        }
        CellGeom c = geom(P, i, l, k);
        if (!in_cell(c, x[j], y[j], z[j])) {   /* Cell::Insert (cell.h:1754-1768) */
            dropped++;
            continue;
        }
        OParticle p;
        p.x = x[j]; p.y = y[j]; p.z = z[j];
        p.pu = px[j]; p.pv = py[j]; p.pw = pz[j];
        p.m = m; p.q_m = q_m;
        p.x1 = p.y1 = p.z1 = 0.0;
        p.direction = 0;
        p.sort = sort;
        P->cells[P->gcn(i, l, k)].push_back(p);
    }
    return dropped;
}

void orc_set_field(orc_plasma *P, int which, const double *a) {
    std::memcpy(P->f[which].data(), a, sizeof(double) * P->G);
}
void orc_get_field(const orc_plasma *P, int which, double *a) {
    std::memcpy(a, P->f[which].data(), sizeof(double) * P->G);
}

/* gpu_plasma.h:524-538 MagneticStageOne -> MagneticFieldTrace (735-767) ->
 * GPU_emh1 / emh1_Element (kernels.cu:1052-1089); shifts from
 * getMagneticFieldTraceShifts (722-733).  Then SetCurrentsToZero
 * (gpu_plasma.h:2058-2069): J := 0. */
void orc_h_half(orc_plasma *P) {
    double c1x = P->tau / P->hx, c1y = P->tau / P->hy, c1z = P->tau / P->hz;
    struct Trace {
        int Q, H, E1, E2, iend, lend, kend;
        double c1, c2;
        int d1[3], d2[3];
    } tr[3] = {
        {ORC_QX, ORC_HX, ORC_EY, ORC_EZ, P->Nx + 1, P->Ny, P->Nz, c1z, c1y, {0, 0, 1}, {0, 1, 0}},
        {ORC_QY, ORC_HY, ORC_EZ, ORC_EX, P->Nx, P->Ny + 1, P->Nz, c1x, c1z, {1, 0, 0}, {0, 0, 1}},
        {ORC_QZ, ORC_HZ, ORC_EX, ORC_EY, P->Nx, P->Ny, P->Nz + 1, c1y, c1x, {0, 1, 0}, {1, 0, 0}},
    };
    for (auto &t : tr) {
        double *Q = P->f[t.Q].data(), *H = P->f[t.H].data();
        const double *E1 = P->f[t.E1].data(), *E2 = P->f[t.E2].data();
        for (int i = 0; i <= t.iend; i++)
            for (int l = 0; l <= t.lend; l++)
                for (int k = 0; k <= t.kend; k++) {
                    long n = P->gcn(i, l, k);
                    long n1 = P->gcn(i + t.d1[0], l + t.d1[1], k + t.d1[2]);
                    long n2 = P->gcn(i + t.d2[0], l + t.d2[1], k + t.d2[2]);
                    double e1_n1 = E1[n1], e1_n = E1[n], e2_n2 = E2[n2], e2_n = E2[n];
                    double q = 0.5 * (t.c1 * (e1_n1 - e1_n) - t.c2 * (e2_n2 - e2_n));
                    Q[n] = q;
                    H[n] += Q[n];
                }
    }
    std::fill(P->f[ORC_JX].begin(), P->f[ORC_JX].end(), 0.0);
    std::fill(P->f[ORC_JY].begin(), P->f[ORC_JY].end(), 0.0);
    std::fill(P->f[ORC_JZ].begin(), P->f[ORC_JZ].end(), 0.0);
}

/* kernels.cu:886-924 GPU_StepAllCells -> MoveParticlesInCell (719-756) ->
 * Cell::MoveSingleParticle (cell.h:1847-1858).  Tiles come from
 * GPU_SetFieldsToCells (kernels.cu:32-49) run after the H half step. */
void orc_push(orc_plasma *P) {
    const double *F[6] = {P->f[ORC_EX].data(), P->f[ORC_EY].data(), P->f[ORC_EZ].data(),
                          P->f[ORC_HX].data(), P->f[ORC_HY].data(), P->f[ORC_HZ].data()};
    long nb = (long)P->Nx * P->Ny * P->Nz;
    parallel_cells(P->nthreads, nb, [&](long b) {
        int i = (int)(b / ((long)P->Ny * P->Nz));
        int l = (int)((b / P->Nz) % P->Ny);
        int k = (int)(b % P->Nz);
        std::vector<OParticle> &v = P->cells[P->gcn(i, l, k)];
        if (v.empty()) return;
        CellGeom c = geom(P, i, l, k);
        OCellTiles T;
        for (int f = 0; f < 6; f++) read_tile(P, F[f], i, l, k, T.E[f]);
        for (auto &p : v) {
            D3 E, H;
            get_field(c, T, p, E, H);
            boris_move(p, E, H, c.tau);
        }
    });
}

/* kernels.cu:927-991 GPU_CurrentsAllCells: per particle index%10 buffer
 * (AccumulateCurrentWithParticlesInCell 811-858, writeCurrentComponent
 * 512-528), buffers summed in order 0..9 (addCellDouble 550-574); then
 * GPU_WriteAllCurrents (kernels.cu:83-112): J[|n|] += tile. */
void orc_deposit(orc_plasma *P) {
    long nb = (long)P->Nx * P->Ny * P->Nz;
    /* every tile of a cell with particles is written whole below; the
     * others are never read (no zero fill of nb x 375 doubles) */
    std::unique_ptr<double[]> tiles(new double[(size_t)nb * 3 * 125]);
    std::vector<char> has(nb, 0);
    parallel_cells(P->nthreads, nb, [&](long b) {
        int i = (int)(b / ((long)P->Ny * P->Nz));
        int l = (int)((b / P->Nz) % P->Ny);
        int k = (int)(b % P->Nz);
        std::vector<OParticle> &v = P->cells[P->gcn(i, l, k)];
        if (v.empty()) return;
        CellGeom c = geom(P, i, l, k);
        static thread_local double buf[10][3][5][5][5];
        std::memset(buf, 0, sizeof(buf));
        for (size_t j = 0; j < v.size(); j++) {
            TTensor t1, t2;
            int cells = current_to_mesh(c, v[j], t1, t2);
            int bi = (int)(j % 10);
            for (int comp = 0; comp < 3; comp++) {
                for (int q = 0; q < 4; q++) {
                    const int *n = t1.J[comp].n[q];
                    buf[bi][comp][n[0]][n[1]][n[2]] += t1.J[comp].t[q];
                }
                if (cells == 2)
                    for (int q = 0; q < 4; q++) {
                        const int *n = t2.J[comp].n[q];
                        buf[bi][comp][n[0]][n[1]][n[2]] += t2.J[comp].t[q];
                    }
            }
        }
        double *tile = &tiles[(size_t)b * 375];
        for (int comp = 0; comp < 3; comp++)
            for (int e = 0; e < 125; e++) {
                double d = 0.0;   /* cell J tile was zeroed (GPU_SetAllCurrentsToZero) */
                for (int bi = 0; bi < 10; bi++) d += (&buf[bi][comp][0][0][0])[e];
                tile[comp * 125 + e] = d;
            }
        has[b] = 1;
    });
    /* tiles into J in cell order.  One thread: J[n] += tile, cell after
     * cell (the reference's serial order).  T threads: each adds the tiles
     * of a contiguous range of cells, in cell order, into a private copy of
     * J starting at zero, and the copies are summed in thread order -- a
     * fixed summation order for a given T (J is zero before the deposit),
     * so the result is deterministic but may differ from one thread's in
     * the last bits.  The parity tests run one thread. */
    const long G = P->G;
    auto scatter = [&](double *const J[3], long b0, long b1) {
        for (long b = b0; b < b1; b++) {
            if (!has[b]) continue;
            int i = (int)(b / ((long)P->Ny * P->Nz));
            int l = (int)((b / P->Nz) % P->Ny);
            int k = (int)(b % P->Nz);
            const double *tile = &tiles[(size_t)b * 375];
            for (int a = 0; a < 5; a++)
                for (int bb = 0; bb < 5; bb++)
                    for (int cc = 0; cc < 5; cc++) {
                        long n = P->gcn(i + a - 2, l + bb - 2, k + cc - 2);
                        if (n < 0) n = -n;
                        int e = (a * 5 + bb) * 5 + cc;
                        for (int comp = 0; comp < 3; comp++) J[comp][n] += tile[comp * 125 + e];
                    }
        }
    };
    double *J[3] = {P->f[ORC_JX].data(), P->f[ORC_JY].data(), P->f[ORC_JZ].data()};
    const int T = P->nthreads;
    if (T <= 1 || nb < 64) {
        scatter(J, 0, nb);
        return;
    }
    std::vector<std::vector<double>> priv((size_t)T * 3);
    parallel_ranges(T, nb, [&](int t, long b0, long b1) {
        double *Jt[3];
        for (int comp = 0; comp < 3; comp++) {
            priv[(size_t)t * 3 + comp].assign(G, 0.0);
            Jt[comp] = priv[(size_t)t * 3 + comp].data();
        }
        scatter(Jt, b0, b1);
    });
    parallel_ranges(T, G, [&](int, long n0, long n1) {
        for (int comp = 0; comp < 3; comp++)
            for (long n = n0; n < n1; n++) {
                double d = J[comp][n];
                for (int t = 0; t < T; t++) d += priv[(size_t)t * 3 + comp][n];
                J[comp][n] = d;
            }
    });
}

/* gpu_plasma.h:2035-2046 reorder_particles:
 *   GPU_MakeDepartureLists (kernels.cu:146-216): flyDirection for every
 *     particle; leavers (!isPointInCell) counted against the 50 cap;
 *   GPU_ArrangeFlights (kernels.cu:316-370): each bulk cell scans its 27
 *     periodic neighbours (ix,iy,iz lexicographic) and Inserts every
 *     particle with direction != 13 that lies in it (as direction 13);
 *   GPU_MakeDepartureLists_rm (kernels.cu:219-242): removes direction != 13
 *     by swap-with-last (cell.h:391-432). */
int orc_migrate(orc_plasma *P) {
    const long G = P->G;
    const int T = P->nthreads;
    std::vector<char> too_many(G, 0);
    parallel_cells(T, G, [&](long n) {
        std::vector<OParticle> &v = P->cells[n];
        if (v.empty()) return;
        int i = (int)(n / ((long)(P->Ny + 2) * (P->Nz + 2)));
        int l = (int)((n / (P->Nz + 2)) % (P->Ny + 2));
        int k = (int)(n % (P->Nz + 2));
        CellGeom c = geom(P, i, l, k);
        int len = 0;
        for (auto &p : v) {
            p.direction = fly_direction(c, p);
            if (!in_cell(c, p.x, p.y, p.z)) {
                if (P->departure_cap > 0 && len == P->departure_cap) {
                    too_many[n] = 1;
                    break;
                }
                len++;
            }
        }
    });
    int status = ORC_OK;
    for (long n = 0; n < G; n++)
        if (too_many[n]) status |= ORC_TOO_MANY_PARTICLES;
    /* GPU_ArrangeFlights.  The reference appends a cell's arrivals while
     * later cells still scan it as a source; arrivals carry direction 13 and
     * are skipped as sources, so every cell's arrivals can be collected
     * first (sources unchanged) and appended afterwards: the same lists in
     * the same order for any thread count. */
    const long nb = (long)P->Nx * P->Ny * P->Nz;
    std::vector<std::vector<OParticle>> arrivals(nb);
    parallel_cells(T, nb, [&](long b) {
        int nx = (int)(b / ((long)P->Ny * P->Nz));
        int ny = (int)((b / P->Nz) % P->Ny);
        int nz = (int)(b % P->Nz);
        std::vector<OParticle> &dst = arrivals[b];
        CellGeom c = geom(P, nx, ny, nz);
        for (int ix = 0; ix < 3; ix++)
            for (int iy = 0; iy < 3; iy++)
                for (int iz = 0; iz < 3; iz++) {
                    long s = P->wrap_cell(nx + ix - 1, ny + iy - 1, nz + iz - 1);
                    const std::vector<OParticle> &src = P->cells[s];
                    for (size_t j = 0; j < src.size(); j++) {
                        OParticle p1 = src[j];
                        if (p1.direction != 13) {
                            p1.direction = 13;
                            if (in_cell(c, p1.x, p1.y, p1.z)) dst.push_back(p1);
                        }
                    }
                }
    });
    parallel_cells(T, nb, [&](long b) {
        int nx = (int)(b / ((long)P->Ny * P->Nz));
        int ny = (int)((b / P->Nz) % P->Ny);
        int nz = (int)(b % P->Nz);
        std::vector<OParticle> &dst = P->cells[P->gcn(nx, ny, nz)];
        dst.insert(dst.end(), arrivals[b].begin(), arrivals[b].end());
    });
    /* GPU_MakeDepartureLists_rm: swap-with-last removal, cell by cell */
    parallel_cells(T, G, [&](long n) {
        std::vector<OParticle> &v = P->cells[n];
        for (long num = 0; num < (long)v.size(); num++) {
            if (v[num].direction != 13) {
                long last = (long)v.size() - 1;
                if (num < last) v[num] = v[last];
                v.pop_back();
                num--;
            }
        }
    });
    return status;
}

/* gpu_plasma.h:926-993 SetPeriodicCurrents -> SetPeriodicCurrentComponent ->
 * GPU_CurrentPeriodic / periodicCurrentElement (kernels.cu:1124-1159);
 * boundary index from getGlobalBoundaryCellNumber (cell.h:885-889). */
void orc_fold_j(orc_plasma *P) {
    int Nx = P->Nx, Ny = P->Ny, Nz = P->Nz;
    auto bnd = [&](int a, int b, int dir, int N) -> long {
        int i1 = (dir == 0) * N + (dir == 1) * a + (dir == 2) * a;
        int l1 = (dir == 0) * a + (dir == 1) * N + (dir == 2) * b;
        int k1 = (dir == 0) * b + (dir == 1) * b + (dir == 2) * N;
        return P->gcn(i1, l1, k1);
    };
    for (int dirE = 0; dirE < 3; dirE++) {
        double *E = P->f[ORC_JX + dirE].data();
        for (int dir = 0; dir < 3; dir++) {
            int na, nbb, N;
            if (dir == 0) { na = Ny + 2; nbb = Nz + 2; N = Nx + 2; }
            else if (dir == 1) { na = Nx + 2; nbb = Nz + 2; N = Ny + 2; }
            else { na = Nx + 2; nbb = Ny + 2; N = Nz + 2; }
            for (int a = 0; a < na; a++)
                for (int b = 0; b < nbb; b++) {
                    long n1 = bnd(a, b, dir, 1);
                    long n_Nm1 = bnd(a, b, dir, N - 1);
                    if (dir != dirE) E[n1] += E[n_Nm1];
                    if (dir != 1 || dirE != 1) E[n_Nm1] = E[n1];
                    long n_Nm2 = bnd(a, b, dir, N - 2);
                    long n0 = bnd(a, b, dir, 0);
                    E[n0] += E[n_Nm2];
                    E[n_Nm2] = E[n0];
                }
        }
    }
}

/* gpu_plasma.h:540-551 MagneticFieldStageTwo -> SimpleMagneticFieldTrace
 * (769-799) -> GPU_emh2 / emh2_Element (kernels.cu:1024-1049): H += Q */
void orc_h2(orc_plasma *P) {
    struct R { int Q, H, iend, lend, kend; } rr[3] = {
        {ORC_QX, ORC_HX, P->Nx + 1, P->Ny, P->Nz},
        {ORC_QY, ORC_HY, P->Nx, P->Ny + 1, P->Nz},
        {ORC_QZ, ORC_HZ, P->Nx, P->Ny, P->Nz + 1},
    };
    for (auto &r : rr) {
        const double *Q = P->f[r.Q].data();
        double *H = P->f[r.H].data();
        for (int i = 0; i <= r.iend; i++)
            for (int l = 0; l <= r.lend; l++)
                for (int k = 0; k <= r.kend; k++) {
                    long n = P->gcn(i, l, k);
                    H[n] += Q[n];
                }
    }
}

/* gpu_plasma.h:455-481 ElectricFieldEvaluate -> ElectricFieldTrace (306-320)
 * -> emeGPUIterate (249-284) -> GPU_eme / emeElement (kernels.cu:1091-1181);
 * starts/shifts from GetElectricFieldStartsDirs (286-304). */
void orc_e(orc_plasma *P) {
    double c1x = P->tau / P->hx, c1y = P->tau / P->hy, c1z = P->tau / P->hz;
    double tau = P->tau;
    struct T {
        int E, H1, H2, J;
        double c1, c2;
        int s[3];
        int d1[3], d2[3];
    } tr[3] = {
        {ORC_EX, ORC_HZ, ORC_HY, ORC_JX, c1y, c1z, {0, 1, 1}, {0, -1, 0}, {0, 0, -1}},
        {ORC_EY, ORC_HX, ORC_HZ, ORC_JY, c1z, c1x, {1, 0, 1}, {0, 0, -1}, {-1, 0, 0}},
        {ORC_EZ, ORC_HY, ORC_HX, ORC_JZ, c1x, c1y, {1, 1, 0}, {-1, 0, 0}, {0, -1, 0}},
    };
    for (auto &t : tr) {
        double *E = P->f[t.E].data();
        const double *H1 = P->f[t.H1].data(), *H2 = P->f[t.H2].data(), *J = P->f[t.J].data();
        for (int i = t.s[0]; i <= P->Nx; i++)
            for (int l = t.s[1]; l <= P->Ny; l++)
                for (int k = t.s[2]; k <= P->Nz; k++) {
                    long n = P->gcn(i, l, k);
                    long n1 = P->gcn(i + t.d1[0], l + t.d1[1], k + t.d1[2]);
                    long n2 = P->gcn(i + t.d2[0], l + t.d2[1], k + t.d2[2]);
                    E[n] += t.c1 * (H1[n] - H1[n1]) - t.c2 * (H2[n] - H2[n2]) - tau * J[n];
                }
    }
}

/* gpu_plasma.h:432-452 ElectricFieldComponentEvaluatePeriodic with the
 * arguments at 486-505 -> PeriodicBoundaries (802-870) -> GPU_periodic /
 * periodicElement (kernels.cu:1104-1122): Ex along z; Ey along x then z;
 * Ez along x.  Each call copies E[0]=E[N] then E[N+1]=E[1]. */
void orc_e_periodic(orc_plasma *P) {
    int Nx = P->Nx, Ny = P->Ny, Nz = P->Nz;
    auto bnd = [&](int a, int b, int dir, int N) -> long {
        int i1 = (dir == 0) * N + (dir == 1) * a + (dir == 2) * a;
        int l1 = (dir == 0) * a + (dir == 1) * N + (dir == 2) * b;
        int k1 = (dir == 0) * b + (dir == 1) * b + (dir == 2) * N;
        return P->gcn(i1, l1, k1);
    };
    auto periodic = [&](double *E, int dir, int s1, int e1, int s2, int e2, int N) {
        for (int a = s1; a <= e1; a++)
            for (int b = s2; b <= e2; b++) E[bnd(a, b, dir, 0)] = E[bnd(a, b, dir, N)];
        for (int a = s1; a <= e1; a++)
            for (int b = s2; b <= e2; b++) E[bnd(a, b, dir, N + 1)] = E[bnd(a, b, dir, 1)];
    };
    /* Ex (dir 0): only the second call: (2, 0,Nx, 0,Ny+1, Nz) */
    periodic(P->f[ORC_EX].data(), 2, 0, Nx, 0, Ny + 1, Nz);
    /* Ey (dir 1): (0, 0,Ny, 1,Nz, Nx) then (2, 0,Nx+1, 0,Ny, Nz) */
    periodic(P->f[ORC_EY].data(), 0, 0, Ny, 1, Nz, Nx);
    periodic(P->f[ORC_EY].data(), 2, 0, Nx + 1, 0, Ny, Nz);
    /* Ez (dir 2): only the first call: (0, 1,Ny, 0,Nz, Nx) */
    periodic(P->f[ORC_EZ].data(), 0, 1, Ny, 0, Nz, Nx);
}

/* gpu_plasma.h:621-657 getElectricEnergy -> GPU_getCellEnergy
 * (wrap_kernel.cu:14-36): sum over all ghosted nodes of Ex^2+Ey^2+Ez^2 */
double orc_energy(const orc_plasma *P) {
    const double *ex = P->f[ORC_EX].data(), *ey = P->f[ORC_EY].data(), *ez = P->f[ORC_EZ].data();
    double s = 0.0;
    for (long n = 0; n < P->G; n++) s += ex[n] * ex[n] + ey[n] * ey[n] + ez[n] * ez[n];
    return s;
}

/* gpu_plasma.h:605-620 Plasma::Step; sumMPI (mpi_shortcut.cxx:21-61) is the
 * identity at one rank and is left out. */
int orc_step(orc_plasma *P, int nt, double *energy) {
    (void)nt;
    orc_h_half(P);
    orc_push(P);
    orc_deposit(P);
    int st = orc_migrate(P);
    orc_fold_j(P);
    orc_h2(P);
    orc_e(P);
    orc_e_periodic(P);
    if (energy) *energy = orc_energy(P);
    return st;
}

long orc_total_particles(const orc_plasma *P) {
    long n = 0;
    for (auto &v : P->cells) n += (long)v.size();
    return n;
}

void orc_cell_counts(const orc_plasma *P, int *counts) {
    for (int i = 0; i < P->Nx; i++)
        for (int l = 0; l < P->Ny; l++)
            for (int k = 0; k < P->Nz; k++)
                counts[((long)i * P->Ny + l) * P->Nz + k] = (int)P->cells[P->gcn(i, l, k)].size();
}

long orc_get_particles(const orc_plasma *P, double *x, double *y, double *z, double *px,
                       double *py, double *pz, double *m, double *q_m, int *sort, int *cell,
                       long cap) {
    long j = 0;
    for (int i = 0; i < P->Nx; i++)
        for (int l = 0; l < P->Ny; l++)
            for (int k = 0; k < P->Nz; k++) {
                for (const auto &p : P->cells[P->gcn(i, l, k)]) {
                    if (j >= cap) return -1;
                    x[j] = p.x; y[j] = p.y; z[j] = p.z;
                    px[j] = p.pu; py[j] = p.pv; pz[j] = p.pw;
                    m[j] = p.m; q_m[j] = p.q_m;
                    sort[j] = p.sort;
                    cell[j] = (int)(((long)i * P->Ny + l) * P->Nz + k);
                    j++;
                }
            }
    return j;
}

void orc_moments(const orc_plasma *P, double *out) {
    std::memset(out, 0, sizeof(double) * 27);
    for (const auto &v : P->cells)
        for (const auto &p : v) {
            double *o = out + 9 * (p.sort < 0 ? 0 : (p.sort > 2 ? 2 : p.sort));
            double g = std::sqrt(1.0 + p.pu * p.pu + p.pv * p.pv + p.pw * p.pw);
            o[0] += 1.0;
            o[1] += p.x; o[2] += p.y; o[3] += p.z;
            o[4] += p.pu; o[5] += p.pv; o[6] += p.pw;
            o[7] += (p.m / p.q_m) * (g - 1.0);
            o[8] += std::fabs(p.pu) + std::fabs(p.pv) + std::fabs(p.pw);
        }
}

/* compare.cxx:9-36: fraction of entries with |a-b| < tol */
double orc_compare(const double *a, const double *b, int num, double tol) {
    double t = 0.0;
    for (int i = 0; i < num; i++)
        if (std::fabs(a[i] - b[i]) < tol) t += 1.0;
    if (num > 0) t /= num;
    return t;
}

/* service_functions.cu:108-126 CheckArraySilent: RMS difference */
double orc_rms(const double *a, const double *b, int size) {
    double diff = 0.0;
    for (int n = 0; n < size; n++) diff += std::pow(a[n] - b[n], 2.0);
    return std::pow(diff / (size), 0.5);
}

int orc_relay(const orc_plasma *P, int i, int l, int k, const double x[3], const double x1[3],
              double x2[3]) {
    const CellGeom c = geom(P, i, l, k);
    const D3 a = {x[0], x[1], x[2]}, b = {x1[0], x1[1], x1[2]};
    I3 i1, i2;
    D3 r = b;
    const int m = relay_point(c, a, b, i1, i2, r);
    x2[0] = m ? r.x : b.x;
    x2[1] = m ? r.y : b.y;
    x2[2] = m ? r.z : b.z;
    return m;
}

void orc_boris(const double E[3], const double H[3], double tau, double q_m, double x[3],
               double p[3], double x1[3]) {
    OParticle q;
    q.x = x[0]; q.y = x[1]; q.z = x[2];
    q.pu = p[0]; q.pv = p[1]; q.pw = p[2];
    q.q_m = q_m;
    q.m = 0;
    D3 e = {E[0], E[1], E[2]}, h = {H[0], H[1], H[2]};
    boris_move(q, e, h, tau);
    p[0] = q.pu; p[1] = q.pv; p[2] = q.pw;
    x1[0] = q.x1; x1[1] = q.y1; x1[2] = q.z1;
}

}  // extern "C"
