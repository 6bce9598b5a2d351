/*
 * k_push_math.h -- per-particle arithmetic of the FP64 step, shared by the
 * push kernel (k_particles_f64.hip) and the particle adjoint
 * (k_adjoint.hip).  The adjoint recomputes the forward step from a snapshot
 * with exactly these functions, so its gather indices, crossing code and
 * relay points are the forward's bit for bit.
 *
 * Both files are built with -ffp-contract=off: every expression keeps the
 * reference's rounding sequence (cell.h, particle.h).
 */
#ifndef K_PUSH_MATH_H
#define K_PUSH_MATH_H

#include "pic_internal.h"

/* Divisors that are constant over a run, with their correctly rounded
 * reciprocals: q = divc(x, d, r) returns the correctly rounded x/d in three
 * instructions (Markstein: q0 = x r, exact remainder x - q0 d by FMA, one
 * FMA correction), bit-identical to IEEE division. */
struct DivConst {
    double rhx, rhy, rhz;     /* 1/hx, 1/hy, 1/hz */
    double r12;               /* 1/12 */
    double Dx, Dy, Dz;        /* (tau hy) hz, (tau hx) hz, (tau hx) hy (cell.h:1578-1580) */
    double rDx, rDy, rDz;
};

inline DivConst make_divconst(const Grid &g) {
    DivConst d;
    d.rhx = 1.0 / g.hx;
    d.rhy = 1.0 / g.hy;
    d.rhz = 1.0 / g.hz;
    d.r12 = 1.0 / 12.0;
    d.Dx = g.tau * g.hy * g.hz;
    d.Dy = g.tau * g.hx * g.hz;
    d.Dz = g.tau * g.hx * g.hy;
    d.rDx = 1.0 / d.Dx;
    d.rDy = 1.0 / d.Dy;
    d.rDz = 1.0 / d.Dz;
    return d;
}

/* pow(v, -0.5) (particle.h:73,116) correctly rounded: r = 1/sqrt(v) refined
 * by one Newton step whose residual 1 - v r^2 is formed exactly with FMAs. */
__device__ __forceinline__ double pow_m05(double v) {
    double s = sqrt(v);
    double r = 1.0 / s;
    double rr = r * r;
    double rr_lo = fma(r, r, -rr);
    double e = fma(-v, rr, 1.0);
    e = fma(-v, rr_lo, e);
    return fma(0.5 * r, e, r);
}

__device__ __forceinline__ double divc(double x, double d, double r) {
    double q0 = x * r;
    double e = fma(-q0, d, x);
    return fma(e, r, q0);
}

/* cell.h:1427-1478 Reflect: (x>xm)(x-xm) + (x<0)(xm+x) + (0<x<xm) x.  The
 * select form returns the same bits in every case, including exact 0 and xm
 * (both map to +0) and the signed zeros of the products. */
__device__ __forceinline__ double reflect1(double x, double xm) {
    return x > xm ? x - xm : (x < 0.0 ? xm + x : ((x > 0 && x < xm) ? x : 0.0));
}

template <int BY, int BZ>
__device__ __forceinline__ double interp(const double *F, int u, int v, int w, double sx,
                                         double sx1, double sy, double sy1, double sz,
                                         double sz1) {
    /* cell.h:1054-1153 Interpolate3D: t2 + t1 */
    const double *p = F + (u * BY + v) * BZ + w;
    double t_ilk = p[0], t_ilk1 = p[1], t_il1k = p[BZ], t_il1k1 = p[BZ + 1];
    double t_i1lk = p[BY * BZ], t_i1lk1 = p[BY * BZ + 1];
    double t_i1l1k = p[BY * BZ + BZ], t_i1l1k1 = p[BY * BZ + BZ + 1];
    double t1 = sx * (sy * (sz * t_ilk + sz1 * t_ilk1) + sy1 * (sz * t_il1k + sz1 * t_il1k1));
    double t2 = sx1 * (sy * (sz * t_i1lk + sz1 * t_i1lk1) + sy1 * (sz * t_i1l1k + sz1 * t_i1l1k1));
    return t2 + t1;
}

struct Geo {
    double hx, hy, hz, x0, y0, z0, tau;
    int ca, cb, cc;   /* owner cell inside the tile */
};

/* GetField (cell.h:1257-1290): tile indices from (x - x0)/h
 * (getCellNumber, getCellNumberCenter: cell.h:527-549), weights from the
 * global x/h (s1..s6_interpolate: cell.h:1193-1236). */
struct GatherIdx {
    int ix, i1x, iy, i1y, iz, i1z;
    int f1[3], f15[3];   /* the weight floors (int)(x/h + 1), (int)(x/h + 1.5) per axis */
    double s1, s2, s3, s4, s5, s6;
    bool ok;          /* every index inside the tile stencil (2..3) */
};

/* the weights of GetField from given floors f1, f15 per axis (the same
 * arithmetic as gather_index): s = f1 - x/h, s' = f15 - 0.5 - x/h */
__device__ __forceinline__ void gather_weights(GatherIdx &G, double xh, double yh, double zh) {
    G.s1 = G.f1[0] - xh;
    G.s2 = G.f15[0] - 0.5 - xh;
    G.s3 = G.f1[1] - yh;
    G.s4 = G.f15[1] - 0.5 - yh;
    G.s5 = G.f1[2] - zh;
    G.s6 = G.f15[2] - 0.5 - zh;
}

__device__ __forceinline__ bool gather_ok(const GatherIdx &G) {
    return !(G.ix < 2 || G.ix > 3 || G.i1x < 2 || G.i1x > 3 || G.iy < 2 || G.iy > 3 ||
             G.i1y < 2 || G.i1y > 3 || G.iz < 2 || G.iz > 3 || G.i1z < 2 || G.i1z > 3);
}

__device__ __forceinline__ GatherIdx gather_index(double x, double y, double z, double frx,
                                                  double fry, double frz, const Grid &g,
                                                  const DivConst &dc) {
    GatherIdx G;
    const double xh = divc(x, g.hx, dc.rhx), yh = divc(y, g.hy, dc.rhy),
                 zh = divc(z, g.hz, dc.rhz);
    G.ix = (int)(frx + 1.0) + 1;
    G.i1x = (int)((frx + 1.0) + 1.5);
    G.iy = (int)(fry + 1.0) + 1;
    G.i1y = (int)((fry + 1.0) + 1.5);
    G.iz = (int)(frz + 1.0) + 1;
    G.i1z = (int)((frz + 1.0) + 1.5);
    G.f1[0] = (int)(xh + 1.0);
    G.f15[0] = (int)(xh + 1.5);
    G.f1[1] = (int)(yh + 1.0);
    G.f15[1] = (int)(yh + 1.5);
    G.f1[2] = (int)(zh + 1.0);
    G.f15[2] = (int)(zh + 1.5);
    gather_weights(G, xh, yh, zh);
    G.ok = gather_ok(G);
    return G;
}

/* Particle::Move (particle.h:160-202) with every intermediate the adjoint
 * needs: pm = p + tau1 E, g1 = pow(pm.pm+1,-1/2), b = tau1 g1 H,
 * s = pm + pm x b, pp = Cayley rotation of s, p1 = pp + tau1 E,
 * g2 = pow(p1.p1+1,-1/2), x1 = x + tau g2 p1. */
struct Boris {
    double pm[3], ps, b[3], s[3], sden, p1[3], g2;
};

__device__ __forceinline__ Boris boris_move(double pu, double pv, double pw, double Ex, double Ey,
                                            double Ez, double Hx, double Hy, double Hz,
                                            double tau1) {
    Boris B;
    pu += tau1 * Ex;
    pv += tau1 * Ey;
    pw += tau1 * Ez;
    B.pm[0] = pu;
    B.pm[1] = pv;
    B.pm[2] = pw;
    double ps = tau1 * pow_m05((pu * pu + pv * pv + pw * pw) * 1. + 1.0);
    B.ps = ps;
    double bx = ps * Hx, by = ps * Hy, bz = ps * Hz;
    B.b[0] = bx;
    B.b[1] = by;
    B.b[2] = bz;
    double su = pu + pv * bz - pw * by;
    double sv = pv + pw * bx - pu * bz;
    double sw = pw + pu * by - pv * bx;
    B.s[0] = su;
    B.s[1] = sv;
    B.s[2] = sw;
    double q1 = bx * bx, q2 = by * by, q3 = bz * bz;
    double q4 = bx * by, q5 = by * bz, q6 = bz * bx;
    double sden = q1 + 1. + q2 + q3;
    B.sden = sden;
    const double rden = 1.0 / sden;   /* one reciprocal for the three quotients */
    double pu1 = divc((q1 + 1.) * su + (q4 + bz) * sv + (q6 - by) * sw, sden, rden);
    double pv1 = divc((q4 - bz) * su + (q2 + 1.) * sv + (q5 + bx) * sw, sden, rden);
    double pw1 = divc((q6 + by) * su + (q5 - bx) * sv + (q3 + 1.) * sw, sden, rden);
    pu = pu1 + Ex * tau1;
    pv = pv1 + Ey * tau1;
    pw = pw1 + Ez * tau1;
    B.p1[0] = pu;
    B.p1[1] = pv;
    B.p1[2] = pw;
    B.g2 = pow_m05(((pu * pu + pv * pv + pw * pw) + 1.0));
    return B;
}

__device__ __forceinline__ double transit_average(double h, int i1, int i2, double x0) {
    return (h * (((double)i1 + (double)i2) * 0.5 - 2.0) + x0);
}

/* CurrentToMesh (cell.h:1306-1422): tile indices of both ends
 * (getCellNumberCenterCurrent, cell.h:560-566), crossing code
 * m = 4|di| + 2|dl| + |dk|, relay point x2 on the crossed dual-cell faces
 * (getCellTransitAverage, cell.h:600-602); the uncrossed coordinates follow
 * the straight path at parameter sr = num/den (getRatioBasedX, cell.h:617). */
struct Relay {
    int i1x, i1y, i1z, i2x, i2y, i2z, m;
    double fx, fy, fz, num, den, sr;
    double x2x, x2y, x2z;
};

/* the relay geometry from given end cells i1 (start) and i2 (end) */
__device__ __forceinline__ Relay relay_from(double x, double y, double z, double x1, double y1,
                                            double z1, const Geo &G, int i1x, int i1y, int i1z,
                                            int i2x, int i2y, int i2z) {
    Relay R;
    R.i1x = i1x;
    R.i1y = i1y;
    R.i1z = i1z;
    R.i2x = i2x;
    R.i2y = i2y;
    R.i2z = i2z;
    int m = 4 * abs(R.i2x - R.i1x) + 2 * abs(R.i2y - R.i1y) + abs(R.i2z - R.i1z);
    if (m > 7) m = 0;
    R.m = m;
    const double dxs = x1 - x, dys = y1 - y, dzs = z1 - z;
    R.fx = transit_average(G.hx, R.i1x, R.i2x, G.x0);
    R.fy = transit_average(G.hy, R.i1y, R.i2y, G.y0);
    R.fz = transit_average(G.hz, R.i1z, R.i2z, G.z0);
    double num = 0.0, den = 1.0;
    switch (m) {
    case 1: num = R.fz - z; den = dzs; break;
    case 2: num = R.fy - y; den = dys; break;
    case 4: num = R.fx - x; den = dxs; break;
    case 3:
        num = (R.fz - z) * dzs + (R.fy - y) * dys;
        den = (z1 - z) * (z1 - z) + (y1 - y) * (y1 - y);
        break;
    case 5:
        num = (R.fz - z) * dzs + (R.fx - x) * dxs;
        den = (z1 - z) * (z1 - z) + (x1 - x) * (x1 - x);
        break;
    case 6:
        num = (R.fy - y) * dys + (R.fx - x) * dxs;
        den = (y1 - y) * (y1 - y) + (x1 - x) * (x1 - x);
        break;
    default: break;
    }
    R.num = num;
    R.den = den;
    R.sr = num / den;
    R.x2x = (m & 4) ? R.fx : (x + dxs * R.sr);
    R.x2y = (m & 2) ? R.fy : (y + dys * R.sr);
    R.x2z = (m & 1) ? R.fz : (z + dzs * R.sr);
    return R;
}

__device__ __forceinline__ Relay relay_point(double x, double y, double z, double x1, double y1,
                                             double z1, double frx, double fry, double frz,
                                             const Geo &G, const DivConst &dc) {
    const double t1x = divc(x1 - G.x0, G.hx, dc.rhx);
    const double t1y = divc(y1 - G.y0, G.hy, dc.rhy);
    const double t1z = divc(z1 - G.z0, G.hz, dc.rhz);
    return relay_from(x, y, z, x1, y1, z1, G, (int)((frx + 1.5) + 1), (int)((fry + 1.5) + 1),
                      (int)((frz + 1.5) + 1), (int)((t1x + 1.5) + 1), (int)((t1y + 1.5) + 1),
                      (int)((t1z + 1.5) + 1));
}

#endif
