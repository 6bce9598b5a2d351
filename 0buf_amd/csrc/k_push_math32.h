/*
 * k_push_math32.h -- per-particle arithmetic of the FP32 precision form,
 * shared by the FP32 push (k_particles_f32.hip) and the FP32 particle
 * adjoint (k_adjoint32.hip), both built with -ffp-contract=fast: the adjoint
 * recomputes the forward's dual cells, crossing code and relay point from
 * the same expressions.
 *
 * Positions are offsets o in [0,1) inside the owning cell (cell units).
 * Node-type axes interpolate between faces i, i+1 with weights (1-o, o);
 * centre-type axes between the two nearest centres, rd = (o >= 1/2), near
 * weight rd ? 3/2 - o : 1/2 - o (rd is also the start's dual cell).
 */
#ifndef K_PUSH_MATH32_H
#define K_PUSH_MATH32_H

#include "pic_internal.h"

/* Trilinear interpolation from the corner node's address p = F + (u BY + v)
 * BZ + w, with the far weights x1, y1, z1 (node-type axes: the offset o;
 * centre-type axes: 1 - the near weight):
 *   each z pair  a = p0 + z1 (p1 - p0),  then  y pairs,  then x,
 * every step one fmaf.  tri_dz reads a box of (value, z difference
 * p[n+1] - p[n]) pairs (the push keeps it in LDS, so a z pair is one fmaf
 * and one 8-byte read); tri() forms the differences from F itself, with the
 * same FP32 subtraction, so both give the same value bit for bit. */
template <int BY, int BZ>
__device__ __forceinline__ float tri_dz(const float2 *p, float x1, float y1, float z1) {
    const float2 q0 = p[0], q1 = p[BZ], q2 = p[BY * BZ], q3 = p[BY * BZ + BZ];
    const float a = fmaf(z1, q0.y, q0.x);
    const float b = fmaf(z1, q1.y, q1.x);
    const float c = fmaf(z1, q2.y, q2.x);
    const float d = fmaf(z1, q3.y, q3.x);
    const float ab = fmaf(y1, b - a, a), cd = fmaf(y1, d - c, c);
    return fmaf(x1, cd - ab, ab);
}

template <int BY, int BZ>
__device__ __forceinline__ float tri(const float *F, int u, int v, int w, float x1, float y1,
                                     float z1) {
    const float *p = F + (u * BY + v) * BZ + w;
    const float a = fmaf(z1, p[1] - p[0], p[0]);
    const float b = fmaf(z1, p[BZ + 1] - p[BZ], p[BZ]);
    const float c = fmaf(z1, p[BY * BZ + 1] - p[BY * BZ], p[BY * BZ]);
    const float d = fmaf(z1, p[BY * BZ + BZ + 1] - p[BY * BZ + BZ], p[BY * BZ + BZ]);
    const float ab = fmaf(y1, b - a, a), cd = fmaf(y1, d - c, c);
    return fmaf(x1, cd - ab, ab);
}

struct Weights32 {
    int rd[3];        /* dual cell of the start (0/1 per axis) */
    float c0[3];      /* centre-type near weights */
    float n0[3];      /* node-type near weights 1 - o */
};

__device__ __forceinline__ Weights32 weights32(float ox, float oy, float oz) {
    Weights32 W;
    W.rd[0] = ox >= 0.5f;
    W.rd[1] = oy >= 0.5f;
    W.rd[2] = oz >= 0.5f;
    W.c0[0] = W.rd[0] ? 1.5f - ox : 0.5f - ox;
    W.c0[1] = W.rd[1] ? 1.5f - oy : 0.5f - oy;
    W.c0[2] = W.rd[2] ? 1.5f - oz : 0.5f - oz;
    W.n0[0] = 1.0f - ox;
    W.n0[1] = 1.0f - oy;
    W.n0[2] = 1.0f - oz;
    return W;
}

/* displacement in cell units, du = (tau/h) gi2 p, each product rounded.
 * Built with -ffp-contract=fast the backend fuses a multiply into a
 * following add whatever the source says, so du leaves through an empty asm
 * statement (no instruction), which it cannot see through: the end offset
 * o + du (end_offset) is then a plain add in every kernel that forms it. */
__device__ __forceinline__ float step_du(float tdh, float gi2, float p) {
    float du = (tdh * gi2) * p;
    asm volatile("" : "+v"(du));
    return du;
}

/* Boris push, Cayley form (particle.h:160-202), FP32, with the
 * intermediates of the adjoint; du = (tau/h) p1 / gamma1 in cell units */
struct Boris32 {
    float pm[3], ps, b[3], s[3], rden, p1[3], gi2, du[3];
};

__device__ __forceinline__ Boris32 boris32(float pu, float pv, float pw, float Ex, float Ey,
                                           float Ez, float Hx, float Hy, float Hz, float tau1,
                                           const float tdh[3]) {
    /* Every multiply-add is an explicit fmaf or a lone product: the particle
     * adjoint recomputes this in another kernel, and its decisions (dual
     * cells, crossing code, relay point) must see the same du bit for bit,
     * whatever the compiler would contract in either kernel. */
    Boris32 B;
    pu = fmaf(tau1, Ex, pu);
    pv = fmaf(tau1, Ey, pv);
    pw = fmaf(tau1, Ez, pw);
    B.pm[0] = pu;
    B.pm[1] = pv;
    B.pm[2] = pw;
    const float ps = tau1 * __frsqrt_rn(fmaf(pw, pw, fmaf(pv, pv, fmaf(pu, pu, 1.0f))));
    B.ps = ps;
    const float bx = ps * Hx, by = ps * Hy, bz = ps * Hz;
    B.b[0] = bx;
    B.b[1] = by;
    B.b[2] = bz;
    const float su = fmaf(-pw, by, fmaf(pv, bz, pu));
    const float sv = fmaf(-pu, bz, fmaf(pw, bx, pv));
    const float sw = fmaf(-pv, bx, fmaf(pu, by, pw));
    B.s[0] = su;
    B.s[1] = sv;
    B.s[2] = sw;
    /* the rotation matrix 1 + b b^T +- [b]x, each entry one fmaf */
    const float m11 = fmaf(bx, bx, 1.0f), m22 = fmaf(by, by, 1.0f), m33 = fmaf(bz, bz, 1.0f);
    const float m12 = fmaf(bx, by, bz), m21 = fmaf(bx, by, -bz);
    const float m23 = fmaf(by, bz, bx), m32 = fmaf(by, bz, -bx);
    const float m31 = fmaf(bz, bx, by), m13 = fmaf(bz, bx, -by);
    const float rden = __builtin_amdgcn_rcpf(fmaf(bz, bz, fmaf(by, by, m11)));
    B.rden = rden;
    pu = fmaf(fmaf(m13, sw, fmaf(m12, sv, m11 * su)), rden, tau1 * Ex);
    pv = fmaf(fmaf(m23, sw, fmaf(m22, sv, m21 * su)), rden, tau1 * Ey);
    pw = fmaf(fmaf(m33, sw, fmaf(m32, sv, m31 * su)), rden, tau1 * Ez);
    B.p1[0] = pu;
    B.p1[1] = pv;
    B.p1[2] = pw;
    const float gi2 = __frsqrt_rn(fmaf(pw, pw, fmaf(pv, pv, fmaf(pu, pu, 1.0f))));
    B.gi2 = gi2;
    B.du[0] = step_du(tdh[0], gi2, pu);
    B.du[1] = step_du(tdh[1], gi2, pv);
    B.du[2] = step_du(tdh[2], gi2, pw);
    return B;
}

/* the end offset o + du.  Every decision of the deposit and the migration
 * (end dual cell, crossing code, receiving cell) is taken from it, and the
 * particle adjoint recomputes them in another kernel: the products of du
 * (step_du) and this sum are rounded one by one in both. */
__device__ __forceinline__ float end_offset(float o, float du) { return o + du; }

/* deposit geometry (CurrentToMesh, cell.h:1306-1422, in cell units): dual
 * cell of each end (floor(o + 1/2)), crossing code m and the relay point on
 * the crossed dual faces (o = 1/2); for two crossed faces the third
 * coordinate follows the path at the parameter of the h^2-weighted
 * projection, as the reference's two-face cases */
struct Relay32 {
    int r1[3], m;
    float num, den, sr, x2[3];
};

__device__ __forceinline__ Relay32 relay32(float ox, float oy, float oz, float o1x, float o1y,
                                           float o1z, const float du[3], const int rd[3],
                                           const float hh[3]) {
    Relay32 R;
    R.r1[0] = o1x >= 0.5f;
    R.r1[1] = o1y >= 0.5f;
    R.r1[2] = o1z >= 0.5f;
    const bool cx = rd[0] != R.r1[0], cy = rd[1] != R.r1[1], cz = rd[2] != R.r1[2];
    const int m = 4 * cx + 2 * cy + cz;
    R.m = m;
    const float dux = du[0], duy = du[1], duz = du[2];
    /* one formula for one and two crossed faces, without branches: the
     * h^2-weighted projection of the crossed faces' parameters; with one
     * crossed face d it is (1/2 - o_d) / du_d exactly in exact arithmetic.
     * No face (m = 0) or three (m = 7, every coordinate of x2 is 1/2) leave
     * sr unused. */
    const float wx = cx ? hh[0] : 0.0f, wy = cy ? hh[1] : 0.0f, wz = cz ? hh[2] : 0.0f;
    const bool used = m != 0 && m != 7;
    /* (explicit fmaf, as in boris32: the adjoint recomputes these) */
    /* both formed unconditionally and then selected: as conditional
     * expressions the compiler branched around them, twice */
    const float num_all = fmaf((0.5f - ox) * dux, wx,
                               fmaf((0.5f - oy) * duy, wy, ((0.5f - oz) * duz) * wz));
    const float den_all = fmaf(dux * dux, wx, fmaf(duy * duy, wy, (duz * duz) * wz));
    const float num = used ? num_all : 0.0f;
    const float den = used ? den_all : 1.0f;
    R.num = num;
    R.den = den;
    const float sr = num * __builtin_amdgcn_rcpf(den);
    R.sr = sr;
    R.x2[0] = (m & 4) ? 0.5f : fmaf(dux, sr, ox);
    R.x2[1] = (m & 2) ? 0.5f : fmaf(duy, sr, oy);
    R.x2[2] = (m & 1) ? 0.5f : fmaf(duz, sr, oz);
    return R;
}

#endif
