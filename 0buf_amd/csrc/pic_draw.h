/*
 * pic_draw.h -- the seeded draws of one particle, shared by the host
 * generator (pic_init.cpp, g++) and the device initializer (k_init.hip):
 * the same source computes the same bits on both sides (integer Wichmann-
 * Hill state, IEEE-rounded divisions and square roots).
 *
 *   wh_uniform   g05cae_ (maxwell.cu:147-179)
 *   wh_normal    g05dde_ (maxwell.cu:77-144), loops bounded so that every
 *                device thread finishes (the bound is never reached by the
 *                generator's sequences; reaching it sets *fail)
 *   mix64        the per-cell stream seeding of stream_mode 1
 *   draw_particle  x, y, z uniform in the cell, ux, uy, uz normal around
 *                the drift, displacement sum_k a_k sin(2 pi k x / Lx),
 *                momentum u / sqrt(1 - u.u) (maxwell.cu:460-467, 521-531)
 */
#ifndef PIC_DRAW_H
#define PIC_DRAW_H

#include <cmath>
#include <cstdint>

#include "pic_amd.h"

#ifdef __HIPCC__
#define PIC_HD __host__ __device__
#else
#define PIC_HD
#endif

#define PIC_NORMAL_MAX_ITER 100000

PIC_HD inline double wh_uniform(pic_rng *r) {
    r->ix = (r->ix - r->ix / 177 * 177) * 171 - (r->ix / 177 << 1);
    r->iy = (r->iy - r->iy / 176 * 176) * 172 - (r->iy / 176 << 1);
    r->iz = (r->iz - r->iz / 178 * 178) * 170 - (r->iz / 178 << 1);
    if (r->ix < 0) r->ix += 30269;
    if (r->iy < 0) r->iy += 30307;
    if (r->iz < 0) r->iz += 30323;
    double ai = (double)r->ix / 30269. + (double)r->iy / 30307. + (double)r->iz / 30323.;
    int ii = (int)ai;
    return ai - ii;
}

/* d[n]: the standard-normal quantile whose two-sided tail mass is 2^-n
 * (d[1] = 0.6745 = Phi^-1(0.75)).  The interval [d[n-1], d[n]) is chosen
 * from the carried uniform store1 by repeated doubling, the deviate inside
 * it by rejection, and the sign by the leftover uniform, which is carried
 * into store1 for the next call. */
PIC_HD inline double wh_normal(pic_rng *r, double a, double b, int *fail) {
    const double d[41] = {
        0., .674489750196082, 1.150349380376008, 1.534120544352546, 1.862731867421652,
        2.153874694061456, 2.417559016236505, 2.66006746861746, 2.885634912426757,
        3.097269078198785, 3.297193345691964, 3.487104104114431, 3.668329285121323,
        3.841930685501911, 4.008772594168585, 4.169569323349106, 4.324919040826046,
        4.475328424654204, 4.621231001499247, 4.763001034267814, 4.900964207963193,
        5.035405969463927, 5.166578119728753, 5.294704084854598, 5.419983174916868,
        5.54259405780294, 5.662697617459439, 5.780439324478935, 5.895951216739571,
        6.009353565530745, 6.120756285971941, 6.230260137989044, 6.33795775455379,
        6.443934526538564, 6.548269367831731, 6.651035379893011, 6.752300431407015,
        6.852127665896068, 6.95057594791675, 7.047700256664409, 7.14355203435219};
    const double half = .5, one = 1.;
    double u = r->store1;
    int n;
    for (n = 1; n <= 39; ++n) {
        if (u > half) break;
        u += u;
    }
    if (n > 39) n = 40;
    const double t = d[n - 1];
    u = wh_uniform(r);
    double w = 0.0, v = 0.0;
    for (int it = 0;; ++it) {
        w = (d[n] - t) * u;
        v = w * (w * half + t);
        int outcome = 0;   /* 1 accept, 2 next interval draw */
        for (int jt = 0; jt < PIC_NORMAL_MAX_ITER; ++jt) {
            u = wh_uniform(r);
            if (v <= u) {
                outcome = 1;
                break;
            }
            v = wh_uniform(r);
            if (u > v) continue;
            outcome = 2;
            break;
        }
        if (outcome == 1) break;
        if (outcome == 0 || it >= PIC_NORMAL_MAX_ITER) {
            if (fail) *fail = 1;
            return a;
        }
        u = (v - u) / (one - u);
    }
    u = (u - v) / (one - v);
    if (u > half) {
        r->store1 = u + u - one;
        return a - b * (w + t);
    }
    r->store1 = u + u;
    return a + b * (w + t);
}

/* splitmix64 finaliser: maps (seed, cell) to per-cell generator seeds */
PIC_HD inline uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

/* the stream of cell c (flat (i*ny + l)*nz + k) and sort in stream_mode 1 */
PIC_HD inline void cell_stream(pic_rng *r, unsigned long long seed, long c, int sort) {
    const uint64_t h = mix64(seed ^ mix64((uint64_t)c * 0x100000001b3ULL + (uint64_t)sort));
    r->ix = 1 + (int)(h % 30268u);
    r->iy = 1 + (int)((h >> 20) % 30306u);
    r->iz = 1 + (int)((h >> 40) % 30322u);
    r->store1 = 1.0;
}

PIC_HD inline void draw_particle(double lx, const pic_species_spec *s, pic_rng *r, int i, int l,
                                 int k, double hx, double hy, double hz, double *o, int *fail) {
    double x = ((double)i + wh_uniform(r)) * hx;
    double y = ((double)l + wh_uniform(r)) * hy;
    double z = ((double)k + wh_uniform(r)) * hz;
    double ux = wh_normal(r, s->drift[0], s->vth[0], fail);
    double uy = wh_normal(r, s->drift[1], s->vth[1], fail);
    double uz = wh_normal(r, s->drift[2], s->vth[2], fail);
    double d = 0.0;
    for (int q = 0; q < 8; q++)
        if (s->perturb[q] != 0.0)
            d += s->perturb[q] * sin(2.0 * 3.14159265358979323846 * (q + 1) * x / lx);
    if (d != 0.0) {
        x += d;
        if (x >= lx) x -= lx;
        if (x < 0.0) x += lx;
    }
    double vb0 = sqrt(1.0 - ux * ux - uy * uy - uz * uz);
    o[0] = x;
    o[1] = y;
    o[2] = z;
    o[3] = ux / vb0;
    o[4] = uy / vb0;
    o[5] = uz / vb0;
}

/* per-(l,k) particle count of a species; host only.  Profile 1 (transverse
 * Gaussian): floor(A exp(-r^2 / 2 sigma^2) + 1/2), r from the box axis in
 * cells, with the axis count A = ppc ny nz / sum exp(...) so that ppc is the
 * mean count per cell */
void pic_species_cell_counts(const pic_params *g, const pic_species_spec *s, int *counts_lk);
/* particle weight m = q n / (count per cell), the axis count for profile 1
 * (the density n is the axis density); host only */
double pic_species_weight(const pic_params *g, const pic_species_spec *s);

#endif
