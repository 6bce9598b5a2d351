/*
 * pic_init.cpp -- seeded initial conditions (host only, never on the step).
 *
 * Restates the reference's random generators from maxwell.cu:
 *   g05cae_  Wichmann-Hill uniform (maxwell.cu:147-179), seeds 1,255,25555
 *            (cag05a_, maxwell.cu:52);
 *   g05dde_  table-driven normal deviate with carried state store1
 *            (maxwell.cu:77-144; cag05b_ starts at store1 = 1, maxwell.cu:47);
 * and the velocity -> momentum conversion u/sqrt(1-u^2) of
 * InitUniformMaxwellianParticles (maxwell.cu:460-467, 521-531), with the
 * weight convention of getMassCharge (maxwell.cu:606-619): m = q n / ppc.
 */
#include "pic_amd.h"

#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

extern "C" {

void pic_rng_seed(pic_rng *r, int ix, int iy, int iz) {
    r->ix = ix;
    r->iy = iy;
    r->iz = iz;
    r->store1 = 1.0;
}

/* maxwell.cu:147-179: three multiplicative congruential generators
 * (171 mod 30269, 172 mod 30307, 170 mod 30323) by Schrage's method; the
 * uniform is the fractional part of the sum of the three ratios. */
double pic_rng_uniform(pic_rng *r) {
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

/* maxwell.cu:77-144: d[n] is the standard-normal quantile whose two-sided
 * tail mass is 2^-n (d[1] = 0.6745 = Phi^-1(0.75)).  The interval
 * [d[n-1], d[n]) is chosen from the carried uniform store1 by repeated
 * doubling, the deviate inside it by rejection, and the sign by the
 * leftover uniform, which is carried into store1 for the next call. */
double pic_rng_normal(pic_rng *r, double a, double b) {
    static const double d[41] = {
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
    double t = d[n - 1];
    u = pic_rng_uniform(r);
    double w, v;
    for (;;) {
        w = (d[n] - t) * u;
        v = w * (w * half + t);
        for (;;) {
            u = pic_rng_uniform(r);
            if (v <= u) goto accept;
            v = pic_rng_uniform(r);
            if (u > v) continue;
            break;
        }
        u = (v - u) / (one - u);
    }
accept:
    u = (u - v) / (one - v);
    if (u > half) {
        r->store1 = u + u - one;
        return a - b * (w + t);
    }
    r->store1 = u + u;
    return a + b * (w + t);
}

}  // extern "C"

namespace {

/* splitmix64 finaliser: maps (seed, cell) to per-cell generator seeds */
inline uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

/* transverse Gaussian occupancy of cell (l,k): peak count at the axis */
inline int cell_count(const pic_params *g, const pic_species_spec *s, int l, int k) {
    if (s->profile == 0) return s->ppc;
    double yc = (l + 0.5) - 0.5 * g->ny, zc = (k + 0.5) - 0.5 * g->nz;
    double sg = s->sigma_cells > 0 ? s->sigma_cells : 1.0;
    double f = std::exp(-(yc * yc + zc * zc) / (2.0 * sg * sg));
    return (int)std::floor(s->ppc * f + 0.5);
}

/* one particle: x,y,z uniform in the cell, then ux,uy,uz normal around the
 * drift; momentum u/sqrt(1-u^2) (maxwell.cu:521-531); x displaced by
 * sum_k a_k sin(2 pi k x / Lx) and wrapped into [0, Lx). */
inline void draw_particle(const pic_params *g, const pic_species_spec *s, pic_rng *r, int i,
                          int l, int k, double hx, double hy, double hz, double *o) {
    double x = ((double)i + pic_rng_uniform(r)) * hx;
    double y = ((double)l + pic_rng_uniform(r)) * hy;
    double z = ((double)k + pic_rng_uniform(r)) * hz;
    double ux = pic_rng_normal(r, s->drift[0], s->vth[0]);
    double uy = pic_rng_normal(r, s->drift[1], s->vth[1]);
    double uz = pic_rng_normal(r, s->drift[2], s->vth[2]);
    double d = 0.0;
    for (int q = 0; q < 8; q++)
        if (s->perturb[q] != 0.0) d += s->perturb[q] * std::sin(2.0 * 3.14159265358979323846 * (q + 1) * x / g->lx);
    if (d != 0.0) {
        x += d;
        if (x >= g->lx) x -= g->lx;
        if (x < 0.0) x += g->lx;
    }
    double vb0 = std::sqrt(1.0 - ux * ux - uy * uy - uz * uz);
    o[0] = x;
    o[1] = y;
    o[2] = z;
    o[3] = ux / vb0;
    o[4] = uy / vb0;
    o[5] = uz / vb0;
}

}  // namespace

extern "C" {

long pic_species_count(const pic_params *g, const pic_species_spec *s) {
    long n = 0;
    for (int l = 0; l < g->ny; l++)
        for (int k = 0; k < g->nz; k++) n += cell_count(g, s, l, k);
    return n * g->nx;
}

int pic_generate_species(const pic_params *g, const pic_species_spec *s, pic_rng *rng,
                         int stream_mode, unsigned long long seed, int nthreads, double *x,
                         double *y, double *z, double *px, double *py, double *pz, long cap,
                         long *n_out, double *m_out) {
    if (!g || !s || g->nx <= 0 || g->ny <= 0 || g->nz <= 0) return PIC_ERR_INVALID;
    long total = pic_species_count(g, s);
    if (n_out) *n_out = total;
    if (total > cap) return PIC_ERR_INVALID;
    double hx = g->lx / (double)g->nx, hy = g->ly / (double)g->ny, hz = g->lz / (double)g->nz;
    /* getMassCharge (maxwell.cu:606-619): weight = q n / ppc; for a profiled
     * species the peak density is carried by the peak count. */
    if (m_out) *m_out = s->charge * s->density / (double)(s->ppc > 0 ? s->ppc : 1);
    /* per-(l,k) counts and offsets of each x-plane */
    std::vector<long> off((size_t)g->ny * g->nz + 1, 0);
    for (int l = 0; l < g->ny; l++)
        for (int k = 0; k < g->nz; k++)
            off[(size_t)l * g->nz + k + 1] = off[(size_t)l * g->nz + k] + cell_count(g, s, l, k);
    long plane = off.back();
    auto gen_cell = [&](pic_rng *r, int i, int l, int k) {
        long base = (long)i * plane + off[(size_t)l * g->nz + k];
        int cnt = cell_count(g, s, l, k);
        double o[6];
        for (int j = 0; j < cnt; j++) {
            draw_particle(g, s, r, i, l, k, hx, hy, hz, o);
            x[base + j] = o[0]; y[base + j] = o[1]; z[base + j] = o[2];
            px[base + j] = o[3]; py[base + j] = o[4]; pz[base + j] = o[5];
        }
    };
    if (stream_mode == 0) {
        pic_rng local;
        pic_rng *r = rng;
        if (!r) {
            pic_rng_seed(&local, 1, 255, 25555);
            r = &local;
        }
        for (int i = 0; i < g->nx; i++)
            for (int l = 0; l < g->ny; l++)
                for (int k = 0; k < g->nz; k++) gen_cell(r, i, l, k);
        return PIC_SUCCESS;
    }
    long ncell = (long)g->nx * g->ny * g->nz;
    int nt = nthreads < 1 ? 1 : nthreads;
    auto worker = [&](int t) {
        for (long c = t; c < ncell; c += nt) {
            uint64_t h = mix64(seed ^ mix64((uint64_t)c * 0x100000001b3ULL + (uint64_t)s->sort));
            pic_rng r;
            pic_rng_seed(&r, 1 + (int)(h % 30268u), 1 + (int)((h >> 20) % 30306u),
                         1 + (int)((h >> 40) % 30322u));
            int i = (int)(c / ((long)g->ny * g->nz));
            int l = (int)((c / g->nz) % g->ny);
            int k = (int)(c % g->nz);
            gen_cell(&r, i, l, k);
        }
    };
    if (nt == 1) {
        worker(0);
    } else {
        std::vector<std::thread> th;
        for (int t = 0; t < nt; t++) th.emplace_back(worker, t);
        for (auto &t : th) t.join();
    }
    return PIC_SUCCESS;
}

}  // extern "C"

/* ---------------------------------------------------------------- params */
#include <cstdio>
#include <cstring>

extern "C" {

int pic_write_param_file(const char *path, const pic_param_file *p) {
    if (!path || !p) return PIC_ERR_INVALID;
    FILE *f = fopen(path, "wt");
    if (!f) return PIC_ERR_INVALID;
    /* maxwell.cu:330-377: value then description, one per line */
    fprintf(f, "%15.5e plasma electron temperature along X \n ", p->tex0);
    fprintf(f, "%15.5e plasma electron temperature along Y \n ", p->tey0);
    fprintf(f, "%15.5e plasma electron temperature along Z \n ", p->tez0);
    fprintf(f, "%15.5e beam impulse\n ", p->rimp);
    fprintf(f, "%15.5e beam velocity dispersion \n ", p->Tb);
    fprintf(f, "%15.5e beam and plasma density ratio \n ", p->rbd);
    fprintf(f, "%15.5e plasma density \n ", p->ni);
    fprintf(f, "%15.5e external magnetic field (along X) \n ", p->B0);
    fprintf(f, "%15.5e domain size X \n ", p->lx);
    fprintf(f, "%15.5e domain size Y \n ", p->ly);
    fprintf(f, "%15.5e domain size Z \n ", p->lz);
    fprintf(f, "%15.5e plasma size Y \n ", p->py);
    fprintf(f, "%15.5e plasma size Z \n ", p->pz);
    fprintf(f, "%15.5e beam size X \n ", p->bx);
    fprintf(f, "%15.5e beam size Y \n ", p->by);
    fprintf(f, "%15.5e beam size Z \n ", p->bz);
    fprintf(f, "%15d   average number of particles in cell \n ", p->lp);
    fprintf(f, "%15d   number of mesh nodes along X \n ", p->nx);
    fprintf(f, "%15d   number of mesh nodes along Y \n ", p->ny);
    fprintf(f, "%15d   number of mesh nodes along Z \n ", p->nz);
    fprintf(f, "%15.5e timestep \n ", p->tau);
    fprintf(f, "%15d   1 if beam-plasma interaction, 0 if beam-beam \n", p->beam_plasma);
    fprintf(f, "%15d   moment to start from saved\n ", p->start_from_file);
    fprintf(f, "%15d   phase to start from save\n ", p->phase);
    fprintf(f, "%15d   total steps \n ", p->ts);
    fprintf(f, "%15d   number of steps between diagnostic files \n ", p->ms);
    fclose(f);
    return PIC_SUCCESS;
}

int pic_read_param_file(const char *path, pic_param_file *p) {
    if (!path || !p) return PIC_ERR_INVALID;
    FILE *f = fopen(path, "rt");
    if (!f) return PIC_ERR_INVALID;
    double v[26];
    char line[512];
    int n = 0;
    while (n < 26 && fgets(line, sizeof(line), f)) {
        char *endp = nullptr;
        double d = strtod(line, &endp);
        if (endp == line) continue;   /* blank or text-only line */
        v[n++] = d;
    }
    fclose(f);
    if (n < 26) return PIC_ERR_INVALID;
    p->tex0 = v[0]; p->tey0 = v[1]; p->tez0 = v[2];
    p->rimp = v[3]; p->Tb = v[4]; p->rbd = v[5]; p->ni = v[6]; p->B0 = v[7];
    p->lx = v[8]; p->ly = v[9]; p->lz = v[10];
    p->py = v[11]; p->pz = v[12];
    p->bx = v[13]; p->by = v[14]; p->bz = v[15];
    p->lp = (int)v[16];
    p->nx = (int)v[17]; p->ny = (int)v[18]; p->nz = (int)v[19];
    p->tau = v[20];
    p->beam_plasma = (int)v[21]; p->start_from_file = (int)v[22]; p->phase = (int)v[23];
    p->ts = (int)v[24]; p->ms = (int)v[25];
    return PIC_SUCCESS;
}

int pic_species_from_param_file(const pic_param_file *pf, pic_params *grid,
                                pic_species_spec species[3]) {
    if (!pf || !grid || !species || pf->lp <= 0) return PIC_ERR_INVALID;
    grid->nx = pf->nx; grid->ny = pf->ny; grid->nz = pf->nz;
    grid->lx = pf->lx; grid->ly = pf->ly; grid->lz = pf->lz;
    grid->tau = pf->tau;
    grid->ni = pf->ni;
    grid->n_per_cell = pf->lp;
    grid->ion_q_m = 1.0 / 1836.0;
    std::memset(species, 0, 3 * sizeof(pic_species_spec));
    pic_species_spec &ion = species[0], &el = species[1], &bm = species[2];
    ion.sort = PIC_ION; ion.ppc = pf->lp; ion.density = pf->ni + pf->rbd; ion.charge = 1.0;
    ion.q_m = 1.0 / 1836.0;
    el.sort = PIC_PLASMA_ELECTRON; el.ppc = 2 * pf->lp; el.density = pf->ni; el.charge = -1.0;
    el.q_m = -1.0;
    el.vth[0] = pf->tex0; el.vth[1] = pf->tey0; el.vth[2] = pf->tez0;
    if (pf->beam_plasma == 1) el.drift[0] = -pf->rbd * pf->rimp;
    bm.sort = PIC_BEAM_ELECTRON; bm.ppc = pf->lp; bm.density = pf->rbd; bm.charge = -1.0;
    bm.q_m = -1.0;
    bm.drift[0] = pf->rimp;
    bm.vth[0] = bm.vth[1] = bm.vth[2] = pf->Tb * pf->rimp;
    return PIC_SUCCESS;
}

}  // extern "C"
