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
#include "pic_draw.h"

#include <algorithm>
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
double pic_rng_uniform(pic_rng *r) { return wh_uniform(r); }

/* maxwell.cu:77-144 (table-driven normal with carried state store1) */
double pic_rng_normal(pic_rng *r, double a, double b) { return wh_normal(r, a, b, nullptr); }

}  // extern "C"

namespace {

/* transverse Gaussian shape of cell (l,k), 1 on the box axis */
inline double profile_shape(const pic_params *g, const pic_species_spec *s, int l, int k) {
    double yc = (l + 0.5) - 0.5 * g->ny, zc = (k + 0.5) - 0.5 * g->nz;
    double sg = s->sigma_cells > 0 ? s->sigma_cells : 1.0;
    return std::exp(-(yc * yc + zc * zc) / (2.0 * sg * sg));
}

/* count on the axis of a profiled species: ppc is the mean over the cross
 * section, so the axis carries ppc * ny nz / sum(shape) */
inline double axis_count(const pic_params *g, const pic_species_spec *s) {
    double sum = 0.0;
    for (int l = 0; l < g->ny; l++)
        for (int k = 0; k < g->nz; k++) sum += profile_shape(g, s, l, k);
    return s->ppc * ((double)g->ny * g->nz) / sum;
}

/* per-(l,k) counts of one x-plane */
std::vector<int> plane_counts(const pic_params *g, const pic_species_spec *s) {
    std::vector<int> c((size_t)g->ny * g->nz, s->ppc);
    if (s->profile == 0) return c;
    const double peak = axis_count(g, s);
    for (int l = 0; l < g->ny; l++)
        for (int k = 0; k < g->nz; k++)
            c[(size_t)l * g->nz + k] = (int)std::floor(peak * profile_shape(g, s, l, k) + 0.5);
    return c;
}

}  // namespace

void pic_species_cell_counts(const pic_params *g, const pic_species_spec *s, int *counts_lk) {
    const std::vector<int> c = plane_counts(g, s);
    std::copy(c.begin(), c.end(), counts_lk);
}

double pic_species_weight(const pic_params *g, const pic_species_spec *s) {
    /* getMassCharge (maxwell.cu:606-619): weight = q n / ppc; a profiled
     * species carries its density n on the axis, with the axis count */
    const double per_cell = s->profile == 0 ? (double)s->ppc : axis_count(g, s);
    return s->charge * s->density / (per_cell > 0 ? per_cell : 1.0);
}

extern "C" {

long pic_species_count(const pic_params *g, const pic_species_spec *s) {
    long n = 0;
    for (int c : plane_counts(g, s)) n += c;
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
    if (m_out) *m_out = pic_species_weight(g, s);
    /* per-(l,k) counts and offsets of each x-plane */
    const std::vector<int> cnt_lk = plane_counts(g, s);
    std::vector<long> off((size_t)g->ny * g->nz + 1, 0);
    for (size_t lk = 0; lk < cnt_lk.size(); lk++) off[lk + 1] = off[lk] + cnt_lk[lk];
    long plane = off.back();
    auto gen_cell = [&](pic_rng *r, int i, int l, int k) {
        long base = (long)i * plane + off[(size_t)l * g->nz + k];
        int cnt = cnt_lk[(size_t)l * g->nz + k];
        double o[6];
        for (int j = 0; j < cnt; j++) {
            draw_particle(g->lx, s, r, i, l, k, hx, hy, hz, o, nullptr);
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
            pic_rng r;
            cell_stream(&r, seed, c, s->sort);
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
