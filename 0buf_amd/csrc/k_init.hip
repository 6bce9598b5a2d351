/*
 * k_init.hip -- device-side seeded initializer: the stream_mode-1 draws of
 * pic_generate_species (one Wichmann-Hill stream per (cell, sort), the
 * reference's g05cae_/g05dde_ from maxwell.cu:77-179) run one lane per cell
 * and write each particle straight into its cell's bucket, behind the
 * particles of the species generated before (the bucket order of the host
 * path: species in load order, then draw order).
 *
 * Every particle is re-checked with the load-time predicates of
 * pic_load_particles (getPointCell, cell.h:593-598, and isPointInCell,
 * cell.h:717); one that would land elsewhere is counted in *misplaced and
 * not stored, and the host reports the whole call as failed.
 */
#include "pic_draw.h"
#include "pic_internal.h"

#pragma clang fp contract(off)

namespace {

struct InitArgs {
    Grid g;
    pic_species_spec spec;
    const int *cnt_lk;        /* per (l,k) count of this species */
    const long *off_lk;       /* exclusive prefix of cnt_lk */
    long plane;               /* particles of this species per x-plane */
    unsigned long long seed;
    unsigned char slot;
    long id_base;             /* id of this species' first particle (load order) */
    unsigned long long *misplaced;
    int *fail;
};

template <class P>
__global__ __launch_bounds__(PIC_BLOCK) void k_generate(InitArgs a, P p) {
    const long c = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    const Grid &g = a.g;
    if (c >= (long)g.Nx * g.Ny * g.Nz) return;
    const int i = (int)(c / ((long)g.Ny * g.Nz));
    const int l = (int)((c / g.Nz) % g.Ny);
    const int k = (int)(c % g.Nz);
    const int lk = l * g.Nz + k;
    const int cnt = a.cnt_lk[lk];
    const int start = p.count[c];
    pic_rng r;
    cell_stream(&r, a.seed, c, a.spec.sort);
    int fail = 0, stored = 0;
    unsigned long long bad = 0;
    const double x0 = (double)i * g.hx, y0 = (double)l * g.hy, z0 = (double)k * g.hz;
    for (int j = 0; j < cnt; j++) {
        double o[6];
        draw_particle(g.Lx, &a.spec, &r, i, l, k, g.hx, g.hy, g.hz, o, &fail);
        /* pic_load_particles' cell choice and acceptance */
        const int ci = (int)((o[0] - 0.0) / g.hx), cl = (int)((o[1] - 0.0) / g.hy),
                  ck = (int)((o[2] - 0.0) / g.hz);
        const bool here = ci == i && cl == l && ck == k && x0 <= o[0] && o[0] < x0 + g.hx &&
                          y0 <= o[1] && o[1] < y0 + g.hy && z0 <= o[2] && o[2] < z0 + g.hz;
        if (!here || start + stored >= p.cap) {
            bad++;
            continue;
        }
        const long e = c * p.cap + start + stored;
        if constexpr (sizeof(*p.p[0]) == 4) {
            /* FP32 form: offsets inside the cell, kept below 1 */
            const double h[3] = {g.hx, g.hy, g.hz}, b[3] = {x0, y0, z0};
            for (int d = 0; d < 3; d++) {
                float of = (float)((o[d] - b[d]) / h[d]);
                if (of >= 1.0f) of = 0.99999994f;
                if (of < 0.0f) of = 0.0f;
                p.p[d][e] = of;
            }
            for (int d = 3; d < 6; d++) p.p[d][e] = (float)o[d];
        } else {
            for (int d = 0; d < 6; d++) p.p[d][e] = o[d];
        }
        p.sp[e] = a.slot;
        stored++;
    }
    p.count[c] = start + stored;
    if (bad) atomicAdd(a.misplaced, bad);
    if (fail) atomicOr(a.fail, 1);
}

/* differentiable form: ids in load order (species-major, then the host
 * generator's cell order i-major with per-(l,k) offsets, then draw order) */
template <class P>
__global__ __launch_bounds__(PIC_BLOCK) void k_generate_ids(InitArgs a, P p,
                                                            const int *start_of_cell) {
    const long c = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    const Grid &g = a.g;
    if (c >= (long)g.Nx * g.Ny * g.Nz) return;
    const int i = (int)(c / ((long)g.Ny * g.Nz));
    const int lk = (int)(c % ((long)g.Ny * g.Nz));
    const int start = start_of_cell[c];
    const int n = p.count[c] - start;
    const long base = a.id_base + (long)i * a.plane + a.off_lk[lk];
    for (int j = 0; j < n; j++) p.id[c * p.cap + start + j] = (int)(base + j);
}

inline unsigned blocks_for(long n) { return (unsigned)((n + PIC_BLOCK - 1) / PIC_BLOCK); }

}  // namespace

void launch_generate_species(const Grid &g, const pic_species_spec &spec, const int *cnt_lk,
                             const long *off_lk, long plane, unsigned long long seed,
                             unsigned char slot, long id_base, PartPtrs *p64, PartPtrs32 *p32,
                             int *start_of_cell, unsigned long long *misplaced, int *fail,
                             hipStream_t s) {
    InitArgs a;
    a.g = g;
    a.spec = spec;
    a.cnt_lk = cnt_lk;
    a.off_lk = off_lk;
    a.plane = plane;
    a.seed = seed;
    a.slot = slot;
    a.id_base = id_base;
    a.misplaced = misplaced;
    a.fail = fail;
    const long ncell = (long)g.Nx * g.Ny * g.Nz;
    if (p64) {
        if (p64->id)
            (void)hipMemcpyAsync(start_of_cell, p64->count, ncell * sizeof(int),
                                 hipMemcpyDeviceToDevice, s);
        hipLaunchKernelGGL(k_generate<PartPtrs>, dim3(blocks_for(ncell)), dim3(PIC_BLOCK), 0, s,
                           a, *p64);
        if (p64->id)
            hipLaunchKernelGGL(k_generate_ids<PartPtrs>, dim3(blocks_for(ncell)), dim3(PIC_BLOCK),
                               0, s, a, *p64, (const int *)start_of_cell);
    } else {
        if (p32->id)
            (void)hipMemcpyAsync(start_of_cell, p32->count, ncell * sizeof(int),
                                 hipMemcpyDeviceToDevice, s);
        hipLaunchKernelGGL(k_generate<PartPtrs32>, dim3(blocks_for(ncell)), dim3(PIC_BLOCK), 0,
                           s, a, *p32);
        if (p32->id)
            hipLaunchKernelGGL(k_generate_ids<PartPtrs32>, dim3(blocks_for(ncell)),
                               dim3(PIC_BLOCK), 0, s, a, *p32, (const int *)start_of_cell);
    }
}
