/*
 * k_arrive.h -- arrival of the departed particles, both precisions: the
 * receiving side of the reference's migration (GPU_ArrangeFlights /
 * Cell::Insert, kernels.cu:316-370), where each leaver joins the bucket of
 * the cell that accepts it.
 *
 * One workgroup per source tile.  Its queue entries carry the destination
 * cell twice: as a bulk cell index (dest) and as an index into the tile's
 * halo box of (TX+2) x (TY+2) x (TZ+2) cells (halo).  Entries bound for the
 * same cell are counted in LDS first, so each destination cell costs one
 * global atomic per source tile instead of one per particle (the beam axis of
 * a skewed workload sends tens of particles per step into the same cells);
 * then every entry takes base + rank as its slot.  A slot past the bucket
 * capacity is kept in q.slot for the regrow path and raises
 * PIC_F_BUCKET_OVERFLOW.  An entry whose halo index or destination cell lies
 * out of range is not placed (slot -1) and raises PIC_F_QUEUE_CORRUPT.
 */
#ifndef K_ARRIVE_H
#define K_ARRIVE_H

#include "k_queue_key.h"
#include "pic_internal.h"

template <int TX, int TY, int TZ, class P, class Q>
__global__ __launch_bounds__(PIC_BLOCK) void k_arrive_tile(Grid g, TileGeom t, P p, Q q,
                                                           DevStatus *st) {
    constexpr int BY = TY + 2, BZ = TZ + 2, NH = (TX + 2) * BY * BZ;
    __shared__ int hcnt[NH];
    __shared__ int hbase[NH];
    const long tile = blockIdx.x;
    const int tid = threadIdx.x;
    const int tz = (int)(tile % t.ntz);
    const int ty = (int)((tile / t.ntz) % t.nty);
    const int tx = (int)(tile / ((long)t.ntz * t.nty));
    int n = q.count[tile];
    n = n < q.qcap ? n : q.qcap;
    const long qb = tile * (long)q.qcap;
    for (int h = tid; h < NH; h += PIC_BLOCK) hcnt[h] = 0;
    __syncthreads();
    /* rank of each entry among the tile's entries for the same cell; an
     * entry whose halo index or cell lies out of range is refused (slot -1) */
    const long ncell = (long)g.Nx * g.Ny * g.Nz;
    bool bad = false;
    for (int j = tid; j < n; j += PIC_BLOCK) {
        const int h = q.halo[qb + j];
        if (!arrival_entry_ok(h, NH, (long)q.dest[qb + j], ncell)) {
            q.slot[qb + j] = -1;
            bad = true;
            continue;
        }
        q.slot[qb + j] = atomicAdd(&hcnt[h], 1);
    }
    __syncthreads();
    /* one claim per destination cell on its bucket count */
    for (int h = tid; h < NH; h += PIC_BLOCK) {
        const int c = hcnt[h];
        if (c == 0) continue;
        int gx = tx * TX - 1 + h / (BY * BZ), gy = ty * TY - 1 + (h / BZ) % BY,
            gz = tz * TZ - 1 + h % BZ;
        if (gx < 0) gx += g.Nx; else if (gx >= g.Nx) gx -= g.Nx;
        if (gy < 0) gy += g.Ny; else if (gy >= g.Ny) gy -= g.Ny;
        if (gz < 0) gz += g.Nz; else if (gz >= g.Nz) gz -= g.Nz;
        hbase[h] = atomicAdd(&p.count[((long)gx * g.Ny + gy) * g.Nz + gz], c);
    }
    __syncthreads();
    bool over = false;
    for (int j = tid; j < n; j += PIC_BLOCK) {
        const long e = qb + j;
        if (q.slot[e] < 0) continue;
        const int slot = hbase[q.halo[e]] + q.slot[e];
        q.slot[e] = slot;
        if (slot >= p.cap) {
            over = true;
            continue;
        }
        const long to = (long)q.dest[e] * p.cap + slot;
        for (int k = 0; k < 6; k++) p.p[k][to] = q.q[k][e];
        p.sp[to] = q.sp[e];
        if (p.id) p.id[to] = q.id[e];
    }
    if (over) atomicOr(&st->flags, (unsigned)PIC_F_BUCKET_OVERFLOW);
    if (bad) atomicOr(&st->flags, (unsigned)PIC_F_QUEUE_CORRUPT);
}

#endif
