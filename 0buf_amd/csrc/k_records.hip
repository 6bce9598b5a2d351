/*
 * k_records.hip -- offsets of the compact per-step records of the
 * differentiable form (no reference counterpart): the push writes tile t's
 * pre-step particles, in its flat order, at off[t] + (flat index), so the
 * record holds exactly the particles and not the bucket capacity.
 *
 * Tiles are numbered (tx * nty + ty) * ntz + tz and hold 2 x 4 x 8 cells
 * (z fastest), as in the push kernels.
 */
#include "pic_internal.h"

namespace {

constexpr int TXC = 2, TYC = 4, TZC = 8, C = TXC * TYC * TZC;

/* particles of each tile (bucket counts capped at the capacity) */
__global__ __launch_bounds__(PIC_BLOCK) void k_tile_sums(Grid g, TileGeom t, const int *count,
                                                         int cap, long *sum) {
    const long tile = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (tile >= t.ntiles) return;
    const int tz = (int)(tile % t.ntz);
    const int ty = (int)((tile / t.ntz) % t.nty);
    const int tx = (int)(tile / ((long)t.ntz * t.nty));
    long n = 0;
    for (int c = 0; c < C; c++) {
        const int gi = tx * TXC + c / (TYC * TZC);
        const int gl = ty * TYC + (c / TZC) % TYC, gk = tz * TZC + c % TZC;
        if (gi >= g.Nx || gl >= g.Ny || gk >= g.Nz) continue;
        const int v = count[((long)gi * g.Ny + gl) * g.Nz + gk];
        n += v < cap ? v : cap;
    }
    sum[tile] = n;
}

/* exclusive scan of the tile sums, one workgroup of SCAN_T threads: each
 * thread scans a contiguous chunk, the chunk totals are scanned in LDS */
constexpr int SCAN_T = 1024;
__global__ __launch_bounds__(SCAN_T) void k_scan_tiles(const long *sum, long n, long *off,
                                                      long room, DevStatus *st) {
    __shared__ long part[SCAN_T];
    const int tid = threadIdx.x;
    const long chunk = (n + SCAN_T - 1) / SCAN_T;
    const long b = tid * chunk, e = b + chunk < n ? b + chunk : n;
    long acc = 0;
    for (long i = b; i < e; i++) acc += sum[i];
    part[tid] = acc;
    __syncthreads();
    if (tid == 0) {
        long run = 0;
        for (int k = 0; k < SCAN_T; k++) {
            const long v = part[k];
            part[k] = run;
            run += v;
        }
        off[n] = run;
        if (run > room) atomicOr(&st->flags, (unsigned)PIC_F_RECORD);
    }
    __syncthreads();
    long run = part[tid];
    for (long i = b; i < e; i++) {
        off[i] = run;
        run += sum[i];
    }
}

/* six arrays of n 4-byte words in one launch (the fields of a step
 * record), blockIdx.y = array */
struct Six {
    const unsigned *src[6];
    unsigned *dst[6];
};
__global__ __launch_bounds__(PIC_BLOCK) void k_copy6(Six a, long n) {
    const unsigned *src = a.src[blockIdx.y];
    unsigned *dst = a.dst[blockIdx.y];
    for (long i = (long)blockIdx.x * PIC_BLOCK + threadIdx.x; i < n;
         i += (long)gridDim.x * PIC_BLOCK)
        dst[i] = src[i];
}

}  // namespace

void launch_copy6(void *const dst[6], const void *const src[6], size_t bytes, hipStream_t s) {
    Six a;
    for (int k = 0; k < 6; k++) {
        a.src[k] = (const unsigned *)src[k];
        a.dst[k] = (unsigned *)dst[k];
    }
    const long n = (long)(bytes / 4);
    long blocks = (n + PIC_BLOCK - 1) / PIC_BLOCK;
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_copy6, dim3((unsigned)blocks, 6), dim3(PIC_BLOCK), 0, s, a, n);
}

void launch_record_offsets(const Grid &g, const TileGeom &t, const int *count, int cap,
                           long *tile_sum, long *off, long room, DevStatus *st, hipStream_t s) {
    hipLaunchKernelGGL(k_tile_sums, dim3((unsigned)((t.ntiles + PIC_BLOCK - 1) / PIC_BLOCK)),
                       dim3(PIC_BLOCK), 0, s, g, t, count, cap, tile_sum);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(SCAN_T), 0, s, (const long *)tile_sum,
                       t.ntiles, off, room, st);
}
