/*
 * pic_internal.h -- context layout and kernel launch interfaces shared by
 * the host API (pic_api.hip) and the kernel files (k_*.hip).  Not part of the
 * public ABI.
 *
 * HBM layout (one GPU):
 *   fields  12 ghosted arrays (nx+2)(ny+2)(nz+2), k fastest, the reference's
 *           flat index n = (i(ny+2) + l)(nz+2) + k (cell.h:796-801);
 *   particles  bucketed SoA: bulk cell c = (i*ny + l)*nz + k owns slots
 *           [c*cap, c*cap + count[c]) of x, y, z, px, py, pz (FP64) and a
 *           one-byte species slot; per-species weight and q/m live in a table;
 *           the FP32 form steps between two sets in a packed per-tile
 *           layout (Layout32) and unpacks into buckets for other readers;
 *   departures  one fixed region of qcap entries per tile of cells.
 */
#ifndef PIC_INTERNAL_H
#define PIC_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pic_amd.h"

#define PIC_MAX_SLOTS 8
#define PIC_BLOCK 256

/* status flags raised on the device */
enum {
    PIC_F_TOO_MANY = 1,
    PIC_F_QUEUE_OVERFLOW = 2,
    PIC_F_BUCKET_OVERFLOW = 4,
    PIC_F_RANGE = 8,
    PIC_F_RECOMPUTE = 16,    /* differentiable form: recompute differs from the first run */
    PIC_F_RECORD = 32,       /* differentiable form: a step record ran out of room */
    PIC_F_QUEUE_CORRUPT = 64, /* an arrival refused a departure record out of range */
    PIC_F_SPILL_OVER = 128    /* FP32 form: a tile's spill list was full (host repairs;
                                 not a status the caller sees) */
};

struct DevStatus {
    unsigned int flags;
    int max_dep;
    unsigned long long departures;
    unsigned long long dropped;
    unsigned long long duplicated;
};

/* geometry shared by every kernel */
struct Grid {
    int Nx, Ny, Nz;      /* bulk cells */
    int sy, sx;          /* ghosted strides: sy = Nz+2, sx = (Ny+2)(Nz+2) */
    long G;              /* ghosted nodes */
    double Lx, Ly, Lz, hx, hy, hz, tau;
    int ghost_mode;      /* pic_params.ghost_mode */
};

struct FieldPtrs {
    double *f[PIC_NFIELDS];
};

/* FP32 precision form: E, H, Q in FP32, J as int64 fixed point with
 * resolution 2^-40 (exact, order-free accumulation) */
#define PIC_J_SCALE 1099511627776.0           /* 2^40 */
#define PIC_J_INV (1.0 / 1099511627776.0)

struct FieldPtrs32 {
    float *E[3], *H[3], *Q[3];
    long long *J[3];
};

/* particle storage (FP64 plain form) */
struct PartPtrs {
    double *p[6];            /* x, y, z, pu, pv, pw */
    unsigned char *sp;       /* species slot */
    int *count;              /* per bulk cell */
    int cap;
    int *id;                 /* differentiable form: load-order particle id, else null */
};

/* particle storage of the FP32 form: x, y, z are offsets in [0,1) inside the
 * owning cell (cell units), px, py, pz the momenta.
 *
 * The live particles also have an inbox per cell: PIC_INBOX records of 32
 * bytes (the departure-queue record format) where the arrival kernel puts
 * the particles that joined the cell this step, one 32-byte write each
 * instead of seven scattered ones.  count is the cell's total; slots
 * [0, bcount) are in the particle arrays, slots bcount + r for r < PIC_INBOX
 * are inbox record r, and slots from bcount + PIC_INBOX on (arrivals past a
 * full inbox) are in the spill list of the cell's tile while the arrays are
 * packed (Layout32 below), in the bucket arrays otherwise.  The next push
 * reads the cell in slot order; the unpack and settle kernels put every
 * slot into the bucket arrays for the other readers.  Checkpoints and
 * staging copies have no inbox (bcount, inbox null). */
constexpr int PIC_INBOX = 32;
struct PartPtrs32 {
    float *p[6];
    unsigned char *sp;
    int *count;
    int cap;
    int *id;                 /* differentiable form only */
    int *bcount;             /* live buckets: slots held by the bucket arrays */
    float4 *inbox;           /* live buckets: PIC_INBOX records (2 float4) per cell */
};

/* The packed layout of the live FP32 particles.  The push reads the live
 * set and writes its stayers into the other set, packed: tile t owns slots
 * [t*tcap, (t+1)*tcap) (tcap = cells per tile x cap), and each wavefront of
 * the tile's workgroup writes the stayers of its cells contiguously in flat
 * order from the flat index of its first particle, so a cell's particles
 * lie at pstart[cell] + [0, bcount[cell]) of its tile's region and share
 * lines with their neighbours (no partial line per cell and array).
 *
 * The push also places its tile's leavers (the arrival): per destination
 * cell one atomic on arr_next claims a run of arrival indices r, and record
 * r goes to the cell's inbox (r < PIC_INBOX, inbox_next) or past it to the
 * destination tile's unsorted spill list (spill_u: scap records of the
 * departure-record format, key = species | local cell << 8 | (r -
 * PIC_INBOX) << 16).  After the push, one workgroup per tile sums each
 * cell's count (bcount + arrivals) and moves its spill records into the
 * sorted list (spill_s): cell c's k-th record at the exclusive prefix over
 * the tile's cells of max(0, arrivals - PIC_INBOX), plus k.  The arrays the
 * push reads (arr, the live inbox, spill_s) hold the previous step's
 * arrivals; the host swaps the per-cell ones with the *_next ones after
 * each push.  A cell holds bcount + arr particles: bcount in the particle
 * arrays, then its arrivals in order.  The host turns the packed set back
 * into buckets (unpack) for every other reader, and when a spill list or a
 * bucket overflows (a spill list then grows). */
struct Layout32 {
    int packed;              /* the set being read is packed (else buckets) */
    int *pstart;             /* per cell: first slot in its tile's region */
    int *arr, *arr_next;     /* per cell: arrivals of the last step / being placed */
    float4 *inbox_next;      /* per cell: PIC_INBOX records being placed */
    float4 *spill_u, *spill_s;  /* per tile: scap records (2 float4 each) */
    int *spill_u_cnt;        /* per tile: records claimed in spill_u (may exceed scap) */
    int scap;
    long tcap;               /* slots per tile region */
};
constexpr int PIC_SPILL_CAP = 256;  /* initial spill records per tile */
constexpr int PIC_SPILL_CAP_MAX = 1 << 16;

/* FP32 departure queue: one 32-byte record per leaver, two float4
 * (ox, oy, oz, pu) and (pv, pw, key, id), with key = species | h << 8 |
 * rank << 16 (h = the destination's index in the source tile's
 * (TX+2)(TY+2)(TZ+2) halo box, rank = the leaver's rank among the tile's
 * leavers bound for h) and id as int bits (differentiable form) */
struct QueuePtrs32 {
    float4 *e;
    int *slot;               /* arrival index r the push gave the record */
    int *count;              /* per tile */
    int qcap;                /* entries per tile */
};

struct QueuePtrs {
    double *q[6];
    unsigned char *sp;
    int *dest;               /* destination bulk cell */
    int *slot;               /* slot claimed by the arrival kernel */
    int *count;              /* per tile */
    int qcap;                /* entries per tile */
    int *id;                 /* differentiable form only */
    unsigned char *halo;     /* destination in the source tile's (TX+2)(TY+2)(TZ+2) halo box */
};

struct TileGeom {
    int TX, TY, TZ;
    int ntx, nty, ntz;
    long ntiles;
};

/* The push kernels number a tile's particles in flat order f = 0..T-1
 * (cell by cell, sOff[c] = first f of local cell c; the step records use
 * this order).  A particle's bucket slot is addressed by a 32-bit byte offset
 * from per-tile base pointers (the host keeps a tile's bucket span below
 * 2^32 bytes): slot j of local cell c is rel = sBase[c] + j, sBase[c] = the
 * first slot of cell c relative to the tile's first cell. */
template <class T>
__device__ __forceinline__ T &slot(T *base, unsigned rel) {
    return *(T *)((char *)base + rel * (unsigned)sizeof(T));
}

/* Stable in-place compaction by wavefronts.  Each wavefront of a push
 * workgroup owns whole cells of the tile (wave_first_cell) and walks them in
 * flat order, one chunk of up to 64 consecutive particles at a time
 * (compose_chunk).  A stayer is stored at slot rank of its cell, rank = the
 * stayers of its cell before it in flat order.  Only slots of particles the
 * wavefront has already loaded are overwritten (rank <= the particle's own
 * slot), so no other wavefront and no later chunk is disturbed, and the
 * buckets need no second pass.
 *
 * stable_rank_runs: the rank of this lane's particle (valid lanes form a prefix
 * of the wavefront and hold consecutive flat indices); the
 * wavefront-uniform carry holds the cell of the previous chunk's last
 * particle and that cell's stayers so far. */
__device__ __forceinline__ int stable_rank_runs(unsigned long long R, int c, bool valid,
                                                bool stays, int &carry_cell, int &carry_cnt) {
    const int lane = (int)(threadIdx.x & 63);
    const unsigned long long S = __ballot(valid && stays);
    const unsigned long long V = __ballot(valid);
    const unsigned long long le = ~0ull >> (63 - lane);
    const int rs = 63 - __clzll(R & le);
    const int base = (rs == 0 && c == carry_cell) ? carry_cnt : 0;
    const int rank = base + __popcll(S & (le >> 1) & (~0ull << rs));
    if (V) {
        const int L = 63 - __clzll(V);
        const int cL = __builtin_amdgcn_readlane(c, L);
        const unsigned long long leL = ~0ull >> (63 - L);
        const int rsL = 63 - __clzll(R & leL);
        const int baseL = (rsL == 0 && cL == carry_cell) ? carry_cnt : 0;
        carry_cnt = baseL + __popcll(S & leL & (~0ull << rsL));
        carry_cell = cL;
    }
    return rank;
}

/* R = the lanes where a run of one cell begins (lane 0 and every lane
 * whose cell differs from the previous lane's) */
__device__ __forceinline__ int stable_rank(int c, bool valid, bool stays, int &carry_cell,
                                           int &carry_cnt) {
    const int lane = (int)(threadIdx.x & 63);
    const int cprev = __shfl_up(c, 1);
    const unsigned long long R = __ballot(valid && (lane == 0 || c != cprev));
    return stable_rank_runs(R, c, valid, stays, carry_cell, carry_cnt);
}

/* Line groups.  A cell's bucket is split into groups of GS slots (GS slots =
 * one 128-byte line of one particle array: 32 FP32 or 16 FP64 values); a
 * group holds the particles of one line, 1..GS of them.  A wavefront walks
 * its groups in order and packs whole groups into chunks of at most 64
 * lanes, so no line is shared by two chunks (a line split across chunks is
 * read twice when it leaves the L2 in between).  Group entry: cell (8 bits)
 * | kind (1: FP32 inbox records, else bucket slots) | first slot (15 bits)
 * | size (8 bits). */
__device__ __forceinline__ unsigned group_entry(int cell, int kind, int start, int size) {
    return (unsigned)cell | ((unsigned)kind << 8) | ((unsigned)start << 9) | ((unsigned)size << 24);
}
__device__ __forceinline__ int group_cell(unsigned e) { return (int)(e & 255u); }
__device__ __forceinline__ int group_kind(unsigned e) { return (int)((e >> 8) & 1u); }
__device__ __forceinline__ int group_start(unsigned e) { return (int)((e >> 9) & 0x7fffu); }
__device__ __forceinline__ int group_size(unsigned e) { return (int)(e >> 24); }

template <int TX, int TY, int TZ, int GS>
__device__ __forceinline__ void build_groups(const int *sCnt, const int *sG, unsigned *sGrp,
                                             int tid, int nthreads) {
    constexpr int C = TX * TY * TZ;
    static_assert(C <= 256, "group entries hold the cell in 8 bits");
    for (int w = tid; w < 4 * C; w += nthreads) {
        const int c = w >> 2, part = w & 3, n = sCnt[c];
        for (int line = part; line < sG[c + 1] - sG[c]; line += 4) {
            const int left = n - line * GS, sz = left < GS ? left : GS;
            sGrp[sG[c] + line] = group_entry(c, 0, line * GS, sz);
        }
    }
}

/* first cell c with sOff[c] >= w*T/nw (C for w = nw): wavefront w owns
 * cells [first(w), first(w+1)) */
template <int C>
__device__ __forceinline__ int wave_first_cell(const int *sOff, int w, int nw) {
    if (w >= nw) return C;
    const int f = (int)(((long)w * sOff[C]) / nw);
    int lo = -1, hi = C;    /* sOff[lo] < f <= sOff[hi] */
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (sOff[mid] >= f) hi = mid; else lo = mid;
    }
    return hi;
}

/* exclusive prefix sum of v over the 64 lanes of a wavefront; *total
 * receives the sum over all lanes */
__device__ __forceinline__ int wave_excl_scan(int v, int &total) {
    const int lane = (int)(threadIdx.x & 63);
    int x = v;
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    total = __shfl(x, 63);
    return x - v;
}

/* chunk composition (wavefront-uniform): whole groups from g0 while they
 * fit 64 lanes; lane start of each group in `starts`; returns the next
 * group */
__device__ __forceinline__ int compose_chunk(const unsigned *sGrp, int g0, int gend,
                                             unsigned long long &starts, int &lanes) {
    int g = g0, n = 0;
    unsigned long long m = 0;
    while (g < gend) {
        const int sz = group_size(__builtin_amdgcn_readfirstlane(sGrp[g]));
        if (n + sz > 64) break;
        m |= 1ull << n;
        n += sz;
        g++;
    }
    starts = m;
    lanes = n;
    return g;
}

/* the same, also marking in `cstarts` the lane where each run of one cell
 * begins (lane 0 and every group whose cell differs from the previous
 * group's): the run starts stable_rank_runs needs, found on the scalar unit
 * from the group entries instead of a lane shuffle */
__device__ __forceinline__ int compose_chunk(const unsigned *sGrp, int g0, int gend,
                                             unsigned long long &starts, int &lanes,
                                             unsigned long long &cstarts) {
    int g = g0, n = 0, prev = -1;
    unsigned long long m = 0, cs = 0;
    while (g < gend) {
        const unsigned e = __builtin_amdgcn_readfirstlane(sGrp[g]);
        const int sz = group_size(e);
        if (n + sz > 64) break;
        m |= 1ull << n;
        if (group_cell(e) != prev) cs |= 1ull << n;
        prev = group_cell(e);
        n += sz;
        g++;
    }
    starts = m;
    lanes = n;
    cstarts = cs;
    return g;
}

/* chunks of exactly 64 particles (fewer only at the end of the wavefront's
 * cells): from group g0, whose first o0 particles an earlier chunk took,
 * whole groups while they fit and then the first part of the next group.
 * starts: the lane where each group's part begins; cstarts: the lane where
 * each run of one cell begins (as above); o_next: how many particles of the
 * returned group this chunk took (0: none, the group is untouched).  A line
 * split between two consecutive chunks is read twice, back to back. */
__device__ __forceinline__ int compose_full_chunk(const unsigned *sGrp, int g0, int o0, int gend,
                                                  unsigned long long &starts, int &lanes,
                                                  unsigned long long &cstarts, int &o_next) {
    int g = g0, n = 0, prev = -1, off = o0;
    unsigned long long m = 0, cs = 0;
    o_next = 0;
    while (g < gend && n < 64) {
        const unsigned e = __builtin_amdgcn_readfirstlane(sGrp[g]);
        const int sz = group_size(e) - off;
        m |= 1ull << n;
        if (group_cell(e) != prev) cs |= 1ull << n;
        prev = group_cell(e);
        if (n + sz > 64) {
            o_next = off + (64 - n);
            n = 64;
            break;
        }
        n += sz;
        off = 0;
        g++;
    }
    starts = m;
    lanes = n;
    cstarts = cs;
    return g;
}

struct SpeciesTable {
    double m[PIC_MAX_SLOTS];
    double q_m[PIC_MAX_SLOTS];
};

/* kernel launchers (k_fields.hip) */
/* the step's fused pass after the push: the current fold, H^{n+1} into
 * h_out (the second H arrays; the caller swaps them with the live ones), the
 * E update and the periodic E ghosts, equal to launch_fold_j, launch_h2,
 * launch_e, launch_e_periodic bit for bit */
void launch_fold_h2_e(const Grid &g, FieldPtrs f, double *const h_out[3], hipStream_t s);
void launch_fold_h2_e32(const Grid &g, FieldPtrs32 f, float *const h_out[3], hipStream_t s);
/* sum of nblocks partials into *out (deterministic tree) */
void launch_reduce_final(const double *partials, int nblocks, double *out, hipStream_t s);
void launch_h_half(const Grid &g, FieldPtrs f, hipStream_t s);
void launch_h2(const Grid &g, FieldPtrs f, hipStream_t s);
void launch_e(const Grid &g, FieldPtrs f, hipStream_t s);
void launch_e_periodic(const Grid &g, FieldPtrs f, hipStream_t s);
void launch_fold_j(const Grid &g, FieldPtrs f, hipStream_t s);
void launch_energy(const Grid &g, FieldPtrs f, double *partials, int nblocks, double *out,
                   hipStream_t s);
void launch_reset_status(DevStatus *st, hipStream_t s);
void launch_h_half32(const Grid &g, FieldPtrs32 f, hipStream_t s);
void launch_h2_32(const Grid &g, FieldPtrs32 f, hipStream_t s);
void launch_e32(const Grid &g, FieldPtrs32 f, hipStream_t s);
void launch_e_periodic32(const Grid &g, FieldPtrs32 f, hipStream_t s);
void launch_fold_j32(const Grid &g, FieldPtrs32 f, hipStream_t s);
void launch_energy32(const Grid &g, FieldPtrs32 f, double *partials, int nblocks, double *out,
                     hipStream_t s);
TileGeom make_tile_geom(int Nx, int Ny, int Nz);

/* Differentiable form of the FP64 step: the push deposits J as int64 fixed
 * point with resolution 2^-48 (order-free, so a recomputed step is bitwise
 * equal to the first run), converted to the FP64 J arrays before the fold;
 * when snap.p[0] is set the push also writes the pre-step particle state into
 * the compact record `snap`: tile t's particles, in the push's flat order,
 * at snap_off[t] + (flat index), and each cell's count in snap.count. */
#define PIC_J_SCALE64 281474976710656.0       /* 2^48 */
struct DiffArgs {
    long long *jfix[3];
    PartPtrs snap;
    const long *snap_off;   /* record of tile t starts at snap_off[t] (flat tile order) */
    long snap_room;         /* entries the record holds */
    /* decision tape (k_tape.h): this step's row, one word per particle id */
    unsigned long long *tape_row;
    long tape_n;            /* ids the row holds */
    int tape_mode;          /* PIC_TAPE_OFF / RECORD / REPLAY */
};

/* particle launchers (k_particles_f64.hip) */
void launch_push_deposit_f64(const Grid &g, const TileGeom &t, FieldPtrs f, PartPtrs p,
                             QueuePtrs q, SpeciesTable sp, int dep_cap, DevStatus *st,
                             const DiffArgs *diff, hipStream_t s);
void launch_jfix_to_f64(const Grid &g, FieldPtrs f, DiffArgs d, hipStream_t s);
/* compact record offsets (k_records.hip): off[t] = particles of the tiles
 * before t (counts capped at cap), off[ntiles] the total; a total beyond
 * `room` raises PIC_F_RECORD (the push writes no record entry past room) */
/* six device arrays of `bytes` (a multiple of 4) each, in one launch */
void launch_copy6(void *const dst[6], const void *const src[6], size_t bytes, hipStream_t s);
void launch_record_offsets(const Grid &g, const TileGeom &t, const int *count, int cap,
                           long *tile_sum, long *off, long room, DevStatus *st, hipStream_t s);
void launch_arrive_f64(const Grid &g, const TileGeom &t, PartPtrs p, QueuePtrs q, DevStatus *st,
                       hipStream_t s);
void launch_scatter_load(const long *offsets, long ncell, const double *src[6],
                         const unsigned char *srcsp, PartPtrs p, hipStream_t s);
void launch_repack(long ncell, PartPtrs from, PartPtrs to, hipStream_t s);
void launch_fix_arrivals(const TileGeom &t, QueuePtrs q, int oldcap, PartPtrs to, hipStream_t s);
void launch_gather_particles(long ncell, const long *offsets, PartPtrs p, double *dst[6],
                             unsigned char *dsp, int *dcell, int *did, hipStream_t s);
void launch_scatter_ids(const long *offsets, long ncell, const int *src, PartPtrs p,
                        hipStream_t s);
void launch_moments(long ncell, PartPtrs p, SpeciesTable sp, const int *slot_sort,
                    double *partials, int nblocks, double *out27, hipStream_t s);

/* FP32 form (k_particles_f32.hip) */
/* diff: differentiable form (ids; compact record of the pre-step state into
 * *diff at snap_off[tile] + flat index when diff->p[0] is set), null
 * otherwise */
/* p: the live set (read; packed or buckets as L.packed says), out: the
 * other set, written packed (p[6], sp, id) */
void launch_push_deposit_f32(const Grid &g, const TileGeom &t, FieldPtrs32 f, PartPtrs32 p,
                             PartPtrs32 out, Layout32 L, QueuePtrs32 q, SpeciesTable sp,
                             int dep_cap, DevStatus *st, const PartPtrs32 *diff,
                             const long *snap_off, long snap_room, hipStream_t s);
/* after the push, per tile: count = bcount + arrivals (PIC_F_BUCKET_OVERFLOW
 * past cap) and the spill records sorted by cell */
void launch_arrivals_done32(const Grid &g, const TileGeom &t, PartPtrs32 p, Layout32 L,
                            DevStatus *st, hipStream_t s);
/* the packed live set p, its inbox records and its arrivals past the inbox
 * (from the spill lists, or from the departure queues of the last push when
 * from_queue is set: the repair after a full spill list) into buckets of
 * `to` (p[6], sp, id, cap >= every count); then bcount = count and no
 * arrival pending */
void launch_unpack32(const Grid &g, const TileGeom &t, PartPtrs32 p, Layout32 L, PartPtrs32 to,
                     const QueuePtrs32 *from_queue, hipStream_t s);
/* buckets loaded whole (bcount = count): no arrival pending */
void launch_clear_arrivals32(const Grid &g, const TileGeom &t, PartPtrs32 p, Layout32 L,
                             hipStream_t s);
void launch_clear_flags32(DevStatus *st, unsigned mask, hipStream_t s);
/* the packed live set with its pending arrivals, compacted cell by cell at
 * offsets[cell] (as launch_gather_raw32 does from buckets); the live set
 * stays as it is */
void launch_gather_packed32(const Grid &g, const TileGeom &t, PartPtrs32 p, Layout32 L,
                            const long *offsets, float *dst[6], unsigned char *dsp, int *did,
                            hipStream_t s);
void launch_scatter_load32(const long *offsets, long ncell, const float *src[6],
                           const unsigned char *srcsp, PartPtrs32 p, hipStream_t s);
void launch_gather_particles32(const Grid &g, long ncell, const long *offsets, PartPtrs32 p,
                               double *dst[6], unsigned char *dsp, int *dcell, int *did,
                               hipStream_t s);
/* compacted copy of the raw FP32 state (offsets, momenta, species, ids) */
void launch_gather_raw32(long ncell, const long *offsets, PartPtrs32 p, float *dst[6],
                         unsigned char *dsp, int *did, hipStream_t s);
void launch_scatter_ids32(const long *offsets, long ncell, const int *src, PartPtrs32 p,
                          hipStream_t s);
void launch_moments32(const Grid &g, long ncell, PartPtrs32 p, SpeciesTable sp,
                      const int *slot_sort, double *partials, int nblocks, double *out27,
                      hipStream_t s);

/* per-cell moments of one sort (k_gauss.hip): op 0 density, 1 kinetic
 * energy, 2 mean velocity (3 x ncell), 3 current (3 x ncell) */
void launch_cell_moments64(long ncell, PartPtrs p, SpeciesTable sp, const int *slot_sort, int sort,
                           int op, double *out, hipStream_t s);
void launch_cell_moments32(long ncell, PartPtrs32 p, SpeciesTable sp, const int *slot_sort,
                           int sort, int op, double *out, hipStream_t s);

/* device-side seeded initializer (k_init.hip): one species into the
 * buckets (p64 or p32), appended behind what the cells already hold */
void launch_generate_species(const Grid &g, const pic_species_spec &spec, const int *cnt_lk,
                             const long *off_lk, long plane, unsigned long long seed,
                             unsigned char slot, long id_base, PartPtrs *p64, PartPtrs32 *p32,
                             int *start_of_cell, unsigned long long *misplaced, int *fail,
                             hipStream_t s);

/* Gauss-law residual div E - rho, FP64 (k_gauss.hip): rho by CIC on the
 * cell-centre lattice into an int64 fixed-point array of G entries, then
 * per bulk centre res = div E - rho; returns max |res - ref| if ref != 0 */
void launch_rho64(const Grid &g, long ncell, PartPtrs p, SpeciesTable sp, long long *rho,
                  hipStream_t s);
void launch_rho32(const Grid &g, long ncell, PartPtrs32 p, SpeciesTable sp, long long *rho,
                  hipStream_t s);
void launch_gauss64(const Grid &g, FieldPtrs f, const long long *rho, double *res,
                    hipStream_t s);
void launch_gauss32(const Grid &g, FieldPtrs32 f, const long long *rho, double *res,
                    hipStream_t s);

/* Differentiable form, adjoint stages (k_adjoint.hip).  Adjoint arrays
 * have the field layout (G entries); the particle adjoints are indexed by
 * particle id. */
struct AdjLam {
    double *E[3], *H[3], *J[3], *Q[3];
};
struct AdjFwd {             /* forward fields the particle stage read */
    const double *E[3];     /* E^n */
    const double *H[3];     /* H^{n+1/2} = H^n + Q */
};
/* the particle adjoints by id, one 64-byte line per id: v[id * PIC_LSTRIDE
 * + k] for k = d/dx, d/dy, d/dz, d/dpu, d/dpv, d/dpw (the reverse sweep
 * reads and writes each particle's six together, at the id's place) */
constexpr int PIC_LSTRIDE = 8;
struct LamParts {
    double *v;
    long n;
};
/* component k of every id between the adjoint lines and a flat array
 * (to_lines: flat -> lines) */
void launch_lam_copy(LamParts lp, double *flat, int k, bool to_lines, hipStream_t s);
void launch_adj_seed_energy(const Grid &g, const double *const E[3], double w, AdjLam l,
                            hipStream_t s);
void launch_adj_e_periodic(const Grid &g, AdjLam l, hipStream_t s);
void launch_adj_curl_e(const Grid &g, AdjLam l, hipStream_t s);
void launch_adj_fold_j(const Grid &g, AdjLam l, hipStream_t s);
void launch_adj_particles_f64(const Grid &g, const TileGeom &t, AdjFwd fwd, PartPtrs snap,
                              const long *snap_off, SpeciesTable sp, AdjLam l, LamParts lp,
                              DevStatus *st, hipStream_t s);
void launch_adj_curl_h(const Grid &g, AdjLam l, hipStream_t s);

/* FP32 form (k_adjoint32.hip): forward fields are floats, the particle
 * adjoints by id are in offset units (d/do = h d/dx) during the sweep */
struct AdjFwd32 {
    const float *E[3];
    const float *H[3];
};
void launch_adj_particles_f32(const Grid &g, const TileGeom &t, AdjFwd32 fwd, PartPtrs32 snap,
                              const long *snap_off, SpeciesTable sp, AdjLam l, LamParts lp,
                              DevStatus *st, hipStream_t s);
void launch_scale_lam(LamParts lp, double sx, double sy, double sz, hipStream_t s);
void launch_adj_seed_energy32(const Grid &g, const float *const E[3], double w, AdjLam l,
                              hipStream_t s);
void launch_copy_parts(long ncell, PartPtrs from, PartPtrs to, hipStream_t s);
void launch_mismatch(const void *a, const void *b, long n, int elem_bytes, DevStatus *st,
                     hipStream_t s);

#endif
