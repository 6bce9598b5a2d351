/*
 * k_particles_f32.hip -- particle stages of the FP32 precision form.
 *
 * Same tile structure as the FP64 plain form (one workgroup per 2x4x8 cells,
 * E/H node box staged in LDS, tile departure regions), with:
 *   * each wavefront owning whole cells of the tile and walking them in
 *     chunks of 64 consecutive particles (a line may be split between two
 *     consecutive chunks); stayers written in flat order into the other
 *     particle set, packed per tile (Layout32 in pic_internal.h), arrivals
 *     into inbox records and spill lists that the next push reads;
 *   * positions held as FP32 offsets in [0,1) inside the owning cell, so the
 *     resolution is h * 2^-24 everywhere in the box;
 *   * FP32 gather, from field boxes staged in LDS beside their z
 *     differences (a z pair of the trilinear interpolation is one fmaf), and
 *     Boris push (the reference's Cayley form);
 *   * the reference's charge-conserving deposit (Villasenor-Buneman segments,
 *     relay on the dual-cell faces as in CurrentToMesh, cell.h:1306-1422)
 *     written in cell units, each contribution rounded once to int64 fixed
 *     point (2^-40) and summed with native 64-bit integer atomics (ds_add_u64
 *     in LDS, then global): the current is exact given the per-particle
 *     contributions and independent of summation order (in LDS as raw
 *     offset words when the launcher's bound allows, see fixmul);
 *   * migration by the cell-local offset: a particle whose offset leaves
 *     [0,1) moves to the neighbouring cell (periodic), exactly one receiver.
 */
#include "pic_internal.h"
#include "k_push_math32.h"
#include "k_queue_key.h"

/* copies of the LDS J box: lane l deposits into copy l % JCOPIES, which
 * divides the same-address conflicts of the 64-bit LDS atomics (consecutive
 * lanes hold particles of one cell) by JCOPIES; the copies are summed before
 * the global flush (integer sums: the result does not depend on it) */
#ifndef PIC_JCOPIES32
#define PIC_JCOPIES32 2
#endif
#ifndef PIC_JPAD32
#define PIC_JPAD32 8
#endif

namespace {

constexpr int TXC = 2, TYC = 4, TZC = 8;
constexpr int GS32 = 32;    /* slots of an FP32 array in one 128-byte line */
constexpr int JCOPIES = PIC_JCOPIES32;

struct PushArgs32 {
    Grid g;
    TileGeom t;
    FieldPtrs32 f;
    PartPtrs32 p;                  /* the live set, read */
    PartPtrs32 o;                  /* the other set, written packed */
    Layout32 L;
    QueuePtrs32 q;
    float qm[PIC_MAX_SLOTS];       /* charge/mass per species slot */
    float cj[PIC_MAX_SLOTS][3];    /* m h_d / tau * 2^40: fixed-point current per
                                      unit cell displacement along d */
    float tdh[3];                  /* tau / h_d */
    float hh[3];                   /* h_d^2, weights of the two-face relay projection */
    float tau;
    int dep_cap;
    int ngrp;                      /* entries of the dynamic group table */
    DevStatus *st;
    PartPtrs32 snap;               /* differentiable form: record of the pre-step state */
    const long *snap_off;          /* record offset of each tile */
    long snap_room;
    int snap_on;
};

/* round-to-nearest int64 of a*b (|a*b| < 2^51) by the 1.5*2^52 shifter: the
 * product of two FP32 values is exact in FP64, so the fma rounds once */
template <bool RAW>
__device__ __forceinline__ long long fixmul(double a, float b) {
    /* d = 1.5 2^52 + round(a b) lies in (2^52, 2^53) for |a b| < 2^51, so
     * its bits are 0x4338000000000000 + round(a b).  RAW keeps the offset:
     * the offset's low 48 bits are 0, so a sum of such words carries the
     * exact sum of the values in its low 48 bits while that sum stays
     * below 2^47 in magnitude (the launcher checks the bound) */
    const double d = fma(a, (double)b, 6755399441055744.0);
    return RAW ? __double_as_longlong(d)
               : (long long)(__double_as_longlong(d) - 0x4338000000000000LL);
}

__device__ __forceinline__ void add64(long long *p, long long v) {
    atomicAdd((unsigned long long *)p, (unsigned long long)v);
}


/* One straight segment a -> b inside dual cell rd (relative to the owner
 * cell; the dual cell spans centres rd-1/2 .. rd+1/2), n = its first node
 * in the J box: the 12 VB node contributions of pqr (cell.h:1511-1693) in
 * cell units,
 *   Jx = (m hx/tau) du (Wy Wz +- dv dw/12)  on (x dual, y/z centres), etc. */
template <int BY, int BZ, int JY0, int JZ0, bool RAW>
__device__ __forceinline__ void vb_seg(long long *sJ, int n, int rdx, int rdy, int rdz, float ax,
                                       float ay, float az, float bx, float by, float bz,
                                       const float *cj) {
    const float dx = bx - ax, dy = by - ay, dz = bz - az;
    const float mx = 0.5f * (ax + bx) - ((float)rdx - 0.5f);
    const float my = 0.5f * (ay + by) - ((float)rdy - 0.5f);
    const float mz = 0.5f * (az + bz) - ((float)rdz - 0.5f);
    const float lx = 1.0f - mx, ly = 1.0f - my, lz = 1.0f - mz;
    const float syz = dy * dz * (1.0f / 12.0f), sxz = dx * dz * (1.0f / 12.0f),
                sxy = dx * dy * (1.0f / 12.0f);
    const double qx = (double)(cj[0] * dx), qy = (double)(cj[1] * dy), qz = (double)(cj[2] * dz);
    long long *Jx = sJ, *Jy = sJ + JY0, *Jz = sJ + JZ0;
    add64(&Jx[n], fixmul<RAW>(qx, ly * lz + syz));
    add64(&Jx[n + 1], fixmul<RAW>(qx, ly * mz - syz));
    add64(&Jx[n + BZ], fixmul<RAW>(qx, my * lz - syz));
    add64(&Jx[n + BZ + 1], fixmul<RAW>(qx, my * mz + syz));
    add64(&Jy[n], fixmul<RAW>(qy, lx * lz + sxz));
    add64(&Jy[n + 1], fixmul<RAW>(qy, lx * mz - sxz));
    add64(&Jy[n + BY * BZ], fixmul<RAW>(qy, mx * lz - sxz));
    add64(&Jy[n + BY * BZ + 1], fixmul<RAW>(qy, mx * mz + sxz));
    add64(&Jz[n], fixmul<RAW>(qz, lx * ly + sxy));
    add64(&Jz[n + BZ], fixmul<RAW>(qz, lx * my - sxy));
    add64(&Jz[n + BY * BZ], fixmul<RAW>(qz, mx * ly - sxy));
    add64(&Jz[n + BY * BZ + BZ], fixmul<RAW>(qz, mx * my + sxy));
}

/* Line groups of one cell, in slot order.  Buckets: lines of [0, bc),
 * inbox records [bc, bc + nin) by IBG (two 128-byte lines of records), bucket
 * slots [bc + PIC_INBOX, cnt) by line (arrivals past a full inbox).
 * Packed: the particle-array part [0, bc), which starts at slot ps of the
 * tile's region, split where the region's lines begin; the inbox records by
 * IBG; the spill records [bc + PIC_INBOX, cnt) by SPG (they lie in cell order
 * in the tile's sorted spill list). */
constexpr int SPG = 128;
constexpr int IBG = 8;    /* inbox records per group: two 128-byte lines */
__device__ __forceinline__ int cell_groups32(int cnt, int bc) {
    const int nin = min(PIC_INBOX, cnt - bc), t0 = bc + PIC_INBOX;
    return (bc + GS32 - 1) / GS32 + (nin + IBG - 1) / IBG +
           (cnt > t0 ? (cnt - 1) / GS32 - t0 / GS32 + 1 : 0);
}
__device__ __forceinline__ int cell_groups_packed(int ps, int bc, int nin, int nsp) {
    const int nb = bc > 0 ? (ps + bc - 1) / GS32 - ps / GS32 + 1 : 0;
    return nb + (nin + IBG - 1) / IBG + (nsp + SPG - 1) / SPG;
}

template <int C>
__device__ __forceinline__ void build_groups32(const int *sCnt, const int *sBc, const int *sBase,
                                               const int *sG, bool packed, unsigned *sGrp,
                                               int tid) {
    for (int w = tid; w < 4 * C; w += PIC_BLOCK) {
        const int c = w >> 2, part = w & 3;
        const int cnt = sCnt[c], bc = sBc[c];
        const int nin = min(PIC_INBOX, cnt - bc), ninb = (nin + IBG - 1) / IBG,
                  t0 = bc + PIC_INBOX;
        const int ng = sG[c + 1] - sG[c];
        const int ps = sBase[c];
        const int nsoa = packed ? (bc > 0 ? (ps + bc - 1) / GS32 - ps / GS32 + 1 : 0)
                                : (bc + GS32 - 1) / GS32;
        for (int k = part; k < ng; k += 4) {
            int kind = 0, start, size;
            if (k < nsoa) {
                if (packed) {
                    const int lo = (ps / GS32 + k) * GS32 - ps;
                    start = max(0, lo);
                    size = min(bc, lo + GS32) - start;
                } else {
                    start = GS32 * k;
                    size = min(GS32, bc - start);
                }
            } else if (k < nsoa + ninb) {
                kind = 1;
                start = bc + IBG * (k - nsoa);
                size = min(IBG, bc + nin - start);
            } else if (packed) {
                kind = 1;
                start = t0 + SPG * (k - nsoa - ninb);
                size = min(SPG, cnt - start);
            } else {
                const int line = t0 / GS32 + (k - nsoa - ninb);
                start = max(t0, GS32 * line);
                size = min(cnt, GS32 * (line + 1)) - start;
            }
            sGrp[sG[c] + k] = group_entry(c, kind, start, size);
        }
    }
}

/* 7 waves per SIMD (at most 72 VGPRs): about 1% faster than 8 at 64 VGPRs,
 * which spill scalar registers */
template <int TX, int TY, int TZ, bool DIFF, bool RAW>
__global__ __launch_bounds__(PIC_BLOCK, 7) void k_push_deposit32(PushArgs32 a) {
    constexpr int C = TX * TY * TZ;
    constexpr int BX = TX + 2, BY = TY + 2, BZ = TZ + 2, NB = BX * BY * BZ;
    /* Field boxes in LDS, one after another: a component interpolated as
     * node-type along x (Ex, Hy, Hz) reads x planes 0..TX only, so its box
     * stops one plane short (NBX nodes); the others take all BX planes.
     * Likewise the current box: Jx (x-dual) takes x planes 0..TX, Jy, Jz
     * all of them.  Fewer LDS bytes per workgroup: 6 per CU. */
    constexpr int NBX = (TX + 1) * BY * BZ;
    constexpr int FB[6] = {0, NBX, NBX + NB, NBX + 2 * NB, NBX + 3 * NB, 2 * NBX + 3 * NB};
    constexpr int NF = 3 * NBX + 3 * NB;
    constexpr int JY0 = NBX, JZ0 = NBX + NB, NJ = NBX + 2 * NB;
    constexpr int NW = PIC_BLOCK / 64;
    __shared__ float2 sF[NF];   /* (value, z difference to the next node) per box node */
    /* J box copies JS elements apart: JS = NJ rounded up to 8 mod 16, so
     * that copy 1 sits 16 (of 32) or 48 (of 64) dword banks from copy 0 and
     * the banks one deposit instruction touches in the two copies (the 8
     * dual-cell offsets of a cell, 2 banks each) do not overlap */
    constexpr int JS = (NJ + 15) / 16 * 16 + PIC_JPAD32;
    __shared__ long long sJ[JCOPIES * JS];
    __shared__ int sOff[C + 1];
    __shared__ int sCnt[C];
    __shared__ int sLeave[C];
    __shared__ int sBase[C];     /* first slot of each cell: relative to the tile's first
                                    cell's bucket, or in the tile's region (packed) */
    __shared__ int sBc[C];       /* slots of each cell held by the particle arrays */
    __shared__ int sIB[C];       /* first inbox record of each cell, relative to the tile's */
    __shared__ int sG[C + 1];    /* first line group of each cell */
    __shared__ int sSpOff[C];    /* packed: first spill record of each cell */
    __shared__ int sFirst[NW + 1];  /* first cell of each wavefront */
    __shared__ int sHalo[NB];    /* leavers bound for each cell of the halo box */
    __shared__ int sQ;
    __shared__ unsigned sFlags;
    __shared__ float4 sSpec[PIC_MAX_SLOTS];  /* (q/m tau/2, cj) per species slot */
    extern __shared__ unsigned sGrp[];    /* line groups of the tile */
    long long *const myJ = sJ + (threadIdx.x % JCOPIES) * JS;

    const Grid &g = a.g;
    const int tid = threadIdx.x;
    const bool snap = DIFF && a.snap_on;
    const bool packed = a.L.packed != 0;
    long tile;
    {
        long nt = a.t.ntiles, b = blockIdx.x;
        long q = nt / 8, r = nt % 8, x = b % 8;
        tile = x * q + (x < r ? x : r) + b / 8;
    }
    const int tz = (int)(tile % a.t.ntz);
    const int ty = (int)((tile / a.t.ntz) % a.t.nty);
    const int tx = (int)(tile / ((long)a.t.ntz * a.t.nty));
    /* workgroup-uniform (the 64-bit divisions above run on the vector
     * unit; readfirstlane keeps the tile's base addresses scalar) */
    const int I0 = __builtin_amdgcn_readfirstlane(tx * TX);
    const int L0 = __builtin_amdgcn_readfirstlane(ty * TY);
    const int K0 = __builtin_amdgcn_readfirstlane(tz * TZ);
    const int cap = a.p.cap;
    const float4 *const SPB = a.L.spill_s + 2 * (tile * (long)a.L.scap);

    /* The field boxes and the cells' counts.  Thread n takes box node n in
     * all six component boxes (one decode, one 32-bit offset: the grid has
     * fewer than 2^31 nodes) and cell n of the tile; every global load is
     * issued before the first LDS write, from uniform base addresses, so
     * the workgroup waits for one memory round trip here. */
    static_assert(NB <= PIC_BLOCK && C <= PIC_BLOCK, "one box node and one cell per thread");
    constexpr int FN[6] = {NBX, NB, NB, NB, NBX, NBX};   /* nodes in each box */
    const float *const Fsrc[6] = {a.f.E[0], a.f.E[1], a.f.E[2], a.f.H[0], a.f.H[1], a.f.H[2]};
    float fv[6], fd[6];   /* node value, next node's value (0 outside the grid) */
#pragma unroll
    for (int q = 0; q < 6; q++) fv[q] = fd[q] = 0.0f;
    {
        const int u = tid / (BY * BZ), v = (tid / BZ) % BY, w = tid % BZ;
        const int gu = I0 + u, gv = L0 + v, gw = K0 + w;
        const bool in_uv = tid < NB && gu <= g.Nx + 1 && gv <= g.Ny + 1;
        const unsigned off = (unsigned)gu * g.sx + (unsigned)gv * g.sy + gw;
        if (in_uv && gw <= g.Nz + 1) {
#pragma unroll
            for (int q = 0; q < 6; q++) fv[q] = Fsrc[q][off];
        }
        if (in_uv && w + 1 < BZ && gw + 1 <= g.Nz + 1) {
#pragma unroll
            for (int q = 0; q < 6; q++) fd[q] = Fsrc[q][off + 1];
        }
    }
    const int ca = tid / (TY * TZ), cb = (tid / TZ) % TY, cc = tid % TZ;
    int cnt = 0, bc = 0, ps = 0;
    {
        const int gi = I0 + ca, gl = L0 + cb, gk = K0 + cc;
        if (tid < C && gi < g.Nx && gl < g.Ny && gk < g.Nz) {
            const long cell = ((long)gi * g.Ny + gl) * g.Nz + gk;
            cnt = a.p.count[cell];
            bc = a.p.bcount[cell];
            if (packed) ps = a.L.pstart[cell];
            cnt = cnt < cap ? cnt : cap;
            bc = max(0, min(bc, cnt));
            if (snap) a.snap.count[cell] = cnt;
        }
    }
    if (tid < NB) {
#pragma unroll
        for (int q = 0; q < 6; q++)
            if (tid < FN[q]) sF[FB[q] + tid] = make_float2(fv[q], fd[q] - fv[q]);
    }
    if (tid < C) {
        sCnt[tid] = cnt;
        sBc[tid] = bc;
        sLeave[tid] = 0;
        sBase[tid] = packed ? ps : ((ca * g.Ny + cb) * g.Nz + cc) * cap;
        sIB[tid] = ((ca * g.Ny + cb) * g.Nz + cc) * PIC_INBOX;
    }
    for (int e = tid; e < JCOPIES * JS; e += PIC_BLOCK) sJ[e] = 0;
    for (int h = tid; h < NB; h += PIC_BLOCK) sHalo[h] = 0;
    /* the per-species constants in LDS: a kernel-argument table indexed by a
     * per-particle value would be a vector memory load in the loop, and its
     * wait would also wait for the next chunk's loads issued ahead of it */
    if (tid < PIC_MAX_SLOTS)
        sSpec[tid] = make_float4(a.qm[tid] * a.tau * 0.5f, a.cj[tid][0], a.cj[tid][1],
                                 a.cj[tid][2]);
    if (tid == 0) {
        sQ = 0;
        sFlags = 0;
    }
    __syncthreads();
    static_assert(C <= 64, "one wavefront scans the tile's cells");
    if (tid < 64) {   /* first particle, line group and spill entry of each cell */
        int cnt = tid < C ? sCnt[tid] : 0, ngc = 0, nspc = 0;
        if (tid < C) {
            const int bc = sBc[tid], nin = min(PIC_INBOX, cnt - bc);
            if (packed) {
                /* the cell's slots past the inbox are its spill records,
                 * sorted by cell in the tile's list */
                nspc = cnt - bc - nin;
                ngc = cell_groups_packed(sBase[tid], bc, nin, nspc);
            } else {
                ngc = cell_groups32(cnt, bc);
            }
        }
        int tot_s, tot_g, tot_p;
        const int so = wave_excl_scan(cnt, tot_s);
        int sg = wave_excl_scan(ngc, tot_g);
        if (tot_g > a.ngrp) {   /* past the table's bound (which holds): no group, fail */
            if (tid == 0) atomicOr(&sFlags, (unsigned)PIC_F_QUEUE_CORRUPT);
            sg = 0;
            tot_g = 0;
        }
        const int spo = wave_excl_scan(nspc, tot_p);
        if (tid == 0 && tot_p > a.L.scap)   /* a full list is repaired before a push */
            atomicOr(&sFlags, (unsigned)PIC_F_QUEUE_CORRUPT);
        if (tid < C) {
            sOff[tid] = so;
            sG[tid] = sg;
            sSpOff[tid] = spo;
        }
        if (tid == 0) {
            sOff[C] = tot_s;
            sG[C] = tot_g;
        }
    }
    __syncthreads();
    build_groups32<C>(sCnt, sBc, sBase, sG, packed, sGrp, tid);
    if (tid <= NW) sFirst[tid] = wave_first_cell<C>(sOff, tid, NW);
    __syncthreads();

    /* per-tile bases of the particle arrays: slot rel of the tile is
     * cell-bucket slot pbase + rel, or slot rel of the tile's region */
    const long pbase = packed ? tile * a.L.tcap : (((long)I0 * g.Ny + L0) * g.Nz + K0) * cap;
    const float *const P0 = a.p.p[0] + pbase, *const P1 = a.p.p[1] + pbase,
                *const P2 = a.p.p[2] + pbase, *const P3 = a.p.p[3] + pbase,
                *const P4 = a.p.p[4] + pbase, *const P5 = a.p.p[5] + pbase;
    const unsigned char *const SP = a.p.sp + pbase;
    const int *const ID = DIFF ? a.p.id + pbase : nullptr;
    const float4 *const IB = a.p.inbox + 2 * ((((long)I0 * g.Ny + L0) * g.Nz + K0) * PIC_INBOX);
    const long obase = tile * a.L.tcap;
    float *const O0 = a.o.p[0] + obase, *const O1 = a.o.p[1] + obase,
                 *const O2 = a.o.p[2] + obase, *const O3 = a.o.p[3] + obase,
                 *const O4 = a.o.p[4] + obase, *const O5 = a.o.p[5] + obase;
    unsigned char *const OSP = a.o.sp + obase;
    int *const OID = DIFF ? a.o.id + obase : nullptr;
    float4 *const QT = a.q.e + 2 * (tile * (long)a.q.qcap);   /* the tile's departure queue */

    /* the wavefront's next chunk is composed and its particles loaded one
     * chunk ahead, so the memory latency overlaps this chunk's arithmetic */
    const int lane = tid & 63, wv = tid >> 6;
    int nfc = 0, nslot = 0;
    /* the next chunk's particles, from the particle arrays or from inbox or
     * spill records.  The paths load into the same registers, so none may
     * compute anything into a register after another path's loads were
     * issued (that would wait for those loads): the addresses are formed
     * before the branch (the empty asm statements keep the compiler from
     * sinking them into it), and the record path loads the species byte of
     * the record's key word by itself rather than masking the loaded word.
     * Each path ends in an empty asm statement with a memory clobber, so the
     * compiler cannot sink the two paths' loads into one shared set: that
     * set took a 64-bit address per lane and load, and split the records'
     * 16- and 8-byte loads into 4-byte ones. */
    float qox = 0.f, qoy = 0.f, qoz = 0.f, qpu = 0.f, qpv = 0.f, qpw = 0.f;
    unsigned char qs = 0;
    int qid = 0;
    bool qin = false;     /* the particle comes from an inbox or spill record */
    auto fetch = [&](int g0, int o0, unsigned long long starts) {
        const unsigned long long m = starts & (~0ull >> (63 - lane));
        const int gi = __popcll(m) - 1;
        const unsigned e = sGrp[g0 + gi];
        const int cf = group_cell(e);
        nfc = cf;
        nslot = group_start(e) + (lane - (63 - __clzll(m))) + (gi == 0 ? o0 : 0);
        unsigned nrel = (unsigned)(sBase[cf] + nslot);
        /* record r of the cell: inbox record r below PIC_INBOX, else spill
         * record r - PIC_INBOX of the cell in the tile's sorted list */
        const int r = nslot - sBc[cf];
        const bool spl = r >= PIC_INBOX;
        unsigned ir = 2u * (unsigned)(spl ? min(sSpOff[cf] + r - PIC_INBOX, a.L.scap - 1)
                                          : sIB[cf] + r);
        const float4 *const RB = spl ? SPB : IB;
        asm volatile("" : "+v"(nrel));
        asm volatile("" : "+v"(ir));
        qin = group_kind(e) != 0;
        if (!qin) {
            qox = slot(P0, nrel); qoy = slot(P1, nrel); qoz = slot(P2, nrel);
            qpu = slot(P3, nrel); qpv = slot(P4, nrel); qpw = slot(P5, nrel);
            qs = slot(SP, nrel);
            if (DIFF) qid = slot(ID, nrel);
            asm volatile("" ::: "memory");
        } else {
            const float4 A = slot(RB, ir);
            const float2 B = *(const float2 *)&slot(RB, ir + 1u);
            qox = A.x; qoy = A.y; qoz = A.z; qpu = A.w; qpv = B.x; qpw = B.y;
            qs = *((const unsigned char *)&slot(RB, ir + 1u) + 8);   /* key & 255 */
            if (DIFF) qid = *((const int *)&slot(RB, ir + 1u) + 3);
            asm volatile("" ::: "memory");
        }
    };
    const int ghi = __builtin_amdgcn_readfirstlane(sG[sFirst[wv + 1]]);
    int ng0 = __builtin_amdgcn_readfirstlane(sG[sFirst[wv]]);
    /* the wavefront's stayers go to its region from the flat index of its
     * first particle on, in flat order */
    const int wout = __builtin_amdgcn_readfirstlane(sOff[sFirst[wv]]);
    int wcarry = 0;
    /* Second deposit segments (particles whose path crosses a dual face)
     * wait in the lane's registers until a later particle of the same lane
     * has one too: the wavefront then deposits every waiting segment at
     * once, so it runs that code in fewer chunks than every one (integer
     * sums: the current does not depend on when a contribution is added).
     * pa: relay point, pb: end offset, pmeta: first node in the J box | the
     * end's dual cell << 8 | species << 11. */
    constexpr bool DEFER_SEG = !DIFF;   /* the recording form has no registers to spare */
    bool pend = false;
    float pa[3] = {0.f, 0.f, 0.f}, pb[3] = {0.f, 0.f, 0.f};
    int pmeta = 0;
    auto deposit_pending = [&]() {
        const float4 sp4 = sSpec[pmeta >> 11];
        const float cjp[3] = {sp4.y, sp4.z, sp4.w};
        vb_seg<BY, BZ, JY0, JZ0, RAW>(myJ, pmeta & 255, (pmeta >> 8) & 1, (pmeta >> 9) & 1,
                           (pmeta >> 10) & 1, pa[0], pa[1], pa[2], pb[0], pb[1], pb[2], cjp);
    };
    unsigned long long nstarts = 0, ncstarts = 0;
    int nlanes = 0, no0 = 0, onext = 0;
    int gnext = compose_full_chunk(sGrp, ng0, no0, ghi, nstarts, nlanes, ncstarts, onext);
    if (lane < nlanes) fetch(ng0, no0, nstarts);
    while (nlanes > 0) {
        const bool valid = lane < nlanes;
        const int c = nfc, j = nslot;
        const float ox0 = qox, oy0 = qoy, oz0 = qoz, pu0 = qpu, pv0 = qpv, pw0 = qpw;
        const unsigned char s = qs;
        const int pid = DIFF ? qid : 0;
        ng0 = gnext;
        no0 = onext;
        gnext = compose_full_chunk(sGrp, ng0, no0, ghi, nstarts, nlanes, ncstarts, onext);
        if (lane < nlanes) fetch(ng0, no0, nstarts);
        bool stays = false;
        float nxo = 0.f, nyo = 0.f, nzo = 0.f, pu = pu0, pv = pv0, pw = pw0;
        if (valid) {
            const int ca = c / (TY * TZ), cb = (c / TZ) % TY, cc = c % TZ;
            const float ox = ox0, oy = oy0, oz = oz0;
            const long r = snap ? a.snap_off[tile] + sOff[c] + j : 0;
            if (snap && r < a.snap_room) {
                a.snap.p[0][r] = ox;
                a.snap.p[1][r] = oy;
                a.snap.p[2][r] = oz;
                a.snap.p[3][r] = pu;
                a.snap.p[4][r] = pv;
                a.snap.p[5][r] = pw;
                a.snap.sp[r] = s;
                a.snap.id[r] = pid;
            }

            /* gather (weights as in k_push_math32.h), Boris push, deposit */
            const Weights32 W = weights32(ox, oy, oz);
            const int rdx = W.rd[0], rdy = W.rd[1], rdz = W.rd[2];
            const float cx0 = W.c0[0], cx1 = 1.0f - cx0, cy0 = W.c0[1], cy1 = 1.0f - cy0,
                        cz0 = W.c0[2], cz1 = 1.0f - cz0;
            /* node of the owner cell in the boxes (24-bit multiplies: full
             * rate, where a 32-bit one is a quarter) and the dual offsets */
            const int n0 = (int)__umul24(__umul24((unsigned)ca, BY) + (unsigned)cb, BZ) + cc;
            const int dX = rdx ? BY * BZ : 0, dY = rdy ? BZ : 0, dZ = rdz;
            const int e0 = FB[0] + n0 + dY + dZ, e1 = FB[1] + n0 + dX + dZ,
                      e2 = FB[2] + n0 + dX + dY;
            const int h0 = FB[3] + n0 + dX, h1 = FB[4] + n0 + dY, h2 = FB[5] + n0 + dZ;
            const float Ex = tri_dz<BY, BZ>(sF + e0, ox, cy1, cz1);
            const float Ey = tri_dz<BY, BZ>(sF + e1, cx1, oy, cz1);
            const float Ez = tri_dz<BY, BZ>(sF + e2, cx1, cy1, oz);
            const float Hx = tri_dz<BY, BZ>(sF + h0, cx1, oy, oz);
            const float Hy = tri_dz<BY, BZ>(sF + h1, ox, cy1, oz);
            const float Hz = tri_dz<BY, BZ>(sF + h2, ox, oy, cz1);

            const float4 spc = sSpec[s];
            const Boris32 B = boris32(pu, pv, pw, Ex, Ey, Ez, Hx, Hy, Hz, spc.x, a.tdh);
            pu = B.p1[0];
            pv = B.p1[1];
            pw = B.p1[2];
            const float o1x = end_offset(ox, B.du[0]), o1y = end_offset(oy, B.du[1]),
                        o1z = end_offset(oz, B.du[2]);
            {
                const Relay32 R = relay32(ox, oy, oz, o1x, o1y, o1z, B.du, W.rd, a.hh);
                const int m = R.m;
                const float cj[3] = {spc.y, spc.z, spc.w};
                const float e1x = m ? R.x2[0] : o1x, e1y = m ? R.x2[1] : o1y, e1z = m ? R.x2[2] : o1z;
                vb_seg<BY, BZ, JY0, JZ0, RAW>(myJ, n0 + dX + dY + dZ, rdx, rdy, rdz, ox, oy, oz, e1x, e1y,
                                   e1z, cj);
                if (m && !DEFER_SEG) {
                    vb_seg<BY, BZ, JY0, JZ0, RAW>(myJ,
                                       n0 + (R.r1[0] ? BY * BZ : 0) + (R.r1[1] ? BZ : 0) + R.r1[2],
                                       R.r1[0], R.r1[1], R.r1[2], R.x2[0], R.x2[1], R.x2[2], o1x,
                                       o1y, o1z, cj);
                } else {
                    /* a lane with a waiting segment and a new one: every
                     * waiting segment of the wavefront is deposited first
                     * (lanes past the chunk's end keep theirs) */
                    if (__ballot(m && pend)) {
                        if (pend) deposit_pending();
                        pend = false;
                    }
                    if (m) {
                        pend = true;
                        pa[0] = R.x2[0]; pa[1] = R.x2[1]; pa[2] = R.x2[2];
                        pb[0] = o1x; pb[1] = o1y; pb[2] = o1z;
                        pmeta = (n0 + (R.r1[0] ? BY * BZ : 0) + (R.r1[1] ? BZ : 0) + R.r1[2]) |
                                (R.r1[0] << 8) | (R.r1[1] << 9) | (R.r1[2] << 10) | ((int)s << 11);
                    }
                }
            }

            /* migration by the cell-local offset */
            int shx = (int)floorf(o1x), shy = (int)floorf(o1y), shz = (int)floorf(o1z);
            nxo = o1x - (float)shx;
            nyo = o1y - (float)shy;
            nzo = o1z - (float)shz;
            if (nxo >= 1.0f) { nxo -= 1.0f; shx += 1; }
            if (nyo >= 1.0f) { nyo -= 1.0f; shy += 1; }
            if (nzo >= 1.0f) { nzo -= 1.0f; shz += 1; }
            stays = (shx | shy | shz) == 0;
            if (!stays) {
                atomicAdd(&sLeave[c], 1);
                /* destination in the halo box (|shift| <= 1 per axis); its
                 * rank and the queue slot are claimed together, so the two
                 * LDS round trips overlap (after a queue overflow, which
                 * fails the step, the halo counts include unqueued leavers) */
                const int h = n0 + (BY * BZ + BZ + 1) + __mul24(shx, BY * BZ) +
                              __mul24(shy, BZ) + shz;
                const int hr = atomicAdd(&sHalo[h], 1);
                const int qs_ = atomicAdd(&sQ, 1);
                if (qs_ < a.q.qcap) {
                    const unsigned key = queue_key(s, (unsigned)h, (unsigned)hr);
                    slot(QT, 2u * (unsigned)qs_) = make_float4(nxo, nyo, nzo, pu);
                    slot(QT, 2u * (unsigned)qs_ + 1u) =
                        make_float4(pv, pw, __uint_as_float(key), __int_as_float(pid));
                }
            }
        }

        /* stayers in flat order into the other set */
        const unsigned long long S = __ballot(valid && stays);
        if (valid && stays) {
            const unsigned dst = (unsigned)(wout + wcarry + __popcll(S & ((1ull << lane) - 1ull)));
            slot(O0, dst) = nxo;
            slot(O1, dst) = nyo;
            slot(O2, dst) = nzo;
            slot(O3, dst) = pu;
            slot(O4, dst) = pv;
            slot(O5, dst) = pw;
            slot(OSP, dst) = s;
            if (DIFF) slot(OID, dst) = pid;
        }
        wcarry += __popcll(S);
    }
    if (pend) deposit_pending();
    __syncthreads();

    int maxout = 0;
    if (tid < 64) {   /* counts and packed starts of the tile's cells */
        const int c = tid;
        const int m = c < C ? sLeave[c] : 0;
        const int nc = c < C ? sCnt[c] - m : 0;
        int tot;
        const int x = wave_excl_scan(nc, tot);
        int w = 0;
        for (int k = 1; k < NW; k++) w += c >= sFirst[k] ? 1 : 0;
        const int f = min(sFirst[w], 63);
        const int xf = __shfl(x, f);
        if (c < C) {
            const int ca = c / (TY * TZ), cb = (c / TZ) % TY, cc = c % TZ;
            const int gi = I0 + ca, gl = L0 + cb, gk = K0 + cc;
            if (gi < g.Nx && gl < g.Ny && gk < g.Nz) {
                const long cell = ((long)gi * g.Ny + gl) * g.Nz + gk;
                a.p.count[cell] = nc;       /* the arrivals are added after the push */
                a.p.bcount[cell] = nc;
                a.L.pstart[cell] = sOff[f] + x - xf;
                a.L.arr[cell] = 0;          /* read above; refilled two steps on */
                maxout = m;
            }
        }
    }

    /* the current boxes: thread n adds box node n of the three components
     * (Jx's box stops one x plane short) */
    if (tid < NB) {
        const int u = tid / (BY * BZ), v = (tid / BZ) % BY, w = tid % BZ;
        const int gu = I0 + u, gv = L0 + v, gw = K0 + w;
        const bool in = gu <= g.Nx + 1 && gv <= g.Ny + 1 && gw <= g.Nz + 1;
        const unsigned off = (unsigned)gu * g.sx + (unsigned)gv * g.sy + gw;
        constexpr int J0[3] = {0, JY0, JZ0};
#pragma unroll
        for (int q = 0; q < 3; q++) {
            if (q == 0 && tid >= NBX) continue;
            const int e = J0[q] + tid;
            long long val = sJ[e];
            for (int r = 1; r < JCOPIES; r++) val += sJ[r * JS + e];
            if (RAW) val = (long long)((unsigned long long)val << 16) >> 16;   /* the low 48 bits */
            if (val != 0 && in) add64(&a.f.J[q][off], val);
        }
    }

    if (maxout > 0) {
        if (a.dep_cap > 0 && maxout > a.dep_cap) atomicOr(&sFlags, (unsigned)PIC_F_TOO_MANY);
        if (maxout > a.st->max_dep) atomicMax(&a.st->max_dep, maxout);
    }

    /* The arrival of the tile's leavers (GPU_ArrangeFlights / Cell::Insert,
     * kernels.cu:316-370), here, where the kernel is bound by its
     * instruction issue rather than by memory: one claim per destination
     * cell on its arrival counter (the leavers were counted and ranked per
     * halo cell above), then each record, written to the tile's queue in
     * the loop and read back from the cache, takes arrival index r =
     * base + rank: inbox record r of its cell while r < PIC_INBOX (one
     * 32-byte write), else the next record of the destination tile's spill
     * list.  The counters and inboxes written here are not the ones this
     * launch reads.  q.slot keeps r for the host's repair of a full spill
     * list. */
    auto dest_of = [&](int h, long &cell, long &dtile, int &lc) {
        int gx = I0 - 1 + h / (BY * BZ), gy = L0 - 1 + (h / BZ) % BY, gz = K0 - 1 + h % BZ;
        if (gx < 0) gx += g.Nx; else if (gx >= g.Nx) gx -= g.Nx;
        if (gy < 0) gy += g.Ny; else if (gy >= g.Ny) gy -= g.Ny;
        if (gz < 0) gz += g.Nz; else if (gz >= g.Nz) gz -= g.Nz;
        cell = ((long)gx * g.Ny + gy) * g.Nz + gz;
        dtile = ((long)(gx / TX) * a.t.nty + gy / TY) * a.t.ntz + gz / TZ;
        lc = ((gx % TX) * TY + gy % TY) * TZ + gz % TZ;
    };
    for (int h = tid; h < NB; h += PIC_BLOCK) {
        const int n = sHalo[h];
        if (n > 0) {
            long cell, dt;
            int lc;
            dest_of(h, cell, dt, lc);
            sHalo[h] = atomicAdd(&a.L.arr_next[cell], n);   /* now the run's first index */
        }
    }
    __syncthreads();
    const int nq = min(sQ, a.q.qcap);
    const float4 *const QE = QT;
    for (int j = tid; j < nq; j += PIC_BLOCK) {
        const float4 A = QE[2 * j];
        float4 B = QE[2 * j + 1];
        const unsigned key = __float_as_uint(B.z);
        const int h = key_halo(key);
        if (h >= NB) {   /* (the loop above writes h < NB only) */
            atomicOr(&sFlags, (unsigned)PIC_F_QUEUE_CORRUPT);
            a.q.slot[tile * (long)a.q.qcap + j] = -1;
            continue;
        }
        const int r = sHalo[h] + key_rank(key);
        a.q.slot[tile * (long)a.q.qcap + j] = r;
        long cell, dt;
        int lc;
        dest_of(h, cell, dt, lc);
        if (r < PIC_INBOX) {
            float4 *d = a.L.inbox_next + 2 * (cell * PIC_INBOX + r);
            d[0] = A;
            d[1] = B;
        } else {
            const int k = atomicAdd(&a.L.spill_u_cnt[dt], 1);
            if (k < a.L.scap) {
                B.z = __uint_as_float((key & 255u) | ((unsigned)lc << 8) |
                                      ((unsigned)(r - PIC_INBOX) << 16));
                float4 *d = a.L.spill_u + 2 * (dt * a.L.scap + k);
                d[0] = A;
                d[1] = B;
            } else {
                atomicOr(&sFlags, (unsigned)PIC_F_SPILL_OVER);
            }
        }
    }
    __syncthreads();
    if (tid == 0) {
        if (sQ > a.q.qcap) atomicOr(&sFlags, (unsigned)PIC_F_QUEUE_OVERFLOW);
        a.q.count[tile] = nq;
        if (sFlags) atomicOr(&a.st->flags, sFlags);
    }
}


__global__ __launch_bounds__(PIC_BLOCK) void k_scatter_load32(const long *off, long ncell,
                                                             const float *s0, const float *s1,
                                                             const float *s2, const float *s3,
                                                             const float *s4, const float *s5,
                                                             const unsigned char *ssp,
                                                             PartPtrs32 p) {
    const long e = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (e >= ncell * p.cap) return;
    const long c = e / p.cap;
    const int j = (int)(e % p.cap);
    const long n = off[c + 1] - off[c];
    if (j == 0) {
        p.count[c] = (int)n;
        if (p.bcount) p.bcount[c] = (int)n;
    }
    if (j >= n) return;
    const long src = off[c] + j;
    p.p[0][e] = s0[src];
    p.p[1][e] = s1[src];
    p.p[2][e] = s2[src];
    p.p[3][e] = s3[src];
    p.p[4][e] = s4[src];
    p.p[5][e] = s5[src];
    p.sp[e] = ssp[src];
}

/* bulk cell of halo-box cell h of a tile (periodic wrap) */
template <int TX, int TY, int TZ>
__device__ __forceinline__ long halo_cell(const Grid &g, const TileGeom &t, long tile, int h) {
    constexpr int BY = TY + 2, BZ = TZ + 2;
    const int tz = (int)(tile % t.ntz), ty = (int)((tile / t.ntz) % t.nty),
              tx = (int)(tile / ((long)t.ntz * t.nty));
    int gx = tx * TX - 1 + h / (BY * BZ), gy = ty * TY - 1 + (h / BZ) % BY,
        gz = tz * TZ - 1 + h % BZ;
    if (gx < 0) gx += g.Nx; else if (gx >= g.Nx) gx -= g.Nx;
    if (gy < 0) gy += g.Ny; else if (gy >= g.Ny) gy -= g.Ny;
    if (gz < 0) gz += g.Nz; else if (gz >= g.Nz) gz -= g.Nz;
    return ((long)gx * g.Ny + gy) * g.Nz + gz;
}

/* one departure record into slot `to` of the arrays of p */
__device__ __forceinline__ void place32_at(PartPtrs32 &p, long to, float4 A, float4 B,
                                           unsigned key) {
    p.p[0][to] = A.x;
    p.p[1][to] = A.y;
    p.p[2][to] = A.z;
    p.p[3][to] = A.w;
    p.p[4][to] = B.x;
    p.p[5][to] = B.y;
    p.sp[to] = (unsigned char)(key & 255u);
    if (p.id) p.id[to] = __float_as_int(B.w);
}
/* ... into bucket slot `slot` of `cell`, or into compact slot off[cell] + slot */
__device__ __forceinline__ void place32(PartPtrs32 &p, long cell, int slot, float4 A, float4 B,
                                        unsigned key, const long *off = nullptr) {
    place32_at(p, off ? off[cell] + slot : cell * p.cap + slot, A, B, key);
}

/* tile and local cell of bulk cell `cell` */
template <int TX, int TY, int TZ>
__device__ __forceinline__ void cell_tile(const Grid &g, const TileGeom &t, long cell, long &tile,
                                          int &lc) {
    const int gi = (int)(cell / ((long)g.Ny * g.Nz)), gl = (int)((cell / g.Nz) % g.Ny),
              gk = (int)(cell % g.Nz);
    tile = ((long)(gi / TX) * t.nty + gl / TY) * t.ntz + gk / TZ;
    lc = ((gi % TX) * TY + gl % TY) * TZ + gk % TZ;
}

/* after a push, one workgroup per tile: each cell holds its stayers and its
 * arrivals (a count past the capacity raises PIC_F_BUCKET_OVERFLOW: the host
 * unpacks into larger buckets before the next push), and the tile's spill
 * records move from the unsorted list into the sorted one: cell c's record
 * with arrival index PIC_INBOX + k at the exclusive prefix over the tile's
 * cells of their spill counts, plus k (no atomics: the position follows
 * from the record) */
template <int TX, int TY, int TZ>
__global__ __launch_bounds__(64) void k_arrivals_done32(Grid g, TileGeom t, PartPtrs32 p,
                                                       Layout32 L, DevStatus *st) {
    constexpr int C = TX * TY * TZ;
    static_assert(C <= 64, "one wavefront scans the tile's cells");
    __shared__ int off[C];
    const long tile = blockIdx.x;
    const int tz = (int)(tile % t.ntz), ty = (int)((tile / t.ntz) % t.nty),
              tx = (int)(tile / ((long)t.ntz * t.nty));
    if (threadIdx.x < 64) {
        const int c = threadIdx.x;
        int nsp = 0;
        bool over = false;
        if (c < C) {
            const int gi = tx * TX + c / (TY * TZ), gl = ty * TY + (c / TZ) % TY,
                      gk = tz * TZ + c % TZ;
            if (gi < g.Nx && gl < g.Ny && gk < g.Nz) {
                const long cell = ((long)gi * g.Ny + gl) * g.Nz + gk;
                const int na = L.arr[cell];
                const int n = p.bcount[cell] + na;
                p.count[cell] = n;
                over = n > p.cap;
                nsp = max(0, na - PIC_INBOX);
            }
        }
        int tot;
        const int o = wave_excl_scan(nsp, tot);
        if (c < C) off[c] = o;
        if (__ballot(over) && c == 0) atomicOr(&st->flags, (unsigned)PIC_F_BUCKET_OVERFLOW);
    }
    __syncthreads();
    const int n = min(L.spill_u_cnt[tile], L.scap);
    const float4 *const U = L.spill_u + 2 * (tile * (long)L.scap);
    float4 *const S = L.spill_s + 2 * (tile * (long)L.scap);
    for (int k = threadIdx.x; k < n; k += 64) {
        const float4 A = U[2 * k], B = U[2 * k + 1];
        const unsigned key = __float_as_uint(B.z);
        const int lc = key_halo(key);
        if (lc >= C) continue;
        const int d = off[lc] + (int)(key >> 16);
        if (d < L.scap) {
            S[2 * d] = A;
            S[2 * d + 1] = B;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) L.spill_u_cnt[tile] = 0;
}

/* the packed live set into buckets of `to` (capacity to.cap >= every
 * count), or into the compact arrays of `to` at off[cell] + slot when off is
 * given: slot j of a cell from its tile's region (j < bcount) or from inbox
 * record j - bcount */
template <int TX, int TY, int TZ>
__global__ __launch_bounds__(PIC_BLOCK) void k_unpack32(Grid g, TileGeom t, long ncell,
                                                       PartPtrs32 p, Layout32 L, PartPtrs32 to,
                                                       const long *off) {
    const long i = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (i >= ncell * to.cap) return;
    const long cell = i / to.cap;
    const int j = (int)(i % to.cap);
    const long e = off ? off[cell] + j : i;
    const int cnt = min(p.count[cell], to.cap), bc = max(0, min(p.bcount[cell], cnt));
    const int nin = min(PIC_INBOX, cnt - bc);
    if (j < bc) {
        long tile;
        int lc;
        cell_tile<TX, TY, TZ>(g, t, cell, tile, lc);
        const long src = tile * L.tcap + L.pstart[cell] + j;
        for (int k = 0; k < 6; k++) to.p[k][e] = p.p[k][src];
        to.sp[e] = p.sp[src];
        if (to.id) to.id[e] = p.id[src];
    } else if (j < bc + nin) {
        const long rec = 2 * (cell * PIC_INBOX + (j - bc));
        const float4 A = p.inbox[rec], B = p.inbox[rec + 1];
        place32_at(to, e, A, B, __float_as_uint(B.z));
    }
}

/* the sorted spill lists into buckets, one workgroup per tile: cell c's
 * k-th spill record (at the exclusive prefix over the tile's cells of their
 * spill counts, plus k) takes slot bcount + PIC_INBOX + k */
template <int TX, int TY, int TZ>
__global__ __launch_bounds__(PIC_BLOCK) void k_unspill32(Grid g, TileGeom t, PartPtrs32 p,
                                                        Layout32 L, PartPtrs32 to,
                                                        const long *coff) {
    constexpr int C = TX * TY * TZ;
    static_assert(C <= 64, "one wavefront scans the tile's cells");
    __shared__ int off[C + 1], bcs[C];
    __shared__ long cellof[C];
    const long tile = blockIdx.x;
    const int tz = (int)(tile % t.ntz), ty = (int)((tile / t.ntz) % t.nty),
              tx = (int)(tile / ((long)t.ntz * t.nty));
    if (threadIdx.x < 64) {
        const int c = threadIdx.x;
        int nsp = 0;
        if (c < C) {
            const int gi = tx * TX + c / (TY * TZ), gl = ty * TY + (c / TZ) % TY,
                      gk = tz * TZ + c % TZ;
            long cell = -1;
            if (gi < g.Nx && gl < g.Ny && gk < g.Nz) {
                cell = ((long)gi * g.Ny + gl) * g.Nz + gk;
                const int cnt = min(p.count[cell], to.cap), bc = max(0, min(p.bcount[cell], cnt));
                nsp = max(0, cnt - bc - PIC_INBOX);
                bcs[c] = bc;
            }
            cellof[c] = cell;
        }
        int tot;
        const int o = wave_excl_scan(nsp, tot);
        if (c < C) off[c] = o;
        if (c == 0) off[C] = tot;
    }
    __syncthreads();
    const int n = min(off[C], L.scap);
    const float4 *const S = L.spill_s + 2 * (tile * (long)L.scap);
    for (int k = threadIdx.x; k < n; k += PIC_BLOCK) {
        int lo = 0, hi = C;   /* the cell whose range holds k: off[lo] <= k < off[lo + 1] */
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (off[mid] <= k) lo = mid; else hi = mid;
        }
        if (cellof[lo] < 0) continue;
        const float4 A = S[2 * k], B = S[2 * k + 1];
        place32(to, cellof[lo], bcs[lo] + PIC_INBOX + (k - off[lo]), A, B, __float_as_uint(B.z),
                coff);
    }
}

/* the arrivals past the inbox from the departure queues of the last push
 * (q.slot = arrival index r) into bucket slots bcount + r of `to`: the
 * repair after a spill list ran full, which reads every such arrival from
 * its queue record */
template <int TX, int TY, int TZ>
__global__ __launch_bounds__(PIC_BLOCK) void k_place_spills32(Grid g, TileGeom t, PartPtrs32 p,
                                                             QueuePtrs32 q, PartPtrs32 to) {
    const long tile = blockIdx.x;
    const int n = min(q.count[tile], q.qcap);
    const long qb = tile * (long)q.qcap;
    for (int j = threadIdx.x; j < n; j += PIC_BLOCK) {
        const int r = q.slot[qb + j];
        if (r < PIC_INBOX) continue;
        const float4 A = q.e[2 * (qb + j)], B = q.e[2 * (qb + j) + 1];
        const unsigned key = __float_as_uint(B.z);
        const long cell = halo_cell<TX, TY, TZ>(g, t, tile, key_halo(key));
        const int slot = p.bcount[cell] + r;
        if (slot < min(p.count[cell], to.cap)) place32(to, cell, slot, A, B, key);
    }
}

/* after an unpack: every particle is in the bucket arrays (bcount =
 * count), no arrival is pending */
__global__ __launch_bounds__(PIC_BLOCK) void k_unpack_done32(long ncell, long ntiles,
                                                            PartPtrs32 p, Layout32 L, int cap) {
    const long e = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (e < ncell) {
        p.bcount[e] = min(p.count[e], cap);
        L.arr[e] = 0;
    }
    if (e < ntiles) L.spill_u_cnt[e] = 0;
}

__global__ void k_clear_flags32(DevStatus *st, unsigned mask) { st->flags &= ~mask; }

/* global coordinates x = (i + o) h in FP64 */
__global__ __launch_bounds__(PIC_BLOCK) void k_gather_particles32(Grid g, long ncell,
                                                                 const long *off, PartPtrs32 p,
                                                                 double *d0, double *d1,
                                                                 double *d2, double *d3,
                                                                 double *d4, double *d5,
                                                                 unsigned char *dsp, int *dcell,
                                                                 int *did) {
    const long e = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (e >= ncell * p.cap) return;
    const long c = e / p.cap;
    const int j = (int)(e % p.cap);
    if (j >= off[c + 1] - off[c]) return;
    const long dst = off[c] + j;
    const int i = (int)(c / ((long)g.Ny * g.Nz)), l = (int)((c / g.Nz) % g.Ny),
              k = (int)(c % g.Nz);
    d0[dst] = ((double)i + (double)p.p[0][e]) * g.hx;
    d1[dst] = ((double)l + (double)p.p[1][e]) * g.hy;
    d2[dst] = ((double)k + (double)p.p[2][e]) * g.hz;
    d3[dst] = p.p[3][e];
    d4[dst] = p.p[4][e];
    d5[dst] = p.p[5][e];
    dsp[dst] = p.sp[e];
    if (dcell) dcell[dst] = (int)c;
    if (did) did[dst] = p.id ? p.id[e] : -1;
}

__global__ __launch_bounds__(PIC_BLOCK) void k_gather_raw32(long ncell, const long *off,
                                                           PartPtrs32 p, float *d0, float *d1,
                                                           float *d2, float *d3, float *d4,
                                                           float *d5, unsigned char *dsp,
                                                           int *did) {
    const long e = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (e >= ncell * p.cap) return;
    const long c = e / p.cap;
    const int j = (int)(e % p.cap);
    if (j >= off[c + 1] - off[c]) return;
    const long dst = off[c] + j;
    d0[dst] = p.p[0][e];
    d1[dst] = p.p[1][e];
    d2[dst] = p.p[2][e];
    d3[dst] = p.p[3][e];
    d4[dst] = p.p[4][e];
    d5[dst] = p.p[5][e];
    dsp[dst] = p.sp[e];
    if (did) did[dst] = p.id[e];
}

__global__ __launch_bounds__(PIC_BLOCK) void k_scatter_ids32(const long *off, long ncell,
                                                            const int *src, PartPtrs32 p) {
    const long e = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (e >= ncell * p.cap) return;
    const long c = e / p.cap;
    const int j = (int)(e % p.cap);
    if (j >= off[c + 1] - off[c]) return;
    p.id[e] = src[off[c] + j];
}

__global__ __launch_bounds__(PIC_BLOCK) void k_moments32(Grid g, long ncell, PartPtrs32 p,
                                                        SpeciesTable sp, const int *slot_sort,
                                                        double *partials) {
    __shared__ double red[PIC_BLOCK];
    double acc[27];
    for (int k = 0; k < 27; k++) acc[k] = 0.0;
    for (long e = (long)blockIdx.x * PIC_BLOCK + threadIdx.x; e < ncell * p.cap;
         e += (long)gridDim.x * PIC_BLOCK) {
        const long c = e / p.cap;
        const int j = (int)(e % p.cap);
        if (j >= p.count[c]) continue;
        const int s = p.sp[e];
        int srt = slot_sort[s];
        srt = srt < 0 ? 0 : (srt > 2 ? 2 : srt);
        const int i = (int)(c / ((long)g.Ny * g.Nz)), l = (int)((c / g.Nz) % g.Ny),
                  k = (int)(c % g.Nz);
        const double px = p.p[3][e], py = p.p[4][e], pz = p.p[5][e];
        const double gm = sqrt(1.0 + px * px + py * py + pz * pz);
        double v[9] = {1.0,
                       ((double)i + (double)p.p[0][e]) * g.hx,
                       ((double)l + (double)p.p[1][e]) * g.hy,
                       ((double)k + (double)p.p[2][e]) * g.hz,
                       px, py, pz,
                       (sp.m[s] / sp.q_m[s]) * (gm - 1.0),
                       fabs(px) + fabs(py) + fabs(pz)};
        for (int q = 0; q < 9; q++) {
            acc[0 * 9 + q] += (srt == 0) ? v[q] : 0.0;
            acc[1 * 9 + q] += (srt == 1) ? v[q] : 0.0;
            acc[2 * 9 + q] += (srt == 2) ? v[q] : 0.0;
        }
    }
    for (int q = 0; q < 27; q++) {
        red[threadIdx.x] = acc[q];
        __syncthreads();
        for (int w = PIC_BLOCK / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) partials[(long)blockIdx.x * 27 + q] = red[0];
        __syncthreads();
    }
}

__global__ void k_moments_final32(const double *partials, int nblocks, double *out27) {
    const int k = threadIdx.x;
    if (k >= 27) return;
    double s = 0.0;
    for (int b = 0; b < nblocks; b++) s += partials[(long)b * 27 + k];
    out27[k] = s;
}

inline unsigned blocks_for(long n) { return (unsigned)((n + PIC_BLOCK - 1) / PIC_BLOCK); }

}  // namespace

static bool raw_current_sums_exact(const PushArgs32 &a, int cap);

void launch_push_deposit_f32(const Grid &g, const TileGeom &t, FieldPtrs32 f, PartPtrs32 p,
                             PartPtrs32 out, Layout32 L, QueuePtrs32 q, SpeciesTable sp,
                             int dep_cap, DevStatus *st, const PartPtrs32 *diff,
                             const long *snap_off, long snap_room, hipStream_t s) {
    PushArgs32 a;
    a.g = g;
    a.t = t;
    a.f = f;
    a.p = p;
    a.o = out;
    a.L = L;
    a.q = q;
    const double h[3] = {g.hx, g.hy, g.hz};
    for (int k = 0; k < PIC_MAX_SLOTS; k++) {
        a.qm[k] = (float)sp.q_m[k];
        for (int d = 0; d < 3; d++) a.cj[k][d] = (float)(sp.m[k] * h[d] / g.tau * PIC_J_SCALE);
    }
    for (int d = 0; d < 3; d++) {
        a.tdh[d] = (float)(g.tau / h[d]);
        a.hh[d] = (float)(h[d] * h[d]);
    }
    a.tau = (float)g.tau;
    a.dep_cap = dep_cap;
    a.st = st;
    a.snap = diff ? *diff : PartPtrs32{};
    a.snap_on = diff && diff->p[0] != nullptr;
    a.snap_off = snap_off;
    a.snap_room = snap_room;
    /* dynamic LDS: one 4-byte entry per group.  Per cell of count <= cap:
     * particle-array lines at most ceil(cap/32) + 1 (buckets: the lines of
     * [0, bc) and of [bc + PIC_INBOX, cnt), together cnt - PIC_INBOX slots
     * when the inbox is full; packed: a run of bc slots that may start
     * inside a line), inbox groups at most PIC_INBOX / IBG, spill groups at
     * most nsp / SPG + 1 (at most scap / SPG + 1 for the tile's
     * cells together beside the per-cell one) */
    const size_t lds = (size_t)4 * (TXC * TYC * TZC * ((p.cap + GS32 - 1) / GS32 + 1 +
                                                     PIC_INBOX / IBG + 1) +
                                    L.scap / SPG + 1);
    a.ngrp = (int)(lds / 4);
    if (diff)
        hipLaunchKernelGGL((k_push_deposit32<TXC, TYC, TZC, true, false>),
                           dim3((unsigned)t.ntiles), dim3(PIC_BLOCK), lds, s, a);
    else if (raw_current_sums_exact(a, p.cap))
        hipLaunchKernelGGL((k_push_deposit32<TXC, TYC, TZC, false, true>),
                           dim3((unsigned)t.ntiles), dim3(PIC_BLOCK), lds, s, a);
    else
        hipLaunchKernelGGL((k_push_deposit32<TXC, TYC, TZC, false, false>),
                           dim3((unsigned)t.ntiles), dim3(PIC_BLOCK), lds, s, a);
}

/* Whether the push may sum its current contributions as raw offset words
 * (fixmul<true>): every sum it forms in the LDS J box must stay below 2^47
 * units of 2^-40 in magnitude.  A node of the box takes contributions from
 * the particles of at most 18 cells of the tile (2 x 3 x 3: the cells whose
 * dual cell at the start or the end covers it), at most cap per cell; one
 * particle's contributions to one node, over its two segments, come to at
 * most |cj| |du| (1 + 1/48) < |cj| / 2 x 13/12 (|du| < 1/2 per axis, a
 * weight at most 1 + 1/48), plus 1/2 unit of rounding for each of the two. */
static bool raw_current_sums_exact(const PushArgs32 &a, int cap) {
    double cjmax = 0.0;
    for (int k = 0; k < PIC_MAX_SLOTS; k++)
        for (int d = 0; d < 3; d++) cjmax = std::max(cjmax, std::fabs((double)a.cj[k][d]));
    const double bound = 18.0 * (double)cap * (0.5 * (13.0 / 12.0) * cjmax + 1.0);
    return bound < 0.99 * 140737488355328.0;   /* 2^47 */
}

void launch_arrivals_done32(const Grid &g, const TileGeom &t, PartPtrs32 p, Layout32 L,
                            DevStatus *st, hipStream_t s) {
    hipLaunchKernelGGL((k_arrivals_done32<TXC, TYC, TZC>), dim3((unsigned)t.ntiles), dim3(64),
                       0, s, g, t, p, L, st);
}

void launch_unpack32(const Grid &g, const TileGeom &t, PartPtrs32 p, Layout32 L, PartPtrs32 to,
                     const QueuePtrs32 *from_queue, hipStream_t s) {
    const long ncell = (long)g.Nx * g.Ny * g.Nz;
    if (from_queue)
        hipLaunchKernelGGL((k_place_spills32<TXC, TYC, TZC>), dim3((unsigned)t.ntiles),
                           dim3(PIC_BLOCK), 0, s, g, t, p, *from_queue, to);
    hipLaunchKernelGGL((k_unpack32<TXC, TYC, TZC>), dim3(blocks_for(ncell * to.cap)),
                       dim3(PIC_BLOCK), 0, s, g, t, ncell, p, L, to, (const long *)nullptr);
    if (!from_queue)
        hipLaunchKernelGGL((k_unspill32<TXC, TYC, TZC>), dim3((unsigned)t.ntiles),
                           dim3(PIC_BLOCK), 0, s, g, t, p, L, to, (const long *)nullptr);
    hipLaunchKernelGGL(k_unpack_done32, dim3(blocks_for(std::max(ncell, t.ntiles))),
                       dim3(PIC_BLOCK), 0, s, ncell, t.ntiles, p, L, to.cap);
}

void launch_gather_packed32(const Grid &g, const TileGeom &t, PartPtrs32 p, Layout32 L,
                            const long *offsets, float *dst[6], unsigned char *dsp, int *did,
                            hipStream_t s) {
    const long ncell = (long)g.Nx * g.Ny * g.Nz;
    PartPtrs32 to{};
    for (int k = 0; k < 6; k++) to.p[k] = dst[k];
    to.sp = dsp;
    to.id = did;
    to.cap = p.cap;
    hipLaunchKernelGGL((k_unpack32<TXC, TYC, TZC>), dim3(blocks_for(ncell * to.cap)),
                       dim3(PIC_BLOCK), 0, s, g, t, ncell, p, L, to, offsets);
    hipLaunchKernelGGL((k_unspill32<TXC, TYC, TZC>), dim3((unsigned)t.ntiles), dim3(PIC_BLOCK),
                       0, s, g, t, p, L, to, offsets);
}

void launch_clear_arrivals32(const Grid &g, const TileGeom &t, PartPtrs32 p, Layout32 L,
                             hipStream_t s) {
    const long ncell = (long)g.Nx * g.Ny * g.Nz;
    hipLaunchKernelGGL(k_unpack_done32, dim3(blocks_for(std::max(ncell, t.ntiles))),
                       dim3(PIC_BLOCK), 0, s, ncell, t.ntiles, p, L, p.cap);
}

void launch_clear_flags32(DevStatus *st, unsigned mask, hipStream_t s) {
    hipLaunchKernelGGL(k_clear_flags32, dim3(1), dim3(1), 0, s, st, mask);
}

void launch_scatter_load32(const long *offsets, long ncell, const float *src[6],
                           const unsigned char *srcsp, PartPtrs32 p, hipStream_t s) {
    hipLaunchKernelGGL(k_scatter_load32, dim3(blocks_for(ncell * p.cap)), dim3(PIC_BLOCK), 0, s,
                       offsets, ncell, src[0], src[1], src[2], src[3], src[4], src[5], srcsp, p);
}

void launch_gather_particles32(const Grid &g, long ncell, const long *offsets, PartPtrs32 p,
                               double *dst[6], unsigned char *dsp, int *dcell, int *did,
                               hipStream_t s) {
    hipLaunchKernelGGL(k_gather_particles32, dim3(blocks_for(ncell * p.cap)), dim3(PIC_BLOCK), 0,
                       s, g, ncell, offsets, p, dst[0], dst[1], dst[2], dst[3], dst[4], dst[5],
                       dsp, dcell, did);
}

void launch_gather_raw32(long ncell, const long *offsets, PartPtrs32 p, float *dst[6],
                         unsigned char *dsp, int *did, hipStream_t s) {
    hipLaunchKernelGGL(k_gather_raw32, dim3(blocks_for(ncell * p.cap)), dim3(PIC_BLOCK), 0, s,
                       ncell, offsets, p, dst[0], dst[1], dst[2], dst[3], dst[4], dst[5], dsp,
                       did);
}

void launch_scatter_ids32(const long *offsets, long ncell, const int *src, PartPtrs32 p,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_scatter_ids32, dim3(blocks_for(ncell * p.cap)), dim3(PIC_BLOCK), 0, s,
                       offsets, ncell, src, p);
}

void launch_moments32(const Grid &g, long ncell, PartPtrs32 p, SpeciesTable sp,
                      const int *slot_sort, double *partials, int nblocks, double *out27,
                      hipStream_t s) {
    hipLaunchKernelGGL(k_moments32, dim3(nblocks), dim3(PIC_BLOCK), 0, s, g, ncell, p, sp,
                       slot_sort, partials);
    hipLaunchKernelGGL(k_moments_final32, dim3(1), dim3(32), 0, s, partials, nblocks, out27);
}
