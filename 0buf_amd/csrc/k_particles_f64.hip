/*
 * k_particles_f64.hip -- particle stages of the FP64 plain form.
 *
 * One workgroup per tile of TX x TY x TZ cells:
 *   1. stage E and H^{n+1/2} of the tile's node box ((TX+2)(TY+2)(TZ+2)
 *      nodes: a cell's stencil reaches global nodes owner+0..owner+2) in LDS;
 *   2. each wavefront owns whole cells of the tile and walks their particles
 *      in flat order, in chunks of whole 128-byte line groups (consecutive
 *      lanes -> consecutive slots of a cell bucket, so loads coalesce and no
 *      line is shared by two chunks); a lane gathers, pushes, and deposits
 *      its charge-conserving current into an LDS copy of the tile's J box
 *      (ds_add_f64);
 *   3. the particle is wrapped, its fly direction and receiving cells are
 *      computed with the reference predicates; leavers go to the tile's
 *      departure region, stayers are compacted in place in flat order;
 *   4. the J box is added to the global arrays with FP64 atomics.
 * The arithmetic mirrors cell.h / particle.h in the reference's evaluation
 * order and is built with -ffp-contract=off.
 */
#include "k_arrive.h"
#include "k_tape.h"
#include "k_push_math.h"

#pragma clang fp contract(off)

namespace {

/* copies of the LDS J box, lane l depositing into copy l % JCOPIES: fewer
 * same-address conflicts of the LDS atomics (consecutive lanes hold
 * particles of one cell); the copies are summed in a fixed order before the
 * global flush */
#ifndef PIC_JCOPIES64
#define PIC_JCOPIES64 2
#endif

constexpr int TXC = 2, TYC = 4, TZC = 8;
constexpr int GS64 = 16;    /* slots of an FP64 array in one 128-byte line */
constexpr int JCOPIES = PIC_JCOPIES64;

struct PushArgs {
    Grid g;
    DivConst dc;
    TileGeom t;
    FieldPtrs f;
    PartPtrs p;
    QueuePtrs q;
    SpeciesTable sp;
    int dep_cap;
    DevStatus *st;
    DiffArgs d;        /* differentiable form (DIFF instantiation only) */
    int snap_on;
};

/* one node contribution to the LDS J box: FP64 atomic add, or in the
 * differentiable form the contribution v 2^48 rounded to an integer with the
 * 1.5 2^52 shifter (exact for |v| < 8, flagged otherwise) and added as int64 */
template <bool DIFF>
__device__ __forceinline__ void jadd(double *sJ, int n, double v, unsigned &bad) {
    if constexpr (DIFF) {
        const double t = fma(v, PIC_J_SCALE64, 6755399441055744.0);
        const long long q = __double_as_longlong(t) - 0x4338000000000000LL;
        if (!(fabs(v) < 8.0)) bad = 1;
        atomicAdd(reinterpret_cast<unsigned long long *>(sJ) + n, (unsigned long long)q);
    } else {
        atomicAdd(&sJ[n], v);
    }
}

/* cell.h:1511-1693 pqr: 12 node contributions of one straight segment,
 * accumulated into the LDS J box at local node (owner + tile index - 2). */
template <int BY, int BZ, int NB, bool DIFF>
__device__ __forceinline__ void pqr_lds(double *sJ, const Geo &c, const DivConst &dc, int ix,
                                        int iy, int iz, double xx, double xy, double xz,
                                        double x1x, double x1y, double x1z, double a,
                                        unsigned &bad, bool active) {
    if (!active) return;
    if (ix < 2 || ix > 3 || iy < 2 || iy > 3 || iz < 2 || iz > 3) {
        bad = 1;
        return;
    }
    double dx = (0.50 * (xx + x1x) - c.hx * ((double)ix - 2.50) - c.x0);
    double dy = (0.50 * (xy + x1y) - c.hy * ((double)iy - 2.50) - c.y0);
    double dz = (0.50 * (xz + x1z) - c.hz * ((double)iz - 2.50) - c.z0);
    double dx1 = c.hx - dx, dy1 = c.hy - dy, dz1 = c.hz - dz;
    double su = x1x - xx, sv = x1y - xy, sw = x1z - xz;
    double s1 = divc(sv * sw, 12.0, dc.r12), s2 = divc(su * sw, 12.0, dc.r12),
           s3 = divc(su * sv, 12.0, dc.r12);
    su = divc(su * a, dc.Dx, dc.rDx);
    sv = divc(sv * a, dc.Dy, dc.rDy);
    sw = divc(sw * a, dc.Dz, dc.rDz);
    const int u = c.ca + ix - 2, v = c.cb + iy - 2, w = c.cc + iz - 2;
    const int n = (u * BY + v) * BZ + w;
    double *Jx = sJ, *Jy = sJ + NB, *Jz = sJ + 2 * NB;
    jadd<DIFF>(Jx, n, su * (dy1 * dz1 + s1), bad);
    jadd<DIFF>(Jx, n + 1, su * (dy1 * dz - s1), bad);
    jadd<DIFF>(Jx, n + BZ, su * (dy * dz1 - s1), bad);
    jadd<DIFF>(Jx, n + BZ + 1, su * (dy * dz + s1), bad);
    jadd<DIFF>(Jy, n, sv * (dx1 * dz1 + s2), bad);
    jadd<DIFF>(Jy, n + 1, sv * (dx1 * dz - s2), bad);
    jadd<DIFF>(Jy, n + BY * BZ, sv * (dx * dz1 - s2), bad);
    jadd<DIFF>(Jy, n + BY * BZ + 1, sv * (dx * dz + s2), bad);
    jadd<DIFF>(Jz, n, sw * (dx1 * dy1 + s3), bad);
    jadd<DIFF>(Jz, n + BZ, sw * (dx1 * dy - s3), bad);
    jadd<DIFF>(Jz, n + BY * BZ, sw * (dx * dy1 - s3), bad);
    jadd<DIFF>(Jz, n + BY * BZ + BZ, sw * (dx * dy + s3), bad);
}

/* 4 waves per SIMD (at most 128 VGPRs): the FP64 chains of the push are
 * latency-bound at 3 */
template <int TX, int TY, int TZ, bool DIFF>
__global__ __launch_bounds__(PIC_BLOCK, 4) void k_push_deposit(PushArgs a) {
    constexpr int C = TX * TY * TZ;
    constexpr int BX = TX + 2, BY = TY + 2, BZ = TZ + 2, NB = BX * BY * BZ;
    __shared__ double sF[6 * NB];
    __shared__ double sJ[JCOPIES * 3 * NB];
    __shared__ int sOff[C + 1];
    __shared__ int sCnt[C];
    __shared__ int sLeave[C];
    __shared__ int sOut[C];
    __shared__ int sBase[C];     /* first slot of each cell, relative to the tile's first cell */
    __shared__ int sG[C + 1];    /* first line group of each cell */
    __shared__ int sQ;
    __shared__ unsigned sFlags;
    __shared__ unsigned long long sDrop, sDup;
    __shared__ double2 sMQ[PIC_MAX_SLOTS];   /* (m, q_m) per species slot */
    extern __shared__ unsigned sGrp[];    /* line groups of the tile */
    double *const myJ = sJ + (threadIdx.x % JCOPIES) * (3 * NB);

    const Grid &g = a.g;
    const int tid = threadIdx.x;
    const bool snap = DIFF && a.snap_on;

    /* XCD-aware tile order: workgroups are dealt to the 8 XCDs round-robin,
     * so each XCD gets a contiguous run of neighbouring tiles and their
     * shared halo nodes stay in that XCD's L2. */
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

    /* Thread n takes box node n of the six field boxes (one decode, one
     * 32-bit offset: the grid has fewer than 2^31 nodes) and cell n of the
     * tile; every global load is issued before the first LDS write. */
    static_assert(NB <= PIC_BLOCK && C <= PIC_BLOCK, "one box node and one cell per thread");
    const int nu = tid / (BY * BZ), nv = (tid / BZ) % BY, nw = tid % BZ;
    const bool node_in = tid < NB && I0 + nu <= g.Nx + 1 && L0 + nv <= g.Ny + 1 &&
                         K0 + nw <= g.Nz + 1;
    const unsigned noff = (unsigned)(I0 + nu) * g.sx + (unsigned)(L0 + nv) * g.sy + (K0 + nw);
    {
        double fv[6];
#pragma unroll
        for (int q = 0; q < 6; q++) fv[q] = 0.0;
        if (node_in) {
#pragma unroll
            for (int q = 0; q < 6; q++) fv[q] = a.f.f[PIC_EX + q][noff];
        }
        const int ca = tid / (TY * TZ), cb = (tid / TZ) % TY, cc = tid % TZ;
        const int gi = I0 + ca, gl = L0 + cb, gk = K0 + cc;
        int cnt = 0;
        if (tid < C && gi < g.Nx && gl < g.Ny && gk < g.Nz) {
            const long cell = ((long)gi * g.Ny + gl) * g.Nz + gk;
            cnt = a.p.count[cell];
            cnt = cnt < cap ? cnt : cap;
            if (snap) a.d.snap.count[cell] = cnt;
        }
        if (tid < NB) {
#pragma unroll
            for (int q = 0; q < 6; q++) sF[q * NB + tid] = fv[q];
        }
        if (tid < C) {
            sCnt[tid] = cnt;
            sLeave[tid] = 0;
            sOut[tid] = 0;
            sBase[tid] = ((ca * g.Ny + cb) * g.Nz + cc) * cap;
        }
    }
    for (int e = tid; e < JCOPIES * 3 * NB; e += PIC_BLOCK) sJ[e] = 0.0;
    /* weight and q/m per species slot: a kernel-argument table indexed by a
     * per-particle value would be a vector memory load in the loop, whose
     * wait would also wait for the particle loads in flight */
    if (tid < PIC_MAX_SLOTS) sMQ[tid] = make_double2(a.sp.m[tid], a.sp.q_m[tid]);
    if (tid == 0) {
        sQ = 0;
        sFlags = 0;
        sDrop = 0;
        sDup = 0;
    }
    __syncthreads();
    static_assert(C <= 64, "one wavefront scans the tile's cells");
    if (tid < 64) {   /* first particle and first line group of each cell */
        const int cnt = tid < C ? sCnt[tid] : 0;
        int tot_s, tot_g;
        const int so = wave_excl_scan(cnt, tot_s),
                  sg = wave_excl_scan((cnt + GS64 - 1) / GS64, tot_g);
        if (tid < C) {
            sOff[tid] = so;
            sG[tid] = sg;
        }
        if (tid == 0) {
            sOff[C] = tot_s;
            sG[C] = tot_g;
        }
    }
    __syncthreads();
    build_groups<TX, TY, TZ, GS64>(sCnt, sG, sGrp, tid, PIC_BLOCK);
    __syncthreads();

    /* per-tile bases of the particle arrays: slot rel of the tile is
     * cell-bucket slot pbase + rel */
    const long pbase = (((long)I0 * g.Ny + L0) * g.Nz + K0) * cap;
    double *const P0 = a.p.p[0] + pbase, *const P1 = a.p.p[1] + pbase,
                  *const P2 = a.p.p[2] + pbase, *const P3 = a.p.p[3] + pbase,
                  *const P4 = a.p.p[4] + pbase, *const P5 = a.p.p[5] + pbase;
    unsigned char *const SP = a.p.sp + pbase;
    int *const ID = DIFF ? a.p.id + pbase : nullptr;

    const double hx = g.hx, hy = g.hy, hz = g.hz, tau = g.tau;
    unsigned bad = 0;
    /* each wavefront owns whole cells and walks their particles in chunks of
     * 64 consecutive ones (a line group may be split between two consecutive
     * chunks); stayers are compacted in place in flat order */
    const int lane = tid & 63, wv = tid >> 6;
    const int ghi = __builtin_amdgcn_readfirstlane(
        sG[wave_first_cell<C>(sOff, wv + 1, PIC_BLOCK / 64)]);
    int g0 = __builtin_amdgcn_readfirstlane(sG[wave_first_cell<C>(sOff, wv, PIC_BLOCK / 64)]);
    int o0 = 0;
    int carry_cell = -1, carry_cnt = 0;
    while (g0 < ghi) {
        unsigned long long starts = 0, runs = 0;
        int lanes = 0, onext = 0;
        const int gnext = compose_full_chunk(sGrp, g0, o0, ghi, starts, lanes, runs, onext);
        const bool valid = lane < lanes;
        int c = 0, j = 0;
        bool stays = false;
        double xn = 0, yn = 0, zn = 0, pu = 0, pv = 0, pw = 0;
        unsigned char s = 0;
        int pid = 0;
        tape_word tout = 0;    /* this step's decisions (decision tape, record mode) */
        if (valid) {
            const unsigned long long msk = starts & (~0ull >> (63 - lane));
            const int gidx = __popcll(msk) - 1;
            const unsigned ge = sGrp[g0 + gidx];
            c = group_cell(ge);
            j = group_start(ge) + (lane - (63 - __clzll(msk))) + (gidx == 0 ? o0 : 0);
            const unsigned rel = (unsigned)(sBase[c] + j);
            Geo G;
            G.ca = c / (TY * TZ);
            G.cb = (c / TZ) % TY;
            G.cc = c % TZ;
            const int gi = I0 + G.ca, gl = L0 + G.cb, gk = K0 + G.cc;
            G.hx = hx; G.hy = hy; G.hz = hz; G.tau = tau;
            G.x0 = (double)gi * hx;
            G.y0 = (double)gl * hy;
            G.z0 = (double)gk * hz;

            const double x = slot(P0, rel), y = slot(P1, rel), z = slot(P2, rel);
            pu = slot(P3, rel);
            pv = slot(P4, rel);
            pw = slot(P5, rel);
            s = slot(SP, rel);
            pid = DIFF ? slot(ID, rel) : 0;
            const double2 mq = sMQ[s];
            const double mass = mq.x, q_m = mq.y;
            /* decision tape (differentiable form): record this step's
             * decisions, or replay a recorded run's */
            const bool trec = DIFF && a.d.tape_mode == PIC_TAPE_RECORD;
            bool trep = DIFF && a.d.tape_mode == PIC_TAPE_REPLAY;
            const bool tin = (trec || trep) && pid >= 0 && pid < a.d.tape_n;
            tape_word tw = 0;
            if (trep) {
                if (tin) tw = a.d.tape_row[pid];
                if (!tin || !((tw >> TAPE_VALID) & 1ull) || ((tw >> TAPE_BAD) & 1ull)) bad = 1;
            }
            const long r = snap ? a.d.snap_off[tile] + sOff[c] + j : 0;
            if (snap && r < a.d.snap_room) {
                a.d.snap.p[0][r] = x;
                a.d.snap.p[1][r] = y;
                a.d.snap.p[2][r] = z;
                a.d.snap.p[3][r] = pu;
                a.d.snap.p[4][r] = pv;
                a.d.snap.p[5][r] = pw;
                a.d.snap.sp[r] = s;
                a.d.snap.id[r] = pid;
            }

            const double frx = divc(x - G.x0, hx, a.dc.rhx);
            const double fry = divc(y - G.y0, hy, a.dc.rhy);
            const double frz = divc(z - G.z0, hz, a.dc.rhz);
            double Ex = 0, Ey = 0, Ez = 0, Hx = 0, Hy = 0, Hz = 0;
            const bool neg = trep ? (tw & 1ull) != 0 : (x < 0 || y < 0 || z < 0);
            if (trec && neg) tout |= 1ull;
            if (!neg) {
                GatherIdx W = gather_index(x, y, z, frx, fry, frz, g, a.dc);
                const int gc[3] = {gi, gl, gk};
                if (trep) {
                    int *ii[6] = {&W.ix, &W.i1x, &W.iy, &W.i1y, &W.iz, &W.i1z};
                    for (int q = 0; q < 6; q++) *ii[q] = tape_get(tw, TAPE_GIDX + q, 1, 2);
                    for (int d = 0; d < 3; d++) {
                        W.f1[d] = gc[d] + tape_get(tw, TAPE_FLOOR + 4 * d, 2, 0);
                        W.f15[d] = gc[d] + tape_get(tw, TAPE_FLOOR + 4 * d + 2, 2, 0);
                    }
                    gather_weights(W, divc(x, g.hx, a.dc.rhx), divc(y, g.hy, a.dc.rhy),
                                   divc(z, g.hz, a.dc.rhz));
                    W.ok = gather_ok(W);
                }
                if (trec) {
                    const int ii[6] = {W.ix, W.i1x, W.iy, W.i1y, W.iz, W.i1z};
                    for (int q = 0; q < 6; q++) tape_put(tout, TAPE_GIDX + q, 1, ii[q], 2);
                    for (int d = 0; d < 3; d++) {
                        tape_put(tout, TAPE_FLOOR + 4 * d, 2, W.f1[d] - gc[d], 0);
                        tape_put(tout, TAPE_FLOOR + 4 * d + 2, 2, W.f15[d] - gc[d], 0);
                    }
                }
                const double s1 = W.s1, s2 = W.s2, s3 = W.s3, s4 = W.s4, s5 = W.s5, s6 = W.s6;
                const double s11 = 1. - s1, s21 = 1. - s2, s31 = 1. - s3, s41 = 1. - s4,
                             s51 = 1. - s5, s61 = 1. - s6;
                if (!W.ok) {
                    bad = 1;
                } else {
                    const int ua = G.ca - 2, va = G.cb - 2, wa = G.cc - 2;
                    Ex = interp<BY, BZ>(sF + 0 * NB, ua + W.ix, va + W.i1y, wa + W.i1z, s1, s11, s4, s41, s6, s61);
                    Ey = interp<BY, BZ>(sF + 1 * NB, ua + W.i1x, va + W.iy, wa + W.i1z, s2, s21, s3, s31, s6, s61);
                    Ez = interp<BY, BZ>(sF + 2 * NB, ua + W.i1x, va + W.i1y, wa + W.iz, s2, s21, s4, s41, s5, s51);
                    Hx = interp<BY, BZ>(sF + 3 * NB, ua + W.i1x, va + W.iy, wa + W.iz, s2, s21, s3, s31, s5, s51);
                    Hy = interp<BY, BZ>(sF + 4 * NB, ua + W.ix, va + W.i1y, wa + W.iz, s1, s11, s4, s41, s5, s51);
                    Hz = interp<BY, BZ>(sF + 5 * NB, ua + W.ix, va + W.iy, wa + W.i1z, s1, s11, s3, s31, s6, s61);
                }
            }

            /* Particle::Move (particle.h:160-202) */
            const Boris B = boris_move(pu, pv, pw, Ex, Ey, Ez, Hx, Hy, Hz, q_m * tau * 0.5);
            pu = B.p1[0];
            pv = B.p1[1];
            pw = B.p1[2];
            const double x1 = x + tau * (B.g2 * pu);
            const double y1 = y + tau * (B.g2 * pv);
            const double z1 = z + tau * (B.g2 * pw);

            /* CurrentToMesh (cell.h:1306-1422): one code path for every lane;
             * without a crossing the first segment is the whole path (x2 = x1)
             * and the second is skipped */
            {
                const Relay R =
                    trep ? relay_from(x, y, z, x1, y1, z1, G, tape_get(tw, TAPE_RI1, 1, 2),
                                      tape_get(tw, TAPE_RI1 + 1, 1, 2),
                                      tape_get(tw, TAPE_RI1 + 2, 1, 2),
                                      tape_get(tw, TAPE_RI2, 2, 1), tape_get(tw, TAPE_RI2 + 2, 2, 1),
                                      tape_get(tw, TAPE_RI2 + 4, 2, 1))
                         : relay_point(x, y, z, x1, y1, z1, frx, fry, frz, G, a.dc);
                if (trec) {
                    tape_put(tout, TAPE_RI1, 1, R.i1x, 2);
                    tape_put(tout, TAPE_RI1 + 1, 1, R.i1y, 2);
                    tape_put(tout, TAPE_RI1 + 2, 1, R.i1z, 2);
                    tape_put(tout, TAPE_RI2, 2, R.i2x, 1);
                    tape_put(tout, TAPE_RI2 + 2, 2, R.i2y, 1);
                    tape_put(tout, TAPE_RI2 + 4, 2, R.i2z, 1);
                }
                const int m = R.m;
                const double e1x = m ? R.x2x : x1, e1y = m ? R.x2y : y1, e1z = m ? R.x2z : z1;
                pqr_lds<BY, BZ, NB, DIFF>(myJ, G, a.dc, R.i1x, R.i1y, R.i1z, x, y, z, e1x, e1y, e1z,
                                          mass, bad, true);
                pqr_lds<BY, BZ, NB, DIFF>(myJ, G, a.dc, R.i2x, R.i2y, R.i2z, R.x2x, R.x2y, R.x2z, x1,
                                          y1, z1, mass, bad, m != 0);
            }

            /* x <- x1, Reflect (cell.h:1410-1478) */
            if (DIFF && (trec || trep)) {
                const double v1[3] = {x1, y1, z1}, Lm[3] = {g.Lx, g.Ly, g.Lz};
                double vn[3];
                for (int d = 0; d < 3; d++) {
                    const int k = trep ? tape_get(tw, TAPE_REFL + 2 * d, 2, 0)
                                       : reflect_case(v1[d], Lm[d]);
                    if (trec) tape_put(tout, TAPE_REFL + 2 * d, 2, k, 0);
                    vn[d] = reflect_apply(v1[d], Lm[d], k);
                }
                xn = vn[0];
                yn = vn[1];
                zn = vn[2];
            } else {
                xn = reflect1(x1, g.Lx);
                yn = reflect1(y1, g.Ly);
                zn = reflect1(z1, g.Lz);
            }

            /* flyDirection (cell.h:694-701) with the owner cell's geometry */
            int ddx = (((xn > G.x0 + hx) && (xn < G.x0 + 2 * hx)) || ((xn < hx) && (gi == g.Nx - 1)))
                          ? 2
                          : (((xn < G.x0) || ((xn > g.Lx - hx) && (gi == 0))) ? 0 : 1);
            int ddy = (((yn > G.y0 + hy) && (yn < G.y0 + 2 * hy)) || ((yn < hy) && (gl == g.Ny - 1)))
                          ? 2
                          : (((yn < G.y0) || ((yn > g.Ly - hy) && (gl == 0))) ? 0 : 1);
            int ddz = (((zn > G.z0 + hz) && (zn < G.z0 + 2 * hz)) || ((zn < hz) && (gk == g.Nz - 1)))
                          ? 2
                          : (((zn < G.z0) || ((zn > g.Lz - hz) && (gk == 0))) ? 0 : 1);
            const int dir = ddx * 9 + ddy * 3 + ddz;
            const bool inside = (G.x0 <= xn) && (xn < G.x0 + hx) && (G.y0 <= yn) &&
                                (yn < G.y0 + hy) && (G.z0 <= zn) && (zn < G.z0 + hz);
            if (!inside) atomicAdd(&sOut[c], 1);

            stays = dir == 13;
            /* a replayed word that cannot be trusted keeps its particle in
             * place (the step reports PIC_F_RANGE) */
            if (trep) stays = ((tw >> TAPE_STAY) & 1ull) || !tin ||
                              !((tw >> TAPE_VALID) & 1ull) || ((tw >> TAPE_BAD) & 1ull);
            if (trec && stays) tout |= 1ull << TAPE_STAY;
            if (!stays) {
                atomicAdd(&sLeave[c], 1);
                /* GPU_ArrangeFlights (kernels.cu:316-370): every bulk cell of the
                 * 27-neighbourhood (getWrapCellNumber) whose isPointInCell
                 * (cell.h:717) holds receives a copy; per axis the test is
                 * separable, so the receivers are a product of per-axis sets. */
                /* per axis, bit d+1 set when neighbour d in {-1,0,1} accepts */
                unsigned bx = 0, by = 0, bz = 0;
                for (int d = -1; d <= 1; d++) {
                    int r = gi + d;
                    if (r >= g.Nx) r -= g.Nx; else if (r < 0) r += g.Nx;
                    double r0 = (double)r * hx;
                    if (r0 <= xn && xn < r0 + hx) bx |= 1u << (d + 1);
                    r = gl + d;
                    if (r >= g.Ny) r -= g.Ny; else if (r < 0) r += g.Ny;
                    r0 = (double)r * hy;
                    if (r0 <= yn && yn < r0 + hy) by |= 1u << (d + 1);
                    r = gk + d;
                    if (r >= g.Nz) r -= g.Nz; else if (r < 0) r += g.Nz;
                    r0 = (double)r * hz;
                    if (r0 <= zn && zn < r0 + hz) bz |= 1u << (d + 1);
                }
                if (trec) {
                    const unsigned bb[3] = {bx, by, bz};
                    for (int d = 0; d < 3; d++) {
                        if (__popc(bb[d]) == 1) tape_put(tout, TAPE_RECV + 2 * d, 2, __ffs(bb[d]) - 1, 0);
                        else tout |= 1ull << TAPE_BAD;
                    }
                }
                if (trep) {
                    /* the recorded receiver, one per axis (0..2) */
                    unsigned *bb[3] = {&bx, &by, &bz};
                    for (int d = 0; d < 3; d++) {
                        const int k = tape_get(tw, TAPE_RECV + 2 * d, 2, 0);
                        if (k > 2) bad = 1;
                        *bb[d] = k > 2 ? 0u : 1u << k;
                    }
                }
                const int ncopy = __popc(bx) * __popc(by) * __popc(bz);
                if (ncopy == 0) atomicAdd(&sDrop, 1ull);
                if (ncopy > 1) atomicAdd(&sDup, (unsigned long long)(ncopy - 1));
                if (ncopy > 0) {
                    int slot = atomicAdd(&sQ, ncopy);
                    const long qb = tile * (long)a.q.qcap;
                    int kq = 0;
                    /* receivers in the reference's scan order (x, then y, then z) */
                    for (int ax = 0; ax < 3; ax++) {
                        if (!((bx >> ax) & 1u)) continue;
                        int rx = gi + ax - 1;
                        if (rx >= g.Nx) rx -= g.Nx; else if (rx < 0) rx += g.Nx;
                        for (int ay = 0; ay < 3; ay++) {
                            if (!((by >> ay) & 1u)) continue;
                            int ry = gl + ay - 1;
                            if (ry >= g.Ny) ry -= g.Ny; else if (ry < 0) ry += g.Ny;
                            for (int az = 0; az < 3; az++) {
                                if (!((bz >> az) & 1u)) continue;
                                int rz = gk + az - 1;
                                if (rz >= g.Nz) rz -= g.Nz; else if (rz < 0) rz += g.Nz;
                                const int idx = slot + kq++;
                                if (idx >= a.q.qcap) continue;
                                const long e = qb + idx;
                                a.q.q[0][e] = xn;
                                a.q.q[1][e] = yn;
                                a.q.q[2][e] = zn;
                                a.q.q[3][e] = pu;
                                a.q.q[4][e] = pv;
                                a.q.q[5][e] = pw;
                                a.q.sp[e] = s;
                                if (DIFF) a.q.id[e] = pid;
                                a.q.dest[e] = (int)(((long)rx * g.Ny + ry) * g.Nz + rz);
                                a.q.halo[e] = (unsigned char)(((G.ca + ax) * BY + (G.cb + ay)) * BZ +
                                                              (G.cc + az));
                            }
                        }
                    }
                }
            }
        }

        if constexpr (DIFF) {
            if (valid && a.d.tape_mode == PIC_TAPE_RECORD && pid >= 0 && pid < a.d.tape_n)
                a.d.tape_row[pid] = tout | (1ull << TAPE_VALID);
        }

        /* stayers compacted in flat order (the reference removes leavers by
         * swap-with-last, removeParticleFromSurfaceDevice, cell.h:391-432;
         * the order inside a cell changes no result but the rounding of the
         * FP64 atomic sums, which have no fixed order either) */
        const int rank = stable_rank_runs(runs, c, valid, stays, carry_cell, carry_cnt);
        if (valid && stays) {
            const unsigned dst = (unsigned)(sBase[c] + rank);
            slot(P0, dst) = xn;
            slot(P1, dst) = yn;
            slot(P2, dst) = zn;
            slot(P3, dst) = pu;
            slot(P4, dst) = pv;
            slot(P5, dst) = pw;
            if (rank != j) {
                slot(SP, dst) = s;
                if (DIFF) slot(ID, dst) = pid;
            }
        }
        g0 = gnext;
        o0 = onext;
    }
    if (bad) atomicOr(&sFlags, (unsigned)PIC_F_RANGE);
    __syncthreads();

    int maxout = 0;
    for (int c = tid; c < C; c += PIC_BLOCK) {
        const int ca = c / (TY * TZ), cb = (c / TZ) % TY, cc = c % TZ;
        const int gi = I0 + ca, gl = L0 + cb, gk = K0 + cc;
        if (gi >= g.Nx || gl >= g.Ny || gk >= g.Nz) continue;
        const long cell = ((long)gi * g.Ny + gl) * g.Nz + gk;
        a.p.count[cell] = sCnt[c] - sLeave[c];
        if (sOut[c] > maxout) maxout = sOut[c];
    }

    /* tile J box -> global J (GPU_WriteAllCurrents, kernels.cu:83-112) */
    if (tid < NB) {   /* thread n adds box node n of the three components */
#pragma unroll
        for (int comp = 0; comp < 3; comp++) {
            const int e = comp * NB + tid;
            if constexpr (DIFF) {
                const unsigned long long *u = reinterpret_cast<const unsigned long long *>(sJ);
                unsigned long long val = u[e];
                for (int r = 1; r < JCOPIES; r++) val += u[r * 3 * NB + e];
                if (val != 0ull && node_in)
                    atomicAdd(reinterpret_cast<unsigned long long *>(a.d.jfix[comp]) + noff, val);
            } else {
                double val = sJ[e];
                for (int r = 1; r < JCOPIES; r++) val += sJ[r * 3 * NB + e];
                if (val != 0.0 && node_in) atomicAdd(&a.f.f[PIC_JX + comp][noff], val);
            }
        }
    }

    if (maxout > 0) {
        if (a.dep_cap > 0 && maxout > a.dep_cap) atomicOr(&sFlags, (unsigned)PIC_F_TOO_MANY);
        if (maxout > a.st->max_dep) atomicMax(&a.st->max_dep, maxout);
    }
    __syncthreads();
    if (tid == 0) {
        int nq = sQ;
        if (nq > a.q.qcap) {
            atomicOr(&sFlags, (unsigned)PIC_F_QUEUE_OVERFLOW);
            nq = a.q.qcap;
        }
        a.q.count[tile] = nq;
        if (sFlags) atomicOr(&a.st->flags, sFlags);
        if (sDrop) atomicAdd(&a.st->dropped, sDrop);
        if (sDup) atomicAdd(&a.st->duplicated, sDup);
    }
}

/* differentiable form: fixed-point J -> the FP64 J arrays (exact for
 * |J| < 2^5), and the fixed-point arrays cleared for the next step */
__global__ __launch_bounds__(PIC_BLOCK) void k_jfix_to_f64(long G, FieldPtrs f, DiffArgs d) {
    const long n = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (n >= G) return;
    for (int c = 0; c < 3; c++) {
        const long long v = d.jfix[c][n];
        f.f[PIC_JX + c][n] = (double)v * (1.0 / PIC_J_SCALE64);
        d.jfix[c][n] = 0;
    }
}


/* initial upload: compact cell-sorted arrays -> buckets */
__global__ __launch_bounds__(PIC_BLOCK) void k_scatter_load(const long *off, long ncell,
                                                           const double *s0, const double *s1,
                                                           const double *s2, const double *s3,
                                                           const double *s4, const double *s5,
                                                           const unsigned char *ssp, PartPtrs p) {
    const long e = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (e >= ncell * p.cap) return;
    const long c = e / p.cap;
    const int j = (int)(e % p.cap);
    const long n = off[c + 1] - off[c];
    if (j == 0) p.count[c] = (int)n;
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

__global__ __launch_bounds__(PIC_BLOCK) void k_repack(long ncell, PartPtrs from, PartPtrs to) {
    const long e = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (e >= ncell * from.cap) return;
    const long c = e / from.cap;
    const int j = (int)(e % from.cap);
    int n = from.count[c];
    if (n > from.cap) n = from.cap;
    if (j >= n) return;
    const long dst = c * to.cap + j;
    for (int k = 0; k < 6; k++) to.p[k][dst] = from.p[k][e];
    to.sp[dst] = from.sp[e];
    if (to.id) to.id[dst] = from.id[e];
}

/* bucket copy with any capacities (checkpoint save and restore): occupied
 * slots, ids and counts */
__global__ __launch_bounds__(PIC_BLOCK) void k_copy_parts(long ncell, PartPtrs from, PartPtrs to) {
    const long e = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (e >= ncell * from.cap) return;
    const long c = e / from.cap;
    const int j = (int)(e % from.cap);
    int n = from.count[c];
    if (n > from.cap) n = from.cap;
    if (j == 0) to.count[c] = n;
    if (j >= n) return;
    const long dst = c * to.cap + j;
    for (int k = 0; k < 6; k++) to.p[k][dst] = from.p[k][e];
    to.sp[dst] = from.sp[e];
    if (to.id) to.id[dst] = from.id[e];
}

__global__ __launch_bounds__(PIC_BLOCK) void k_scatter_ids(const long *off, long ncell,
                                                          const int *src, PartPtrs p) {
    const long e = (long)blockIdx.x * PIC_BLOCK + threadIdx.x;
    if (e >= ncell * p.cap) return;
    const long c = e / p.cap;
    const int j = (int)(e % p.cap);
    if (j >= off[c + 1] - off[c]) return;
    p.id[e] = src[off[c] + j];
}

__global__ __launch_bounds__(PIC_BLOCK) void k_fix_arrivals(QueuePtrs q, int oldcap, PartPtrs to) {
    const long tile = blockIdx.x;
    const int n = q.count[tile];
    const long qb = tile * (long)q.qcap;
    for (int j = threadIdx.x; j < n; j += PIC_BLOCK) {
        const long e = qb + j;
        const int slot = q.slot[e];
        if (slot < oldcap || slot >= to.cap) continue;
        const long dst = (long)q.dest[e] * to.cap + slot;
        for (int k = 0; k < 6; k++) to.p[k][dst] = q.q[k][e];
        to.sp[dst] = q.sp[e];
        if (to.id) to.id[dst] = q.id[e];
    }
}

__global__ __launch_bounds__(PIC_BLOCK) void k_gather_particles(long ncell, const long *off,
                                                               PartPtrs p, double *d0, double *d1,
                                                               double *d2, double *d3, double *d4,
                                                               double *d5, unsigned char *dsp,
                                                               int *dcell, int *did) {
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
    if (dcell) dcell[dst] = (int)c;
    if (did) did[dst] = p.id ? p.id[e] : -1;
}

/* FP64 moments per sort: N, sum x,y,z, sum p, sum (m/q_m)(gamma-1), sum |p| */
__global__ __launch_bounds__(PIC_BLOCK) void k_moments(long ncell, PartPtrs p, SpeciesTable sp,
                                                      const int *slot_sort, double *partials) {
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
        const double px = p.p[3][e], py = p.p[4][e], pz = p.p[5][e];
        const double gm = sqrt(1.0 + px * px + py * py + pz * pz);
        double v[9] = {1.0, p.p[0][e], p.p[1][e], p.p[2][e], px, py, pz,
                       (sp.m[s] / sp.q_m[s]) * (gm - 1.0), fabs(px) + fabs(py) + fabs(pz)};
        for (int k = 0; k < 9; k++) {
            double add = v[k];
            acc[0 * 9 + k] += (srt == 0) ? add : 0.0;
            acc[1 * 9 + k] += (srt == 1) ? add : 0.0;
            acc[2 * 9 + k] += (srt == 2) ? add : 0.0;
        }
    }
    for (int k = 0; k < 27; k++) {
        red[threadIdx.x] = acc[k];
        __syncthreads();
        for (int w = PIC_BLOCK / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) partials[(long)blockIdx.x * 27 + k] = red[0];
        __syncthreads();
    }
}

__global__ void k_moments_final(const double *partials, int nblocks, double *out27) {
    const int k = threadIdx.x;
    if (k >= 27) return;
    double s = 0.0;
    for (int b = 0; b < nblocks; b++) s += partials[(long)b * 27 + k];
    out27[k] = s;
}

inline unsigned blocks_for(long n) { return (unsigned)((n + PIC_BLOCK - 1) / PIC_BLOCK); }

}  // namespace


void launch_push_deposit_f64(const Grid &g, const TileGeom &t, FieldPtrs f, PartPtrs p,
                             QueuePtrs q, SpeciesTable sp, int dep_cap, DevStatus *st,
                             const DiffArgs *diff, hipStream_t s) {
    PushArgs a;
    a.g = g;
    a.dc = make_divconst(g);
    a.t = t;
    a.f = f;
    a.p = p;
    a.q = q;
    a.sp = sp;
    a.dep_cap = dep_cap;
    a.st = st;
    a.d = diff ? *diff : DiffArgs{};
    a.snap_on = diff && diff->snap.p[0] != nullptr;
    /* dynamic LDS: one 4-byte entry per line group */
    const size_t lds = (size_t)4 * TXC * TYC * TZC * ((p.cap + GS64 - 1) / GS64);
    if (diff)
        hipLaunchKernelGGL((k_push_deposit<TXC, TYC, TZC, true>), dim3((unsigned)t.ntiles),
                           dim3(PIC_BLOCK), lds, s, a);
    else
        hipLaunchKernelGGL((k_push_deposit<TXC, TYC, TZC, false>), dim3((unsigned)t.ntiles),
                           dim3(PIC_BLOCK), lds, s, a);
}

void launch_jfix_to_f64(const Grid &g, FieldPtrs f, DiffArgs d, hipStream_t s) {
    hipLaunchKernelGGL(k_jfix_to_f64, dim3(blocks_for(g.G)), dim3(PIC_BLOCK), 0, s, g.G, f, d);
}

void launch_arrive_f64(const Grid &g, const TileGeom &t, PartPtrs p, QueuePtrs q, DevStatus *st,
                       hipStream_t s) {
    hipLaunchKernelGGL((k_arrive_tile<TXC, TYC, TZC, PartPtrs, QueuePtrs>), dim3((unsigned)t.ntiles),
                       dim3(PIC_BLOCK), 0, s, g, t, p, q, st);
}

void launch_scatter_load(const long *offsets, long ncell, const double *src[6],
                         const unsigned char *srcsp, PartPtrs p, hipStream_t s) {
    hipLaunchKernelGGL(k_scatter_load, dim3(blocks_for(ncell * p.cap)), dim3(PIC_BLOCK), 0, s,
                       offsets, ncell, src[0], src[1], src[2], src[3], src[4], src[5], srcsp, p);
}

void launch_repack(long ncell, PartPtrs from, PartPtrs to, hipStream_t s) {
    hipLaunchKernelGGL(k_repack, dim3(blocks_for(ncell * from.cap)), dim3(PIC_BLOCK), 0, s, ncell,
                       from, to);
}

void launch_fix_arrivals(const TileGeom &t, QueuePtrs q, int oldcap, PartPtrs to, hipStream_t s) {
    hipLaunchKernelGGL(k_fix_arrivals, dim3((unsigned)t.ntiles), dim3(PIC_BLOCK), 0, s, q, oldcap,
                       to);
}

void launch_gather_particles(long ncell, const long *offsets, PartPtrs p, double *dst[6],
                             unsigned char *dsp, int *dcell, int *did, hipStream_t s) {
    hipLaunchKernelGGL(k_gather_particles, dim3(blocks_for(ncell * p.cap)), dim3(PIC_BLOCK), 0, s,
                       ncell, offsets, p, dst[0], dst[1], dst[2], dst[3], dst[4], dst[5], dsp,
                       dcell, did);
}

void launch_scatter_ids(const long *offsets, long ncell, const int *src, PartPtrs p,
                        hipStream_t s) {
    hipLaunchKernelGGL(k_scatter_ids, dim3(blocks_for(ncell * p.cap)), dim3(PIC_BLOCK), 0, s,
                       offsets, ncell, src, p);
}

void launch_copy_parts(long ncell, PartPtrs from, PartPtrs to, hipStream_t s) {
    hipLaunchKernelGGL(k_copy_parts, dim3(blocks_for(ncell * from.cap)), dim3(PIC_BLOCK), 0, s,
                       ncell, from, to);
}

void launch_moments(long ncell, PartPtrs p, SpeciesTable sp, const int *slot_sort,
                    double *partials, int nblocks, double *out27, hipStream_t s) {
    hipLaunchKernelGGL(k_moments, dim3(nblocks), dim3(PIC_BLOCK), 0, s, ncell, p, sp, slot_sort,
                       partials);
    hipLaunchKernelGGL(k_moments_final, dim3(1), dim3(32), 0, s, partials, nblocks, out27);
}

TileGeom make_tile_geom(int Nx, int Ny, int Nz) {
    TileGeom t;
    t.TX = TXC;
    t.TY = TYC;
    t.TZ = TZC;
    t.ntx = (Nx + TXC - 1) / TXC;
    t.nty = (Ny + TYC - 1) / TYC;
    t.ntz = (Nz + TZC - 1) / TZC;
    t.ntiles = (long)t.ntx * t.nty * t.ntz;
    return t;
}
