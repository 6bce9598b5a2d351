/*
 * pic_api.hip -- host side of libpic_amd.so: the C ABI declared in
 * include/pic_amd.h.  Plain HIP runtime calls; kernels live in k_*.hip.
 *
 * Step orchestration follows Plasma::Step (gpu_plasma.h:605-620): H half
 * step (+ J := 0), fused gather/push/deposit/departures, arrivals, periodic
 * J fold, H second half, E update, periodic E ghosts, energy.  The whole step
 * is enqueued on the context stream and synchronised once; the reference
 * synchronises five times and copies stage arrays to the host every step.
 */
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "pic_ctx.h"
#include "pic_draw.h"

namespace picx {

thread_local char g_hip_err[256] = "";

int record_hip(hipError_t e, const char *where) {
    snprintf(g_hip_err, sizeof(g_hip_err), "%s: %s (%d)", where, hipGetErrorString(e), (int)e);
    return e == hipErrorOutOfMemory ? PIC_ERR_NOMEM : PIC_ERR_HIP;
}


template <class P>
void free_parts(P &p) {
    for (int k = 0; k < 6; k++)
        if (p.p[k]) (void)hipFree(p.p[k]);
    if (p.sp) (void)hipFree(p.sp);
    if (p.count) (void)hipFree(p.count);
    if (p.id) (void)hipFree(p.id);
    if constexpr (std::is_same<P, PartPtrs32>::value) {
        if (p.bcount) (void)hipFree(p.bcount);
        if (p.inbox) (void)hipFree(p.inbox);
    }
    p = P{};
}

/* buckets of ncell x cap slots; ids only for the differentiable form */
template <class P>
int alloc_parts(P &p, long ncell, int cap, bool with_count, bool with_id) {
    using T = std::remove_pointer_t<std::remove_reference_t<decltype(p.p[0])>>;
    p = P{};
    p.cap = cap;
    size_t n = (size_t)ncell * cap;
    for (int k = 0; k < 6; k++) HIPCHK(hipMalloc(&p.p[k], n * sizeof(T)));
    HIPCHK(hipMalloc(&p.sp, n));
    if (with_count) HIPCHK(hipMalloc(&p.count, ncell * sizeof(int)));
    if (with_id) HIPCHK(hipMalloc(&p.id, n * sizeof(int)));
    return PIC_SUCCESS;
}

template int alloc_parts<PartPtrs>(PartPtrs &, long, int, bool, bool);
template void free_parts<PartPtrs>(PartPtrs &);
template int alloc_parts<PartPtrs32>(PartPtrs32 &, long, int, bool, bool);
template void free_parts<PartPtrs32>(PartPtrs32 &);

/* FP32 form: slots of one particle set, enough for the buckets (cell * cap)
 * and for the packed tile regions (tile * tcap) */
long live_slots32(const pic_ctx *c) {
    return std::max(c->NC, c->t.ntiles * (long)(c->t.TX * c->t.TY * c->t.TZ));
}

/* FP32 form: the unsorted and sorted spill lists of every tile, scap
 * records each (empty: no arrival pending) */
int alloc_spill32(pic_ctx *c, int scap) {
    Layout32 &L = c->lay32;
    if (L.spill_u) (void)hipFree(L.spill_u);
    if (L.spill_s) (void)hipFree(L.spill_s);
    L.spill_u = L.spill_s = nullptr;
    const size_t bytes = (size_t)c->t.ntiles * scap * 2 * sizeof(float4);
    HIPCHK(hipMalloc(&L.spill_u, bytes));
    HIPCHK(hipMalloc(&L.spill_s, bytes));
    L.scap = scap;
    return PIC_SUCCESS;
}

/* FP32 form, once per context: the per-cell counts, both inboxes and
 * arrival counters, the packed starts and the spill lists per tile */
int alloc_cells32(pic_ctx *c) {
    PartPtrs32 &p = c->p32;
    Layout32 &L = c->lay32;
    if (p.count) return PIC_SUCCESS;
    const long NC = c->NC, nt = c->t.ntiles;
    const size_t inbox = (size_t)NC * PIC_INBOX * 2 * sizeof(float4);
    HIPCHK(hipMalloc(&p.count, NC * sizeof(int)));
    HIPCHK(hipMalloc(&p.bcount, NC * sizeof(int)));
    HIPCHK(hipMalloc(&p.inbox, inbox));
    HIPCHK(hipMalloc(&L.inbox_next, inbox));
    HIPCHK(hipMalloc(&L.pstart, NC * sizeof(int)));
    HIPCHK(hipMalloc(&L.arr, NC * sizeof(int)));
    HIPCHK(hipMalloc(&L.arr_next, NC * sizeof(int)));
    HIPCHK(hipMemset(L.arr, 0, NC * sizeof(int)));
    HIPCHK(hipMemset(L.arr_next, 0, NC * sizeof(int)));
    HIPCHK(hipMalloc(&L.spill_u_cnt, nt * sizeof(int)));
    HIPCHK(hipMemset(L.spill_u_cnt, 0, nt * sizeof(int)));
    L.packed = 0;
    return alloc_spill32(c, PIC_SPILL_CAP);
}

/* FP32 form: the particle arrays of one set (p, sp, id), capacity cap */
int alloc_set32(pic_ctx *c, PartPtrs32 &s, int cap) {
    const size_t n = (size_t)live_slots32(c) * cap;
    for (int k = 0; k < 6; k++) HIPCHK(hipMalloc(&s.p[k], n * sizeof(float)));
    HIPCHK(hipMalloc(&s.sp, n));
    if (c->diff) HIPCHK(hipMalloc(&s.id, n * sizeof(int)));
    s.cap = cap;
    return PIC_SUCCESS;
}

void free_set32(PartPtrs32 &s) {
    for (int k = 0; k < 6; k++) {
        if (s.p[k]) (void)hipFree(s.p[k]);
        s.p[k] = nullptr;
    }
    if (s.sp) (void)hipFree(s.sp);
    if (s.id) (void)hipFree(s.id);
    s.sp = nullptr;
    s.id = nullptr;
}

/* the live set and the set the push writes, both of capacity cap (the
 * per-cell arrays on first use) */
int alloc_live32(pic_ctx *c, int cap) {
    int rc = alloc_cells32(c);
    if (rc) return rc;
    free_set32(c->p32);
    free_set32(c->alt32);
    if ((rc = alloc_set32(c, c->p32, cap)) || (rc = alloc_set32(c, c->alt32, cap))) return rc;
    c->lay32.tcap = (long)c->t.TX * c->t.TY * c->t.TZ * cap;
    c->lay32.packed = 0;
    return PIC_SUCCESS;
}

void free_layout32(Layout32 &L) {
    for (void *q : {(void *)L.pstart, (void *)L.arr, (void *)L.arr_next, (void *)L.inbox_next,
                    (void *)L.spill_u, (void *)L.spill_s, (void *)L.spill_u_cnt})
        if (q) (void)hipFree(q);
    L = Layout32{};
}

/* the arrays the next push reads take the place of those this push wrote */
void swap_arrivals32(pic_ctx *c) {
    Layout32 &L = c->lay32;
    std::swap(c->p32.inbox, L.inbox_next);
    std::swap(L.arr, L.arr_next);
}

/* the live and the other FP32 set trade places (per-cell arrays stay) */
void swap_sets32(pic_ctx *c) {
    for (int k = 0; k < 6; k++) std::swap(c->p32.p[k], c->alt32.p[k]);
    std::swap(c->p32.sp, c->alt32.sp);
    std::swap(c->p32.id, c->alt32.id);
}

template <class Q>
void free_queue(Q &q) {
    for (int k = 0; k < 6; k++)
        if (q.q[k]) (void)hipFree(q.q[k]);
    if (q.sp) (void)hipFree(q.sp);
    if (q.dest) (void)hipFree(q.dest);
    if (q.slot) (void)hipFree(q.slot);
    if (q.count) (void)hipFree(q.count);
    if (q.id) (void)hipFree(q.id);
    if (q.halo) (void)hipFree(q.halo);
    q = Q{};
}

template <class Q>
int alloc_queue(Q &q, const TileGeom &t, int qcap, bool with_id) {
    using T = std::remove_pointer_t<std::remove_reference_t<decltype(q.q[0])>>;
    q = Q{};
    q.qcap = qcap;
    size_t n = (size_t)t.ntiles * qcap;
    for (int k = 0; k < 6; k++) HIPCHK(hipMalloc(&q.q[k], n * sizeof(T)));
    HIPCHK(hipMalloc(&q.sp, n));
    HIPCHK(hipMalloc(&q.dest, n * sizeof(int)));
    HIPCHK(hipMalloc(&q.slot, n * sizeof(int)));
    HIPCHK(hipMalloc(&q.count, t.ntiles * sizeof(int)));
    HIPCHK(hipMemset(q.count, 0, t.ntiles * sizeof(int)));
    if (with_id) HIPCHK(hipMalloc(&q.id, n * sizeof(int)));
    HIPCHK(hipMalloc(&q.halo, n));
    return PIC_SUCCESS;
}

void free_queue(QueuePtrs32 &q) {
    if (q.e) (void)hipFree(q.e);
    if (q.slot) (void)hipFree(q.slot);
    if (q.count) (void)hipFree(q.count);
    q = QueuePtrs32{};
}

int alloc_queue(QueuePtrs32 &q, const TileGeom &t, int qcap, bool) {
    q = QueuePtrs32{};
    q.qcap = qcap;
    const size_t n = (size_t)t.ntiles * qcap;
    HIPCHK(hipMalloc(&q.e, 2 * n * sizeof(float4)));
    HIPCHK(hipMalloc(&q.slot, n * sizeof(int)));
    HIPCHK(hipMalloc(&q.count, t.ntiles * sizeof(int)));
    HIPCHK(hipMemset(q.count, 0, t.ntiles * sizeof(int)));
    return PIC_SUCCESS;
}

int queue_capacity(const TileGeom &t, int cap) {
    long c = (long)t.TX * t.TY * t.TZ * cap / 2;
    c = (c + 63) / 64 * 64;
    return (int)std::max(64L, c);
}

/* slots per bucket: 1.5 x the fullest cell + 8, rounded up to 32 so that every
 * bucket starts on a 128-byte line in the FP32 arrays (and on a line pair
 * in the FP64 ones): a cell's particles then occupy whole lines and a
 * wavefront's loads and in-place stores touch no partial line at a bucket
 * start */
#ifndef PIC_CAP_ALIGN
#define PIC_CAP_ALIGN 32
#endif
/* the push kernels address a tile's particle slots (and FP32 inbox records)
 * by 32-bit byte offsets from the bucket of its first cell: the tile's span
 * (from that cell to the end of its last cell's bucket) must stay below
 * 2^32 bytes; a bucket holds at most 256 lines of 128 bytes (the group
 * entries keep a slot in 15 bits) */
bool push_span_ok(const pic_ctx *c, int cap) {
    const TileGeom &t = c->t;
    const int elem = c->fp32 ? 4 : 8;
    const long cells = (long)(t.TX - 1) * c->g.Ny * c->g.Nz + (long)(t.TY - 1) * c->g.Nz + t.TZ;
    const long span = cells * cap * elem;
    const long inbox = c->fp32 ? cells * PIC_INBOX * 2 * (long)sizeof(float4) : 0;
    /* the FP32 queue records rank a leaver among its tile's leavers in 16 bits */
    const bool ranks_ok = !c->fp32 || queue_capacity(t, cap) < (1 << 16);
    return span < (1L << 32) && inbox < (1L << 32) && (long)cap * elem <= 256L * 128 && ranks_ok;
}

int auto_capacity(int maxc) {
    int cap = (int)(maxc * 1.5) + 8;
    cap = (cap + PIC_CAP_ALIGN - 1) / PIC_CAP_ALIGN * PIC_CAP_ALIGN;
    return std::max(cap, PIC_CAP_ALIGN);
}

/* upload staged particles into cell buckets (cell-sorted, stable) */
template <class P, class Q>
int commit_into(pic_ctx *c, P &parts, Q &queue, int cap, const std::vector<long> &off) {
    using T = std::remove_pointer_t<std::remove_reference_t<decltype(parts.p[0])>>;
    const long NC = c->NC;
    const long P_ = (long)c->staged.cell.size();
    std::vector<T> sorted[6];
    std::vector<unsigned char> ssp(P_);
    std::vector<int> sid(c->diff ? P_ : 0);
    for (int k = 0; k < 6; k++) sorted[k].resize(P_);
    {
        const Grid &g = c->g;
        std::vector<long> pos(off.begin(), off.end() - 1);
        for (long j = 0; j < P_; j++) {
            const int cell = c->staged.cell[j];
            const long d = pos[cell]++;
            if (std::is_same<T, float>::value) {
                /* FP32 form: offset inside the owning cell, kept below 1 */
                const int ci = cell / (g.Ny * g.Nz), cl = (cell / g.Nz) % g.Ny, ck = cell % g.Nz;
                const double hh[3] = {g.hx, g.hy, g.hz};
                const int cc[3] = {ci, cl, ck};
                for (int k = 0; k < 3; k++) {
                    float o = (float)((c->staged.v[k][j] - (double)cc[k] * hh[k]) / hh[k]);
                    if (o >= 1.0f) o = 0.99999994f;
                    if (o < 0.0f) o = 0.0f;
                    sorted[k][d] = (T)o;
                }
                for (int k = 3; k < 6; k++) sorted[k][d] = (T)c->staged.v[k][j];
            } else {
                for (int k = 0; k < 6; k++) sorted[k][d] = (T)c->staged.v[k][j];
            }
            ssp[d] = c->staged.sp[j];
            if (c->diff) sid[d] = c->staged.id[j];
        }
    }
    int rc;
    if constexpr (std::is_same<P, PartPtrs32>::value) rc = alloc_live32(c, cap);
    else rc = alloc_parts(parts, NC, cap, true, c->diff);
    if (rc) return rc;
    rc = alloc_queue(queue, c->t, queue_capacity(c->t, cap), c->diff);
    if (rc) return rc;
    long *d_off = nullptr;
    T *d_src[6] = {nullptr};
    unsigned char *d_sp = nullptr;
    size_t Pn = (size_t)std::max(P_, 1L);
    HIPCHK(hipMalloc(&d_off, (NC + 1) * sizeof(long)));
    for (int k = 0; k < 6; k++) HIPCHK(hipMalloc(&d_src[k], Pn * sizeof(T)));
    HIPCHK(hipMalloc(&d_sp, Pn));
    HIPCHK(hipMemcpy(d_off, off.data(), (NC + 1) * sizeof(long), hipMemcpyHostToDevice));
    if (P_ > 0) {
        for (int k = 0; k < 6; k++)
            HIPCHK(hipMemcpy(d_src[k], sorted[k].data(), P_ * sizeof(T), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_sp, ssp.data(), P_, hipMemcpyHostToDevice));
    }
    const T *src[6] = {d_src[0], d_src[1], d_src[2], d_src[3], d_src[4], d_src[5]};
    if constexpr (std::is_same<T, float>::value)
        launch_scatter_load32(d_off, NC, src, d_sp, parts, c->stream);
    else
        launch_scatter_load(d_off, NC, src, d_sp, parts, c->stream);
    int *d_id = nullptr;
    if (c->diff) {
        HIPCHK(hipMalloc(&d_id, Pn * sizeof(int)));
        if (P_ > 0) HIPCHK(hipMemcpy(d_id, sid.data(), P_ * sizeof(int), hipMemcpyHostToDevice));
        if constexpr (std::is_same<P, PartPtrs>::value)
            launch_scatter_ids(d_off, NC, d_id, parts, c->stream);
        else
            launch_scatter_ids32(d_off, NC, d_id, parts, c->stream);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    if (d_id) (void)hipFree(d_id);
    (void)hipFree(d_off);
    for (int k = 0; k < 6; k++) (void)hipFree(d_src[k]);
    (void)hipFree(d_sp);
    return PIC_SUCCESS;
}

int commit(pic_ctx *c) {
    if (c->committed) return PIC_SUCCESS;
    const long NC = c->NC;
    const long P = (long)c->staged.cell.size();
    if (c->tape_mode == PIC_TAPE_RECORD) {
        c->tape_cells.assign(c->n_ids, -1);
        for (long j = 0; j < P; j++) c->tape_cells[c->staged.id[j]] = c->staged.cell[j];
        c->tape_cells_set = true;
    } else if (c->tape_mode == PIC_TAPE_REPLAY) {
        /* every loaded particle starts in its recorded cell */
        long kept = 0;
        for (long k = 0; k < (long)c->tape_cells.size(); k++) kept += c->tape_cells[k] >= 0;
        if (kept != P || (long)c->tape_cells.size() != c->n_ids) {
            snprintf(g_hip_err, sizeof(g_hip_err),
                     "decision tape: %ld particles recorded in cells, %ld loaded", kept, P);
            return PIC_ERR_STATE;
        }
        for (long j = 0; j < P; j++) {
            const int cell = c->tape_cells[c->staged.id[j]];
            if (cell < 0 || cell >= NC) {
                snprintf(g_hip_err, sizeof(g_hip_err), "decision tape: particle %d was dropped",
                         c->staged.id[j]);
                return PIC_ERR_STATE;
            }
            c->staged.cell[j] = cell;
        }
    }
    std::vector<long> off(NC + 1, 0);
    for (long j = 0; j < P; j++) off[c->staged.cell[j] + 1]++;
    int maxc = 0;
    for (long k = 0; k < NC; k++) maxc = std::max(maxc, (int)off[k + 1]);
    for (long k = 0; k < NC; k++) off[k + 1] += off[k];
    int cap = c->prm.bucket_capacity > 0 ? c->prm.bucket_capacity : auto_capacity(maxc);
    if (maxc > cap) cap = auto_capacity(maxc);
    if (!push_span_ok(c, cap)) return PIC_ERR_INVALID;
    int rc = c->fp32 ? commit_into(c, c->p32, c->q32, cap, off)
                     : commit_into(c, c->p, c->q, cap, off);
    if (rc) return rc;
    Staged empty;
    std::swap(c->staged, empty);
    c->committed = true;
    c->last.bucket_capacity = cap;
    return PIC_SUCCESS;
}

/* bucket overflow repair: grow every bucket, re-place the arrivals whose
 * claimed slot did not fit (FP64 form) */
template <class P, class Q>
int regrow_parts(pic_ctx *c, P &parts, Q &queue) {
    const long NC = c->NC;
    std::vector<int> cnt(NC);
    HIPCHK(hipMemcpy(cnt.data(), parts.count, NC * sizeof(int), hipMemcpyDeviceToHost));
    int maxc = 0;
    for (long k = 0; k < NC; k++) maxc = std::max(maxc, cnt[k]);
    const int oldcap = parts.cap;
    const int cap = std::max(auto_capacity(maxc),
                             (oldcap + 8 + PIC_CAP_ALIGN - 1) / PIC_CAP_ALIGN * PIC_CAP_ALIGN);
    if (!push_span_ok(c, cap)) return PIC_ERR_INVALID;
    P np;
    int rc = alloc_parts(np, NC, cap, false, c->diff);
    if (rc) return rc;
    np.count = parts.count;
    launch_repack(NC, parts, np, c->stream);
    launch_fix_arrivals(c->t, queue, oldcap, np, c->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    parts.count = nullptr;
    free_parts(parts);
    parts = np;
    Q nq;
    rc = alloc_queue(nq, c->t, queue_capacity(c->t, cap), c->diff);
    if (rc) return rc;
    free_queue(queue);
    queue = nq;
    c->regrows++;
    c->last.bucket_capacity = cap;
    c->last.bucket_regrows = c->regrows;
    return PIC_SUCCESS;
}

/* FP32 form: a count past the capacity after the arrivals of a push: the
 * packed set, its inboxes and its arrivals past the inbox (from the spill
 * lists, or from the departure queues after a full spill list) unpacked
 * into buckets of a larger capacity */
int regrow32(pic_ctx *c, bool from_queue) {
    const long NC = c->NC;
    if (!c->lay32.packed) {
        snprintf(g_hip_err, sizeof(g_hip_err), "bucket overflow of an unpacked particle set");
        return PIC_ERR_STATE;
    }
    std::vector<int> cnt(NC);
    HIPCHK(hipMemcpy(cnt.data(), c->p32.count, NC * sizeof(int), hipMemcpyDeviceToHost));
    int maxc = 0;
    for (long k = 0; k < NC; k++) maxc = std::max(maxc, cnt[k]);
    const int oldcap = c->p32.cap;
    const int cap = std::max(auto_capacity(maxc),
                             (oldcap + 8 + PIC_CAP_ALIGN - 1) / PIC_CAP_ALIGN * PIC_CAP_ALIGN);
    if (!push_span_ok(c, cap)) return PIC_ERR_INVALID;
    PartPtrs32 na{}, nb{};
    int rc;
    if ((rc = alloc_set32(c, na, cap)) || (rc = alloc_set32(c, nb, cap))) return rc;
    launch_unpack32(c->g, c->t, c->p32, c->lay32, na, from_queue ? &c->q32 : nullptr, c->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    free_set32(c->p32);
    free_set32(c->alt32);
    for (int k = 0; k < 6; k++) c->p32.p[k] = na.p[k];
    c->p32.sp = na.sp;
    c->p32.id = na.id;
    c->p32.cap = cap;
    c->alt32 = nb;
    c->lay32.tcap = (long)c->t.TX * c->t.TY * c->t.TZ * cap;
    c->lay32.packed = 0;
    QueuePtrs32 nq;
    if ((rc = alloc_queue(nq, c->t, queue_capacity(c->t, cap), c->diff))) return rc;
    free_queue(c->q32);
    c->q32 = nq;
    c->regrows++;
    c->last.bucket_capacity = cap;
    c->last.bucket_regrows = c->regrows;
    return PIC_SUCCESS;
}

int *count_ptr(pic_ctx *c) { return c->fp32 ? c->p32.count : c->p.count; }
int bucket_cap(pic_ctx *c) { return c->fp32 ? c->p32.cap : c->p.cap; }

/* FP32 form: the packed live set into buckets of the other set, which
 * becomes the live one */
int unpack_live32(pic_ctx *c, bool from_queue) {
    PartPtrs32 to = c->alt32;
    to.cap = c->p32.cap;
    launch_unpack32(c->g, c->t, c->p32, c->lay32, to, from_queue ? &c->q32 : nullptr, c->stream);
    HIPCHK(hipGetLastError());
    swap_sets32(c);
    c->lay32.packed = 0;
    return PIC_SUCCESS;
}

int settle_noalloc(pic_ctx *c) {
    if (!c->fp32 || !c->committed || !c->lay32.packed) return PIC_SUCCESS;
    if (c->arrival_pending) {
        snprintf(g_hip_err, sizeof(g_hip_err),
                 "the last arrival's status is not examined yet (pic_sync first)");
        return PIC_ERR_STATE;
    }
    return unpack_live32(c, false);
}

int settle(pic_ctx *c) {
    if (c->fp32 && c->arrival_pending) {
        const int rc = finish_step(c);
        if (rc) return rc;
    }
    return settle_noalloc(c);
}

int despill(pic_ctx *c) { return unpack_live32(c, true); }
int *qcount_ptr(pic_ctx *c) { return c->fp32 ? c->q32.count : c->q.count; }

/* stage launchers dispatching on the form */
void enq_h_half(pic_ctx *c) {
    if (c->fp32) launch_h_half32(c->g, c->f32, c->stream);
    else launch_h_half(c->g, c->f, c->stream);
}
void enq_push(pic_ctx *c) {
    if (c->fp32) {
        launch_push_deposit_f32(c->g, c->t, c->f32, c->p32, c->alt32, c->lay32, c->q32, c->sp,
                                c->prm.departure_cap, c->d_st,
                                c->diff ? &c->rec_snap32 : nullptr, c->rec_off, c->rec_room,
                                c->stream);
        swap_sets32(c);
        swap_arrivals32(c);
        c->lay32.packed = 1;
    } else if (c->diff) {
        DiffArgs d{};
        for (int k = 0; k < 3; k++) d.jfix[k] = c->d_jfix[k];
        d.snap = c->rec_snap;
        d.snap_off = c->rec_off;
        d.snap_room = c->rec_room;
        if (c->tape_mode != PIC_TAPE_OFF && c->tape_step < c->tape_nsteps) {
            d.tape_mode = c->tape_mode;
            d.tape_row = c->d_tape + (size_t)c->tape_step * c->tape_nids;
            d.tape_n = c->tape_nids;
            if (++c->tape_step >= c->tape_nsteps) c->tape_mode = PIC_TAPE_OFF;
        }
        launch_push_deposit_f64(c->g, c->t, c->f, c->p, c->q, c->sp, c->prm.departure_cap,
                                c->d_st, &d, c->stream);
        launch_jfix_to_f64(c->g, c->f, d, c->stream);
    } else {
        launch_push_deposit_f64(c->g, c->t, c->f, c->p, c->q, c->sp, c->prm.departure_cap,
                                c->d_st, nullptr, c->stream);
    }
}
void enq_arrive(pic_ctx *c, hipStream_t st) {
    if (c->fp32) {
        /* the push placed the arrivals; the counts take them in */
        launch_arrivals_done32(c->g, c->t, c->p32, c->lay32, c->d_st, st);
        c->arrival_pending = true;
    }
    else launch_arrive_f64(c->g, c->t, c->p, c->q, c->d_st, st);
}
void enq_arrive(pic_ctx *c) { enq_arrive(c, c->stream); }
void enq_fold(pic_ctx *c) {
    if (c->fp32) launch_fold_j32(c->g, c->f32, c->stream);
    else launch_fold_j(c->g, c->f, c->stream);
}
void enq_h2(pic_ctx *c) {
    if (c->fp32) launch_h2_32(c->g, c->f32, c->stream);
    else launch_h2(c->g, c->f, c->stream);
}
void enq_e(pic_ctx *c) {
    if (c->fp32) launch_e32(c->g, c->f32, c->stream);
    else launch_e(c->g, c->f, c->stream);
}
void enq_e_periodic(pic_ctx *c) {
    if (c->fp32) launch_e_periodic32(c->g, c->f32, c->stream);
    else launch_e_periodic(c->g, c->f, c->stream);
}
/* the step's field update after the push in one pass: current fold, H
 * second half, E update and periodic E ghosts (equal to enq_fold, enq_h2,
 * enq_e, enq_e_periodic bit for bit), H^{n+1} landing in the second H
 * arrays, which then become the live ones */
void enq_fold_h2_e(pic_ctx *c) {
    if (c->fp32) {
        launch_fold_h2_e32(c->g, c->f32, c->h_alt32, c->stream);
        for (int k = 0; k < 3; k++) std::swap(c->f32.H[k], c->h_alt32[k]);
    } else {
        launch_fold_h2_e(c->g, c->f, c->h_alt, c->stream);
        for (int k = 0; k < 3; k++) std::swap(c->f.f[PIC_HX + k], c->h_alt[k]);
    }
}
void enq_energy(pic_ctx *c) {
    if (c->fp32) launch_energy32(c->g, c->f32, c->d_part, c->npart, c->d_red, c->stream);
    else launch_energy(c->g, c->f, c->d_part, c->npart, c->d_red, c->stream);
}

int check_ready(pic_ctx *c) {
    if (!c) return PIC_ERR_INVALID;
    if (!c->committed) return commit(c);
    return PIC_SUCCESS;
}

/* stage timing: an event at the start of each stage that has kernels of
 * its own; a stage whose work a fused pass does is not marked and counts 0 */
inline void mark(pic_ctx *c, int k) {
    if (!c->prof) return;
    (void)hipEventRecord(c->ev[k], c->stream);
    c->marked |= 1u << k;
}

void accumulate_profile(pic_ctx *c) {
    if (!c->prof) return;
    for (int k = 0; k < PIC_NSTAGES; k++) {
        if (!(c->marked & (1u << k))) {
            c->prof_n[k] += 1;
            continue;
        }
        int j = k + 1;
        while (j < PIC_NSTAGES && !(c->marked & (1u << j))) j++;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->ev[k], c->ev[j]) == hipSuccess) {
            c->prof_ms[k] += ms;
            c->prof_n[k] += 1;
        }
    }
    c->marked = 0;
}

/* one whole step enqueued on the context stream (Plasma::Step,
 * gpu_plasma.h:605-620); the energy lands in h_red[0] after the sync */
int enq_step(pic_ctx *c) {
    launch_reset_status(c->d_st, c->stream);
    mark(c, PIC_STAGE_H_HALF);
    enq_h_half(c);
    mark(c, PIC_STAGE_PUSH_DEPOSIT);
    enq_push(c);
    mark(c, PIC_STAGE_MIGRATE);
    /* the arrival (particles only) beside the field stages (fields only);
     * with stage timing on, one after the other so that each stage's events
     * bracket its own kernels */
    const bool overlap = c->side && !c->prof;
    if (overlap) {
        HIPCHK(hipEventRecord(c->ev_dep, c->stream));
        HIPCHK(hipStreamWaitEvent(c->side, c->ev_dep, 0));
        enq_arrive(c, c->side);
        HIPCHK(hipEventRecord(c->ev_join, c->side));
    } else {
        enq_arrive(c, c->stream);
    }
    /* current fold, H second half, E update and the periodic E ghosts in
     * one pass: its time is the E stage's (FOLD_J, H2, E_PERIODIC: 0) */
    mark(c, PIC_STAGE_E);
    enq_fold_h2_e(c);
    mark(c, PIC_STAGE_ENERGY);
    enq_energy(c);
    mark(c, PIC_NSTAGES);
    if (overlap) HIPCHK(hipStreamWaitEvent(c->stream, c->ev_join, 0));
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->h_red, c->d_red, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    return PIC_SUCCESS;
}

int finish_step(pic_ctx *c) {
    HIPCHK(hipMemcpyAsync(c->h_st, c->d_st, sizeof(DevStatus), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const DevStatus st = *c->h_st;
    const bool arrived = c->arrival_pending;
    c->arrival_pending = false;
    c->last.dropped = (long long)st.dropped;
    c->last.duplicated = (long long)st.duplicated;
    c->last.max_cell_departures = st.max_dep;
    c->last.flags = (int)(st.flags & ~(unsigned)PIC_F_SPILL_OVER);
    c->last.bucket_capacity = bucket_cap(c);
    c->last.bucket_regrows = c->regrows;
    c->last.spill_repairs = (int)c->spill_repairs;
    if (st.flags & PIC_F_QUEUE_CORRUPT) return PIC_ERR_QUEUE_CORRUPT;
    /* FP32 form: a count past the capacity (larger buckets), arrivals that
     * found no room in their tile's spill list (still in the departure
     * queues): every particle into buckets */
    if (c->fp32 && arrived && c->lay32.packed) {
        const bool over = st.flags & PIC_F_SPILL_OVER;
        int rc = PIC_SUCCESS;
        if (st.flags & PIC_F_BUCKET_OVERFLOW) rc = regrow32(c, over);
        else if (over) rc = despill(c);
        if (rc) return rc;
        if (over) {
            /* the lists are empty now; larger ones for the steps to come
             * (and for a recompute of these steps) */
            c->spill_repairs++;
            HIPCHK(hipStreamSynchronize(c->stream));
            if (c->lay32.scap < PIC_SPILL_CAP_MAX &&
                (rc = alloc_spill32(c, std::min(4 * c->lay32.scap, PIC_SPILL_CAP_MAX))))
                return rc;
        }
        c->last.spill_repairs = (int)c->spill_repairs;
    } else if (!c->fp32 && (st.flags & PIC_F_BUCKET_OVERFLOW)) {
        int rc = regrow_parts(c, c->p, c->q);
        if (rc) return rc;
    }
    if (st.flags & PIC_F_QUEUE_OVERFLOW) return PIC_ERR_QUEUE_OVERFLOW;
    if (st.flags & PIC_F_RANGE) return PIC_ERR_CELL_RANGE;
    if (st.flags & PIC_F_RECORD) {
        snprintf(g_hip_err, sizeof(g_hip_err), "step record holds fewer particles than the box");
        return PIC_ERR_STATE;
    }
    if (st.flags & PIC_F_TOO_MANY) return PIC_ERR_TOO_MANY_PARTICLES;
    return PIC_SUCCESS;
}

}  // namespace picx

using namespace picx;

extern "C" {

/* ------------------------------------------------------------------ archAPI */
int MemoryCopy(void *dst, void *src, size_t size, int dir) {
    hipMemcpyKind k;
    switch (dir) {
    case HOST_TO_DEVICE: k = hipMemcpyHostToDevice; break;
    case HOST_TO_HOST: k = hipMemcpyHostToHost; break;
    case DEVICE_TO_HOST: k = hipMemcpyDeviceToHost; break;
    case DEVICE_TO_DEVICE: k = hipMemcpyDeviceToDevice; break;
    default: return (int)hipErrorInvalidValue;
    }
    return (int)hipMemcpy(dst, src, size, k);
}
int MemoryAllocate(void **dst, size_t size) { return (int)hipMalloc(dst, size); }
int MemoryFree(void *p) { return (int)hipFree(p); }
int GetDeviceMemory(size_t *m_free, size_t *m_total) { return (int)hipMemGetInfo(m_free, m_total); }
int MemorySet(void *s, int c, size_t n) { return (int)hipMemset(s, c, n); }
int ThreadSynchronize(void) { return (int)hipDeviceSynchronize(); }
int DeviceSynchronize(void) { return (int)hipDeviceSynchronize(); }
int getLastError(void) { return (int)hipGetLastError(); }
const char *getErrorString(int err) { return hipGetErrorString((hipError_t)err); }
int SetDevice(int n) { return (int)hipSetDevice(n); }

/* ------------------------------------------------------------------ context */
void pic_params_default(pic_params *p) {
    std::memset(p, 0, sizeof(*p));
    p->nx = 32; p->ny = 8; p->nz = 8;
    p->lx = 3.2; p->ly = 0.8; p->lz = 0.8;
    p->ni = 1.0;
    p->n_per_cell = 16;
    p->ion_q_m = 1.0 / 1836.0;
    p->tau = 0.05;
    p->precision = PIC_FP64;
    p->departure_cap = 50;
    p->bucket_capacity = 0;
    p->device = 0;
}

const char *pic_error_string(int s) {
    switch (s) {
    case PIC_SUCCESS: return "success";
    case PIC_ERR_INVALID: return "invalid argument";
    case PIC_ERR_HIP: return "HIP runtime error";
    case PIC_ERR_NO_DEVICE: return "no HIP device";
    case PIC_ERR_TOO_MANY_PARTICLES: return "too many particles leaving one cell";
    case PIC_ERR_QUEUE_OVERFLOW: return "departure queue overflow";
    case PIC_ERR_NOMEM: return "device out of memory";
    case PIC_ERR_STATE: return "call out of order";
    case PIC_ERR_CELL_RANGE: return "particle stencil outside its tile";
    case PIC_ERR_QUEUE_CORRUPT: return "departure record out of range (not placed)";
    default: return "unknown status";
    }
}

const char *pic_last_hip_error(void) { return g_hip_err; }

int pic_create(const pic_params *prm, pic_ctx **out) {
    if (!prm || !out) return PIC_ERR_INVALID;
    *out = nullptr;
    if (prm->nx < 3 || prm->ny < 3 || prm->nz < 3) return PIC_ERR_INVALID;
    /* the field kernels index the ghosted nodes in 32 bits */
    if ((long)(prm->nx + 2) * (prm->ny + 2) * (prm->nz + 2) >= (1L << 31)) return PIC_ERR_INVALID;
    if (!(prm->lx > 0 && prm->ly > 0 && prm->lz > 0 && prm->tau > 0)) return PIC_ERR_INVALID;
    if (prm->precision != PIC_FP64 && prm->precision != PIC_FP32) return PIC_ERR_INVALID;
    if (prm->ghost_mode != 0 && prm->ghost_mode != 1) return PIC_ERR_INVALID;
    /* The FP32 push takes a particle's end to lie in the start's dual cell
     * or the next one and its new cell to be the owner or a face neighbour:
     * a displacement below half a cell per axis.  With |v| < c = 1 that holds
     * for every particle when tau <= h/2 on every axis. */
    if (prm->precision == PIC_FP32 &&
        (prm->tau > 0.5 * (prm->lx / prm->nx) || prm->tau > 0.5 * (prm->ly / prm->ny) ||
         prm->tau > 0.5 * (prm->lz / prm->nz)))
        return PIC_ERR_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PIC_ERR_NO_DEVICE;
    if (prm->device < 0 || prm->device >= ndev) return PIC_ERR_INVALID;
    HIPCHK(hipSetDevice(prm->device));
    if (int rc0 = noalloc_init(prm->device)) return rc0;
    pic_ctx *c = new pic_ctx();
    c->prm = *prm;
    c->dev = prm->device;
    Grid &g = c->g;
    g.Nx = prm->nx; g.Ny = prm->ny; g.Nz = prm->nz;
    g.sy = g.Nz + 2;
    g.sx = (g.Ny + 2) * (g.Nz + 2);
    g.G = (long)(g.Nx + 2) * g.sx;
    g.Lx = prm->lx; g.Ly = prm->ly; g.Lz = prm->lz;
    /* Cell ctor (cell.h:731-737): h = L / N */
    g.hx = prm->lx / (double)prm->nx;
    g.hy = prm->ly / (double)prm->ny;
    g.hz = prm->lz / (double)prm->nz;
    g.tau = prm->tau;
    g.ghost_mode = prm->ghost_mode;
    c->NC = (long)g.Nx * g.Ny * g.Nz;
    c->t = make_tile_geom(g.Nx, g.Ny, g.Nz);
    int rc = PIC_SUCCESS;
    auto fail = [&](int r) {
        pic_destroy(c);
        return r;
    };
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(record_hip(e, "hipStreamCreate"));
    /* both streams at the default priority */
    e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(record_hip(e, "hipStreamCreate"));
    e = hipEventCreateWithFlags(&c->ev_dep, hipEventDisableTiming);
    if (e != hipSuccess) return fail(record_hip(e, "hipEventCreate"));
    e = hipEventCreate(&c->ev_join);
    if (e != hipSuccess) return fail(record_hip(e, "hipEventCreate"));
    c->fp32 = prm->precision == PIC_FP32;
    for (int k = 0; k < PIC_NFIELDS; k++) {
        void **ptr;
        size_t esz = sizeof(double);
        if (!c->fp32) {
            ptr = (void **)&c->f.f[k];
        } else if (k >= PIC_JX && k <= PIC_JZ) {
            ptr = (void **)&c->f32.J[k - PIC_JX];
            esz = sizeof(long long);
        } else {
            esz = sizeof(float);
            ptr = k < PIC_HX ? (void **)&c->f32.E[k] : k < PIC_JX ? (void **)&c->f32.H[k - PIC_HX]
                                                               : (void **)&c->f32.Q[k - PIC_QX];
        }
        e = hipMalloc(ptr, g.G * esz);
        if (e != hipSuccess) return fail(record_hip(e, "hipMalloc fields"));
        e = hipMemset(*ptr, 0, g.G * esz);
        if (e != hipSuccess) return fail(record_hip(e, "hipMemset fields"));
    }
    for (int k = 0; k < 3; k++) {
        void **ptr = c->fp32 ? (void **)&c->h_alt32[k] : (void **)&c->h_alt[k];
        const size_t esz = c->fp32 ? sizeof(float) : sizeof(double);
        e = hipMalloc(ptr, g.G * esz);
        if (e != hipSuccess) return fail(record_hip(e, "hipMalloc fields"));
        e = hipMemset(*ptr, 0, g.G * esz);
        if (e != hipSuccess) return fail(record_hip(e, "hipMemset fields"));
    }
    c->diff = prm->differentiable != 0;
    for (int k = 0; c->diff && !c->fp32 && k < 3; k++) {
        if ((e = hipMalloc(&c->d_jfix[k], g.G * sizeof(long long))) != hipSuccess)
            return fail(record_hip(e, "hipMalloc jfix"));
        if ((e = hipMemset(c->d_jfix[k], 0, g.G * sizeof(long long))) != hipSuccess)
            return fail(record_hip(e, "hipMemset jfix"));
    }
    if ((e = hipMalloc(&c->d_st, sizeof(DevStatus))) != hipSuccess) return fail(record_hip(e, "status"));
    if ((e = hipMemset(c->d_st, 0, sizeof(DevStatus))) != hipSuccess) return fail(record_hip(e, "status"));
    if ((e = hipHostMalloc(&c->h_st, sizeof(DevStatus))) != hipSuccess) return fail(record_hip(e, "status"));
    if ((e = hipMalloc(&c->d_part, (size_t)c->npart * 27 * sizeof(double))) != hipSuccess)
        return fail(record_hip(e, "partials"));
    if ((e = hipMalloc(&c->d_red, 64 * sizeof(double))) != hipSuccess) return fail(record_hip(e, "red"));
    if ((e = hipHostMalloc(&c->h_red, 64 * sizeof(double))) != hipSuccess) return fail(record_hip(e, "red"));
    if ((e = hipMalloc(&c->d_slot_sort, PIC_MAX_SLOTS * sizeof(int))) != hipSuccess)
        return fail(record_hip(e, "slots"));
    (void)rc;
    *out = c;
    return PIC_SUCCESS;
}

int pic_destroy(pic_ctx *c) {
    if (!c) return PIC_SUCCESS;
    hipSetDevice(c->dev);
    if (c->stream) hipStreamSynchronize(c->stream);
    for (int k = 0; k < PIC_NFIELDS; k++)
        if (c->f.f[k]) (void)hipFree(c->f.f[k]);
    for (int k = 0; k < 3; k++) {
        if (c->f32.E[k]) (void)hipFree(c->f32.E[k]);
        if (c->f32.H[k]) (void)hipFree(c->f32.H[k]);
        if (c->f32.Q[k]) (void)hipFree(c->f32.Q[k]);
        if (c->f32.J[k]) (void)hipFree(c->f32.J[k]);
    }
    for (int k = 0; k < 3; k++) {
        if (c->h_alt[k]) (void)hipFree(c->h_alt[k]);
        if (c->h_alt32[k]) (void)hipFree(c->h_alt32[k]);
    }
    free_adjoint(c);
    for (int k = 0; k < 3; k++)
        if (c->d_jfix[k]) (void)hipFree(c->d_jfix[k]);
    free_parts(c->p);
    free_queue(c->q);
    free_parts(c->p32);
    free_parts(c->alt32);
    free_layout32(c->lay32);
    free_queue(c->q32);
    if (c->d_rho) (void)hipFree(c->d_rho);
    if (c->d_gres) (void)hipFree(c->d_gres);
    if (c->d_st) hipFree(c->d_st);
    if (c->h_st) hipHostFree(c->h_st);
    if (c->d_part) hipFree(c->d_part);
    if (c->d_red) hipFree(c->d_red);
    if (c->h_red) hipHostFree(c->h_red);
    if (c->d_slot_sort) hipFree(c->d_slot_sort);
    if (c->d_tape) hipFree(c->d_tape);
    for (int k = 0; k <= PIC_NSTAGES; k++)
        if (c->ev[k]) hipEventDestroy(c->ev[k]);
    if (c->ev_dep) hipEventDestroy(c->ev_dep);
    if (c->ev_join) hipEventDestroy(c->ev_join);
    if (c->side) hipStreamDestroy(c->side);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return PIC_SUCCESS;
}

int pic_get_params(pic_ctx *c, pic_params *out) {
    if (!c || !out) return PIC_ERR_INVALID;
    *out = c->prm;
    return PIC_SUCCESS;
}

int pic_load_particles(pic_ctx *c, int sort, const double *x, const double *y, const double *z,
                       const double *px, const double *py, const double *pz, long n, double m,
                       double q_m, long *n_dropped) {
    if (!c || n < 0 || (n > 0 && (!x || !y || !z || !px || !py || !pz))) return PIC_ERR_INVALID;
    if (sort < 0 || sort > 2 || q_m == 0.0) return PIC_ERR_INVALID;
    if (c->committed) return PIC_ERR_STATE;
    int slot = -1;
    for (int s = 0; s < c->nslots; s++)
        if (c->slot_sort[s] == sort && c->sp.m[s] == m && c->sp.q_m[s] == q_m) slot = s;
    if (slot < 0) {
        if (c->nslots >= PIC_MAX_SLOTS) return PIC_ERR_INVALID;
        slot = c->nslots++;
        c->slot_sort[slot] = sort;
        c->sp.m[slot] = m;
        c->sp.q_m[slot] = q_m;
        HIPCHK(hipMemcpy(c->d_slot_sort, c->slot_sort, sizeof(c->slot_sort), hipMemcpyHostToDevice));
    }
    const Grid &g = c->g;
    long dropped = 0;
    Staged &S = c->staged;
    for (int k = 0; k < 6; k++) S.v[k].reserve(S.v[k].size() + n);
    S.sp.reserve(S.sp.size() + n);
    S.cell.reserve(S.cell.size() + n);
    if (c->diff) S.id.reserve(S.id.size() + n);
    for (long j = 0; j < n; j++) {
        /* getPointCell (cell.h:593-598): (int)((x - 0.0)/h) per axis */
        int i = (int)((x[j] - 0.0) / g.hx);
        int l = (int)((y[j] - 0.0) / g.hy);
        int k = (int)((z[j] - 0.0) / g.hz);
        if (i < 0 || l < 0 || k < 0 || i >= g.Nx || l >= g.Ny || k >= g.Nz) {
            dropped++;
            continue;
        }
        /* Cell::Insert -> isPointInCell (cell.h:717, 1754-1768) */
        double x0 = (double)i * g.hx, y0 = (double)l * g.hy, z0 = (double)k * g.hz;
        if (!((x0 <= x[j]) && (x[j] < x0 + g.hx) && (y0 <= y[j]) && (y[j] < y0 + g.hy) &&
              (z0 <= z[j]) && (z[j] < z0 + g.hz))) {
            dropped++;
            continue;
        }
        S.v[0].push_back(x[j]);
        S.v[1].push_back(y[j]);
        S.v[2].push_back(z[j]);
        S.v[3].push_back(px[j]);
        S.v[4].push_back(py[j]);
        S.v[5].push_back(pz[j]);
        S.sp.push_back((unsigned char)slot);
        S.cell.push_back((int)(((long)i * g.Ny + l) * g.Nz + k));
        if (c->diff) S.id.push_back((int)(c->n_ids + j));
    }
    c->n_ids += n;
    if (n_dropped) *n_dropped = dropped;
    return PIC_SUCCESS;
}

int pic_generate_device(pic_ctx *c, const pic_species_spec *specs, int nspec,
                        unsigned long long seed, long *n_out) {
    if (!c || !specs || nspec < 1 || nspec > PIC_MAX_SLOTS) return PIC_ERR_INVALID;
    if (c->committed || !c->staged.cell.empty()) return PIC_ERR_STATE;
    const pic_params &gp = c->prm;
    const Grid &g = c->g;
    const long NC = c->NC, NLK = (long)g.Ny * g.Nz;
    std::vector<std::vector<int>> cnt(nspec, std::vector<int>(NLK));
    std::vector<std::vector<long>> off(nspec, std::vector<long>(NLK + 1, 0));
    std::vector<int> tot(NLK, 0);
    for (int s = 0; s < nspec; s++) {
        const pic_species_spec &sp = specs[s];
        if (sp.sort < 0 || sp.sort > 2 || sp.q_m == 0.0) return PIC_ERR_INVALID;
        for (int q = 0; q < 8; q++)
            if (sp.perturb[q] != 0.0) return PIC_ERR_INVALID;   /* moves particles between cells */
        pic_species_cell_counts(&gp, &sp, cnt[s].data());
        for (long j = 0; j < NLK; j++) {
            off[s][j + 1] = off[s][j] + cnt[s][j];
            tot[j] += cnt[s][j];
        }
    }
    int maxc = 0;
    for (long j = 0; j < NLK; j++) maxc = std::max(maxc, tot[j]);
    int cap = c->prm.bucket_capacity > 0 ? c->prm.bucket_capacity : auto_capacity(maxc);
    if (maxc > cap) cap = auto_capacity(maxc);
    if (!push_span_ok(c, cap)) return PIC_ERR_INVALID;
    HIPCHK(hipSetDevice(c->dev));
    int rc = c->fp32 ? alloc_live32(c, cap) : alloc_parts(c->p, NC, cap, true, c->diff);
    if (rc) return rc;
    rc = c->fp32 ? alloc_queue(c->q32, c->t, queue_capacity(c->t, cap), c->diff)
                 : alloc_queue(c->q, c->t, queue_capacity(c->t, cap), c->diff);
    if (rc) return rc;
    HIPCHK(hipMemsetAsync(count_ptr(c), 0, NC * sizeof(int), c->stream));
    int *d_cnt = nullptr, *d_start = nullptr, *d_fail = nullptr;
    long *d_off = nullptr;
    unsigned long long *d_bad = nullptr;
    HIPCHK(hipMalloc(&d_cnt, NLK * sizeof(int)));
    HIPCHK(hipMalloc(&d_off, (NLK + 1) * sizeof(long)));
    HIPCHK(hipMalloc(&d_start, NC * sizeof(int)));
    HIPCHK(hipMalloc(&d_bad, sizeof(unsigned long long)));
    HIPCHK(hipMalloc(&d_fail, sizeof(int)));
    HIPCHK(hipMemsetAsync(d_bad, 0, sizeof(unsigned long long), c->stream));
    HIPCHK(hipMemsetAsync(d_fail, 0, sizeof(int), c->stream));
    long total = 0;
    for (int s = 0; s < nspec; s++) {
        const pic_species_spec &sp = specs[s];
        /* getMassCharge (maxwell.cu:606-619): weight = q n / ppc */
        const double m = pic_species_weight(&c->prm, &sp);
        int slot = -1;
        for (int q = 0; q < c->nslots; q++)
            if (c->slot_sort[q] == sp.sort && c->sp.m[q] == m && c->sp.q_m[q] == sp.q_m) slot = q;
        if (slot < 0) {
            if (c->nslots >= PIC_MAX_SLOTS) return PIC_ERR_INVALID;
            slot = c->nslots++;
            c->slot_sort[slot] = sp.sort;
            c->sp.m[slot] = m;
            c->sp.q_m[slot] = sp.q_m;
        }
        HIPCHK(hipMemcpyAsync(d_cnt, cnt[s].data(), NLK * sizeof(int), hipMemcpyHostToDevice,
                              c->stream));
        HIPCHK(hipMemcpyAsync(d_off, off[s].data(), (NLK + 1) * sizeof(long),
                              hipMemcpyHostToDevice, c->stream));
        const long plane = off[s][NLK];
        launch_generate_species(g, sp, d_cnt, d_off, plane, seed, (unsigned char)slot,
                                c->n_ids + total, c->fp32 ? nullptr : &c->p,
                                c->fp32 ? &c->p32 : nullptr, d_start, d_bad, d_fail, c->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));   /* the host count vectors are reused */
        total += plane * g.Nx;
    }
    HIPCHK(hipMemcpy(c->d_slot_sort, c->slot_sort, sizeof(c->slot_sort), hipMemcpyHostToDevice));
    if (c->fp32)
        HIPCHK(hipMemcpy(c->p32.bcount, c->p32.count, NC * sizeof(int), hipMemcpyDeviceToDevice));
    unsigned long long bad = 0;
    int fail = 0;
    HIPCHK(hipMemcpy(&bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&fail, d_fail, sizeof(fail), hipMemcpyDeviceToHost));
    (void)hipFree(d_cnt);
    (void)hipFree(d_off);
    (void)hipFree(d_start);
    (void)hipFree(d_bad);
    (void)hipFree(d_fail);
    c->n_ids += total;
    c->committed = true;
    c->last.bucket_capacity = cap;
    if (n_out) *n_out = total - (long)bad;
    if (bad || fail) {
        snprintf(g_hip_err, sizeof(g_hip_err),
                 "device initializer: %llu particles outside their cell, generator fail %d", bad,
                 fail);
        return PIC_ERR_STATE;
    }
    return PIC_SUCCESS;
}

int pic_load_field(pic_ctx *c, int which, const double *host) {
    if (!c || which < 0 || which >= PIC_NFIELDS || !host) return PIC_ERR_INVALID;
    HIPCHK(hipSetDevice(c->dev));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!c->fp32) {
        HIPCHK(hipMemcpy(c->f.f[which], host, c->g.G * sizeof(double), hipMemcpyHostToDevice));
    } else if (which >= PIC_JX && which <= PIC_JZ) {
        std::vector<long long> t(c->g.G);
        for (long n = 0; n < c->g.G; n++) t[n] = (long long)llrint(host[n] * PIC_J_SCALE);
        HIPCHK(hipMemcpy(c->f32.J[which - PIC_JX], t.data(), c->g.G * sizeof(long long),
                         hipMemcpyHostToDevice));
    } else {
        std::vector<float> t(c->g.G);
        for (long n = 0; n < c->g.G; n++) t[n] = (float)host[n];
        float *dst = which < PIC_HX ? c->f32.E[which] : which < PIC_JX ? c->f32.H[which - PIC_HX]
                                                                       : c->f32.Q[which - PIC_QX];
        HIPCHK(hipMemcpy(dst, t.data(), c->g.G * sizeof(float), hipMemcpyHostToDevice));
    }
    return PIC_SUCCESS;
}

int pic_get_field(pic_ctx *c, int which, double *host) {
    if (!c || which < 0 || which >= PIC_NFIELDS || !host) return PIC_ERR_INVALID;
    HIPCHK(hipSetDevice(c->dev));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!c->fp32) {
        HIPCHK(hipMemcpy(host, c->f.f[which], c->g.G * sizeof(double), hipMemcpyDeviceToHost));
    } else if (which >= PIC_JX && which <= PIC_JZ) {
        std::vector<long long> t(c->g.G);
        HIPCHK(hipMemcpy(t.data(), c->f32.J[which - PIC_JX], c->g.G * sizeof(long long),
                         hipMemcpyDeviceToHost));
        for (long n = 0; n < c->g.G; n++) host[n] = (double)t[n] * PIC_J_INV;
    } else {
        std::vector<float> t(c->g.G);
        const float *src = which < PIC_HX ? c->f32.E[which]
                           : which < PIC_JX ? c->f32.H[which - PIC_HX] : c->f32.Q[which - PIC_QX];
        HIPCHK(hipMemcpy(t.data(), src, c->g.G * sizeof(float), hipMemcpyDeviceToHost));
        for (long n = 0; n < c->g.G; n++) host[n] = (double)t[n];
    }
    return PIC_SUCCESS;
}

/* ------------------------------------------------------------------ stages */
int pic_yee_h_half(pic_ctx *c) {
    int rc = check_ready(c);
    if (rc) return rc;
    launch_reset_status(c->d_st, c->stream);
    enq_h_half(c);
    HIPCHK(hipGetLastError());
    return PIC_SUCCESS;
}

int pic_push_deposit(pic_ctx *c) {
    int rc = check_ready(c);
    if (rc) return rc;
    /* an arrival whose status was not examined (no pic_sync since): its
     * repair comes first */
    if (c->arrival_pending && (rc = finish_step(c))) return rc;
    enq_push(c);
    HIPCHK(hipGetLastError());
    return PIC_SUCCESS;
}

int pic_migrate(pic_ctx *c) {
    int rc = check_ready(c);
    if (rc) return rc;
    enq_arrive(c);
    HIPCHK(hipGetLastError());
    return PIC_SUCCESS;
}

int pic_fold_j(pic_ctx *c) {
    if (!c) return PIC_ERR_INVALID;
    enq_fold(c);
    HIPCHK(hipGetLastError());
    return PIC_SUCCESS;
}

int pic_yee_h2(pic_ctx *c) {
    if (!c) return PIC_ERR_INVALID;
    enq_h2(c);
    HIPCHK(hipGetLastError());
    return PIC_SUCCESS;
}

int pic_yee_e(pic_ctx *c) {
    if (!c) return PIC_ERR_INVALID;
    enq_e(c);
    HIPCHK(hipGetLastError());
    return PIC_SUCCESS;
}

int pic_e_periodic(pic_ctx *c) {
    if (!c) return PIC_ERR_INVALID;
    enq_e_periodic(c);
    HIPCHK(hipGetLastError());
    return PIC_SUCCESS;
}

int pic_energy(pic_ctx *c, double *energy) {
    if (!c) return PIC_ERR_INVALID;
    enq_energy(c);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->h_red, c->d_red, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (energy) *energy = c->h_red[0];
    return PIC_SUCCESS;
}

int pic_sync(pic_ctx *c) {
    if (!c) return PIC_ERR_INVALID;
    if (!c->committed) return PIC_SUCCESS;
    return finish_step(c);
}

int pic_step(pic_ctx *c, int nt, double *energy) {
    (void)nt;
    int rc = check_ready(c);
    if (rc) return rc;
    HIPCHK(hipSetDevice(c->dev));
    if (c->arrival_pending && (rc = finish_step(c))) return rc;
    rc = enq_step(c);
    if (rc) return rc;
    rc = finish_step(c);
    accumulate_profile(c);
    if (energy) *energy = c->h_red[0];
    return rc;
}

int pic_run(pic_ctx *c, int nt0, int nsteps, double *energies) {
    for (int s = 0; s < nsteps; s++) {
        int rc = pic_step(c, nt0 + s, energies ? &energies[s] : nullptr);
        if (rc) return rc;
    }
    return PIC_SUCCESS;
}

int pic_get_status(pic_ctx *c, pic_step_status *st) {
    if (!c || !st) return PIC_ERR_INVALID;
    *st = c->last;
    if (c->committed) {
        HIPCHK(hipStreamSynchronize(c->stream));
        std::vector<int> qc(c->t.ntiles);
        HIPCHK(hipMemcpy(qc.data(), qcount_ptr(c), c->t.ntiles * sizeof(int), hipMemcpyDeviceToHost));
        long long s = 0;
        for (int v : qc) s += v;
        st->departures = s;
    }
    return PIC_SUCCESS;
}

int pic_profile(pic_ctx *c, int enable) {
    if (!c) return PIC_ERR_INVALID;
    if (enable && !c->ev[0])
        for (int k = 0; k <= PIC_NSTAGES; k++) HIPCHK(hipEventCreate(&c->ev[k]));
    c->prof = enable != 0;
    for (int k = 0; k < PIC_NSTAGES; k++) {
        c->prof_ms[k] = 0.0;
        c->prof_n[k] = 0;
    }
    return PIC_SUCCESS;
}

int pic_profile_read(pic_ctx *c, double *ms, long *launches, int nstages) {
    if (!c || nstages < 0) return PIC_ERR_INVALID;
    for (int k = 0; k < nstages && k < PIC_NSTAGES; k++) {
        if (ms) ms[k] = c->prof_ms[k];
        if (launches) launches[k] = c->prof_n[k];
    }
    return PIC_SUCCESS;
}

long pic_particle_count(pic_ctx *c) {
    if (!c) return -1;
    if (!c->committed) return (long)c->staged.cell.size();
    std::vector<int> cnt(c->NC);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return -1;
    if (hipMemcpy(cnt.data(), count_ptr(c), c->NC * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
        return -1;
    long n = 0;
    for (int v : cnt) n += v;
    return n;
}

int pic_get_cell_counts(pic_ctx *c, int *counts) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (!counts) return PIC_ERR_INVALID;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(counts, count_ptr(c), c->NC * sizeof(int), hipMemcpyDeviceToHost));
    return PIC_SUCCESS;
}

int pic_get_particles(pic_ctx *c, double *x, double *y, double *z, double *px, double *py,
                      double *pz, double *m, double *q_m, int *sort, int *cell, long cap,
                      long *n_out) {
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = settle(c))) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    const long NC = c->NC;
    std::vector<int> cnt(NC);
    HIPCHK(hipMemcpy(cnt.data(), count_ptr(c), NC * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<long> off(NC + 1, 0);
    for (long k = 0; k < NC; k++) off[k + 1] = off[k] + std::min(cnt[k], bucket_cap(c));
    const long P = off[NC];
    if (n_out) *n_out = P;
    if (P > cap) return PIC_ERR_INVALID;
    size_t Pn = (size_t)std::max(P, 1L);
    long *d_off = nullptr;
    double *d[6] = {nullptr};
    unsigned char *d_sp = nullptr;
    int *d_cell = nullptr;
    HIPCHK(hipMalloc(&d_off, (NC + 1) * sizeof(long)));
    for (int k = 0; k < 6; k++) HIPCHK(hipMalloc(&d[k], Pn * sizeof(double)));
    HIPCHK(hipMalloc(&d_sp, Pn));
    HIPCHK(hipMalloc(&d_cell, Pn * sizeof(int)));
    HIPCHK(hipMemcpy(d_off, off.data(), (NC + 1) * sizeof(long), hipMemcpyHostToDevice));
    if (c->fp32)
        launch_gather_particles32(c->g, NC, d_off, c->p32, d, d_sp, d_cell, nullptr, c->stream);
    else launch_gather_particles(NC, d_off, c->p, d, d_sp, d_cell, nullptr, c->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    double *outs[6] = {x, y, z, px, py, pz};
    for (int k = 0; k < 6; k++)
        if (outs[k] && P) HIPCHK(hipMemcpy(outs[k], d[k], P * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<unsigned char> sp(Pn);
    if (P) HIPCHK(hipMemcpy(sp.data(), d_sp, P, hipMemcpyDeviceToHost));
    if (cell && P) HIPCHK(hipMemcpy(cell, d_cell, P * sizeof(int), hipMemcpyDeviceToHost));
    for (long j = 0; j < P; j++) {
        int s = sp[j];
        if (sort) sort[j] = c->slot_sort[s];
        if (m) m[j] = c->sp.m[s];
        if (q_m) q_m[j] = c->sp.q_m[s];
    }
    hipFree(d_off);
    for (int k = 0; k < 6; k++) hipFree(d[k]);
    hipFree(d_sp);
    hipFree(d_cell);
    return PIC_SUCCESS;
}

int pic_moments(pic_ctx *c, double *out27) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (!out27) return PIC_ERR_INVALID;
    if ((rc = settle(c))) return rc;
    if (c->fp32)
        launch_moments32(c->g, c->NC, c->p32, c->sp, c->d_slot_sort, c->d_part, c->npart,
                         c->d_red, c->stream);
    else
        launch_moments(c->NC, c->p, c->sp, c->d_slot_sort, c->d_part, c->npart, c->d_red,
                       c->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->h_red, c->d_red, 27 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    std::memcpy(out27, c->h_red, 27 * sizeof(double));
    return PIC_SUCCESS;
}

int pic_cell_moments(pic_ctx *c, int sort, int op, double *out) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (!out || sort < 0 || sort > 2 || op < PIC_DENSITY || op > PIC_CURRENT) return PIC_ERR_INVALID;
    if ((rc = settle(c))) return rc;
    const long n = (op >= PIC_VELOCITY ? 3 : 1) * c->NC;
    double *d = nullptr;
    HIPCHK(hipMalloc(&d, n * sizeof(double)));
    if (c->fp32)
        launch_cell_moments32(c->NC, c->p32, c->sp, c->d_slot_sort, sort, op, d, c->stream);
    else
        launch_cell_moments64(c->NC, c->p, c->sp, c->d_slot_sort, sort, op, d, c->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    rc = PIC_SUCCESS;
    if (hipMemcpy(out, d, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) rc = PIC_ERR_HIP;
    (void)hipFree(d);
    return rc;
}

/* ------------------------------------------------------- decision tape */
static int tape_alloc(pic_ctx *c, int nsteps) {
    const size_t words = (size_t)nsteps * (size_t)c->n_ids;
    if (c->d_tape && c->tape_nsteps == nsteps && c->tape_nids == c->n_ids) return PIC_SUCCESS;
    if (c->d_tape) (void)hipFree(c->d_tape);
    c->d_tape = nullptr;
    hipError_t e = hipMalloc(&c->d_tape, std::max(words, (size_t)1) * sizeof(unsigned long long));
    if (e != hipSuccess) return e == hipErrorOutOfMemory ? PIC_ERR_NOMEM : record_hip(e, "tape");
    /* words never written (ids dropped at load) read as "not recorded" */
    HIPCHK(hipMemset(c->d_tape, 0, std::max(words, (size_t)1) * sizeof(unsigned long long)));
    c->tape_nsteps = nsteps;
    c->tape_nids = c->n_ids;
    return PIC_SUCCESS;
}

static int tape_usable(pic_ctx *c, int nsteps) {
    if (!c || nsteps < 1) return PIC_ERR_INVALID;
    if (c->fp32 || !c->diff) {
        snprintf(g_hip_err, sizeof(g_hip_err), "decision tape: FP64 differentiable form only");
        return PIC_ERR_INVALID;
    }
    return PIC_SUCCESS;
}

int pic_tape(pic_ctx *c, int mode, int nsteps) {
    if (mode == PIC_TAPE_OFF) {
        if (!c) return PIC_ERR_INVALID;
        c->tape_mode = PIC_TAPE_OFF;
        return PIC_SUCCESS;
    }
    int rc = tape_usable(c, nsteps);
    if (rc) return rc;
    if (mode != PIC_TAPE_RECORD && mode != PIC_TAPE_REPLAY) return PIC_ERR_INVALID;
    /* the initial cells are taken (record) or imposed (replay) at the first
     * step, when the loaded particles are committed to their buckets */
    if (c->committed || c->n_ids == 0) {
        snprintf(g_hip_err, sizeof(g_hip_err),
                 "decision tape: needs loaded particles that have not stepped yet");
        return PIC_ERR_STATE;
    }
    if (mode == PIC_TAPE_RECORD) {
        if ((rc = tape_alloc(c, nsteps))) return rc;
        c->tape_cells_set = false;
    } else if (!c->tape_cells_set || !c->d_tape || c->tape_nsteps < nsteps ||
               c->tape_nids != c->n_ids) {
        snprintf(g_hip_err, sizeof(g_hip_err), "decision tape: no tape set for these particles");
        return PIC_ERR_STATE;
    }
    c->tape_mode = mode;
    c->tape_step = 0;
    c->tape_nsteps = nsteps;
    return PIC_SUCCESS;
}

int pic_tape_get(pic_ctx *c, unsigned long long *words, int *cells, long nids, int nsteps) {
    int rc = tape_usable(c, nsteps);
    if (rc) return rc;
    if (!c->d_tape || !c->tape_cells_set || nids != c->tape_nids || nsteps > c->tape_nsteps)
        return PIC_ERR_STATE;
    if (c->stream) HIPCHK(hipStreamSynchronize(c->stream));
    if (words)
        HIPCHK(hipMemcpy(words, c->d_tape, (size_t)nsteps * nids * sizeof(unsigned long long),
                         hipMemcpyDeviceToHost));
    if (cells) std::memcpy(cells, c->tape_cells.data(), nids * sizeof(int));
    return PIC_SUCCESS;
}

int pic_tape_set(pic_ctx *c, const unsigned long long *words, const int *cells, long nids,
                 int nsteps) {
    int rc = tape_usable(c, nsteps);
    if (rc) return rc;
    if (!words || !cells || nids != c->n_ids) return PIC_ERR_INVALID;
    if (c->committed) return PIC_ERR_STATE;
    if ((rc = tape_alloc(c, nsteps))) return rc;
    HIPCHK(hipMemcpy(c->d_tape, words, (size_t)nsteps * nids * sizeof(unsigned long long),
                     hipMemcpyHostToDevice));
    c->tape_cells.assign(cells, cells + nids);
    c->tape_cells_set = true;
    return PIC_SUCCESS;
}

int pic_gauss_residual(pic_ctx *c, double *res, double *max_drift, int set_reference) {
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = settle(c))) return rc;
    const Grid &g = c->g;
    if (!c->d_rho) HIPCHK(hipMalloc(&c->d_rho, g.G * sizeof(long long)));
    if (!c->d_gres) HIPCHK(hipMalloc(&c->d_gres, c->NC * sizeof(double)));
    HIPCHK(hipMemsetAsync(c->d_rho, 0, g.G * sizeof(long long), c->stream));
    if (c->fp32) {
        launch_rho32(g, c->NC, c->p32, c->sp, c->d_rho, c->stream);
        launch_gauss32(g, c->f32, c->d_rho, c->d_gres, c->stream);
    } else {
        launch_rho64(g, c->NC, c->p, c->sp, c->d_rho, c->stream);
        launch_gauss64(g, c->f, c->d_rho, c->d_gres, c->stream);
    }
    HIPCHK(hipGetLastError());
    std::vector<double> h(c->NC);
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(h.data(), c->d_gres, c->NC * sizeof(double), hipMemcpyDeviceToHost));
    if (res) std::memcpy(res, h.data(), c->NC * sizeof(double));
    if (set_reference || c->gres_ref.size() != h.size()) c->gres_ref = h;
    double d = 0.0;
    for (long k = 0; k < c->NC; k++) d = std::max(d, std::fabs(h[k] - c->gres_ref[k]));
    if (max_drift) *max_drift = d;
    return PIC_SUCCESS;
}

}  // extern "C"
