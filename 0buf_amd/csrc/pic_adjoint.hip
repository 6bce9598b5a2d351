/*
 * pic_adjoint.hip -- host side of the differentiable form (the reference has
 * no counterpart): forward run with checkpoints, segment recompute with
 * per-step records, and the reverse sweep through the adjoint kernels of
 * k_adjoint.hip.
 *
 * Both precisions: the FP64 plain form's differentiable variant and the
 * FP32 precision form (FP32 forward and particle adjoint, field adjoints
 * accumulated in FP64, k_adjoint32.hip).
 *
 * Memory (one GPU): a checkpoint is the compacted particle state (six FP64
 * arrays, or the FP32 form's offsets and momenta, species and id of the
 * occupied slots, with per-cell offsets and counts) plus E and H; checkpoints are kept every K steps and at the end of
 * the run.  During the reverse sweep one segment of
 * K steps is recomputed from its checkpoint and each step's pre-push
 * particle state (written by the push kernel itself into a compact record,
 * tile by tile in the push's order) and E^n, H^{n+1/2} are recorded; the
 * segment is then reversed step by step.  The J deposit of
 * both forms is order-free (int64 fixed point), so the recompute is bitwise
 * the first forward run.
 */
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <type_traits>
#include <vector>

#include "pic_ctx.h"

namespace picx {
void enq_h_half(pic_ctx *c);
void enq_push(pic_ctx *c);
void enq_arrive(pic_ctx *c);
void enq_fold(pic_ctx *c);
void enq_h2(pic_ctx *c);
void enq_e(pic_ctx *c);
void enq_e_periodic(pic_ctx *c);
void enq_fold_h2_e(pic_ctx *c);
}  // namespace picx

using namespace picx;

/* E, H of one time level, in the form's own precision (8 or 4 bytes) */
struct FieldSet {
    void *f[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
};

/* compacted particle state: cell c's particles at [off[c], off[c+1]);
 * v[] holds FP64 positions/momenta, or the FP32 form's raw offsets/momenta */
struct Compact {
    long P = 0;
    long *off = nullptr;      /* NC + 1 */
    int *count = nullptr;     /* NC */
    void *v[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    unsigned char *sp = nullptr;
    int *id = nullptr;
};

struct AdjointStore {
    int nsteps = 0, K = 0;
    bool recorded = false;
    long long anomalies = 0;          /* dropped + duplicated leavers in the forward run */
    std::vector<Compact> ck;          /* checkpoints at 0, K, 2K, ... and at nsteps */
    std::vector<FieldSet> ckf;
    /* per-step particle records of one segment, compact: tile t's particles
     * in the push's flat order at recoff[r][t] (FP64 or FP32 form) */
    std::vector<PartPtrs> rec;
    std::vector<PartPtrs32> rec32;
    std::vector<long *> recoff;       /* ntiles + 1 offsets per record */
    long room = 0;                    /* particles a record holds */
    long *tile_sum = nullptr;         /* scratch of the offset scan */
    std::vector<FieldSet> recf;       /* E^n, H^{n+1/2} of each recorded step */
    void *Eend[3] = {nullptr, nullptr, nullptr};   /* E at the end of the segment */
    AdjLam lam{};
    LamParts lp{};
    double *lp_flat = nullptr;        /* one component of every id (host transfers) */
    long lam_ni = 0;                  /* particle adjoint entries allocated */
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    double ms[3] = {0, 0, 0};
};

namespace {

size_t elem(const pic_ctx *c) { return c->fp32 ? sizeof(float) : sizeof(double); }

/* live E (k < 3) and H (k >= 3) arrays of the context */
void *live_field(pic_ctx *c, int k) {
    if (c->fp32) return k < 3 ? (void *)c->f32.E[k] : (void *)c->f32.H[k - 3];
    return c->f.f[PIC_EX + k];
}

int *live_count(pic_ctx *c) { return c->fp32 ? c->p32.count : c->p.count; }
int live_cap(const pic_ctx *c) { return c->fp32 ? c->p32.cap : c->p.cap; }

void free_fields(FieldSet &f) {
    for (int k = 0; k < 6; k++)
        if (f.f[k]) (void)hipFree(f.f[k]);
    f = FieldSet{};
}

int alloc_fields(FieldSet &f, long G, size_t es) {
    for (int k = 0; k < 6; k++) HIPCHK(hipMalloc(&f.f[k], G * es));
    return PIC_SUCCESS;
}

void free_lam(AdjointStore *s) {
    for (int k = 0; k < 3; k++) {
        if (s->lam.E[k]) (void)hipFree(s->lam.E[k]);
        if (s->lam.H[k]) (void)hipFree(s->lam.H[k]);
        if (s->lam.J[k]) (void)hipFree(s->lam.J[k]);
        if (s->lam.Q[k]) (void)hipFree(s->lam.Q[k]);
        if (s->Eend[k]) (void)hipFree(s->Eend[k]);
        s->Eend[k] = nullptr;
    }
    if (s->lp.v) (void)hipFree(s->lp.v);
    if (s->lp_flat) (void)hipFree(s->lp_flat);
    s->lp_flat = nullptr;
    s->lam = AdjLam{};
    s->lp = LamParts{};
}

void free_compact(Compact &k) {
    if (k.off) (void)hipFree(k.off);
    if (k.count) (void)hipFree(k.count);
    for (int j = 0; j < 6; j++)
        if (k.v[j]) (void)hipFree(k.v[j]);
    if (k.sp) (void)hipFree(k.sp);
    if (k.id) (void)hipFree(k.id);
    k = Compact{};
}

/* device memory of a checkpoint of P particles */
int reserve_state(pic_ctx *c, Compact &dst, FieldSet &f, long P) {
    const size_t es = elem(c);
    if (dst.off && dst.P < P) free_compact(dst);
    if (!dst.off) {
        const size_t Pn = (size_t)std::max(P, 1L);
        HIPCHK(hipMalloc(&dst.off, (c->NC + 1) * sizeof(long)));
        HIPCHK(hipMalloc(&dst.count, c->NC * sizeof(int)));
        for (int j = 0; j < 6; j++) HIPCHK(hipMalloc(&dst.v[j], Pn * es));
        HIPCHK(hipMalloc(&dst.sp, Pn));
        HIPCHK(hipMalloc(&dst.id, Pn * sizeof(int)));
        dst.P = P;
    }
    if (!f.f[0]) return alloc_fields(f, c->g.G, es);
    return PIC_SUCCESS;
}

int save_state(pic_ctx *c, Compact &dst, FieldSet &f) {
    const long NC = c->NC;
    const size_t es = elem(c);
    std::vector<int> cnt(NC);
    /* FP32 form: the packed set is read where it lies (no unpack), once the
     * last arrival's status is examined */
    int rc0 = c->fp32 ? (c->arrival_pending ? finish_step(c) : PIC_SUCCESS) : settle(c);
    if (rc0) return rc0;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(cnt.data(), live_count(c), NC * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<long> off(NC + 1, 0);
    for (long k = 0; k < NC; k++) off[k + 1] = off[k] + std::min(cnt[k], live_cap(c));
    const long P = off[NC];
    int rc = reserve_state(c, dst, f, P);
    if (rc) return rc;
    HIPCHK(hipMemcpy(dst.off, off.data(), (NC + 1) * sizeof(long), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpyAsync(dst.count, live_count(c), NC * sizeof(int), hipMemcpyDeviceToDevice,
                          c->stream));
    if (c->fp32) {
        float *v[6];
        for (int j = 0; j < 6; j++) v[j] = (float *)dst.v[j];
        if (c->lay32.packed)
            launch_gather_packed32(c->g, c->t, c->p32, c->lay32, dst.off, v, dst.sp, dst.id,
                                   c->stream);
        else
            launch_gather_raw32(NC, dst.off, c->p32, v, dst.sp, dst.id, c->stream);
    } else {
        double *v[6];
        for (int j = 0; j < 6; j++) v[j] = (double *)dst.v[j];
        launch_gather_particles(NC, dst.off, c->p, v, dst.sp, nullptr, dst.id, c->stream);
    }
    const size_t bytes = c->g.G * es;
    for (int k = 0; k < 6; k++)
        HIPCHK(hipMemcpyAsync(f.f[k], live_field(c, k), bytes, hipMemcpyDeviceToDevice,
                              c->stream));
    HIPCHK(hipGetLastError());
    return PIC_SUCCESS;
}

int restore_state(pic_ctx *c, const Compact &src, const FieldSet &f) {
    if (c->fp32) {
        const float *v[6];
        for (int j = 0; j < 6; j++) v[j] = (const float *)src.v[j];
        launch_scatter_load32(src.off, c->NC, v, src.sp, c->p32, c->stream);   /* bcount = count */
        launch_scatter_ids32(src.off, c->NC, src.id, c->p32, c->stream);
        launch_clear_arrivals32(c->g, c->t, c->p32, c->lay32, c->stream);
        c->lay32.packed = 0;
        c->arrival_pending = false;
    } else {
        const double *v[6];
        for (int j = 0; j < 6; j++) v[j] = (const double *)src.v[j];
        launch_scatter_load(src.off, c->NC, v, src.sp, c->p, c->stream);
        launch_scatter_ids(src.off, c->NC, src.id, c->p, c->stream);
    }
    const size_t bytes = c->g.G * elem(c);
    for (int k = 0; k < 6; k++)
        HIPCHK(hipMemcpyAsync(live_field(c, k), f.f[k], bytes, hipMemcpyDeviceToDevice,
                              c->stream));
    HIPCHK(hipGetLastError());
    return PIC_SUCCESS;
}

/* one forward step that records the pre-push particle state into record r
 * of the segment (through the push kernel) and E^n, H^{n+1/2} into `f`.
 * The recompute repeats steps the first forward has already run with the
 * same counts and capacity, so it enqueues its steps without waiting: the
 * status accumulates over the segment and is checked once at its end
 * (check_segment). */
int recorded_step(pic_ctx *c, AdjointStore *s, int r) {
    FieldSet &f = s->recf[r];
    const size_t bytes = c->g.G * elem(c);
    enq_h_half(c);
    const void *src[6];
    for (int k = 0; k < 6; k++) src[k] = live_field(c, k);
    launch_copy6(f.f, src, bytes, c->stream);
    launch_record_offsets(c->g, c->t, live_count(c), live_cap(c), s->tile_sum, s->recoff[r],
                          s->room, c->d_st, c->stream);
    if (c->fp32) c->rec_snap32 = s->rec32[r];
    else c->rec_snap = s->rec[r];
    c->rec_off = s->recoff[r];
    c->rec_room = s->room;
    enq_push(c);
    c->rec_snap = PartPtrs{};
    c->rec_snap32 = PartPtrs32{};
    c->rec_off = nullptr;
    c->rec_room = 0;
    enq_arrive(c);
    if (c->fp32) {
        /* the status is checked at the segment's end; only where the first
         * forward needed the repair of a full spill list is each step's
         * arrival examined (and repaired) here */
        c->arrival_pending = false;
        if (c->spill_repairs > 0) {
            HIPCHK(hipMemcpyAsync(c->h_st, c->d_st, sizeof(DevStatus), hipMemcpyDeviceToHost,
                                  c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            if (c->lay32.packed && (c->h_st->flags & PIC_F_SPILL_OVER)) {
                int rc = despill(c);
                if (rc) return rc;
                launch_clear_flags32(c->d_st, PIC_F_SPILL_OVER, c->stream);
            }
        }
    }
    enq_fold_h2_e(c);
    HIPCHK(hipGetLastError());
    return PIC_SUCCESS;
}

/* end of a recomputed segment: any flag of its steps fails the sweep (the
 * first forward ran the same steps without one; the recompute sees the
 * forward's own counts, which the live capacity, the largest of the run,
 * already holds) */
int check_segment(pic_ctx *c) {
    HIPCHK(hipMemcpyAsync(c->h_st, c->d_st, sizeof(DevStatus), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->h_st->flags) {
        snprintf(g_hip_err, sizeof(g_hip_err), "recomputed segment raised status flags 0x%x",
                 c->h_st->flags);
        return PIC_ERR_STATE;
    }
    return PIC_SUCCESS;
}

/* a compact record of `room` particles with per-cell counts */
template <class P>
int alloc_record(P &r, long ncell, long room, int cap) {
    using T = std::remove_pointer_t<std::remove_reference_t<decltype(r.p[0])>>;
    r = P{};
    r.cap = cap;
    const size_t n = (size_t)std::max(room, 1L);
    for (int k = 0; k < 6; k++) HIPCHK(hipMalloc(&r.p[k], n * sizeof(T)));
    HIPCHK(hipMalloc(&r.sp, n));
    HIPCHK(hipMalloc(&r.id, n * sizeof(int)));
    HIPCHK(hipMalloc(&r.count, ncell * sizeof(int)));
    return PIC_SUCCESS;
}

int zero_async(void *p, size_t bytes, hipStream_t s) {
    HIPCHK(hipMemsetAsync(p, 0, bytes, s));
    return PIC_SUCCESS;
}

}  // namespace

namespace picx {
void free_adjoint(pic_ctx *c) {
    AdjointStore *s = c->adj;
    if (!s) return;
    for (auto &p : s->ck) free_compact(p);
    for (auto &f : s->ckf) free_fields(f);
    for (auto &p : s->rec) free_parts(p);
    for (auto &p : s->rec32) free_parts(p);
    for (auto *o : s->recoff)
        if (o) (void)hipFree(o);
    if (s->tile_sum) (void)hipFree(s->tile_sum);
    for (auto &f : s->recf) free_fields(f);
    free_lam(s);
    for (int k = 0; k < 3; k++)
        if (s->ev[k]) (void)hipEventDestroy(s->ev[k]);
    delete s;
    c->adj = nullptr;
}
}  // namespace picx

extern "C" {

long pic_particle_id_count(pic_ctx *c) { return c ? c->n_ids : -1; }

int pic_get_particle_ids(pic_ctx *c, int *id, long cap, long *n_out) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (!id) return PIC_ERR_INVALID;
    if (c->fp32 ? !c->p32.id : !c->p.id) {
        const long n = pic_particle_count(c);
        if (n_out) *n_out = n;
        if (n > cap) return PIC_ERR_INVALID;
        for (long j = 0; j < n; j++) id[j] = -1;
        return PIC_SUCCESS;
    }
    HIPCHK(hipSetDevice(c->dev));
    if ((rc = settle(c))) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    const long NC = c->NC;
    std::vector<int> cnt(NC);
    const int bcap = c->fp32 ? c->p32.cap : c->p.cap;
    HIPCHK(hipMemcpy(cnt.data(), c->fp32 ? c->p32.count : c->p.count, NC * sizeof(int),
                     hipMemcpyDeviceToHost));
    std::vector<long> off(NC + 1, 0);
    for (long k = 0; k < NC; k++) off[k + 1] = off[k] + std::min(cnt[k], bcap);
    const long P = off[NC];
    if (n_out) *n_out = P;
    if (P > cap) return PIC_ERR_INVALID;
    const size_t Pn = (size_t)std::max(P, 1L);
    long *d_off = nullptr;
    double *d[6] = {nullptr};
    unsigned char *d_sp = nullptr;
    int *d_cell = nullptr, *d_id = nullptr;
    HIPCHK(hipMalloc(&d_off, (NC + 1) * sizeof(long)));
    for (int k = 0; k < 6; k++) HIPCHK(hipMalloc(&d[k], Pn * sizeof(double)));
    HIPCHK(hipMalloc(&d_sp, Pn));
    HIPCHK(hipMalloc(&d_cell, Pn * sizeof(int)));
    HIPCHK(hipMalloc(&d_id, Pn * sizeof(int)));
    HIPCHK(hipMemcpy(d_off, off.data(), (NC + 1) * sizeof(long), hipMemcpyHostToDevice));
    if (c->fp32)
        launch_gather_particles32(c->g, NC, d_off, c->p32, d, d_sp, d_cell, d_id, c->stream);
    else
        launch_gather_particles(NC, d_off, c->p, d, d_sp, d_cell, d_id, c->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    if (P) HIPCHK(hipMemcpy(id, d_id, P * sizeof(int), hipMemcpyDeviceToHost));
    (void)hipFree(d_off);
    for (int k = 0; k < 6; k++) (void)hipFree(d[k]);
    (void)hipFree(d_sp);
    (void)hipFree(d_cell);
    (void)hipFree(d_id);
    return PIC_SUCCESS;
}

int pic_forward_record(pic_ctx *c, int nsteps, int ckpt_every, double *energy) {
    if (!c || nsteps < 1 || ckpt_every < 1) return PIC_ERR_INVALID;
    if (!c->diff) return PIC_ERR_STATE;
    int rc = check_ready(c);
    if (rc) return rc;
    HIPCHK(hipSetDevice(c->dev));
    /* a store of the same shape is reused (its checkpoints and records keep
     * their device memory); any other is replaced */
    AdjointStore *s = c->adj;
    if (!s || s->nsteps != nsteps || s->K != ckpt_every) {
        free_adjoint(c);
        s = new AdjointStore();
        c->adj = s;
        s->nsteps = nsteps;
        s->K = ckpt_every;
        for (int k = 0; k < 3; k++) HIPCHK(hipEventCreate(&s->ev[k]));
    }
    s->recorded = false;
    s->anomalies = 0;
    const int nck = (nsteps - 1) / ckpt_every + 1;
    s->ck.resize(nck + 1);
    s->ckf.resize(nck + 1);
    /* checkpoint memory up front, outside the timed run */
    const long P = pic_particle_count(c);
    for (int k = 0; k <= nck; k++)
        if ((rc = reserve_state(c, s->ck[k], s->ckf[k], P))) return rc;
    HIPCHK(hipEventRecord(s->ev[0], c->stream));
    for (int t = 0; t < nsteps; t++) {
        if (t % ckpt_every == 0) {
            rc = save_state(c, s->ck[t / ckpt_every], s->ckf[t / ckpt_every]);
            if (rc) return rc;
        }
        rc = enq_step(c);
        if (rc) return rc;
        rc = finish_step(c);
        if (rc) return rc;
        s->anomalies += c->last.dropped + c->last.duplicated;
        if (energy) energy[t] = c->h_red[0];
    }
    rc = save_state(c, s->ck[nck], s->ckf[nck]);
    if (rc) return rc;
    HIPCHK(hipEventRecord(s->ev[1], c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, s->ev[0], s->ev[1]));
    s->ms[0] = ms;
    s->ms[1] = s->ms[2] = 0.0;
    s->recorded = true;
    return PIC_SUCCESS;
}

int pic_adjoint(pic_ctx *c, const double *weights, const pic_adjoint_io *seed,
                pic_adjoint_io *grad) {
    if (!c) return PIC_ERR_INVALID;
    AdjointStore *s = c->adj;
    if (!c->diff || !s || !s->recorded) return PIC_ERR_STATE;
    /* a leaver taken by two cells (or by none) has no single adjoint path */
    if (s->anomalies) return PIC_ERR_STATE;
    HIPCHK(hipSetDevice(c->dev));
    const Grid &g = c->g;
    const long G = g.G;
    const long NI = std::max(c->n_ids, 1L);
    hipStream_t st = c->stream;
    int rc;

    /* adjoint arrays: kept between calls of the same particle count */
    if (!s->lam.E[0] || s->lam_ni != NI) {
        free_lam(s);
        for (int k = 0; k < 3; k++) {
            HIPCHK(hipMalloc(&s->lam.E[k], G * sizeof(double)));
            HIPCHK(hipMalloc(&s->lam.H[k], G * sizeof(double)));
            HIPCHK(hipMalloc(&s->lam.J[k], G * sizeof(double)));
            HIPCHK(hipMalloc(&s->lam.Q[k], G * sizeof(double)));
            HIPCHK(hipMalloc(&s->Eend[k], G * elem(c)));
        }
        HIPCHK(hipMalloc(&s->lp.v, NI * PIC_LSTRIDE * sizeof(double)));
        HIPCHK(hipMalloc(&s->lp_flat, NI * sizeof(double)));
        s->lam_ni = NI;
    }
    for (int k = 0; k < 3; k++) {
        if ((rc = zero_async(s->lam.E[k], G * sizeof(double), st))) return rc;
        if ((rc = zero_async(s->lam.H[k], G * sizeof(double), st))) return rc;
    }
    if ((rc = zero_async(s->lp.v, NI * PIC_LSTRIDE * sizeof(double), st))) return rc;
    s->lp.n = c->n_ids;
    if (seed) {
        const double *ps[6] = {seed->x, seed->y, seed->z, seed->px, seed->py, seed->pz};
        for (int k = 0; k < 6; k++)
            if (ps[k] && c->n_ids) {
                HIPCHK(hipMemcpyAsync(s->lp_flat, ps[k], c->n_ids * sizeof(double),
                                      hipMemcpyHostToDevice, st));
                launch_lam_copy(s->lp, s->lp_flat, k, true, st);
            }
        for (int k = 0; k < 6; k++)
            if (seed->field[k])
                HIPCHK(hipMemcpyAsync(k < 3 ? s->lam.E[k] : s->lam.H[k - 3], seed->field[k],
                                      G * sizeof(double), hipMemcpyHostToDevice, st));
    }
    /* the FP32 sweep keeps the position adjoints in offset units: d/do = h d/dx */
    if (c->fp32) launch_scale_lam(s->lp, g.hx, g.hy, g.hz, st);
    const int K = s->K;
    const int cap = live_cap(c);
    /* records hold the largest particle count of the run (constant without
     * dropped or duplicated leavers) */
    long room = 0;
    for (const Compact &k : s->ck) room = std::max(room, k.P);
    if (room != s->room) {
        for (auto &p : s->rec) free_parts(p);
        for (auto &p : s->rec32) free_parts(p);
        s->room = room;
    }
    s->recf.resize(K);
    s->recoff.resize(K, nullptr);
    if (c->fp32) s->rec32.resize(K);
    else s->rec.resize(K);
    if (!s->tile_sum) HIPCHK(hipMalloc(&s->tile_sum, c->t.ntiles * sizeof(long)));
    for (int r = 0; r < K; r++) {
        if (c->fp32) {
            PartPtrs32 &q = s->rec32[r];
            if (q.p[0] && q.cap != cap) free_parts(q);
            if (!q.p[0] && (rc = alloc_record(q, c->NC, room, cap))) return rc;
        } else {
            PartPtrs &q = s->rec[r];
            if (q.p[0] && q.cap != cap) free_parts(q);
            if (!q.p[0] && (rc = alloc_record(q, c->NC, room, cap))) return rc;
        }
        if (!s->recoff[r]) HIPCHK(hipMalloc(&s->recoff[r], (c->t.ntiles + 1) * sizeof(long)));
        if (!s->recf[r].f[0] && (rc = alloc_fields(s->recf[r], G, elem(c)))) return rc;
    }

    s->ms[1] = s->ms[2] = 0.0;
    const int nseg = (s->nsteps - 1) / K + 1;
    /* with no seed on s_T the adjoint state stays exactly zero back to the
     * last step whose energy weight is nonzero: segments after it are not
     * recomputed */
    int t_last = s->nsteps;
    if (!seed) {
        t_last = 0;
        for (int t = s->nsteps; t >= 1 && weights; t--)
            if (weights[t - 1] != 0.0) {
                t_last = t;
                break;
            }
    }
    for (int seg = nseg - 1; seg >= 0; seg--) {
        if (seg * K >= t_last) continue;
        const int t0 = seg * K, t1 = std::min(t0 + K, s->nsteps);
        HIPCHK(hipEventRecord(s->ev[0], st));
        if ((rc = restore_state(c, s->ck[seg], s->ckf[seg]))) return rc;
        launch_reset_status(c->d_st, st);
        for (int t = t0 + 1; t <= t1; t++)
            if ((rc = recorded_step(c, s, t - t0 - 1))) return rc;
        if ((rc = check_segment(c))) return rc;
        for (int k = 0; k < 3; k++)
            HIPCHK(hipMemcpyAsync(s->Eend[k], live_field(c, k), G * elem(c),
                                  hipMemcpyDeviceToDevice, st));
        HIPCHK(hipEventRecord(s->ev[1], st));
        launch_reset_status(c->d_st, st);
        /* the recomputed segment must end on the stored state, bit for bit:
         * fields and per-cell counts */
        for (int k = 0; k < 6; k++)
            launch_mismatch(live_field(c, k), s->ckf[seg + 1].f[k], G, (int)elem(c), c->d_st,
                            st);
        launch_mismatch(live_count(c), s->ck[seg + 1].count, c->NC, 4, c->d_st, st);
        for (int t = t1; t > t0; t--) {
            const int r = t - t0 - 1;
            const void *Et[3];
            for (int k = 0; k < 3; k++) Et[k] = (t == t1) ? s->Eend[k] : s->recf[r + 1].f[k];
            if (weights && weights[t - 1] != 0.0) {
                if (c->fp32) {
                    const float *E32[3] = {(const float *)Et[0], (const float *)Et[1],
                                           (const float *)Et[2]};
                    launch_adj_seed_energy32(g, E32, weights[t - 1], s->lam, st);
                } else {
                    const double *E64[3] = {(const double *)Et[0], (const double *)Et[1],
                                            (const double *)Et[2]};
                    launch_adj_seed_energy(g, E64, weights[t - 1], s->lam, st);
                }
            }
            launch_adj_e_periodic(g, s->lam, st);
            launch_adj_curl_e(g, s->lam, st);
            launch_adj_fold_j(g, s->lam, st);
            if (c->fp32) {
                AdjFwd32 fwd;
                for (int k = 0; k < 3; k++) {
                    fwd.E[k] = (const float *)s->recf[r].f[k];
                    fwd.H[k] = (const float *)s->recf[r].f[3 + k];
                }
                launch_adj_particles_f32(g, c->t, fwd, s->rec32[r], s->recoff[r], c->sp, s->lam,
                                         s->lp, c->d_st, st);
            } else {
                AdjFwd fwd;
                for (int k = 0; k < 3; k++) {
                    fwd.E[k] = (const double *)s->recf[r].f[k];
                    fwd.H[k] = (const double *)s->recf[r].f[3 + k];
                }
                launch_adj_particles_f64(g, c->t, fwd, s->rec[r], s->recoff[r], c->sp, s->lam,
                                         s->lp, c->d_st, st);
            }
            launch_adj_curl_h(g, s->lam, st);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(s->ev[2], st));
        HIPCHK(hipMemcpyAsync(c->h_st, c->d_st, sizeof(DevStatus), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (c->h_st->flags & PIC_F_RANGE) return PIC_ERR_CELL_RANGE;
        if (c->h_st->flags & PIC_F_RECOMPUTE) {
            snprintf(g_hip_err, sizeof(g_hip_err),
                     "recomputed segment %d differs from its checkpoint", seg);
            return PIC_ERR_STATE;
        }
        float a = 0.f, b = 0.f;
        HIPCHK(hipEventElapsedTime(&a, s->ev[0], s->ev[1]));
        HIPCHK(hipEventElapsedTime(&b, s->ev[1], s->ev[2]));
        s->ms[1] += a;
        s->ms[2] += b;
    }
    if (c->fp32) launch_scale_lam(s->lp, 1.0 / g.hx, 1.0 / g.hy, 1.0 / g.hz, st);
    if (grad) {
        double *pg[6] = {grad->x, grad->y, grad->z, grad->px, grad->py, grad->pz};
        for (int k = 0; k < 6; k++)
            if (pg[k] && c->n_ids) {
                launch_lam_copy(s->lp, s->lp_flat, k, false, st);
                HIPCHK(hipMemcpyAsync(pg[k], s->lp_flat, c->n_ids * sizeof(double),
                                      hipMemcpyDeviceToHost, st));
            }
        for (int k = 0; k < 6; k++)
            if (grad->field[k])
                HIPCHK(hipMemcpyAsync(grad->field[k], k < 3 ? s->lam.E[k] : s->lam.H[k - 3],
                                      G * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    /* leave the context at s_nsteps */
    if ((rc = restore_state(c, s->ck.back(), s->ckf.back()))) return rc;
    HIPCHK(hipStreamSynchronize(st));
    return PIC_SUCCESS;
}

int pic_rewind(pic_ctx *c) {
    if (!c) return PIC_ERR_INVALID;
    AdjointStore *s = c->adj;
    if (!c->diff || !s || !s->recorded) return PIC_ERR_STATE;
    HIPCHK(hipSetDevice(c->dev));
    int rc = restore_state(c, s->ck[0], s->ckf[0]);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return PIC_SUCCESS;
}

int pic_adjoint_timing(pic_ctx *c, double *out3) {
    if (!c || !out3) return PIC_ERR_INVALID;
    if (!c->adj) return PIC_ERR_STATE;
    for (int k = 0; k < 3; k++) out3[k] = c->adj->ms[k];
    return PIC_SUCCESS;
}

}  // extern "C"
