/*
 * pic_ctx.h -- the object behind the opaque pic_ctx handle, shared by the
 * host files pic_api.hip (plain and precision forms) and pic_adjoint.hip
 * (differentiable form).  Not part of the ABI.
 */
#ifndef PIC_CTX_H
#define PIC_CTX_H

#include <vector>

#include "pic_internal.h"

#define HIPCHK(x)                                          \
    do {                                                   \
        hipError_t e_ = (x);                               \
        if (e_ != hipSuccess) return picx::record_hip(e_, #x);   \
    } while (0)

struct Staged {
    std::vector<double> v[6];
    std::vector<unsigned char> sp;
    std::vector<int> cell;
    std::vector<int> id;     /* differentiable form: load-order id */
};

struct AdjointStore;   /* pic_adjoint.hip */

struct pic_ctx {
    pic_params prm;
    Grid g;
    TileGeom t;
    long NC = 0;
    int dev = 0;
    hipStream_t stream = nullptr;
    /* a step runs the arrival on `side` while the field stages run on
     * `stream` (the two touch disjoint data); ev_dep orders it after the
     * push, ev_join joins it before the step's end.  With stage timing on,
     * the step stays on `stream` alone. */
    hipStream_t side = nullptr;
    hipEvent_t ev_dep = nullptr, ev_join = nullptr;
    bool fp32 = false;
    FieldPtrs f{};          /* FP64 plain form */
    PartPtrs p{};
    QueuePtrs q{};
    FieldPtrs32 f32{};      /* FP32 precision form */
    /* second H arrays: the step's fused H / E pass writes H^{n+1} there and
     * swaps them with the live ones (f.f[PIC_HX..] or f32.H) */
    double *h_alt[3] = {nullptr, nullptr, nullptr};
    float *h_alt32[3] = {nullptr, nullptr, nullptr};
    PartPtrs32 p32{};       /* the live set (count, bcount, inbox: per cell) */
    PartPtrs32 alt32{};     /* the other set the push writes (p, sp, id) */
    Layout32 lay32{};       /* packed layout state; lay32.packed: p32 holds it */
    QueuePtrs32 q32{};
    long long *d_rho = nullptr;   /* Gauss residual scratch */
    double *d_gres = nullptr;
    std::vector<double> gres_ref;
    /* differentiable form: particle ids, fixed-point current (the FP64 form's
     * int64 J arrays; the FP32 form's J is fixed point already), adjoint state */
    bool diff = false;
    long long *d_jfix[3] = {nullptr, nullptr, nullptr};
    long n_ids = 0;
    AdjointStore *adj = nullptr;
    /* decision tape (FP64 differentiable form): tape_nsteps rows of tape_nids
     * words; the push of step tape_step uses row tape_step while tape_mode is
     * on; tape_cells: the bulk cell of each id at load (-1 dropped) */
    unsigned long long *d_tape = nullptr;
    long tape_nids = 0;
    int tape_nsteps = 0, tape_step = 0, tape_mode = 0;
    bool tape_cells_set = false;
    std::vector<int> tape_cells;
    PartPtrs rec_snap{};      /* record target of the next push (adjoint recompute) */
    PartPtrs32 rec_snap32{};
    const long *rec_off = nullptr;   /* tile offsets of that record */
    long rec_room = 0;               /* entries that record holds */
    SpeciesTable sp{};
    int slot_sort[PIC_MAX_SLOTS] = {0};
    int nslots = 0;
    int *d_slot_sort = nullptr;
    DevStatus *d_st = nullptr;
    DevStatus *h_st = nullptr;
    double *d_part = nullptr;
    int npart = 1024;
    double *d_red = nullptr;
    double *h_red = nullptr;
    Staged staged;
    bool committed = false;
    /* FP32 form: an arrival ran whose status no host call has examined yet
     * (a full spill list or bucket needs the host's repair before the next
     * push or any reader) */
    bool arrival_pending = false;
    long spill_repairs = 0;   /* steps whose spill list ran full (repaired) */
    int regrows = 0;
    pic_step_status last{};
    /* stage timing */
    bool prof = false;
    hipEvent_t ev[PIC_NSTAGES + 1] = {};
    double prof_ms[PIC_NSTAGES] = {0};
    long prof_n[PIC_NSTAGES] = {0};
    unsigned marked = 0;    /* stages marked in the step being timed */
};

namespace picx {
extern thread_local char g_hip_err[256];
int record_hip(hipError_t e, const char *where);
int commit(pic_ctx *c);
int check_ready(pic_ctx *c);
int enq_step(pic_ctx *c);
int finish_step(pic_ctx *c);
/* FP32 form: put every particle into the bucket arrays (unpack the packed
 * set with its pending arrivals) before anything other than the push reads
 * the buckets (no-op otherwise); examines a pending arrival's status first */
int settle(pic_ctx *c);
/* the same without any allocation: PIC_ERR_STATE while an arrival's status
 * is pending (its repair may allocate) */
int settle_noalloc(pic_ctx *c);
/* FP32 form: the arrival's status after a full spill list: every particle
 * into the bucket arrays of the other set, from the packed set, its inbox
 * records and the departure queues (the set becomes the live one) */
int despill(pic_ctx *c);
template <class P>
int alloc_parts(P &p, long ncell, int cap, bool with_count, bool with_id);
template <class P>
void free_parts(P &p);
void free_adjoint(pic_ctx *c);   /* pic_adjoint.hip */
/* pic_noalloc.hip: device address of this device's scratch array, taken once
 * per device with the device current */
int noalloc_init(int dev);
}  // namespace picx

#endif
