/*
 * pic_amd.h -- C ABI of the MI355X-native EM-PIC time step (libpic_amd.so).
 *
 * Drop-in boundary for the reference's Plasma::Step (gpu_plasma.h:605-620).
 * Plain C types only: pointers, sizes, ints and doubles.  Every entry point
 * returns an int status (0 = success) and never exits the process.
 *
 * Ownership: a pic_ctx owns all device memory it allocates and frees it in
 * pic_destroy.  Host buffers passed in or out belong to the caller; copies
 * are synchronous.  One host thread per context; a context is not
 * re-entrant.  All device work of a context runs on its own HIP stream.
 */
#ifndef PIC_AMD_H
#define PIC_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PIC_API_VERSION 1

/* ------------------------------------------------------------------------
 * archAPI: the reference's backend seam, same names, constants and int codes
 * (archAPI.h:8-11, 92-125), implemented directly on the HIP runtime.
 * ---------------------------------------------------------------------- */
#define HOST_TO_DEVICE   -131313   /* archAPI.h:8  */
#define HOST_TO_HOST     -131314   /* archAPI.h:9  */
#define DEVICE_TO_HOST   -131315   /* archAPI.h:10 */
#define DEVICE_TO_DEVICE -131316   /* archAPI.h:11 */

/* archAPI.h:103 / archAPI.cu:37-55 -- returns the HIP error code (0 ok) */
int MemoryCopy(void *dst, void *src, size_t size, int dir);
/* archAPI.h:107 / archAPI.cu:58-63 -- device allocation */
int MemoryAllocate(void **dst, size_t size);
/* archAPI.h:109 -- free and total device memory in bytes */
int GetDeviceMemory(size_t *m_free, size_t *m_total);
/* archAPI.h:111 -- device memset */
int MemorySet(void *s, int c, size_t n);
/* archAPI.h:114 / archAPI.cu:96-100 */
int ThreadSynchronize(void);
/* archAPI.h:117 -- host-side device synchronize */
int DeviceSynchronize(void);
/* archAPI.h:118 / archAPI.cu:111-120 */
int getLastError(void);
/* archAPI.h:92 / archAPI.cu:142-146 */
const char *getErrorString(int err);
/* archAPI.h:94 / archAPI.cu:15-19 */
int SetDevice(int n);
/* frees memory from MemoryAllocate (the reference never frees: gpu_plasma.h:1629) */
int MemoryFree(void *p);

/* ------------------------------------------------------------------------
 * Status codes.  Positive codes are this library's; a HIP runtime error is
 * reported as PIC_ERR_HIP and its text by pic_last_hip_error().
 * ---------------------------------------------------------------------- */
enum pic_status {
    PIC_SUCCESS = 0,
    PIC_ERR_INVALID = 1,            /* bad argument or unsupported shape */
    PIC_ERR_HIP = 2,                /* HIP runtime call failed */
    PIC_ERR_NO_DEVICE = 3,          /* no GPU visible */
    PIC_ERR_TOO_MANY_PARTICLES = 4, /* > departure_cap leavers in one cell: the
                                       reference's TOO_MANY_PARTICLES exit
                                       (cell.h:27, kernels.cu:193-203,
                                       gpu_plasma.h:1953-1957) */
    PIC_ERR_QUEUE_OVERFLOW = 5,     /* departure queue of a tile overflowed */
    PIC_ERR_NOMEM = 6,              /* device allocation failed */
    PIC_ERR_STATE = 7,              /* call out of order (e.g. no particles) */
    PIC_ERR_CELL_RANGE = 8,         /* a particle's stencil left its tile box */
    PIC_ERR_QUEUE_CORRUPT = 9       /* a departure record named a destination
                                       outside its tile's halo box, or a slot
                                       no claim covers; it was not placed */
};

/* field ids: the 12 ghosted arrays of init.cu:310-331 */
enum pic_field {
    PIC_EX = 0, PIC_EY, PIC_EZ, PIC_HX, PIC_HY, PIC_HZ,
    PIC_JX, PIC_JY, PIC_JZ, PIC_QX, PIC_QY, PIC_QZ, PIC_NFIELDS
};

/* particle sorts, particle_types.h:11 */
enum pic_sort { PIC_ION = 0, PIC_PLASMA_ELECTRON = 1, PIC_BEAM_ELECTRON = 2 };

enum pic_precision {
    PIC_FP64 = 0,   /* plain form: reproduces the reference's FP64 step */
    PIC_FP32 = 1    /* precision form: FP32 push/gather/curl, int64 J */
};

/* Constructor arguments of Plasma (gpu_plasma.h:1574-1589) plus the build's
 * own layout and precision controls. */
typedef struct pic_params {
    int nx, ny, nz;          /* bulk cells per axis (Plasma ctor nx,ny,nz) */
    double lx, ly, lz;       /* domain size (Plasma ctor lx,ly,lz) */
    double ni;               /* plasma density (Plasma ctor ni) */
    int n_per_cell;          /* particles per cell (Plasma ctor n_per_cell) */
    double ion_q_m;          /* ion charge/mass (Plasma ctor q_m) */
    double tau;              /* time step (Plasma ctor TAU) */
    int precision;           /* enum pic_precision */
    int departure_cap;       /* PARTICLES_FLYING_ONE_DIRECTION (cell.h:25) = 50;
                                <= 0 disables the check */
    int bucket_capacity;     /* particle slots per cell; 0 = automatic.  Loading
                                or a regrow fails with PIC_ERR_INVALID when a
                                capacity leaves the push kernels' addressing:
                                a tile's buckets (TX-1 planes + TY-1 rows +
                                TZ cells) over 4 GiB, a bucket over 256 lines
                                of 128 bytes (FP64: 4096 slots), or in the
                                FP32 form 2048 slots or more (its departure
                                records rank leavers in 16 bits) */
    int device;              /* HIP device ordinal */
    int differentiable;      /* 1: differentiable form: particles carry their
                                load-order id, J is accumulated in int64 fixed
                                point (FP64 form 2^-48, FP32 form its usual
                                2^-40; order-free, so a recomputed step is
                                bitwise the first run) and pic_forward_record /
                                pic_adjoint are enabled.  With PIC_FP32 the
                                forward and the particle adjoint run in FP32,
                                field adjoints and gradients accumulate in FP64 */
    int ghost_mode;          /* 0: the reference's periodic E ghosts (only Ex along z,
                                Ey along x and z, Ez along x: gpu_plasma.h:486-505);
                                1: additionally Ex and Ez along y, so every ghost
                                node a particle interpolates from is a periodic
                                image and the field seen across the y faces is
                                continuous */
} pic_params;

typedef struct pic_ctx pic_ctx;

/* Step diagnostics, filled by pic_get_status after each step. */
typedef struct pic_step_status {
    long long departures;     /* leavers queued this step (copies included) */
    long long dropped;        /* leavers accepted by no cell (reference drops
                                 them: kernels.cu:316-370 + Cell::Insert) */
    long long duplicated;     /* extra copies of leavers accepted by 2+ cells */
    int max_cell_departures;  /* max leavers of one cell this step */
    int bucket_regrows;       /* bucket capacity increases so far */
    int bucket_capacity;      /* current particle slots per cell */
    int flags;                /* bit0 too many, bit1 queue overflow,
                                 bit2 bucket overflow (repaired), bit3 range,
                                 bit6 departure record refused */
    int spill_repairs;        /* FP32 form: steps so far whose arrivals overflowed a
                                 tile's spill list (repaired by the host) */
} pic_step_status;

void pic_params_default(pic_params *p);
int pic_create(const pic_params *p, pic_ctx **out);
int pic_destroy(pic_ctx *ctx);
const char *pic_error_string(int status);
const char *pic_last_hip_error(void);

/* Plasma::addParticleListToCells (init.cu:602-630): each particle goes to
 * the cell getPointCell gives (cell.h:593-598) if Cell::Insert accepts it
 * (isPointInCell, cell.h:717); otherwise it is dropped and counted.
 * m is the charge weight (Particle::m), q_m the charge/mass ratio. */
int pic_load_particles(pic_ctx *ctx, int sort, const double *x, const double *y,
                       const double *z, const double *px, const double *py,
                       const double *pz, long n, double m, double q_m,
                       long *n_dropped);
/* Load one ghosted field array of (nx+2)(ny+2)(nz+2) doubles, k fastest
 * (cell.h:796-801).  Replaces InitFields / copyFieldsToGPU
 * (init.cu:355-372, service_functions.cu:221-337). */
int pic_load_field(pic_ctx *ctx, int which, const double *host);
int pic_get_field(pic_ctx *ctx, int which, double *host);

/* One full time step: Plasma::Step(nt) (gpu_plasma.h:605-620).  Returns
 * after the stream completes; *energy (may be NULL) receives the
 * getElectricEnergy() value that Diagnose(nt) writes (gpu_plasma.h:621-686). */
int pic_step(pic_ctx *ctx, int nt, double *energy);
/* nsteps consecutive pic_step calls from nt0; energies[i] (may be NULL)
 * receives the energy after step nt0+i.  Stops at the first error. */
int pic_run(pic_ctx *ctx, int nt0, int nsteps, double *energies);

/* Stage entry points: the launch seam of Plasma::Step, in this order.
 * Each enqueues on the context stream and returns without synchronising. */
int pic_yee_h_half(pic_ctx *ctx);   /* MagneticStageOne + zero J
                                       (gpu_plasma.h:524-538, 2058-2069) */
int pic_push_deposit(pic_ctx *ctx); /* tile gather + Boris push + current
                                       deposit + tile->grid J + departures
                                       (gpu_plasma.h:1006-1036, 1757-1855,
                                       1857-1967) */
int pic_migrate(pic_ctx *ctx);      /* arrivals (gpu_plasma.h:1969-2033) */
int pic_fold_j(pic_ctx *ctx);       /* SetPeriodicCurrents (gpu_plasma.h:980-993) */
int pic_yee_h2(pic_ctx *ctx);       /* MagneticFieldStageTwo (gpu_plasma.h:540-551) */
int pic_yee_e(pic_ctx *ctx);        /* ElectricFieldTrace x3 (gpu_plasma.h:455-481) */
int pic_e_periodic(pic_ctx *ctx);   /* periodic E ghosts (gpu_plasma.h:432-452, 486-505) */
int pic_energy(pic_ctx *ctx, double *energy); /* getElectricEnergy, synchronises */
int pic_sync(pic_ctx *ctx);         /* wait for the stream; check step status */

int pic_get_status(pic_ctx *ctx, pic_step_status *st);

/* Stage timing with HIP events recorded on the context stream around every
 * stage that pic_step launches (ids below).  pic_profile(ctx, 1) enables and
 * clears it; pic_profile_read fills total milliseconds and launch counts per
 * stage for the steps run since. */
enum pic_stage_id {
    PIC_STAGE_H_HALF = 0, PIC_STAGE_PUSH_DEPOSIT, PIC_STAGE_MIGRATE, PIC_STAGE_FOLD_J,
    PIC_STAGE_H2, PIC_STAGE_E, PIC_STAGE_E_PERIODIC, PIC_STAGE_ENERGY, PIC_NSTAGES
};
int pic_profile(pic_ctx *ctx, int enable);
int pic_profile_read(pic_ctx *ctx, double *ms, long *launches, int nstages);
long pic_particle_count(pic_ctx *ctx);
/* per bulk cell particle counts, index (i*ny + l)*nz + k */
int pic_get_cell_counts(pic_ctx *ctx, int *counts);
/* particles in cell order; any output pointer may be NULL.  cell is the bulk
 * index of the owning cell.  *n_out receives the count (error if > cap). */
int pic_get_particles(pic_ctx *ctx, double *x, double *y, double *z, double *px,
                      double *py, double *pz, double *m, double *q_m, int *sort,
                      int *cell, long cap, long *n_out);
/* Gauss-law residual div E - rho at every bulk cell centre, in FP64 (both
 * forms).  rho is the cloud-in-cell charge on the cell-centre lattice that
 * the charge-conserving deposit (cell.h:1511-1693) conserves; the Yee update
 * then keeps div E - rho constant up to round-off.  res (may be NULL)
 * receives nx*ny*nz values; the first call (or set_reference != 0) stores the
 * reference, and *max_drift receives max |res - reference|.  No counterpart
 * in the reference, which never fills its Rho arrays (cell.h:2085). */
int pic_gauss_residual(pic_ctx *ctx, double *res, double *max_drift, int set_reference);
/* per sort (3) x 9 moments, reduced in FP64: N, sum x, y, z, sum px, py, pz,
 * sum (m/q_m)(gamma-1), sum |px|+|py|+|pz| */
int pic_moments(pic_ctx *ctx, double *out27);

/* pic_moments without allocation or synchronisation.  No counterpart in the
 * reference: this entry point is for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  It enqueues the same moment kernels on the context stream and
 * writes the 27 moments to d_out27, a caller-owned DEVICE array of 27
 * doubles, valid after pic_sync (or any wait on the context stream).  The
 * reduction's per-block partial sums live in a static device array of the
 * library, PIC_NOALLOC_MAX_BLOCKS x 27 doubles (432 KiB) per GPU, held until
 * the process ends; nblocks (1..PIC_NOALLOC_MAX_BLOCKS, else
 * PIC_ERR_INVALID) is the number of partial blocks, and with nblocks = 1024
 * the result is bit for bit pic_moments'.  The context's particles must be
 * committed already (a step, pic_moments or a readback has run; else
 * PIC_ERR_STATE).  All calls on one GPU share the array: calls that overlap
 * on one GPU (two contexts' streams, two host threads) race. */
#define PIC_NOALLOC_MAX_BLOCKS 2048
int pic_moments_noalloc(pic_ctx *ctx, int nblocks, double *d_out27);

/* the parameters the context was created with */
int pic_get_params(pic_ctx *ctx, pic_params *out);

/* Per-cell moments of one sort, reduced in FP64 over each cell's particles:
 * the reference's diagnostics_operations (particle_types.h:15-16, declared
 * and never implemented).  out: nx*ny*nz values for PIC_DENSITY (sum of the
 * charge weights m) and PIC_KINETIC_ENERGY (sum of (m/q_m)(gamma-1), with
 * gamma-1 formed as p.p/(gamma+1)),
 * 3*nx*ny*nz (x block, y block, z block) for PIC_VELOCITY (mean of p/gamma,
 * 0 in empty cells) and PIC_CURRENT (sum of m p/gamma). */
enum pic_moment_op { PIC_DENSITY = 0, PIC_KINETIC_ENERGY = 1, PIC_VELOCITY = 2, PIC_CURRENT = 3 };
int pic_cell_moments(pic_ctx *ctx, int sort, int op, double *out);

/* ------------------------------------------------------------------------
 * State dumps in the reference's format (mumu<num:3><nt:8>.dat, Fortran
 * unformatted records; readControlPoint gpu_plasma.h:1088-1125,
 * load_data.cu:35-343): 12 field records Ex..Qz of (nx+2)(ny+2)(nz+2)
 * doubles, then per sort ION, PLASMA_ELECTRON, BEAM_ELECTRON a record
 * (count, q_m, m) and six records x, y, z, px, py, pz.
 * ---------------------------------------------------------------------- */
typedef struct pic_dump_sort {
    long n;                  /* particles (read: filled in) */
    double q_m, m;
    double *x, *y, *z, *px, *py, *pz;
    long cap;                /* read: capacity of the six arrays */
} pic_dump_sort;

/* "mumu%03d%08d.dat" (gpu_plasma.h:1101) */
int pic_dump_file_name(int num, int nt, char *buf, size_t len);
/* host only; a null field pointer writes zeros / skips on read.  pic_dump_read
 * returns PIC_ERR_NOMEM (counts filled in) when a sort exceeds its cap. */
int pic_dump_write(const char *path, long G, const double *const fields[12],
                   const pic_dump_sort sorts[3]);
int pic_dump_read(const char *path, long G, double *const fields[12], pic_dump_sort sorts[3]);
/* a context's fields and particles to a dump / a dump into an uncommitted
 * context (LoadParticleData, load_data.cu:320-343) */
int pic_save_state_dump(pic_ctx *ctx, const char *path);
int pic_load_state_dump(pic_ctx *ctx, const char *path, long *n_dropped);
/* checkControlPoint (gpu_plasma.h:1176-1276): out13[0..11] the RMS
 * difference (CheckArraySilent, service_functions.cu:108-126) of Ex..Qz
 * against the dump, out13[12] the RMS difference of the particles (each sort
 * ordered by position; HUGE_VAL if a sort's count differs); report_path, if
 * set, gets the control_points.dat line appended */
int pic_control_point(pic_ctx *ctx, const char *path, int nt, int num, double *out13,
                      const char *report_path);

/* ------------------------------------------------------------------------
 * Differentiable form.  No reference counterpart: the reference has no
 * adjoint.  The state after step t is s_t = (E, H, particles); a run of
 * nsteps from s_0 gives the energy history W_t = sum over all ghosted nodes
 * of Ex^2 + Ey^2 + Ez^2 (the reference's Diagnose, gpu_plasma.h:621-686).
 * pic_adjoint returns the vector-Jacobian product
 *     d/ds_0 [ sum_t w_t W_t + <seed, s_nsteps> ]
 * through hand-written adjoint kernels; the cell index of a particle (every
 * floor, crossing code and periodic wrap) is held fixed.
 * ---------------------------------------------------------------------- */

/* host arrays for particle adjoints (indexed by particle id: the position in
 * load order over all pic_load_particles calls, dropped particles included)
 * and field adjoints (ghosted layout, G entries each); any pointer may be
 * null */
typedef struct pic_adjoint_io {
    double *x, *y, *z, *px, *py, *pz;
    double *field[6];        /* Ex, Ey, Ez, Hx, Hy, Hz */
} pic_adjoint_io;

/* number of particle ids handed out so far (particles offered to
 * pic_load_particles, dropped ones included) */
long pic_particle_id_count(pic_ctx *ctx);
/* ids in the order pic_get_particles returns particles (-1 without ids) */
int pic_get_particle_ids(pic_ctx *ctx, int *id, long cap, long *n_out);
/* run nsteps from the current state, keeping checkpoints of the state every
 * ckpt_every steps in device memory; energy[t-1] = W_t */
int pic_forward_record(pic_ctx *ctx, int nsteps, int ckpt_every, double *energy);
/* after pic_forward_record: recompute each checkpoint segment (recording the
 * per-step particle and field state) and run the adjoint step by step back
 * to s_0.  weights: nsteps doubles (w_t = weights[t-1]) or null; seed: the
 * adjoint of s_nsteps or null; grad receives the adjoint of s_0.  Without a
 * seed, segments after the last nonzero weight are skipped (their adjoint is
 * zero).  The context is left at s_nsteps. */
int pic_adjoint(pic_ctx *ctx, const double *weights, const pic_adjoint_io *seed,
                pic_adjoint_io *grad);
/* back to s_0 of the last pic_forward_record (its first checkpoint); a later
 * pic_forward_record of the same nsteps and ckpt_every reuses the
 * checkpoint and record memory */
int pic_rewind(pic_ctx *ctx);
/* device time of the last pic_forward_record and pic_adjoint calls (ms):
 * out[0] forward with checkpoints, out[1] recompute, out[2] adjoint steps */
int pic_adjoint_timing(pic_ctx *ctx, double *out3);

/* Decision tape of the FP64 differentiable form.  No counterpart in the
 * reference; for checking gradients against finite differences.  The step
 * is piecewise smooth in the particle state: the integers it derives from
 * positions (the cell a particle is loaded into; per step the gather's tile
 * indices and weight floors, the deposit's relay cells, the periodic wraps
 * and the receiving cell) are constant on the pieces, and pic_adjoint
 * differentiates the piece its run lies on.  A run that records these
 * integers and a run from a perturbed state that replays them stay on the
 * same piece, so central differences of the replaying runs converge to the
 * adjoint at second order in the step.
 *   pic_tape(ctx, PIC_TAPE_RECORD, n): the next n steps (the context holds
 *     loaded, not yet stepped particles) record their decisions;
 *   pic_tape_get: the recorded tape to the host: words (n x ids, step-major,
 *     indexed by particle id) and cells (ids: the bulk cell each particle was
 *     loaded into, -1 if dropped);
 *   pic_tape_set + pic_tape(ctx, PIC_TAPE_REPLAY, n): the next n steps of a
 *     context with the same particles (ids) loaded replay that tape; the
 *     particles start in the recorded cells.  A step whose tape word is not
 *     replayable (a decision outside its field, or a leaver with other than
 *     one receiver when it was recorded) ends with PIC_ERR_CELL_RANGE.
 * FP64 differentiable form only (PIC_ERR_INVALID otherwise). */
enum pic_tape_mode { PIC_TAPE_OFF = 0, PIC_TAPE_RECORD = 1, PIC_TAPE_REPLAY = 2 };
int pic_tape(pic_ctx *ctx, int mode, int nsteps);
int pic_tape_get(pic_ctx *ctx, unsigned long long *words, int *cells, long nids, int nsteps);
int pic_tape_set(pic_ctx *ctx, const unsigned long long *words, const int *cells, long nids,
                 int nsteps);

/* ------------------------------------------------------------------------
 * Seeded initial conditions (host only; the step never calls these).
 * Generators restate the reference's maxwell.cu: Wichmann-Hill uniform
 * g05cae_ (147-179) and the table-driven normal g05dde_ (77-144).
 * ---------------------------------------------------------------------- */
typedef struct pic_rng {
    int ix, iy, iz;     /* Wichmann-Hill state (cag05a_, maxwell.cu:52) */
    double store1;      /* g05dde_ carry (cag05b_, maxwell.cu:47) */
} pic_rng;

void pic_rng_seed(pic_rng *r, int ix, int iy, int iz);    /* store1 = 1.0 */
double pic_rng_uniform(pic_rng *r);                       /* g05cae_ */
double pic_rng_normal(pic_rng *r, double mean, double sd);/* g05dde_ */

typedef struct pic_species_spec {
    int sort;                 /* enum pic_sort */
    int ppc;                  /* particles per cell (mean count if profiled) */
    double density;           /* |charge density| n of the species (on the box axis
                                 if profiled) */
    double charge;            /* +1 ions, -1 electrons */
    double q_m;               /* charge/mass */
    double drift[3];          /* mean velocity (units of c) */
    double vth[3];            /* velocity spread (units of c) */
    int profile;              /* 0 uniform; 1 transverse Gaussian in y,z */
    double sigma_cells;       /* Gaussian width in cells (profile 1) */
    double perturb[8];        /* amplitudes a_k of x-displacement
                                 a_k sin(2 pi k x / Lx), k = 1..8 */
} pic_species_spec;

/* Number of particles pic_generate_species will produce for this grid. */
long pic_species_count(const pic_params *grid, const pic_species_spec *s);
/* Fills x..pz (capacity cap) for one species and returns its weight m in
 * *m_out.  stream_mode 0: one Wichmann-Hill stream continued through *rng in
 * cell order (i-major, k fastest), per particle x,y,z then ux,uy,uz;
 * stream_mode 1: an independent stream per cell seeded from (seed, cell),
 * generated on nthreads host threads (same result for any thread count). */
int pic_generate_species(const pic_params *grid, const pic_species_spec *s,
                         pic_rng *rng, int stream_mode, unsigned long long seed,
                         int nthreads, double *x, double *y, double *z,
                         double *px, double *py, double *pz, long cap,
                         long *n_out, double *m_out);

/* Device-side seeded initializer: the stream_mode-1 draws of
 * pic_generate_species for nspec species, run on the GPU (one lane per cell)
 * and written straight into the context's cell buckets, equal bit for bit to
 * pic_generate_species(stream_mode 1, seed) of each species followed by
 * pic_load_particles in species order (ids included).  The context must
 * hold no particles yet; species with displacement amplitudes are refused
 * (a displaced particle can leave its cell).  A particle that the load-time
 * predicates would place elsewhere fails the call with PIC_ERR_STATE. */
int pic_generate_device(pic_ctx *ctx, const pic_species_spec *specs, int nspec,
                        unsigned long long seed, long *n_out);

/* ------------------------------------------------------------------------
 * Parameter file: the 000_params.dat format that writeParamsFile writes
 * (maxwell.cu:330-377) and readParameterFile (params.h:11-20, declared but
 * never defined in the reference) would read: one value per line, in order
 * tex0 tey0 tez0 rimp Tb rbd ni B0 lx ly lz py pz bx by bz lp nx ny nz tau
 * beam_plasma start_from_file phase ts ms.
 * ---------------------------------------------------------------------- */
typedef struct pic_param_file {
    double tex0, tey0, tez0;   /* plasma electron thermal velocities */
    double rimp, Tb;           /* beam velocity and relative spread */
    double rbd, ni;            /* beam/plasma density ratio, plasma density */
    double B0;                 /* external magnetic field along x */
    double lx, ly, lz;         /* domain size */
    double py, pz;             /* plasma size in y, z */
    double bx, by, bz;         /* beam size */
    int lp;                    /* average particles per cell */
    int nx, ny, nz;            /* cells */
    double tau;                /* time step */
    int beam_plasma, start_from_file, phase;
    int ts;                    /* total steps (run_control.h TOTAL_STEPS) */
    int ms;                    /* steps between diagnostics */
} pic_param_file;

int pic_read_param_file(const char *path, pic_param_file *out);
int pic_write_param_file(const char *path, const pic_param_file *in);
/* The reference's species from a parameter file (maxwell.cu:379-555,
 * 606-619): ions lp per cell of density ni + rbd (q/m = 1/1836, cold);
 * plasma electrons 2 lp per cell of density ni with thermal velocities
 * tex0..tez0 and, when beam_plasma is set, the return-current drift
 * -rbd * rimp; beam electrons lp per cell of density rbd drifting at rimp
 * along x with spread Tb * rimp.  Also fills the grid of *grid. */
int pic_species_from_param_file(const pic_param_file *pf, pic_params *grid,
                                pic_species_spec species[3]);

#ifdef __cplusplus
}
#endif
#endif /* PIC_AMD_H */
