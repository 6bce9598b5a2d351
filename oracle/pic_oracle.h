/*
 * pic_oracle.h -- CPU restatement of the reference EM-PIC time step (FP64).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path in
 * 0buf_amd/.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product library never links, loads or calls it.
 *
 * It restates Plasma::Step (gpu_plasma.h:605-620) operation for operation with
 * the reference's per-cell AoS particle lists, 5x5x5 field tiles, 10-buffer
 * current deposit, 27-neighbour arrival scan and swap-with-last removal.
 * Every function in pic_oracle.cpp cites the reference file:line it follows.
 *
 * Pins (no function is left unpinned): the Boris push and the seeded-input
 * RNG bit-exact against the reference's own particle.h and maxwell.cxx built
 * unmodified (tests/test_oracle_golden.py); gather, deposit (straight VB
 * segments, continuity for every crossing case, relay points by geometry),
 * Yee, migration and energy by closed forms and invariants
 * (tests/test_oracle_physics.py).
 */
#ifndef PIC_ORACLE_H
#define PIC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_plasma orc_plasma;

/* field ids: same order as the product header (pic_amd.h) */
enum { ORC_EX = 0, ORC_EY, ORC_EZ, ORC_HX, ORC_HY, ORC_HZ,
       ORC_JX, ORC_JY, ORC_JZ, ORC_QX, ORC_QY, ORC_QZ, ORC_NFIELDS };

/* status bits returned by orc_migrate / orc_step */
enum { ORC_OK = 0, ORC_TOO_MANY_PARTICLES = 1 };

orc_plasma *orc_create(int nx, int ny, int nz, double lx, double ly, double lz,
                       double tau, int departure_cap);
void orc_destroy(orc_plasma *p);
void orc_set_threads(orc_plasma *p, int nthreads);

/* Plasma::addParticleListToCells (init.cu:602-630): cell by getPointCell,
 * accepted iff Cell::Insert (isPointInCell).  Returns the number dropped. */
long orc_load_particles(orc_plasma *p, int sort, const double *x, const double *y,
                        const double *z, const double *px, const double *py,
                        const double *pz, long n, double m, double q_m);

void orc_set_field(orc_plasma *p, int which, const double *a);
void orc_get_field(const orc_plasma *p, int which, double *a);

/* stages of Plasma::Step (gpu_plasma.h:605-620) */
void orc_h_half(orc_plasma *p);        /* MagneticStageOne + SetCurrentsToZero */
void orc_push(orc_plasma *p);          /* GPU_StepAllCells */
void orc_deposit(orc_plasma *p);       /* GPU_CurrentsAllCells + GPU_WriteAllCurrents */
int  orc_migrate(orc_plasma *p);       /* reorder_particles */
void orc_fold_j(orc_plasma *p);        /* SetPeriodicCurrents */
void orc_h2(orc_plasma *p);            /* MagneticFieldStageTwo */
void orc_e(orc_plasma *p);             /* ElectricFieldTrace x3 */
void orc_e_periodic(orc_plasma *p);    /* ElectricFieldComponentEvaluatePeriodic x3 */
double orc_energy(const orc_plasma *p);/* getElectricEnergy */
int  orc_step(orc_plasma *p, int nt, double *energy);

long orc_total_particles(const orc_plasma *p);
/* per bulk cell counts, index (i*Ny + l)*Nz + k */
void orc_cell_counts(const orc_plasma *p, int *counts);
/* dumps particles cell by cell; cell = bulk index of the owning cell */
long orc_get_particles(const orc_plasma *p, double *x, double *y, double *z,
                       double *px, double *py, double *pz, double *m,
                       double *q_m, int *sort, int *cell, long cap);
/* per sort (3) x 9 moments: N, sum x,y,z, sum px,py,pz, sum (m/q_m)(gamma-1),
 * sum |px|+|py|+|pz| */
void orc_moments(const orc_plasma *p, double *out);

/* compare.cxx:9-36 and service_functions.cu:108-126 */
double orc_compare(const double *a, const double *b, int num, double tol);
double orc_rms(const double *a, const double *b, int size);

/* the zig-zag relay point of CurrentToMesh (cell.h:1306-1403) for a particle
 * of cell (i, l, k) moving x -> x1: returns the crossing code m and writes
 * the relay point (x1 itself when m = 0); for the relay-point tests */
int orc_relay(const orc_plasma *P, int i, int l, int k, const double x[3], const double x1[3],
              double x2[3]);

/* one Boris push on a single particle (particle.h:160-202); for golden tests */
void orc_boris(const double E[3], const double H[3], double tau, double q_m,
               double x[3], double p[3], double x1[3]);

#ifdef __cplusplus
}
#endif
#endif
