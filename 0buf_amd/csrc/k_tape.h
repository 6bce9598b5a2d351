/*
 * k_tape.h -- one particle's discrete decisions in one step of the FP64
 * differentiable form, packed in a 64-bit word (the decision tape).
 *
 * The step is piecewise smooth in the particle state: every integer the
 * push derives from a position (the gather's tile indices and weight
 * floors, the relay cells of the deposit, the periodic wraps, the receiving
 * cell) is constant on the pieces.  The adjoint differentiates the piece the
 * run lies on.  A run that reads these integers from a tape recorded by
 * another run stays on that run's piece even where its own positions would
 * cross to the next one: its finite differences then converge to the
 * adjoint at second order.
 *
 * Fields (bit offsets):
 *   0       the zero-field case of a negative coordinate (cell.h:1257-1290)
 *   1..6    gather tile indices ix, i1x, iy, i1y, iz, i1z minus 2 (0 or 1)
 *   7..18   weight floors (int)(x/h + 1) and (int)(x/h + 1.5) minus the
 *           owner cell's index, per axis (x, y, z), 2 bits each
 *   19..21  relay start cells i1 minus 2, per axis
 *   22..27  relay end cells i2 minus 1, per axis, 2 bits each
 *   28..33  Reflect case per axis: 0 inside, 1 above L, 2 below 0, 3 zero
 *   34      the particle stays (flyDirection 13)
 *   35..40  its receiving cell per axis, offset + 1 from the owner (leavers)
 *   41      the word holds a recorded step
 *   42      some decision fell outside its field's range, or the leaver had
 *           no single receiver: the word cannot be replayed
 *
 * Host and device: the CPU test suite compiles this header on its own.
 */
#ifndef K_TAPE_H
#define K_TAPE_H

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PIC_TD __host__ __device__ __forceinline__
#else
#define PIC_TD inline
#endif

typedef unsigned long long tape_word;

enum {
    TAPE_NEG = 0, TAPE_GIDX = 1, TAPE_FLOOR = 7, TAPE_RI1 = 19, TAPE_RI2 = 22, TAPE_REFL = 28,
    TAPE_STAY = 34, TAPE_RECV = 35, TAPE_VALID = 41, TAPE_BAD = 42
};

PIC_TD tape_word tape_bits(tape_word w, int at, int n) { return (w >> at) & ((1ull << n) - 1ull); }

/* put value v (expected in [lo, lo + 2^n)) at bit `at`; out of range marks
 * the word bad */
PIC_TD void tape_put(tape_word &w, int at, int n, int v, int lo) {
    const int u = v - lo;
    if (u < 0 || u >= (1 << n)) {
        w |= 1ull << TAPE_BAD;
        return;
    }
    w |= (tape_word)u << at;
}

PIC_TD int tape_get(tape_word w, int at, int n, int lo) { return (int)tape_bits(w, at, n) + lo; }

/* Reflect case of x in [0, L] (cell.h:1427-1478 as reflect1 selects it) */
PIC_TD int reflect_case(double x, double xm) {
    return x > xm ? 1 : (x < 0.0 ? 2 : ((x > 0 && x < xm) ? 0 : 3));
}
PIC_TD double reflect_apply(double x, double xm, int k) {
    return k == 1 ? x - xm : (k == 2 ? xm + x : (k == 0 ? x : 0.0));
}

#endif
