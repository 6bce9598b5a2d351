// Host checks of the decision-tape word (0buf_amd/csrc/k_tape.h), built by
// tests/test_tape_word.py with g++.  Prints "ok" or the first failure.
#include <cmath>
#include <cstdio>

#include "k_tape.h"

#define FAIL(...) do { std::printf(__VA_ARGS__); std::printf("\n"); return 1; } while (0)

// Reflect as the FP64 push writes it (cell.h:1427-1478)
static double reflect_ref(double x, double xm) {
    return x > xm ? x - xm : (x < 0.0 ? xm + x : ((x > 0 && x < xm) ? x : 0.0));
}

int main() {
    // every field at every value it can hold, alone and all together
    struct F { int at, n, lo; } f[] = {
        {TAPE_NEG, 1, 0}, {TAPE_GIDX, 1, 2}, {TAPE_GIDX + 5, 1, 2}, {TAPE_FLOOR, 2, 0},
        {TAPE_FLOOR + 10, 2, 0}, {TAPE_RI1, 1, 2}, {TAPE_RI1 + 2, 1, 2}, {TAPE_RI2, 2, 1},
        {TAPE_RI2 + 4, 2, 1}, {TAPE_REFL, 2, 0}, {TAPE_REFL + 4, 2, 0}, {TAPE_STAY, 1, 0},
        {TAPE_RECV, 2, 0}, {TAPE_RECV + 4, 2, 0}};
    const int nf = sizeof(f) / sizeof(f[0]);
    for (int a = 0; a < nf; a++)
        for (int v = f[a].lo; v < f[a].lo + (1 << f[a].n); v++) {
            tape_word w = 0;
            tape_put(w, f[a].at, f[a].n, v, f[a].lo);
            if (tape_get(w, f[a].at, f[a].n, f[a].lo) != v) FAIL("field %d value %d", a, v);
            if ((w >> TAPE_BAD) & 1ull) FAIL("field %d value %d marked bad", a, v);
            for (int b = 0; b < nf; b++)   // no other field disturbed
                if (b != a && tape_bits(w, f[b].at, f[b].n) != 0) FAIL("field %d leaks", a);
            tape_word w2 = 0;   // out of range marks the word bad, leaves the field
            tape_put(w2, f[a].at, f[a].n, f[a].lo + (1 << f[a].n), f[a].lo);
            tape_put(w2, f[a].at, f[a].n, f[a].lo - 1, f[a].lo);
            if (!((w2 >> TAPE_BAD) & 1ull) || tape_bits(w2, f[a].at, f[a].n) != 0)
                FAIL("field %d out of range accepted", a);
        }
    // fields do not overlap and stay below the flag bits
    tape_word all = 0;
    for (int a = 0; a < nf; a++) {
        tape_word m = ((1ull << f[a].n) - 1ull) << f[a].at;
        if (all & m) FAIL("field %d overlaps", a);
        all |= m;
    }
    // the tables list every field: gather 6 x 1 bit, floors 6 x 2, relay 3 + 3 x 2,
    // reflect 3 x 2, stay 1, receiver 3 x 2: the full layout ends below TAPE_VALID
    if (TAPE_RECV + 6 != TAPE_VALID || TAPE_VALID + 1 != TAPE_BAD) FAIL("layout");
    // Reflect: case + apply equals the push's select form, edge cases included
    const double L = 3.2;
    const double xs[] = {-1e-300, -0.0, 0.0, 1e-300, 1.0, L - 1e-15, L, L + 1e-15, 4.0, -0.5,
                         std::nextafter(0.0, 1.0), std::nextafter(L, 0.0)};
    for (double x : xs) {
        const double a = reflect_apply(x, L, reflect_case(x, L)), b = reflect_ref(x, L);
        if (!(a == b && std::signbit(a) == std::signbit(b))) FAIL("reflect %g: %g vs %g", x, a, b);
    }
    std::printf("ok\n");
    return 0;
}
