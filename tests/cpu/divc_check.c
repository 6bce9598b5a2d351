/* Check that the FMA-corrected quotient q = fma(x - q0*h, r, q0), q0 = x*r,
 * r = RN(1/h), equals the correctly rounded x/h (Markstein's theorem), for
 * the divisors the FP64 push uses and for random divisors.  Prints the
 * number of mismatches.  Build: gcc -O2 -ffp-contract=off divc_check.c -lm
 * (tests/test_divc.py builds and runs it; -DUNCORRECTED checks that the probe
 * sees the mismatches of the uncorrected product x * r). */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

static uint64_t s = 88172645463325252ULL;
static double u01(void) {
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    return (double)(s >> 11) * (1.0 / 9007199254740992.0);
}
static double divc(double x, double h, double r) {
    double q0 = x * r;
#ifdef UNCORRECTED
    return q0;   /* the product alone: must show mismatches */
#endif
    double e = fma(-q0, h, x);
    return fma(e, r, q0);
}
int main(int argc, char **argv) {
    long n = argc > 1 ? atol(argv[1]) : 2000000;
    long bad = 0, total = 0;
    double fixed[] = {0.1, 3.2 / 32, 0.8 / 8, 25.6 / 256, 6.4 / 64, 12.8 / 128, 1.5 / 15, 12.0,
                      (0.05 * 0.1) * 0.1, 0.3 / 3, 7.0 / 3.0, 1.0 / 3.0};
    for (unsigned d = 0; d < sizeof(fixed) / sizeof(fixed[0]); d++) {
        double h = fixed[d], r = 1.0 / h;
        for (long i = 0; i < n; i++) {
            double x = (u01() - 0.3) * 60.0 * pow(2.0, (int)(u01() * 40) - 20);
            total++;
            if (divc(x, h, r) != x / h) bad++;
            double x2 = h * (double)(i % 4096) + (u01() - 0.5) * 1e-12;  /* near multiples */
            total++;
            if (divc(x2, h, r) != x2 / h) bad++;
        }
    }
    for (long i = 0; i < 4 * n; i++) {   /* random divisors */
        double h = (u01() + 0.5) * pow(2.0, (int)(u01() * 20) - 10);
        double r = 1.0 / h;
        double x = (u01() - 0.5) * pow(2.0, (int)(u01() * 40) - 20);
        total++;
        if (divc(x, h, r) != x / h) bad++;
    }
    printf("%ld %ld\n", bad, total);
    return bad != 0;
}
