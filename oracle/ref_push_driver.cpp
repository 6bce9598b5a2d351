/*
 * ref_push_driver.cpp -- golden-vector generator for the Boris push.
 *
 * TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile against the
 * reference's own particle.h (Particle::Move, particle.h:160-202) where it
 * lies under the reference checkout; the binary goes to oracle/_ref/.
 *
 * stdin:  n, then n records of 14 doubles: E[3] H[3] x[3] p[3] q_m tau
 * stdout: n records of 6 doubles: p'[3] x1[3]   (raw little-endian)
 */
#include <cstdio>
#include <vector>
#include "particle.h"

int main() {
    long n = 0;
    if (fread(&n, sizeof(long), 1, stdin) != 1) return 1;
    std::vector<double> in(14 * n), out(6 * n);
    if (fread(in.data(), sizeof(double), 14 * n, stdin) != (size_t)(14 * n)) return 2;
    for (long j = 0; j < n; j++) {
        const double *r = &in[14 * j];
        Particle p(r[6], r[7], r[8], r[9], r[10], r[11], 1.0, r[12]);
        Field f;
        f.E.x = r[0]; f.E.y = r[1]; f.E.z = r[2];
        f.H.x = r[3]; f.H.y = r[4]; f.H.z = r[5];
        p.Move(f, r[13]);
        double *o = &out[6 * j];
        o[0] = p.pu; o[1] = p.pv; o[2] = p.pw;
        o[3] = p.x1; o[4] = p.y1; o[5] = p.z1;
    }
    fwrite(out.data(), sizeof(double), 6 * n, stdout);
    return 0;
}
