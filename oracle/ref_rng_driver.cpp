/*
 * ref_rng_driver.cpp -- golden-vector generator for the seeded-input RNG.
 *
 * TEST INFRASTRUCTURE ONLY.  Linked by oracle/Makefile with the reference's
 * own maxwell.cxx (Wichmann-Hill g05cae_ at maxwell.cu:147-179, normal
 * g05dde_ at maxwell.cu:77-144, seeds 1,255,25555 at maxwell.cu:52).
 *
 * argv: n_uniform n_normal mean sigma
 * stdout: n_uniform doubles from rnd_uniform, then n_normal doubles from
 * rnd_gaussian(mean, sigma), drawn from the one shared generator state in
 * that order (raw little-endian).
 *
 * argv: params   -- calls the reference's writeParamsFile
 * (maxwell.cu:330-377) with fixed sample values; it writes 000_params.dat in
 * the current directory.
 */
#include <cstdio>
#include <cstdlib>
#include <vector>

double rnd_uniform(int dbg_print);
int writeParamsFile(double tex0, double tey0, double tez0, double Tb, double rimp, double rbd,
                    double ni, double lx, double ly, double lz, int lp, int nx, int ny, int nz,
                    double tau, double B0, double bx, double by, double bz, double py, double pz,
                    int beam_plasma, int start_from_file, int ts, int ms, int phase);
double rnd_gaussian(double a, double b, int dbg_print);

int main(int argc, char **argv) {
    if (argc == 2 && argv[1][0] == 'p')
        return writeParamsFile(1e-3, 2e-3, 3e-3, 0.14, 0.2, 2e-3, 1.0, 1.1424, 0.05, 0.06, 1000,
                               100, 4, 5, 1e-3, 0.5, 1.1424, 0.05, 0.06, 0.04, 0.05, 1, 0, 5, 2,
                               0);
    if (argc < 5) return 1;
    long nu = atol(argv[1]), nn = atol(argv[2]);
    double a = atof(argv[3]), b = atof(argv[4]);
    std::vector<double> out;
    for (long j = 0; j < nu; j++) out.push_back(rnd_uniform(0));
    for (long j = 0; j < nn; j++) out.push_back(rnd_gaussian(a, b, 0));
    fwrite(out.data(), sizeof(double), out.size(), stdout);
    return 0;
}
