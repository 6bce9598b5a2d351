/*
 * plasma_main.cpp -- the reference's main.cu driver (main.cu:8-25) on the
 * C++ facade: construct Plasma, Initialize() from 000_params.dat, Compute().
 *
 *   plasma_main [param_file [first last [nx ny nz lx ly lz ni ppc q_m tau]]]
 *
 * Defaults are main.cu:14's Plasma(100,4,4,1.1424,0.05,0.05,1.0,2000,1.0,0.001)
 * and run_control.h's steps 1..5.  Writes energy.dat like Diagnose.  Exit
 * status: the library's int status of the first failing call (0 = success).
 */
#include <cstdlib>

#include "plasma.hpp"

int main(int argc, char **argv) {
    const char *pfile = argc > 1 ? argv[1] : "000_params.dat";
    int first = argc > 3 ? std::atoi(argv[2]) : START_STEP_NUMBER;
    int last = argc > 3 ? std::atoi(argv[3]) : TOTAL_STEPS;
    int nx = 100, ny = 4, nz = 4, ppc = 2000;
    double lx = 1.1424, ly = 0.05, lz = 0.05, ni = 1.0, q_m = 1.0, tau = 0.001;
    if (argc > 13) {
        nx = std::atoi(argv[4]);
        ny = std::atoi(argv[5]);
        nz = std::atoi(argv[6]);
        lx = std::atof(argv[7]);
        ly = std::atof(argv[8]);
        lz = std::atof(argv[9]);
        ni = std::atof(argv[10]);
        ppc = std::atoi(argv[11]);
        q_m = std::atof(argv[12]);
        tau = std::atof(argv[13]);
    }
    Plasma *plasma = new Plasma(nx, ny, nz, lx, ly, lz, ni, ppc, q_m, tau);
    int rc = plasma->last_status();
    if (rc == 0) rc = plasma->Initialize(pfile);
    if (rc == 0) rc = plasma->Compute(first, last);
    delete plasma;
    return rc;
}
