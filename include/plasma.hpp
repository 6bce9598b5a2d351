/*
 * plasma.hpp -- header-only C++ facade with the reference's host interface
 * (class Plasma of gpu_plasma.h) over the C ABI of libpic_amd.so.
 *
 * A reference caller keeps its code: construct, Initialize(), Compute() or
 * Step(nt) in a loop, getElectricEnergy(), Diagnose(nt) appending
 * "%10d %25.15e \n" to energy.dat.  The differences are the ones of the
 * C ABI: errors are returned and kept (last_status()), never exit(); the
 * state lives in device memory owned by the context.
 */
#ifndef PIC_PLASMA_HPP
#define PIC_PLASMA_HPP

#include <cstdio>
#include <string>
#include <vector>

#include "pic_amd.h"

/* run_control.h:10,45 */
#ifndef START_STEP_NUMBER
#define START_STEP_NUMBER 1
#endif
#ifndef TOTAL_STEPS
#define TOTAL_STEPS 5
#endif

class Plasma {
public:
    /* gpu_plasma.h:1574 Plasma(nx,ny,nz,lx,ly,lz,ni1,n_per_cell1,q_m,TAU) */
    Plasma(int nx, int ny, int nz, double lx, double ly, double lz, double ni1, int n_per_cell1,
           double q_m, double TAU) {
        pic_params_default(&prm_);
        prm_.nx = nx;
        prm_.ny = ny;
        prm_.nz = nz;
        prm_.lx = lx;
        prm_.ly = ly;
        prm_.lz = lz;
        prm_.ni = ni1;
        prm_.n_per_cell = n_per_cell1;
        prm_.ion_q_m = q_m;
        prm_.tau = TAU;
        status_ = pic_create(&prm_, &ctx_);
        report(status_, "pic_create");
    }
    virtual ~Plasma() {
        if (ctx_) pic_destroy(ctx_);
    }
    Plasma(const Plasma &) = delete;
    Plasma &operator=(const Plasma &) = delete;

    /* init.cu:141-155 Initialize -> InitializeCPU: the seeded beam plasma of
     * the parameter file (maxwell.cu:330-377 readParams), three species in
     * the reference's enum order, then the particles go to their cells */
    int Initialize(const char *param_file = "000_params.dat") {
        if (!ctx_) return status_;
        pic_param_file pf;
        int rc = pic_read_param_file(param_file, &pf);
        if (rc) return fail(rc, "pic_read_param_file");
        pic_params grid;
        pic_species_spec sp[3];
        if ((rc = pic_species_from_param_file(&pf, &grid, sp))) return fail(rc, "species");
        pic_rng rng;
        pic_rng_seed(&rng, 1, 255, 25555);
        for (int s = 0; s < 3; s++) {
            long n = pic_species_count(&prm_, &sp[s]), nout = 0;
            std::vector<double> v[6];
            for (auto &a : v) a.resize(n > 0 ? n : 1);
            double m = 0.0;
            rc = pic_generate_species(&prm_, &sp[s], &rng, 0, 0, 1, v[0].data(), v[1].data(),
                                      v[2].data(), v[3].data(), v[4].data(), v[5].data(), n,
                                      &nout, &m);
            if (rc) return fail(rc, "pic_generate_species");
            long dropped = 0;
            rc = pic_load_particles(ctx_, sp[s].sort, v[0].data(), v[1].data(), v[2].data(),
                                    v[3].data(), v[4].data(), v[5].data(), nout, m, sp[s].q_m,
                                    &dropped);
            if (rc) return fail(rc, "pic_load_particles");
        }
        return status_ = PIC_SUCCESS;
    }

    /* gpu_plasma.h:605-620 Step: field half step, push, second half, energy */
    void Step(int nt) {
        if (!ctx_) return;
        double ee = 0.0;
        int rc = pic_step(ctx_, nt, &ee);
        if (rc) {
            fail(rc, "pic_step");
            return;
        }
        energy_ = ee;
        Diagnose(nt);
    }

    /* gpu_plasma.h:621-657: sum over all ghosted nodes of Ex^2+Ey^2+Ez^2 */
    virtual double getElectricEnergy() {
        double ee = 0.0;
        if (ctx_) {
            int rc = pic_energy(ctx_, &ee);
            if (rc) fail(rc, "pic_energy");
        }
        return ee;
    }

    /* gpu_plasma.h:658-686: energy.dat, truncated on the first call */
    void Diagnose(int nt) {
        std::FILE *f = std::fopen(energy_file_.c_str(), first_diag_ ? "wt" : "at");
        first_diag_ = false;
        if (!f) return;
        std::fprintf(f, "%10d %25.15e \n", nt, energy_);
        std::fclose(f);
    }

    /* gpu_plasma.h:2329-2347 Compute: steps START_STEP_NUMBER..TOTAL_STEPS */
    int Compute(int first = START_STEP_NUMBER, int last = TOTAL_STEPS) {
        for (int nt = first; nt <= last && status_ == PIC_SUCCESS; nt++) Step(nt);
        return status_;
    }

    /* Cell::number_of_particles (cell.h:187) for every bulk cell */
    std::vector<int> number_of_particles() {
        std::vector<int> c((size_t)prm_.nx * prm_.ny * prm_.nz);
        if (ctx_) {
            int rc = pic_get_cell_counts(ctx_, c.data());
            if (rc) fail(rc, "pic_get_cell_counts");
        }
        return c;
    }

    void set_energy_file(const std::string &path) { energy_file_ = path; }
    int last_status() const { return status_; }
    pic_ctx *context() { return ctx_; }

private:
    int fail(int rc, const char *where) {
        status_ = rc;
        report(rc, where);
        return rc;
    }
    static void report(int rc, const char *where) {
        if (rc)
            std::fprintf(stderr, "%s: %s %s\n", where, pic_error_string(rc), pic_last_hip_error());
    }

    pic_params prm_{};
    pic_ctx *ctx_ = nullptr;
    int status_ = PIC_SUCCESS;
    double energy_ = 0.0;
    bool first_diag_ = true;
    std::string energy_file_ = "energy.dat";
};

#endif
