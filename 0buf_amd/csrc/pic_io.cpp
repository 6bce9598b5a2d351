/*
 * pic_io.cpp -- state dumps in the reference's Fortran-record format and the
 * control-point report, written against the public C ABI (host only).
 *
 * A dump file mumu<num:3><nt:8>.dat (readControlPoint, gpu_plasma.h:1088-
 * 1125) is a sequence of Fortran unformatted records (int32 byte count,
 * payload, int32 byte count: readFortranBinaryArray, load_data.cu:35-64):
 *   12 field records Ex Ey Ez Hx Hy Hz Jx Jy Jz Qx Qy Qz, (nx+2)(ny+2)(nz+2)
 *      doubles each in the ghosted layout (readPreliminary3Darrays,
 *      load_data.cu:66-130);
 *   for each sort ION, PLASMA_ELECTRON, BEAM_ELECTRON (readBinaryParticles-
 *      AllSorts, load_data.cu:307-317): one record of three doubles
 *      (particle count, q_m, m: readBinaryParticleArraysOneSort,
 *      load_data.cu:160-201), then six records x, y, z, px, py, pz.
 */
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "pic_amd.h"

namespace {

bool write_record(std::FILE *f, const void *p, size_t bytes) {
    const int n = (int)bytes;
    return std::fwrite(&n, sizeof(int), 1, f) == 1 &&
           (bytes == 0 || std::fwrite(p, 1, bytes, f) == bytes) &&
           std::fwrite(&n, sizeof(int), 1, f) == 1;
}

/* reads one record into p (cap bytes); returns its payload size or -1 */
long read_record(std::FILE *f, void *p, size_t cap) {
    int n = 0, m = 0;
    if (std::fread(&n, sizeof(int), 1, f) != 1 || n < 0) return -1;
    if ((size_t)n > cap) return -1;
    if (n > 0 && std::fread(p, 1, (size_t)n, f) != (size_t)n) return -1;
    if (std::fread(&m, sizeof(int), 1, f) != 1 || m != n) return -1;
    return n;
}

}  // namespace

extern "C" {

int pic_dump_file_name(int num, int nt, char *buf, size_t len) {
    if (!buf || len < 16) return PIC_ERR_INVALID;
    std::snprintf(buf, len, "mumu%03d%08d.dat", num, nt);
    return PIC_SUCCESS;
}

int pic_dump_write(const char *path, long G, const double *const fields[12],
                   const pic_dump_sort sorts[3]) {
    if (!path || G <= 0 || !fields || !sorts) return PIC_ERR_INVALID;
    std::FILE *f = std::fopen(path, "wb");
    if (!f) return PIC_ERR_INVALID;
    bool ok = true;
    for (int k = 0; k < 12 && ok; k++) {
        if (!fields[k]) {
            std::vector<double> z(G, 0.0);
            ok = write_record(f, z.data(), G * sizeof(double));
        } else {
            ok = write_record(f, fields[k], G * sizeof(double));
        }
    }
    for (int s = 0; s < 3 && ok; s++) {
        const pic_dump_sort &d = sorts[s];
        const double head[3] = {(double)d.n, d.q_m, d.m};
        ok = write_record(f, head, sizeof(head));
        const double *a[6] = {d.x, d.y, d.z, d.px, d.py, d.pz};
        for (int k = 0; k < 6 && ok; k++) {
            if (d.n > 0 && !a[k]) ok = false;
            else ok = write_record(f, a[k], d.n * sizeof(double));
        }
    }
    ok = (std::fclose(f) == 0) && ok;
    return ok ? PIC_SUCCESS : PIC_ERR_INVALID;
}

int pic_dump_read(const char *path, long G, double *const fields[12], pic_dump_sort sorts[3]) {
    if (!path || G <= 0 || !sorts) return PIC_ERR_INVALID;
    std::FILE *f = std::fopen(path, "rb");
    if (!f) return PIC_ERR_INVALID;
    std::vector<double> scratch;
    bool ok = true;
    for (int k = 0; k < 12 && ok; k++) {
        double *dst = fields ? fields[k] : nullptr;
        if (!dst) {
            scratch.resize(G);
            dst = scratch.data();
        }
        ok = read_record(f, dst, G * sizeof(double)) == (long)(G * sizeof(double));
    }
    bool fits = true;
    for (int s = 0; s < 3 && ok; s++) {
        pic_dump_sort &d = sorts[s];
        double head[3];
        ok = read_record(f, head, sizeof(head)) == (long)sizeof(head);
        if (!ok) break;
        const long n = (long)head[0];
        d.n = n;
        d.q_m = head[1];
        d.m = head[2];
        double *a[6] = {d.x, d.y, d.z, d.px, d.py, d.pz};
        for (int k = 0; k < 6 && ok; k++) {
            if (n > d.cap || !a[k]) {
                /* skip the payload; the caller learns n and sizes its buffers */
                fits = false;
                scratch.resize(std::max<long>(n, 1));
                ok = read_record(f, scratch.data(), n * sizeof(double)) == (long)(n * sizeof(double));
            } else {
                ok = read_record(f, a[k], n * sizeof(double)) == (long)(n * sizeof(double));
            }
        }
    }
    std::fclose(f);
    if (!ok) return PIC_ERR_INVALID;
    return fits ? PIC_SUCCESS : PIC_ERR_NOMEM;
}

/* the context's state as a dump: fields, then the particles grouped by sort
 * in cell order; a sort loaded with two different weights or q/m cannot be
 * written (the format has one of each per sort) */
int pic_save_state_dump(pic_ctx *ctx, const char *path) {
    if (!ctx || !path) return PIC_ERR_INVALID;
    pic_params prm;
    int rc = pic_get_params(ctx, &prm);
    if (rc) return rc;
    const long G = (long)(prm.nx + 2) * (prm.ny + 2) * (prm.nz + 2);
    std::vector<double> F(12 * G);
    const double *fp[12];
    for (int k = 0; k < 12; k++) {
        if ((rc = pic_get_field(ctx, k, F.data() + k * G))) return rc;
        fp[k] = F.data() + k * G;
    }
    const long n = pic_particle_count(ctx);
    if (n < 0) return PIC_ERR_HIP;
    const size_t nn = (size_t)std::max(n, 1L);
    std::vector<double> v[6], m(nn), qm(nn);
    for (auto &a : v) a.resize(nn);
    std::vector<int> sort(nn);
    long got = 0;
    if ((rc = pic_get_particles(ctx, v[0].data(), v[1].data(), v[2].data(), v[3].data(),
                                v[4].data(), v[5].data(), m.data(), qm.data(), sort.data(),
                                nullptr, n, &got)))
        return rc;
    std::vector<double> by[3][6];
    pic_dump_sort ds[3];
    for (int s = 0; s < 3; s++) {
        std::memset(&ds[s], 0, sizeof(ds[s]));
        bool first = true;
        for (long j = 0; j < got; j++) {
            if (sort[j] != s) continue;
            if (first) {
                ds[s].m = m[j];
                ds[s].q_m = qm[j];
                first = false;
            } else if (m[j] != ds[s].m || qm[j] != ds[s].q_m) {
                return PIC_ERR_INVALID;
            }
            for (int k = 0; k < 6; k++) by[s][k].push_back(v[k][j]);
        }
        ds[s].n = (long)by[s][0].size();
        ds[s].cap = ds[s].n;
        double **dst[6] = {&ds[s].x, &ds[s].y, &ds[s].z, &ds[s].px, &ds[s].py, &ds[s].pz};
        for (int k = 0; k < 6; k++) *dst[k] = by[s][k].data();
    }
    return pic_dump_write(path, G, fp, ds);
}

/* LoadParticleData (load_data.cu:320-343): fields and the three sorts of a
 * dump into a context whose particles are not yet committed (the particles
 * go through pic_load_particles, so Cell::Insert's drop rule applies) */
int pic_load_state_dump(pic_ctx *ctx, const char *path, long *n_dropped) {
    if (!ctx || !path) return PIC_ERR_INVALID;
    pic_params prm;
    int rc = pic_get_params(ctx, &prm);
    if (rc) return rc;
    const long G = (long)(prm.nx + 2) * (prm.ny + 2) * (prm.nz + 2);
    pic_dump_sort ds[3];
    std::memset(ds, 0, sizeof(ds));
    rc = pic_dump_read(path, G, nullptr, ds);           /* counts only */
    if (rc != PIC_SUCCESS && rc != PIC_ERR_NOMEM) return rc;
    std::vector<double> F(12 * G);
    double *fp[12];
    for (int k = 0; k < 12; k++) fp[k] = F.data() + k * G;
    std::vector<double> buf[3][6];
    for (int s = 0; s < 3; s++) {
        ds[s].cap = ds[s].n;
        double **dst[6] = {&ds[s].x, &ds[s].y, &ds[s].z, &ds[s].px, &ds[s].py, &ds[s].pz};
        for (int k = 0; k < 6; k++) {
            buf[s][k].resize(std::max<long>(ds[s].n, 1));
            *dst[k] = buf[s][k].data();
        }
    }
    if ((rc = pic_dump_read(path, G, fp, ds))) return rc;
    for (int k = 0; k < 12; k++)
        if ((rc = pic_load_field(ctx, k, fp[k]))) return rc;
    long dropped = 0;
    for (int s = 0; s < 3; s++) {
        if (ds[s].n == 0) continue;
        long d = 0;
        if ((rc = pic_load_particles(ctx, s, ds[s].x, ds[s].y, ds[s].z, ds[s].px, ds[s].py,
                                     ds[s].pz, ds[s].n, ds[s].m, ds[s].q_m, &d)))
            return rc;
        dropped += d;
    }
    if (n_dropped) *n_dropped = dropped;
    return PIC_SUCCESS;
}

/* checkControlPoint (gpu_plasma.h:1176-1276): the RMS difference
 * CheckArraySilent (service_functions.cu:108-126) of each of the 12 arrays
 * against the dump, out[12] the same measure over the particles (each sort's
 * coordinates and momenta, both sides ordered by position); with report_path
 * the reference's control_points.dat line is appended */
int pic_control_point(pic_ctx *ctx, const char *path, int nt, int num, double *out13,
                      const char *report_path) {
    if (!ctx || !path || !out13) return PIC_ERR_INVALID;
    pic_params prm;
    int rc = pic_get_params(ctx, &prm);
    if (rc) return rc;
    const long G = (long)(prm.nx + 2) * (prm.ny + 2) * (prm.nz + 2);
    pic_dump_sort ds[3];
    std::memset(ds, 0, sizeof(ds));
    rc = pic_dump_read(path, G, nullptr, ds);
    if (rc != PIC_SUCCESS && rc != PIC_ERR_NOMEM) return rc;
    std::vector<double> F(12 * G), mine(G);
    double *fp[12];
    for (int k = 0; k < 12; k++) fp[k] = F.data() + k * G;
    std::vector<double> buf[3][6];
    for (int s = 0; s < 3; s++) {
        ds[s].cap = ds[s].n;
        double **dst[6] = {&ds[s].x, &ds[s].y, &ds[s].z, &ds[s].px, &ds[s].py, &ds[s].pz};
        for (int k = 0; k < 6; k++) {
            buf[s][k].resize(std::max<long>(ds[s].n, 1));
            *dst[k] = buf[s][k].data();
        }
    }
    if ((rc = pic_dump_read(path, G, fp, ds))) return rc;
    for (int k = 0; k < 12; k++) {
        if ((rc = pic_get_field(ctx, k, mine.data()))) return rc;
        double d = 0.0;
        for (long n = 0; n < G; n++) d += std::pow(mine[n] - fp[k][n], 2.0);
        out13[k] = std::pow(d / (double)G, 0.5);
    }
    /* particles: per sort, both sets ordered by (x, y, z) */
    const long n = pic_particle_count(ctx);
    const size_t nn = (size_t)std::max(n, 1L);
    std::vector<double> v[6];
    for (auto &a : v) a.resize(nn);
    std::vector<int> sort(nn);
    long got = 0;
    if ((rc = pic_get_particles(ctx, v[0].data(), v[1].data(), v[2].data(), v[3].data(),
                                v[4].data(), v[5].data(), nullptr, nullptr, sort.data(), nullptr,
                                n, &got)))
        return rc;
    double sum = 0.0;
    long cnt = 0;
    bool same = true;
    for (int s = 0; s < 3; s++) {
        std::vector<long> a, b;
        for (long j = 0; j < got; j++)
            if (sort[j] == s) a.push_back(j);
        for (long j = 0; j < ds[s].n; j++) b.push_back(j);
        if ((long)a.size() != ds[s].n) {
            same = false;
            continue;
        }
        const double *db[6] = {ds[s].x, ds[s].y, ds[s].z, ds[s].px, ds[s].py, ds[s].pz};
        auto less_mine = [&](long p, long q) {
            for (int k = 0; k < 3; k++)
                if (v[k][p] != v[k][q]) return v[k][p] < v[k][q];
            return false;
        };
        auto less_dump = [&](long p, long q) {
            for (int k = 0; k < 3; k++)
                if (db[k][p] != db[k][q]) return db[k][p] < db[k][q];
            return false;
        };
        std::sort(a.begin(), a.end(), less_mine);
        std::sort(b.begin(), b.end(), less_dump);
        for (size_t j = 0; j < a.size(); j++)
            for (int k = 0; k < 6; k++) {
                const double d = v[k][a[j]] - db[k][b[j]];
                sum += d * d;
                cnt++;
            }
    }
    out13[12] = same ? (cnt ? std::sqrt(sum / (double)cnt) : 0.0) : HUGE_VAL;
    if (report_path) {
        std::FILE *f = std::fopen(report_path, "at");
        if (!f) return PIC_ERR_INVALID;
        std::fprintf(f,
                     "nt %5d num %3d Ex %15.5e Ey %15.5e Ez %15.5e Hx %15.5e Hy %15.5e Hz %15.5e "
                     "Jx %15.5e Jy %15.5e Jz %15.5e Qx %15.5e Qy %15.5e Qz %15.5e particles "
                     "%15.5e\n",
                     nt, num, out13[0], out13[1], out13[2], out13[3], out13[4], out13[5],
                     out13[6], out13[7], out13[8], out13[9], out13[10], out13[11], out13[12]);
        std::fclose(f);
    }
    return PIC_SUCCESS;
}

}  // extern "C"
