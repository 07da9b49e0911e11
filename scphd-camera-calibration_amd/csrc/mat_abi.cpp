/*
 * mat_abi.cpp -- MAT-file entry points of the C ABI: the measurement
 * sequence reader of the reference main (readMatfile,
 * cheesinglee/scphd-camera-calibration src/main.cpp:19-49) and the per-step
 * output writer (writeMat, src/scphdcameracalibration.cu:2200-2373).
 */
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "matfile.h"
#include "scphd.h"

extern "C" {

int scphd_mat_read_measurements(const char* path, const char* varname, int32_t* n_steps, int64_t* n_total,
                                int32_t* counts, double* u, double* v) {
    if (!path || !varname || !n_steps || !n_total) return SCPHD_ERR_ARG;
    std::map<std::string, MatVar> vars;
    std::string err;
    if (!mat_read(path, &vars, &err)) return SCPHD_ERR_CONFIG;
    auto it = vars.find(varname);
    if (it == vars.end() || it->second.mclass != 1) return SCPHD_ERR_CONFIG;
    const std::vector<MatVar>& cells = it->second.cells;
    int64_t tot = 0;
    for (const MatVar& c : cells) {
        const int64_t n = c.dims.size() >= 2 ? c.dims[1] : 0;
        if (n > 0 && (c.dims[0] != 2 || (int64_t)c.data.size() < 2 * n)) return SCPHD_ERR_CONFIG;
        tot += n;
    }
    if (counts) {
        if (*n_steps < (int32_t)cells.size() || *n_total < tot || (tot > 0 && (!u || !v))) return SCPHD_ERR_ARG;
        int64_t o = 0;
        for (size_t k = 0; k < cells.size(); ++k) {
            const MatVar& c = cells[k];
            const int64_t n = c.dims.size() >= 2 ? c.dims[1] : 0;
            counts[k] = (int32_t)n;
            for (int64_t i = 0; i < n; ++i, ++o) {
                u[o] = c.data[2 * i];
                v[o] = c.data[2 * i + 1];
            }
        }
    }
    *n_steps = (int32_t)cells.size();
    *n_total = tot;
    return SCPHD_OK;
}

int scphd_write_mat(scphd_t* f, const char* path) {
    if (!f || !path) return SCPHD_ERR_ARG;
    scphd_counts c;
    int rc = scphd_get_counts(f, &c);
    if (rc) return rc;
    const int Np = (int)c.n_particles, S = (int)c.particles_per_feature;
    std::vector<double> st(12 * (size_t)Np), w(Np);
    if ((rc = scphd_get_parents(f, st.data(), w.data()))) return rc;
    /* materialises a deferred predict, which can change nothing but the samples */
    if ((rc = scphd_get_maps(f, nullptr, nullptr, nullptr))) return rc;
    if ((rc = scphd_get_counts(f, &c))) return rc;
    std::vector<int32_t> gc(Np);
    std::vector<double> wg((size_t)c.n_gaussians + 1), smp((size_t)c.n_samples * 7 + 1);
    if ((rc = scphd_get_maps(f, gc.data(), wg.data(), smp.data()))) return rc;
    int best = 0; /* the first highest-weight parent */
    for (int p = 1; p < Np; ++p)
        if (w[p] > w[best]) best = p;
    int64_t g0 = 0;
    for (int p = 0; p < best; ++p) g0 += gc[p];
    const int64_t G = gc[best], n = c.n_samples;
    std::vector<double> mw(wg.begin() + g0, wg.begin() + g0 + G);
    std::vector<double> mp((size_t)G * S * 6), mpw((size_t)G * S);
    for (int64_t i = 0; i < G * S; ++i) {
        for (int k = 0; k < 6; ++k) mp[(size_t)i * 6 + k] = smp[(size_t)k * n + g0 * S + i];
        mpw[(size_t)i] = smp[(size_t)6 * n + g0 * S + i];
    }
    MatWriter mat;
    mat.add_matrix("particles", 12, Np, st); /* 12 x N_p, column p = parent p */
    mat.add_matrix("weights", 1, Np, w);
    mat.add_matrix("map_weights", 1, G, mw);
    mat.add_matrix("map_particles", 6, G * S, mp);
    mat.add_matrix("map_particle_weights", 1, G * S, mpw);
    std::string err;
    return mat.write(path, &err) ? SCPHD_OK : SCPHD_ERR_ARG;
}

}  /* extern "C" */
