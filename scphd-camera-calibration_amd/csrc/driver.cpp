/*
 * driver.cpp -- SCPHDCameraCalibration over the C ABI (reference interface
 * src/scphdcameracalibration.cuh:29-122 of cheesinglee/scphd-camera-calibration).
 */
#include "scphdcameracalibration.h"

#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "matfile.h"

SCPHDCameraCalibration::SCPHDCameraCalibration(const char* config_file) {
    char err[512] = {0};
    if (scphd_params_load(config_file, &params_, err, sizeof err) != SCPHD_OK)
        throw std::runtime_error(std::string("configuration: ") + err);
    check(scphd_create(&params_, &h_), "scphd_create");
}

SCPHDCameraCalibration::SCPHDCameraCalibration(const scphd_params& params) : params_(params) {
    check(scphd_create(&params_, &h_), "scphd_create");
}

SCPHDCameraCalibration::~SCPHDCameraCalibration() { scphd_destroy(h_); }

void SCPHDCameraCalibration::check(int rc, const char* what) {
    if (rc == SCPHD_OK) return;
    std::string msg = std::string(what) + " failed (" + std::to_string(rc) + ")";
    if (h_) msg += ": " + std::string(scphd_last_error(h_));
    throw std::runtime_error(msg);
}

void SCPHDCameraCalibration::predict(double dt) { check(scphd_predict(h_, dt), "predict"); }

void SCPHDCameraCalibration::update(std::vector<double> u, std::vector<double> v, bool fixed_camera) {
    if (u.size() != v.size()) throw std::invalid_argument("update: u and v differ in length");
    check(scphd_update(h_, u.data(), v.data(), (int32_t)u.size(), fixed_camera ? 1 : 0), "update");
}

void SCPHDCameraCalibration::resample() { check(scphd_resample(h_, nullptr), "resample"); }

void SCPHDCameraCalibration::checkStuff() {
    const int rc = scphd_check(h_);
    if (rc == SCPHD_ERR_NAN) {
        std::cout << "nan weight detected" << std::endl;
        std::exit(2);
    }
    check(rc, "checkStuff");
}

DisparityMeasurementModel SCPHDCameraCalibration::measurement_model() const {
    return DisparityMeasurementModel(params_.fx, params_.fy, params_.u0, params_.v0, params_.std_u, params_.std_v,
                                     params_.pd, params_.lambda);
}

void SCPHDCameraCalibration::setMeasurement_model(const DisparityMeasurementModel& m) {
    /* Takes effect from the next constructed filter state: the model lives
     * in the handle's parameters. */
    scphd_params p = params_;
    p.fx = m.fx();
    p.fy = m.fy();
    p.u0 = m.u0();
    p.v0 = m.v0();
    p.std_u = m.std_u();
    p.std_v = m.std_v();
    p.pd = m.pd();
    p.lambda = m.kappa() * (2 * m.v0()) * (2 * m.v0());
    std::vector<double> st(12 * (size_t)p.n_particles), w(p.n_particles);
    scphd_counts c;
    check(scphd_get_counts(h_, &c), "get_counts");
    std::vector<int32_t> gc(p.n_particles);
    std::vector<double> wg((size_t)c.n_gaussians), smp((size_t)c.n_samples * 7);
    check(scphd_get_parents(h_, st.data(), w.data()), "get_parents");
    check(scphd_get_maps(h_, gc.data(), wg.data(), smp.data()), "get_maps");
    scphd_t* nh = nullptr;
    check(scphd_create(&p, &nh), "scphd_create");
    const int rc = scphd_set_state(nh, st.data(), w.data(), gc.data(), c.n_gaussians, wg.data(), smp.data(),
                                   c.n_predict, c.n_update, c.n_resample);
    if (rc != SCPHD_OK) {
        scphd_destroy(nh);
        check(rc, "set_state");
    }
    scphd_destroy(h_);
    h_ = nh;
    params_ = p;
}

void SCPHDCameraCalibration::writeMat(const char* filename) {
    check(scphd_write_mat(h_, filename), "writeMat");
}

void SCPHDCameraCalibration::transformTest() {
    const int Np = params_.n_particles;
    scphd_counts c;
    check(scphd_get_counts(h_, &c), "get_counts");
    std::vector<double> st(12 * (size_t)Np), w(Np);
    check(scphd_get_parents(h_, st.data(), w.data()), "get_parents");
    std::vector<int32_t> gc(Np);
    std::vector<double> wg((size_t)c.n_gaussians), smp((size_t)c.n_samples * 7);
    check(scphd_get_maps(h_, gc.data(), wg.data(), smp.data()), "get_maps");
    const DisparityMeasurementModel m = measurement_model();
    const int64_t n = c.n_samples;
    int64_t i = 0;
    for (int p = 0; p < Np; ++p) {
        const double* e = &st[(size_t)p * 12];
        Extrinsics x{{e[0], e[1], e[2], e[3], e[4], e[5]}, {e[6], e[7], e[8], e[9], e[10], e[11]}};
        for (int64_t k = 0; k < (int64_t)gc[p] * params_.particles_per_feature; ++k, ++i) {
            EuclideanPoint pw{smp[i], smp[n + i], smp[2 * n + i], smp[3 * n + i], smp[4 * n + i], smp[5 * n + i]};
            for (int fixed = 1; fixed >= 0; --fixed) {
                Extrinsics ee = fixed ? Extrinsics{{0, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0}} : x;
                DisparityPoint d = m.computeMeasurement(pw, ee);
                EuclideanPoint b = m.invertMeasurement(d, ee);
                std::printf("%d %lld %s world %g %g %g disparity %g %g %g back %g %g %g\n", p, (long long)k,
                            fixed ? "fixed" : "displaced", pw.x, pw.y, pw.z, d.u, d.v, d.d, b.x, b.y, b.z);
            }
        }
    }
}
