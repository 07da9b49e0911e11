/*
 * scphdcameracalibration.h -- C++ driver class with the interface of the
 * reference SCPHDCameraCalibration (cheesinglee/scphd-camera-calibration
 * src/scphdcameracalibration.cuh:29-122), implemented over the C ABI of
 * scphd.h.  A program written against the reference class compiles against
 * this one; the filter runs on the GPU through hand-written gfx950 kernels.
 *
 * Behaviour kept from the reference: the constructor reads the libconfig
 * file (an unreadable file or missing key throws std::runtime_error, where
 * the reference lets a libconfig exception escape); checkStuff() prints
 * "nan weight detected" and exits with status 2 on a NaN parent weight
 * (src/scphdcameracalibration.cu:812-821).  Other failures throw
 * std::runtime_error with the library's message.
 */
#ifndef SCPHDCAMERACALIBRATION_H
#define SCPHDCAMERACALIBRATION_H

#include <string>
#include <vector>

#include "scphd.h"
#include "scphd_models.h"

class SCPHDCameraCalibration {
public:
    explicit SCPHDCameraCalibration(const char* config_file);
    explicit SCPHDCameraCalibration(const scphd_params& params);
    ~SCPHDCameraCalibration();
    SCPHDCameraCalibration(const SCPHDCameraCalibration&) = delete;
    SCPHDCameraCalibration& operator=(const SCPHDCameraCalibration&) = delete;

    /* src/scphdcameracalibration.cu:1034 */
    void predict(double dt);
    /* src/scphdcameracalibration.cu:1576 */
    void update(std::vector<double> u, std::vector<double> v, bool fixed_camera);
    /* src/scphdcameracalibration.cu:2097-2373: writes a MAT v5 file with
     * the variables particles (12 x N_p), weights (1 x N_p) and the
     * highest-weight parent's map: map_weights (1 x G), map_particles
     * (6 x G*S) and map_particle_weights (1 x G*S). */
    void writeMat(const char* filename);
    /* src/scphdcameracalibration.cu:2651 */
    void resample();
    /* src/scphdcameracalibration.cu:812 */
    void checkStuff();
    /* src/scphdcameracalibration.cu:2822-2830 */
    DisparityMeasurementModel measurement_model() const;
    void setMeasurement_model(const DisparityMeasurementModel& measurement_model);
    /* src/scphdcameracalibration.cu:735: world -> disparity -> world round
     * trip of the current map for manual inspection, printed to stdout. */
    void transformTest();

    /* additions */
    const scphd_params& params() const { return params_; }
    scphd_t* handle() const { return h_; }

private:
    void check(int rc, const char* what);
    scphd_params params_;
    scphd_t* h_ = nullptr;
};

#endif /* SCPHDCAMERACALIBRATION_H */
