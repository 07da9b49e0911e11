/*
 * main.cpp -- command-line driver with the behaviour of the reference main
 * (cheesinglee/scphd-camera-calibration src/main.cpp:53-99):
 *   scphd_camera_calibration <configuration> <data.mat> [output_dir]
 * reads cells Z1 and Z2 (2 x n pixel measurements per frame), writes
 * prior.mat, then for k < min(frames, max_steps): predict(0.5), update with
 * Z1{k} on the fixed camera for even k and Z2{k} on the moving camera for
 * odd k (Z2 every frame when slam is set), writes k.mat, checks for NaN
 * weights and resamples.
 */
#include <algorithm>
#include <cstdio>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#include "matfile.h"
#include "scphdcameracalibration.h"

namespace {

struct ImageMeasurement {
    std::vector<double> u, v;
};

std::vector<ImageMeasurement> cells(const std::map<std::string, MatVar>& vars, const char* name) {
    auto it = vars.find(name);
    if (it == vars.end() || it->second.mclass != 1) throw std::runtime_error(std::string("no cell array ") + name);
    std::vector<ImageMeasurement> Z;
    for (const MatVar& c : it->second.cells) {
        ImageMeasurement m;
        const int64_t n = c.dims.size() >= 2 ? c.dims[1] : 0;
        if (n > 0 && c.dims[0] != 2) throw std::runtime_error(std::string("cell of ") + name + " is not 2 x n");
        for (int64_t i = 0; i < n; ++i) {
            m.u.push_back(c.data[2 * i]);
            m.v.push_back(c.data[2 * i + 1]);
        }
        Z.push_back(m);
    }
    return Z;
}

}  // namespace

int main(int argc, char* argv[]) {
    if (argc < 3) {
        std::cout << "Usage: scphd_camera_calibration [configuration] [data] [output_dir]" << std::endl;
        return 0;
    }
    try {
        std::map<std::string, MatVar> vars;
        std::string err;
        if (!mat_read(argv[2], &vars, &err)) {
            std::cerr << err << std::endl;
            return 1;
        }
        const std::vector<ImageMeasurement> Z1 = cells(vars, "Z1"), Z2 = cells(vars, "Z2");
        const std::string dir = argc > 3 ? std::string(argv[3]) + "/" : std::string();
        std::cout << "read " << Z1.size() << " measurements" << std::endl;
        SCPHDCameraCalibration calibration(argv[1]);
        calibration.writeMat((dir + "prior.mat").c_str());
        const int n_steps = (int)std::min(Z1.size(), Z2.size());
        const int kmax = std::min(n_steps, calibration.params().max_steps);
        const bool slam = calibration.params().slam != 0;
        for (int k = 0; k < kmax; ++k) {
            std::cout << "k = " << k << std::endl;
            calibration.predict(0.5);
            if (slam)
                calibration.update(Z2[k].u, Z2[k].v, false);
            else if (k % 2 == 0)
                calibration.update(Z1[k].u, Z1[k].v, true);
            else
                calibration.update(Z2[k].u, Z2[k].v, false);
            calibration.writeMat((dir + std::to_string(k) + ".mat").c_str());
            calibration.checkStuff();
            calibration.resample();
        }
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
