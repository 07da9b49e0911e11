/*
 * matfile.h -- MAT-file level 5 reader and writer for the measurement
 * sequences and per-step outputs of the calibration filter.
 *
 * Replaces the matio calls of the reference: readMatfile
 * (cheesinglee/scphd-camera-calibration src/main.cpp:19-49, cells Z1/Z2 of
 * 2 x n double matrices) and writeMat (src/scphdcameracalibration.cu:2200-2373,
 * which writes MAT 7.3/HDF5; level 5 is written here because HDF5 is not
 * available).  Reads numeric matrices (any numeric storage type, converted
 * to double), cell arrays of them, and zlib-compressed elements; writes
 * uncompressed double matrices.
 */
#ifndef SCPHD_MATFILE_H
#define SCPHD_MATFILE_H

#include <cstdint>
#include <map>
#include <string>
#include <vector>

struct MatVar {
    int mclass = 0;              /* MATLAB array class (6 = double, 1 = cell) */
    std::vector<int64_t> dims;
    std::vector<double> data;    /* numeric data, column-major */
    std::vector<MatVar> cells;   /* cell contents, column-major */
    int64_t numel() const {
        int64_t n = 1;
        for (int64_t d : dims) n *= d;
        return dims.empty() ? 0 : n;
    }
};

/* Reads every variable it understands; returns false with *err on a
 * malformed file. */
bool mat_read(const std::string& path, std::map<std::string, MatVar>* vars, std::string* err);

class MatWriter {
public:
    void add_matrix(const std::string& name, int64_t rows, int64_t cols, const std::vector<double>& colmajor);
    bool write(const std::string& path, std::string* err) const;

private:
    std::vector<uint8_t> body_;
};

#endif
