/*
 * sanitize_check.cpp -- drives the host C++ of the library (configuration
 * parser, MAT-file reader and writer) and the FP64 oracle through their
 * ordinary and malformed inputs, for a build with AddressSanitizer and
 * UndefinedBehaviorSanitizer (tests/test_sanitizers.py).  Exits nonzero on
 * a wrong result; a sanitizer finding aborts the run.
 *
 *   sanitize_check <calibration.cfg> <scratch directory> [sequence.mat ...]
 *
 * Each sequence file (cell arrays of measurement matrices, possibly
 * compressed) is read whole and then at every truncation and with flipped
 * bytes in its first elements.
 */
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "matfile.h"
#include "scphd.h"

extern "C" {
void* orc_create(const void* params);
void orc_destroy(void* h);
int orc_set_threads(int n);
int orc_predict(void* h, double dt);
int orc_update(void* h, const double* u, const double* v, int M, int fixed);
int orc_resample(void* h, int32_t* resampled);
int orc_check(void* h);
void orc_get_counts(void* h, int64_t* out);
void orc_get_maps(void* h, int32_t* gcount, double* wg, double* samples);
void orc_get_parents(void* h, double* states, double* weights);
}

static int failures = 0;

#define EXPECT(cond, what)                                            \
    do {                                                              \
        if (!(cond)) {                                                \
            std::fprintf(stderr, "FAILED: %s (%s)\n", what, #cond);   \
            ++failures;                                               \
        }                                                             \
    } while (0)

static std::string slurp(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    return std::string((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

static void spill(const std::string& path, const std::string& text) {
    std::ofstream f(path, std::ios::binary);
    f << text;
}

/* the parser on the reference's configuration and on broken variants of it */
static void check_config(const std::string& cfg, const std::string& dir, scphd_params* out) {
    char err[512];
    EXPECT(scphd_params_load(cfg.c_str(), out, err, sizeof err) == SCPHD_OK, "golden configuration loads");
    const std::string text = slurp(cfg);
    const std::string bad = dir + "/bad.cfg";
    /* every prefix of the file: a missing key, an unterminated group or
     * string, or a clean parse -- never a crash */
    for (size_t n = 0; n < text.size(); n += 7) {
        spill(bad, text.substr(0, n));
        scphd_params p;
        (void)scphd_params_load(bad.c_str(), &p, err, sizeof err);
    }
    const char* broken[] = {"n_particles = = 3\n", "a = \"unterminated\n", "g : { x = 1; \n", "@include \"nope.cfg\"\n",
                            "x = 0x;\n", "y = 1e999999;\n", "z = [1, 2,\n", "w = (1, \"a\", {q = 2;});\n"};
    for (const char* b : broken) {
        spill(bad, b);
        scphd_params p;
        EXPECT(scphd_params_load(bad.c_str(), &p, err, sizeof err) != SCPHD_OK, "a broken file is refused");
    }
    scphd_params p;
    EXPECT(scphd_params_load((dir + "/absent.cfg").c_str(), &p, nullptr, 0) != SCPHD_OK, "a missing file");
}

/* writer -> reader round trip, then every truncation of the file */
static void check_mat(const std::string& dir) {
    const std::string path = dir + "/seq.mat";
    MatWriter w;
    std::vector<double> a(2 * 37), b(2 * 5);
    for (size_t i = 0; i < a.size(); ++i) a[i] = 0.25 * (double)i - 3.0;
    for (size_t i = 0; i < b.size(); ++i) b[i] = std::sqrt((double)i);
    w.add_matrix("A", 2, 37, a);
    w.add_matrix("B", 2, 5, b);
    std::string err;
    EXPECT(w.write(path, &err), "MAT file written");
    std::map<std::string, MatVar> vars;
    EXPECT(mat_read(path, &vars, &err), "MAT file read back");
    EXPECT(vars.count("A") && vars["A"].data == a, "matrix A round trip");
    EXPECT(vars.count("B") && vars["B"].data == b, "matrix B round trip");
    const std::string text = slurp(path);
    const std::string cut = dir + "/cut.mat";
    for (size_t n = 0; n < text.size(); n += 3) {
        spill(cut, text.substr(0, n));
        std::map<std::string, MatVar> v;
        (void)mat_read(cut, &v, &err);
    }
    /* flipped bytes in the header and the first element tags */
    for (size_t i = 116; i < std::min<size_t>(text.size(), 200); ++i) {
        std::string t = text;
        t[i] = (char)(t[i] ^ 0x5A);
        spill(cut, t);
        std::map<std::string, MatVar> v;
        (void)mat_read(cut, &v, &err);
    }
}

/* a sequence file as it is, truncated, and with flipped bytes */
static void check_sequence(const std::string& path, const std::string& dir) {
    std::map<std::string, MatVar> vars;
    std::string err;
    EXPECT(mat_read(path, &vars, &err), "sequence file reads");
    const std::string text = slurp(path);
    const std::string cut = dir + "/seqcut.mat";
    const size_t step = std::max<size_t>(1, text.size() / 400);
    for (size_t n = 0; n < text.size(); n += step) {
        spill(cut, text.substr(0, n));
        std::map<std::string, MatVar> v;
        (void)mat_read(cut, &v, &err);
    }
    for (size_t i = 128; i < std::min<size_t>(text.size(), 512); i += 3) {
        std::string t = text;
        t[i] = (char)(t[i] ^ 0xA5);
        spill(cut, t);
        std::map<std::string, MatVar> v;
        (void)mat_read(cut, &v, &err);
    }
}

/* a few frames of the oracle on synthetic measurements */
static void check_oracle(scphd_params p) {
    p.n_particles = 16;
    p.particles_per_feature = 32;
    p.seed = 12345;
    orc_set_threads(2);
    void* h = orc_create(&p);
    EXPECT(h != nullptr, "oracle created");
    if (!h) return;
    uint64_t s = 0x9E3779B97F4A7C15ull;
    auto uni = [&s]() {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return (double)(s >> 11) * 0x1.0p-53;
    };
    for (int k = 0; k < 5; ++k) {
        const int M = k == 3 ? 0 : 12 + 3 * k; /* one empty set */
        std::vector<double> u(M + 1), v(M + 1);
        for (int m = 0; m < M; ++m) {
            u[m] = 640.0 * uni();
            v[m] = 480.0 * uni();
        }
        EXPECT(orc_predict(h, 1.0 / 30.0) == 0, "predict");
        EXPECT(orc_update(h, u.data(), v.data(), M, k % 2 == 0) == 0, "update");
        EXPECT(orc_check(h) == 0, "finite weights");
        int32_t res = 0;
        EXPECT(orc_resample(h, &res) == 0, "resample");
    }
    int64_t c[8];
    orc_get_counts(h, c);
    std::vector<int32_t> gc(c[0]);
    std::vector<double> wg(std::max<int64_t>(c[2], 1)), smp(7 * std::max<int64_t>(c[3], 1));
    std::vector<double> st(12 * c[0]), w(c[0]);
    orc_get_maps(h, gc.data(), wg.data(), smp.data());
    orc_get_parents(h, st.data(), w.data());
    int64_t ng = 0;
    for (int32_t g : gc) ng += g;
    EXPECT(ng == c[2], "map table adds up");
    double ws = 0.0;
    for (double x : w) ws += x;
    EXPECT(std::fabs(ws - 1.0) < 1e-12, "parent weights sum to one");
    orc_destroy(h);
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: sanitize_check <calibration.cfg> <scratch directory> [sequence.mat ...]\n");
        return 2;
    }
    scphd_params p;
    check_config(argv[1], argv[2], &p);
    check_mat(argv[2]);
    for (int a = 3; a < argc; ++a) check_sequence(argv[a], argv[2]);
    check_oracle(p);
    std::printf("sanitize_check: %d failure(s)\n", failures);
    return failures ? 1 : 0;
}
