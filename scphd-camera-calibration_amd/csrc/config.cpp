/*
 * config.cpp -- reader for the libconfig files of the calibration filter.
 *
 * The reference reads its parameters through libconfig++
 * (cheesinglee/scphd-camera-calibration src/scphdcameracalibration.cu:827-1005,
 * :1145, :1246-1247, :1343, :1744, :1920, :2596-2610, :2655; src/main.cpp:74,77),
 * which is not available here.  This parser covers the libconfig grammar
 * those files use: settings `name = value` or `name : value` ended by `;`,
 * `,` or nothing; groups `{ ... }`; arrays `[ ... ]` and lists `( ... )` of
 * scalars; integers (decimal, hex, L suffix), floats, booleans, strings;
 * `#`, `//` and C comments; `@include "file"` relative to the including
 * file.  Paths are dotted (prior.cartesian.x.a).
 *
 * Schema: the keys the reference code reads are authoritative.  The shipped
 * config/calibration.cfg gives prior.*.{mean,std} instead of prior.*.{a,b};
 * those map to the uniform distribution with the same mean and standard
 * deviation, a = mean - sqrt(3) std, b = mean + sqrt(3) std.  `slam`
 * (src/main.cpp:77) defaults to false when absent.  Keys of this
 * implementation with no reference counterpart: seed, max_gaussians,
 * storage ("fp32" | "fp64" | "fp32-arith", or 0 | 1 | 2).
 */
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "scphd.h"

namespace {

struct Value {
    enum Kind { NUM, BOOL, STR, GROUP, LIST } kind = NUM;
    double num = 0.0;
    bool is_int = false; /* an integer literal: exact value in u (two's complement when negative) */
    uint64_t u = 0;
    bool b = false;
    std::string str;
    std::vector<Value> items;
};

struct Parser {
    std::string src, file;
    size_t i = 0;
    int line = 1;
    std::string err;
    std::map<std::string, Value> out;
    int depth = 0;

    bool fail(const std::string& m) {
        if (err.empty()) err = file + ":" + std::to_string(line) + ": " + m;
        return false;
    }
    void skip() {
        for (;;) {
            while (i < src.size() && isspace((unsigned char)src[i])) {
                if (src[i] == '\n') ++line;
                ++i;
            }
            if (i < src.size() && src[i] == '#') {
                while (i < src.size() && src[i] != '\n') ++i;
            } else if (i + 1 < src.size() && src[i] == '/' && src[i + 1] == '/') {
                while (i < src.size() && src[i] != '\n') ++i;
            } else if (i + 1 < src.size() && src[i] == '/' && src[i + 1] == '*') {
                i += 2;
                while (i + 1 < src.size() && !(src[i] == '*' && src[i + 1] == '/')) {
                    if (src[i] == '\n') ++line;
                    ++i;
                }
                i = std::min(src.size(), i + 2);
            } else {
                return;
            }
        }
    }
    bool name(std::string& n) {
        skip();
        const size_t s = i;
        if (i < src.size() && (isalpha((unsigned char)src[i]) || src[i] == '*')) {
            ++i;
            while (i < src.size() && (isalnum((unsigned char)src[i]) || src[i] == '_' || src[i] == '-' || src[i] == '*'))
                ++i;
        }
        if (i == s) return fail("expected a setting name");
        n = src.substr(s, i - s);
        return true;
    }
    bool string_lit(std::string& s) {
        /* adjacent string literals concatenate */
        s.clear();
        while (i < src.size() && src[i] == '"') {
            ++i;
            while (i < src.size() && src[i] != '"') {
                if (src[i] == '\\' && i + 1 < src.size()) {
                    ++i;
                    const char c = src[i];
                    s += c == 'n' ? '\n' : c == 't' ? '\t' : c;
                } else {
                    s += src[i];
                }
                ++i;
            }
            if (i >= src.size()) return fail("unterminated string");
            ++i;
            skip();
        }
        return true;
    }
    bool scalar(Value& v) {
        skip();
        if (i >= src.size()) return fail("unexpected end of file");
        if (src[i] == '"') {
            v.kind = Value::STR;
            return string_lit(v.str);
        }
        size_t s = i;
        while (i < src.size() && (isalnum((unsigned char)src[i]) || src[i] == '.' || src[i] == '-' || src[i] == '+' ||
                                  src[i] == '_'))
            ++i;
        std::string tok = src.substr(s, i - s);
        std::string low;
        for (char c : tok) low += (char)tolower((unsigned char)c);
        if (low == "true" || low == "false") {
            v.kind = Value::BOOL;
            v.b = low == "true";
            v.num = v.b ? 1.0 : 0.0;
            return true;
        }
        if (tok.empty()) return fail("expected a value");
        std::string t = tok;
        while (!t.empty() && (t.back() == 'L' || t.back() == 'l')) t.pop_back();
        char* end = nullptr;
        double x;
        bool integer = false;
        uint64_t ui = 0;
        if (t.size() > 2 && t[0] == '0' && (t[1] == 'x' || t[1] == 'X')) {
            ui = strtoull(t.c_str(), &end, 16);
            x = (double)ui;
            integer = true;
        } else {
            size_t d = (t[0] == '-' || t[0] == '+') ? 1 : 0;
            integer = d < t.size() && t.find_first_not_of("0123456789", d) == std::string::npos;
            if (integer) {
                char* e2 = nullptr;
                ui = t[0] == '-' ? (uint64_t)strtoll(t.c_str(), &e2, 10) : strtoull(t.c_str(), &e2, 10);
            }
            x = strtod(t.c_str(), &end);
        }
        if (!end || *end != '\0') return fail("malformed number '" + tok + "'");
        v.kind = Value::NUM;
        v.num = x;
        v.is_int = integer;
        v.u = ui;
        return true;
    }
    bool value(Value& v, const std::string& path) {
        skip();
        if (i >= src.size()) return fail("unexpected end of file");
        if (src[i] == '{') {
            ++i;
            v.kind = Value::GROUP;
            if (++depth > 64) return fail("nesting too deep");
            if (!settings(path + ".", '}')) return false;
            --depth;
            return true;
        }
        if (src[i] == '[' || src[i] == '(') {
            const char close = src[i] == '[' ? ']' : ')';
            ++i;
            v.kind = Value::LIST;
            skip();
            int k = 0;
            while (i < src.size() && src[i] != close) {
                Value x;
                const std::string ip = path + ".[" + std::to_string(k++) + "]";
                if (!value(x, ip)) return false;
                if (x.kind != Value::GROUP) out[ip] = x;
                v.items.push_back(x);
                skip();
                if (i < src.size() && src[i] == ',') {
                    ++i;
                    skip();
                }
            }
            if (i >= src.size()) return fail("unterminated list");
            ++i;
            return true;
        }
        return scalar(v);
    }
    bool include(const std::string& prefix) {
        /* @include "path" */
        i += 8;
        skip();
        std::string path;
        if (!string_lit(path)) return false;
        std::string dir = file;
        const size_t sl = dir.find_last_of('/');
        dir = sl == std::string::npos ? std::string() : dir.substr(0, sl + 1);
        const std::string full = (!path.empty() && path[0] == '/') ? path : dir + path;
        std::ifstream in(full);
        if (!in) return fail("cannot open include file " + full);
        std::stringstream ss;
        ss << in.rdbuf();
        Parser sub;
        sub.src = ss.str();
        sub.file = full;
        if (!sub.settings(prefix, 0)) {
            err = sub.err;
            return false;
        }
        for (auto& kv : sub.out) out[kv.first] = kv.second;
        return true;
    }
    bool settings(const std::string& prefix, char close) {
        for (;;) {
            skip();
            if (i >= src.size()) {
                if (close) return fail("missing closing brace");
                return true;
            }
            if (close && src[i] == close) {
                ++i;
                return true;
            }
            if (src.compare(i, 8, "@include") == 0) {
                if (!include(prefix)) return false;
                continue;
            }
            std::string n;
            if (!name(n)) return false;
            skip();
            if (i >= src.size() || (src[i] != '=' && src[i] != ':')) return fail("expected '=' or ':' after " + n);
            ++i;
            Value v;
            const std::string path = prefix + n;
            if (!value(v, path)) return false;
            if (v.kind != Value::GROUP) out[path] = v;
            else out[path] = v;
            skip();
            if (i < src.size() && (src[i] == ';' || src[i] == ',')) ++i;
        }
    }
};

struct Lookup {
    const std::map<std::string, Value>& m;
    std::string err;
    bool has(const std::string& k) const { return m.count(k) != 0; }
    bool num(const std::string& k, double& x) {
        auto it = m.find(k);
        if (it == m.end()) {
            if (err.empty()) err = "setting not found: " + k;
            return false;
        }
        if (it->second.kind != Value::NUM && it->second.kind != Value::BOOL) {
            if (err.empty()) err = "setting is not a number: " + k;
            return false;
        }
        x = it->second.num;
        return true;
    }
    /* an unsigned 64-bit setting, exact for integer literals */
    bool u64(const std::string& k, uint64_t& x) {
        auto it = m.find(k);
        if (it == m.end() || it->second.kind != Value::NUM) {
            if (err.empty()) err = "setting is not a number: " + k;
            return false;
        }
        const Value& v = it->second;
        if (v.is_int) {
            x = v.u;
            return true;
        }
        if (v.num != std::floor(v.num) || v.num < 0 || v.num >= 9007199254740992.0) {
            if (err.empty()) err = "setting is not an integer: " + k;
            return false;
        }
        x = (uint64_t)v.num;
        return true;
    }
    bool i32(const std::string& k, int32_t& x) {
        double d;
        if (!num(k, d)) return false;
        if (d != std::floor(d) || std::fabs(d) > 2147483647.0) {
            if (err.empty()) err = "setting is not an integer: " + k;
            return false;
        }
        x = (int32_t)d;
        return true;
    }
    void opt_i32(const std::string& k, int32_t& x) {
        if (has(k)) i32(k, x);
    }
};

void copy_err(const std::string& e, char* buf, size_t n) {
    if (!buf || n == 0) return;
    std::snprintf(buf, n, "%s", e.c_str());
}

}  // namespace

extern "C" {

int scphd_params_init(scphd_params* p) {
    if (!p) return SCPHD_ERR_ARG;
    std::memset(p, 0, sizeof(*p));
    /* config/calibration.cfg of the reference */
    p->verbosity = 0;
    p->save_all_maps = 0;
    p->max_steps = 100;
    p->n_particles = 100;
    p->particles_per_feature = 100;
    p->min_weight = 0.0001;
    p->min_separation = 6.0;
    p->max_velocity = 5.0;
    p->min_range = 0.5;
    p->min_neff = 0.75;
    p->ps = 0.9;
    const double mean[12] = {0.0, -50.0, -3.0, 0.0, 0.0, 0.0, 0.3927, 0.1, 0.0, 0.0, 0.0, 0.0};
    const double sd[12] = {0.0, 0.0, 3.0, 0.0, 0.0, 0.0, 0.0, 0.1, 0.0, 0.0, 0.0, 0.0};
    for (int c = 0; c < 12; ++c) {
        p->prior_a[c] = mean[c] - std::sqrt(3.0) * sd[c];
        p->prior_b[c] = mean[c] + std::sqrt(3.0) * sd[c];
    }
    const double cam[6] = {0.0, 0.0, 0.15, 0.0, 0.001, 0.0};
    for (int c = 0; c < 6; ++c) p->cam_accel_std[c] = cam[c];
    p->feat_accel_std[0] = 0.05;
    p->feat_accel_std[1] = 0.05;
    p->feat_accel_std[2] = 0.1;
    p->fx = -895.6561;
    p->fy = -891.2656;
    p->u0 = 400.0;
    p->v0 = 300.0;
    p->std_u = 5.0;
    p->std_v = 5.0;
    p->lambda = 10.0001;
    p->pd = 0.95;
    p->birth_w0 = 0.000001;
    p->birth_d0 = 8.0;
    p->birth_vu0 = 0.0;
    p->birth_vv0 = 0.0;
    p->birth_vd0 = 0.0;
    p->birth_var_d0 = 2.0;
    p->birth_var_vu0 = 6.0;
    p->birth_var_vv0 = 6.0;
    p->birth_var_vd0 = 0.05;
    p->slam = 0;
    p->max_gaussians = 0;
    p->storage = SCPHD_STORAGE_FP32;
    p->seed = 0;
    return SCPHD_OK;
}

int scphd_params_load(const char* cfg_path, scphd_params* p, char* errbuf, size_t errlen) {
    if (!cfg_path || !p) return SCPHD_ERR_ARG;
    std::ifstream in(cfg_path);
    if (!in) {
        copy_err(std::string("cannot open ") + cfg_path, errbuf, errlen);
        return SCPHD_ERR_CONFIG;
    }
    std::stringstream ss;
    ss << in.rdbuf();
    Parser ps;
    ps.src = ss.str();
    ps.file = cfg_path;
    if (!ps.settings("", 0)) {
        copy_err(ps.err, errbuf, errlen);
        return SCPHD_ERR_CONFIG;
    }
    scphd_params q;
    scphd_params_init(&q);
    Lookup L{ps.out, ""};
    bool ok = true;
    ok &= L.i32("verbosity", q.verbosity);
    ok &= L.i32("n_particles", q.n_particles);
    ok &= L.i32("particles_per_feature", q.particles_per_feature);
    const char* comp[6] = {"x", "y", "z", "vx", "vy", "vz"};
    const char* grp[2] = {"cartesian", "angular"};
    for (int g = 0; g < 2; ++g)
        for (int c = 0; c < 6; ++c) {
            const std::string base = std::string("prior.") + grp[g] + "." + comp[c] + ".";
            const int k = 6 * g + c;
            if (L.has(base + "a") || L.has(base + "b")) {
                ok &= L.num(base + "a", q.prior_a[k]);
                ok &= L.num(base + "b", q.prior_b[k]);
            } else {
                double m = 0, sd = 0;
                ok &= L.num(base + "mean", m);
                ok &= L.num(base + "std", sd);
                q.prior_a[k] = m - std::sqrt(3.0) * sd;
                q.prior_b[k] = m + std::sqrt(3.0) * sd;
            }
        }
    const char* ax[3] = {"ax", "ay", "az"};
    for (int g = 0; g < 2; ++g)
        for (int c = 0; c < 3; ++c)
            ok &= L.num(std::string("camera_motion_model.") + grp[g] + "." + ax[c], q.cam_accel_std[3 * g + c]);
    for (int c = 0; c < 3; ++c) ok &= L.num(std::string("feature_motion_model.") + ax[c], q.feat_accel_std[c]);
    ok &= L.num("measurement_model.fx", q.fx);
    ok &= L.num("measurement_model.fy", q.fy);
    ok &= L.num("measurement_model.u0", q.u0);
    ok &= L.num("measurement_model.v0", q.v0);
    ok &= L.num("measurement_model.std_u", q.std_u);
    ok &= L.num("measurement_model.std_v", q.std_v);
    ok &= L.num("measurement_model.pd", q.pd);
    ok &= L.num("measurement_model.lambda", q.lambda);
    ok &= L.num("ps", q.ps);
    ok &= L.num("max_velocity", q.max_velocity);
    ok &= L.num("min_range", q.min_range);
    ok &= L.num("min_weight", q.min_weight);
    double sam = 0;
    ok &= L.num("save_all_maps", sam);
    q.save_all_maps = sam != 0;
    ok &= L.num("birth.w0", q.birth_w0);
    ok &= L.num("birth.d0", q.birth_d0);
    ok &= L.num("birth.vu0", q.birth_vu0);
    ok &= L.num("birth.vv0", q.birth_vv0);
    ok &= L.num("birth.vd0", q.birth_vd0);
    ok &= L.num("birth.var_d0", q.birth_var_d0);
    ok &= L.num("birth.var_vu0", q.birth_var_vu0);
    ok &= L.num("birth.var_vv0", q.birth_var_vv0);
    ok &= L.num("birth.var_vd0", q.birth_var_vd0);
    ok &= L.num("min_separation", q.min_separation);
    ok &= L.num("min_neff", q.min_neff);
    ok &= L.i32("max_steps", q.max_steps);
    if (L.has("slam")) {
        double sl = 0;
        ok &= L.num("slam", sl);
        q.slam = sl != 0;
    }
    L.opt_i32("max_gaussians", q.max_gaussians);
    if (L.has("seed")) ok &= L.u64("seed", q.seed);
    if (L.has("storage")) {
        const Value& v = ps.out.at("storage");
        if (v.kind == Value::STR)
            q.storage = v.str == "fp64"         ? SCPHD_STORAGE_FP64
                        : v.str == "fp32-arith" ? SCPHD_STORAGE_FP32_ARITH
                                                : SCPHD_STORAGE_FP32;
        else
            q.storage = v.num == 2 ? SCPHD_STORAGE_FP32_ARITH : v.num != 0 ? SCPHD_STORAGE_FP64 : SCPHD_STORAGE_FP32;
    }
    if (!ok) {
        copy_err(L.err, errbuf, errlen);
        return SCPHD_ERR_CONFIG;
    }
    *p = q;
    return SCPHD_OK;
}

}  /* extern "C" */
