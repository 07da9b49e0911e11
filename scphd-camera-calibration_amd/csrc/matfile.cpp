/*
 * matfile.cpp -- MAT-file level 5 reader/writer (see matfile.h).
 * Format: 128-byte header, then data elements (type, byte count, payload
 * padded to 8 bytes; "small" elements pack type and count in one word).
 * miCOMPRESSED (15) wraps a zlib stream of one element; miMATRIX (14) holds
 * array flags, dimensions, name and the real part.
 */
#include "matfile.h"

#include <zlib.h>

#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

namespace {

enum { miINT8 = 1, miUINT8, miINT16, miUINT16, miINT32, miUINT32, miSINGLE, miDOUBLE = 9, miINT64 = 12,
       miUINT64 = 13, miMATRIX = 14, miCOMPRESSED = 15 };
enum { mxCELL = 1, mxDOUBLE = 6 };

struct Cursor {
    const uint8_t* p;
    size_t n, i = 0;
    bool get32(uint32_t& v) {
        if (i + 4 > n) return false;
        std::memcpy(&v, p + i, 4);
        i += 4;
        return true;
    }
};

bool read_tag(Cursor& c, uint32_t& type, uint32_t& nbytes, const uint8_t*& data) {
    uint32_t w0;
    if (!c.get32(w0)) return false;
    if (w0 >> 16) { /* small data element */
        type = w0 & 0xffff;
        nbytes = w0 >> 16;
        if (c.i + 4 > c.n || nbytes > 4) return false;
        data = c.p + c.i;
        c.i += 4;
        return true;
    }
    type = w0;
    if (!c.get32(nbytes)) return false;
    if (c.i + nbytes > c.n) return false;
    data = c.p + c.i;
    c.i += ((size_t)nbytes + 7) & ~(size_t)7;
    if (c.i > c.n) c.i = c.n;
    return true;
}

bool to_doubles(uint32_t type, const uint8_t* d, uint32_t nbytes, std::vector<double>& out) {
    out.clear();
    auto conv = [&](auto tag) {
        using T = decltype(tag);
        const size_t k = nbytes / sizeof(T);
        out.resize(k);
        for (size_t j = 0; j < k; ++j) {
            T x;
            std::memcpy(&x, d + j * sizeof(T), sizeof(T));
            out[j] = (double)x;
        }
    };
    switch (type) {
        case miDOUBLE: conv(double()); return true;
        case miSINGLE: conv(float()); return true;
        case miINT8: conv(int8_t()); return true;
        case miUINT8: conv(uint8_t()); return true;
        case miINT16: conv(int16_t()); return true;
        case miUINT16: conv(uint16_t()); return true;
        case miINT32: conv(int32_t()); return true;
        case miUINT32: conv(uint32_t()); return true;
        case miINT64: conv(int64_t()); return true;
        case miUINT64: conv(uint64_t()); return true;
        default: return false;
    }
}

bool parse_matrix(const uint8_t* d, uint32_t nbytes, MatVar& v, std::string& name, int depth) {
    if (depth > 16) return false;
    Cursor c{d, nbytes};
    uint32_t t, nb;
    const uint8_t* q;
    if (!read_tag(c, t, nb, q) || t != miUINT32 || nb < 8) return false; /* array flags */
    uint32_t flags;
    std::memcpy(&flags, q, 4);
    v.mclass = flags & 0xff;
    /* only cells and numeric arrays are read; other classes (char, struct,
     * object, sparse, opaque, ...) are skipped whole */
    if (v.mclass != mxCELL && !(v.mclass >= 6 && v.mclass <= 15)) {
        v.mclass = 0;
        name.clear();
        return true;
    }
    if (!read_tag(c, t, nb, q) || t != miINT32) return false; /* dimensions */
    v.dims.resize(nb / 4);
    for (size_t j = 0; j < v.dims.size(); ++j) {
        int32_t x;
        std::memcpy(&x, q + 4 * j, 4);
        v.dims[j] = x;
    }
    if (!read_tag(c, t, nb, q)) return false; /* name */
    name.assign((const char*)q, nb);
    if (v.mclass == mxCELL) {
        const int64_t n = v.numel();
        if (n < 0 || n > (1 << 26)) return false;
        v.cells.resize((size_t)n);
        for (int64_t j = 0; j < n; ++j) {
            if (!read_tag(c, t, nb, q) || t != miMATRIX) return false;
            std::string cn;
            if (nb > 0 && !parse_matrix(q, nb, v.cells[(size_t)j], cn, depth + 1)) return false;
        }
        return true;
    }
    if (v.mclass >= 6 && v.mclass <= 15) { /* numeric classes */
        if (v.numel() == 0) return true;
        if (!read_tag(c, t, nb, q)) return false;
        if (!to_doubles(t, q, nb, v.data)) return false;
        return (int64_t)v.data.size() == v.numel();
    }
    return true;
}

bool parse_elements(const uint8_t* d, size_t n, std::map<std::string, MatVar>* vars, std::string* err) {
    Cursor c{d, n};
    while (c.i + 8 <= c.n) {
        uint32_t t, nb;
        const uint8_t* q;
        if (!read_tag(c, t, nb, q)) {
            *err = "truncated data element";
            return false;
        }
        if (t == miCOMPRESSED) {
            c.i = (size_t)(q - c.p) + nb; /* compressed elements carry no padding */
            std::vector<uint8_t> out;
            z_stream zs;
            std::memset(&zs, 0, sizeof zs);
            if (inflateInit(&zs) != Z_OK) {
                *err = "zlib init";
                return false;
            }
            zs.next_in = const_cast<uint8_t*>(q);
            zs.avail_in = nb;
            uint8_t buf[1 << 16];
            int rc;
            do {
                zs.next_out = buf;
                zs.avail_out = sizeof buf;
                rc = inflate(&zs, Z_NO_FLUSH);
                if (rc != Z_OK && rc != Z_STREAM_END) {
                    inflateEnd(&zs);
                    *err = "zlib inflate failed";
                    return false;
                }
                out.insert(out.end(), buf, buf + (sizeof buf - zs.avail_out));
            } while (rc != Z_STREAM_END && zs.avail_in > 0);
            inflateEnd(&zs);
            if (!parse_elements(out.data(), out.size(), vars, err)) return false;
        } else if (t == miMATRIX) {
            MatVar v;
            std::string name;
            if (!parse_matrix(q, nb, v, name, 0)) {
                *err = "malformed matrix element";
                return false;
            }
            if (v.mclass != 0) (*vars)[name] = std::move(v);
        }
    }
    return true;
}

void put32(std::vector<uint8_t>& b, uint32_t v) {
    const uint8_t* p = (const uint8_t*)&v;
    b.insert(b.end(), p, p + 4);
}

void pad8(std::vector<uint8_t>& b) {
    while (b.size() % 8) b.push_back(0);
}

}  // namespace

bool mat_read(const std::string& path, std::map<std::string, MatVar>* vars, std::string* err) {
    std::ifstream in(path, std::ios::binary);
    if (!in) {
        *err = "cannot open " + path;
        return false;
    }
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string s = ss.str();
    if (s.size() < 128 || s[126] != 'I' || s[127] != 'M') {
        *err = "not a little-endian MAT level 5 file";
        return false;
    }
    return parse_elements((const uint8_t*)s.data() + 128, s.size() - 128, vars, err);
}

void MatWriter::add_matrix(const std::string& name, int64_t rows, int64_t cols, const std::vector<double>& cm) {
    std::vector<uint8_t> m;
    put32(m, miUINT32);
    put32(m, 8);
    put32(m, mxDOUBLE);
    put32(m, 0);
    put32(m, miINT32);
    put32(m, 8);
    put32(m, (uint32_t)rows);
    put32(m, (uint32_t)cols);
    put32(m, miINT8);
    put32(m, (uint32_t)name.size());
    m.insert(m.end(), name.begin(), name.end());
    pad8(m);
    put32(m, miDOUBLE);
    put32(m, (uint32_t)(rows * cols * 8));
    const uint8_t* p = (const uint8_t*)cm.data();
    m.insert(m.end(), p, p + rows * cols * 8);
    pad8(m);
    put32(body_, miMATRIX);
    put32(body_, (uint32_t)m.size());
    body_.insert(body_.end(), m.begin(), m.end());
}

bool MatWriter::write(const std::string& path, std::string* err) const {
    char hdr[128];
    std::memset(hdr, ' ', 116);
    const char* text = "MATLAB 5.0 MAT-file, written by the scphd gfx950 calibration filter";
    std::memcpy(hdr, text, std::strlen(text));
    std::memset(hdr + 116, 0, 8);
    hdr[124] = 0x00;
    hdr[125] = 0x01;
    hdr[126] = 'I';
    hdr[127] = 'M';
    std::ofstream out(path, std::ios::binary);
    if (!out) {
        *err = "cannot create " + path;
        return false;
    }
    out.write(hdr, 128);
    out.write((const char*)body_.data(), (std::streamsize)body_.size());
    return (bool)out;
}
