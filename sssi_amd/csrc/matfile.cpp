// matfile.cpp -- MAT-file (Level 5) reading and writing of real double
// arrays, the durable format of the reference's per-shot results: its RTM
// apps save dataTrue / dataSmooth / rtmsnapshot per shot and the stacked image
// with `save(..., '-v7.3')` when the file does not exist yet, and `load` them
// back (applications/rtm_2d_acoustic/mainRtmTimeCpmlFor2dAw.m:170-184,
// :247-264, mainRtmTimeSpongeFor2dAw.m:165-179, :255-259, :343-346).
//
// -v7.3 files are HDF5 containers, which need libhdf5 (not available to this
// build).  Level 5 is the documented binary MAT format that MATLAB's `load`
// reads as well (MathWorks, "MAT-File Format"): a 128-byte header, then one
// tagged data element per variable.  Files are written uncompressed (as
// `save -v6` does), to a temporary name that is flushed and renamed over the
// target, so a reader never sees a partial file.  The reader also accepts
// zlib-compressed elements (`save -v7`) and any numeric storage type of a
// double array.  A variable is limited to 2^31 bytes (the Level 5 limit).
#include <zlib.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <unistd.h>

#include "common.h"

namespace sssi {

namespace {

enum : uint32_t {
    miINT8 = 1, miUINT8 = 2, miINT16 = 3, miUINT16 = 4, miINT32 = 5, miUINT32 = 6,
    miSINGLE = 7, miDOUBLE = 9, miINT64 = 12, miUINT64 = 13, miMATRIX = 14, miCOMPRESSED = 15
};
constexpr uint32_t mxDOUBLE_CLASS = 6;

size_t pad8(size_t n) { return (n + 7) / 8 * 8; }

void put32(std::vector<uint8_t> &b, uint32_t v) {
    const size_t o = b.size();
    b.resize(o + 4);
    std::memcpy(b.data() + o, &v, 4);
}
void put_bytes(std::vector<uint8_t> &b, const void *p, size_t n) {
    const size_t o = b.size();
    b.resize(pad8(o + n));
    std::memcpy(b.data() + o, p, n);
}

// Everything of a miMATRIX element but the real part's payload.
std::vector<uint8_t> matrix_head(const char *name, int ndims, const int64_t *dims,
                                 size_t payload) {
    std::vector<uint8_t> s;
    put32(s, miUINT32);  // array flags: class double, real, not global/logical
    put32(s, 8);
    put32(s, mxDOUBLE_CLASS);
    put32(s, 0);
    put32(s, miINT32);   // dimensions (at least 2)
    const int nd = std::max(ndims, 2);
    put32(s, (uint32_t)(4 * nd));
    std::vector<int32_t> d(nd, 1);
    for (int k = 0; k < ndims; ++k) d[k] = (int32_t)dims[k];
    put_bytes(s, d.data(), 4 * (size_t)nd);
    const size_t nl = std::strlen(name);
    put32(s, miINT8);    // name
    put32(s, (uint32_t)nl);
    put_bytes(s, name, nl);
    put32(s, miDOUBLE);  // real part tag; the payload follows
    put32(s, (uint32_t)payload);
    return s;
}

struct Var {
    std::string name;
    std::vector<int64_t> dims;
    std::vector<double> data;
};

template <class T>
void widen(const uint8_t *p, size_t n, std::vector<double> &out) {
    out.resize(n);
    for (size_t q = 0; q < n; ++q) {
        T v;
        std::memcpy(&v, p + q * sizeof(T), sizeof(T));
        out[q] = (double)v;
    }
}

// One data element at p (n bytes available): its type, payload and total size.
bool element(const uint8_t *p, size_t n, uint32_t &type, const uint8_t *&data, size_t &bytes,
             size_t &total) {
    if (n < 8) return false;
    uint32_t t, b;
    std::memcpy(&t, p, 4);
    if (t >> 16) {  // small data element: the payload shares the 8-byte tag
        type = t & 0xffff;
        bytes = t >> 16;
        if (bytes > 4) return false;
        data = p + 4;
        total = 8;
        return true;
    }
    std::memcpy(&b, p + 4, 4);
    type = t;
    bytes = b;
    if (8 + (size_t)b > n) return false;
    data = p + 8;
    total = type == miCOMPRESSED ? 8 + b : 8 + pad8(b);
    if (total > n) total = n;
    return true;
}

// A miMATRIX element's payload -> Var (real double arrays; false otherwise).
bool parse_matrix(const uint8_t *p, size_t n, Var &v) {
    uint32_t type;
    const uint8_t *d;
    size_t bytes, total;
    // flags
    if (!element(p, n, type, d, bytes, total) || type != miUINT32 || bytes < 4) return false;
    uint32_t flags;
    std::memcpy(&flags, d, 4);
    if ((flags & 0xff) != mxDOUBLE_CLASS || (flags & 0x0800)) return false;  // double, real
    p += total; n -= total;
    // dims
    if (!element(p, n, type, d, bytes, total) || type != miINT32) return false;
    v.dims.resize(bytes / 4);
    size_t count = 1;
    for (size_t k = 0; k < v.dims.size(); ++k) {
        int32_t x;
        std::memcpy(&x, d + 4 * k, 4);
        v.dims[k] = x;
        count *= (size_t)std::max(x, 0);
    }
    p += total; n -= total;
    // name
    if (!element(p, n, type, d, bytes, total) || type != miINT8) return false;
    v.name.assign((const char *)d, bytes);
    p += total; n -= total;
    // real part, in any numeric storage type
    if (!element(p, n, type, d, bytes, total)) return false;
    size_t w = 0;
    switch (type) {
        case miDOUBLE: w = 8; break;
        case miSINGLE: case miINT32: case miUINT32: w = 4; break;
        case miINT16: case miUINT16: w = 2; break;
        case miINT8: case miUINT8: w = 1; break;
        case miINT64: case miUINT64: w = 8; break;
        default: return false;
    }
    if (bytes != count * w) return false;
    switch (type) {
        case miDOUBLE: widen<double>(d, count, v.data); break;
        case miSINGLE: widen<float>(d, count, v.data); break;
        case miINT32: widen<int32_t>(d, count, v.data); break;
        case miUINT32: widen<uint32_t>(d, count, v.data); break;
        case miINT16: widen<int16_t>(d, count, v.data); break;
        case miUINT16: widen<uint16_t>(d, count, v.data); break;
        case miINT8: widen<int8_t>(d, count, v.data); break;
        case miUINT8: widen<uint8_t>(d, count, v.data); break;
        case miINT64: widen<int64_t>(d, count, v.data); break;
        case miUINT64: widen<uint64_t>(d, count, v.data); break;
    }
    return true;
}

std::vector<uint8_t> read_file(const char *path) {
    FILE *f = std::fopen(path, "rb");
    if (!f) fail(SSSI_EINVAL, "cannot open %s", path);
    std::vector<uint8_t> b;
    uint8_t buf[1 << 16];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof(buf), f)) > 0) b.insert(b.end(), buf, buf + k);
    std::fclose(f);
    return b;
}

// The variable `name` of a Level 5 MAT-file.
bool find_var(const char *path, const char *name, Var &out) {
    std::vector<uint8_t> b = read_file(path);
    if (b.size() < 128) fail(SSSI_EINVAL, "%s is not a MAT-file", path);
    if (!(b[126] == 'I' && b[127] == 'M'))
        fail(SSSI_EINVAL, "%s is not a little-endian Level 5 MAT-file (an HDF5-based -v7.3 "
             "file needs libhdf5)", path);
    size_t o = 128;
    while (o + 8 <= b.size()) {
        uint32_t type;
        const uint8_t *d;
        size_t bytes, total;
        if (!element(b.data() + o, b.size() - o, type, d, bytes, total)) break;
        std::vector<uint8_t> inflated;
        if (type == miCOMPRESSED) {
            // the payload inflates to one element (normally a miMATRIX)
            z_stream zs;
            std::memset(&zs, 0, sizeof(zs));
            if (inflateInit(&zs) != Z_OK) fail(SSSI_EINVAL, "%s: zlib init failed", path);
            zs.next_in = const_cast<uint8_t *>(d);
            zs.avail_in = (uInt)bytes;
            std::vector<uint8_t> chunk(1 << 20);
            int rc;
            do {
                zs.next_out = chunk.data();
                zs.avail_out = (uInt)chunk.size();
                rc = inflate(&zs, Z_NO_FLUSH);
                if (rc != Z_OK && rc != Z_STREAM_END) {
                    inflateEnd(&zs);
                    fail(SSSI_EINVAL, "%s: corrupt compressed element", path);
                }
                inflated.insert(inflated.end(), chunk.data(),
                                chunk.data() + (chunk.size() - zs.avail_out));
            } while (rc != Z_STREAM_END);
            inflateEnd(&zs);
            uint32_t it;
            const uint8_t *id;
            size_t ib, itot;
            if (element(inflated.data(), inflated.size(), it, id, ib, itot) && it == miMATRIX) {
                Var v;
                if (parse_matrix(id, ib, v) && v.name == name) {
                    out = std::move(v);
                    return true;
                }
            }
        } else if (type == miMATRIX) {
            Var v;
            if (parse_matrix(d, bytes, v) && v.name == name) {
                out = std::move(v);
                return true;
            }
        }
        o += total;
    }
    return false;
}

}  // namespace

void mat_write(const char *path, int nvar, const char *const *names, const double *const *data,
               const int *ndims, const int64_t *dims) {
    if (!path || nvar < 1 || !names || !data || !ndims || !dims)
        fail(SSSI_EINVAL, "mat_write: bad arguments");
    const std::string tmp = std::string(path) + ".tmp" + std::to_string((long long)getpid());
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) fail(SSSI_EINVAL, "cannot create %s", tmp.c_str());
    bool ok = true;
    char head[128];
    std::memset(head, ' ', sizeof(head));
    const char text[] = "MATLAB 5.0 MAT-file, written by libsssi";
    std::memcpy(head, text, sizeof(text) - 1);
    std::memset(head + 116, 0, 8);          // no subsystem data
    const uint16_t ver = 0x0100;
    std::memcpy(head + 124, &ver, 2);
    head[126] = 'I';                          // 'MI' read as a little-endian int16
    head[127] = 'M';
    ok = ok && std::fwrite(head, 1, 128, f) == 128;
    const int64_t *dp = dims;
    for (int v = 0; v < nvar && ok; ++v) {
        if (!names[v] || !*names[v] || std::strlen(names[v]) > 63 || ndims[v] < 1 ||
            ndims[v] > 8 || (!data[v] && ndims[v] > 0)) {
            ok = false;
            break;
        }
        size_t count = 1;
        for (int k = 0; k < ndims[v]; ++k) {
            if (dp[k] < 0 || dp[k] > INT32_MAX) ok = false;
            count *= (size_t)std::max<int64_t>(dp[k], 0);
        }
        const size_t payload = count * 8;
        std::vector<uint8_t> sub = matrix_head(names[v], ndims[v], dp, payload);
        const size_t elem = sub.size() + pad8(payload);
        if (!ok || elem >= ((size_t)1 << 31)) {
            ok = false;
            break;
        }
        const uint32_t tag[2] = {miMATRIX, (uint32_t)elem};
        ok = ok && std::fwrite(tag, 4, 2, f) == 2;
        ok = ok && std::fwrite(sub.data(), 1, sub.size(), f) == sub.size();
        ok = ok && (payload == 0 || std::fwrite(data[v], 8, count, f) == count);
        const uint8_t zeros[8] = {0};
        const size_t padn = pad8(payload) - payload;
        ok = ok && (padn == 0 || std::fwrite(zeros, 1, padn, f) == padn);
        dp += ndims[v];
    }
    ok = ok && std::fflush(f) == 0 && fsync(fileno(f)) == 0;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        fail(SSSI_EINVAL, "cannot write %s (variable names 1..63 chars, 1..8 dims, under 2 GiB "
             "per variable)", path);
    }
}

bool mat_read(const char *path, const char *name, int *ndims, int64_t *dims, double *out,
              int64_t n) {
    if (!path || !name) fail(SSSI_EINVAL, "mat_read: path and name must not be NULL");
    Var v;
    if (!find_var(path, name, v)) return false;
    if (ndims) *ndims = (int)v.dims.size();
    if (dims)
        for (size_t k = 0; k < v.dims.size() && k < 8; ++k) dims[k] = v.dims[k];
    if (out) {
        if (n != (int64_t)v.data.size())
            fail(SSSI_EINVAL, "%s: variable %s has %zu values, not %lld", path, name,
                 v.data.size(), (long long)n);
        std::memcpy(out, v.data.data(), v.data.size() * sizeof(double));
    }
    return true;
}

bool file_exists(const char *path) { return access(path, F_OK) == 0; }

}  // namespace sssi

using namespace sssi;

extern "C" {

int sssi_mat_write(const char *path, int nvar, const char *const *names,
                   const double *const *data, const int *ndims, const int64_t *dims) {
    try {
        mat_write(path, nvar, names, data, ndims, dims);
        set_last_error("");
        return SSSI_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return SSSI_EINVAL;
    }
}

int sssi_mat_read(const char *path, const char *name, int *ndims, int64_t *dims, double *out,
                  int64_t n) {
    try {
        if (!mat_read(path, name, ndims, dims, out, n))
            fail(SSSI_EINVAL, "%s holds no real double variable %s", path ? path : "", name);
        set_last_error("");
        return SSSI_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return SSSI_EINVAL;
    }
}

}  // extern "C"
