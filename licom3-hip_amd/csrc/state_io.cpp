// state_io.cpp -- host-side state I/O of the LICOM3 tracer engine: the
// restart stream (inirun.F90:598-700 reads it, ssave-cdf-instant.F90:100-830
// writes it) and the EOS coefficient table dncoef.h1 (constant_mod.F90:331-364).
//
// The restart is a Fortran sequential unformatted stream: each record is a
// 4-byte length, the payload, and the length again.  The reference compiles
// every Fortran unit with -convert big_endian (Makefile:4), so its record
// lengths, doubles and integers are big-endian; that is the default here,
// and licom_restart.little_endian selects the native order of a build
// without that flag.  Records,
// in order (inirun.F90:612-686): h0; u, v, T, S, ws one record per level
// (km each); su, sv, swv, sshf, lthf, fresh; then (number_month, number_day)
// as two default integers.  Every 2-D record is the global grid nx x ny in
// double precision, i fastest, j in LICOM's north-to-south order.  The
// reference writes these records into fort.22 (h0, u), fort.23 (v),
// fort.24 (T), fort.25 (S), fort.26 (ws) and fort.27 (surface fields and the
// date); their concatenation is the stream the reader takes.
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "licom_tracer.h"
#include "tracer_internal.h"

namespace licom {
namespace {

struct File {
    FILE *f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

// Byte order of one stream: big-endian unless the caller asks for native
// little-endian (the host is little-endian x86-64).
struct Order {
    bool swap;
    int32_t i32(int32_t v) const { return swap ? (int32_t)__builtin_bswap32((uint32_t)v) : v; }
    void doubles(double *p, size_t n) const {
        if (!swap) return;
        for (size_t q = 0; q < n; ++q) {
            uint64_t u;
            std::memcpy(&u, p + q, 8);
            u = __builtin_bswap64(u);
            std::memcpy(p + q, &u, 8);
        }
    }
    void ints(int32_t *p, size_t n) const {
        for (size_t q = 0; q < n; ++q) p[q] = i32(p[q]);
    }
};

int io_error(const char *what, const char *path) {
    set_error("%s: %s (%s)", what, path, errno ? std::strerror(errno) : "format");
    return LICOM_EINVAL;
}

// One record: payload of `bytes` bytes into dst (NULL: skipped), in the
// file's byte order (the caller converts the payload).
int read_record(FILE *f, const char *path, void *dst, size_t bytes, int index, const Order &o) {
    int32_t head = 0, tail = 0;
    errno = 0;
    if (std::fread(&head, 4, 1, f) != 1) {
        set_error("restart %s: record %d missing (file ends early)", path, index);
        return LICOM_EINVAL;
    }
    head = o.i32(head);
    if (head < 0 || (size_t)head != bytes) {
        set_error("restart %s: record %d holds %d bytes, expected %zu", path, index, (int)head, bytes);
        return LICOM_EINVAL;
    }
    if (dst) {
        if (std::fread(dst, 1, bytes, f) != bytes) return io_error("restart: short record", path);
    } else if (std::fseek(f, (long)bytes, SEEK_CUR) != 0) {
        return io_error("restart: seek", path);
    }
    if (std::fread(&tail, 4, 1, f) != 1 || o.i32(tail) != head) {
        set_error("restart %s: record %d has a bad trailing length", path, index);
        return LICOM_EINVAL;
    }
    return LICOM_OK;
}

// src is already in the file's byte order
int write_record(FILE *f, const char *path, const void *src, size_t bytes, const Order &o) {
    const int32_t n = o.i32((int32_t)bytes);
    if (std::fwrite(&n, 4, 1, f) != 1 || std::fwrite(src, 1, bytes, f) != bytes ||
        std::fwrite(&n, 4, 1, f) != 1)
        return io_error("restart: write", path);
    return LICOM_OK;
}

bool bad_dims(int nx, int ny, int km) {
    return nx <= 0 || ny <= 0 || km <= 0 || km > LICOM_MAX_KM ||
           (size_t)nx * (size_t)ny * 8 > (size_t)INT32_MAX;
}

// ---- list-directed reading of dncoef.h1 (Fortran READ(u,*)) ----
// A READ with an empty list skips one line; a READ of n values takes the
// next n values, continuing over lines, and the following READ starts on a
// new line.  Values are separated by blanks or commas; a Fortran exponent
// letter D or d is read as E; r*c repeats c r times.
struct ListReader {
    FILE *f;
    std::string line;
    size_t pos = 0;
    bool have = false;
    bool next_line() {
        line.clear();
        pos = 0;
        int ch;
        bool any = false;
        while ((ch = std::fgetc(f)) != EOF) {
            any = true;
            if (ch == '\n') break;
            line.push_back((char)ch);
        }
        have = any;
        return any;
    }
    bool skip_record() { return next_line(); }
    // n values into out; false when the file ends or a value does not parse
    bool values(int n, double *out) {
        int got = 0;
        if (!next_line()) return false;
        while (got < n) {
            while (pos < line.size() && (line[pos] == ' ' || line[pos] == '\t' || line[pos] == ',' ||
                                         line[pos] == '\r'))
                ++pos;
            if (pos >= line.size()) {
                if (!next_line()) return false;
                continue;
            }
            size_t end = pos;
            while (end < line.size() && line[end] != ' ' && line[end] != '\t' && line[end] != ',' &&
                   line[end] != '\r')
                ++end;
            std::string tok = line.substr(pos, end - pos);
            pos = end;
            long rep = 1;
            const size_t star = tok.find('*');
            if (star != std::string::npos) {
                char *e = nullptr;
                rep = std::strtol(tok.c_str(), &e, 10);
                if (e != tok.c_str() + star || rep <= 0) return false;
                tok = tok.substr(star + 1);
            }
            for (char &c : tok)
                if (c == 'D' || c == 'd') c = 'E';
            char *e = nullptr;
            const double v = std::strtod(tok.c_str(), &e);
            if (tok.empty() || *e != '\0') return false;
            for (long r = 0; r < rep && got < n; ++r) out[got++] = v;
        }
        return true;
    }
};

}  // namespace
}  // namespace licom

extern "C" {

int licom_restart_read(const char *path, int nx, int ny, int km, licom_restart *r) {
    using namespace licom;
    if (!path || !r || bad_dims(nx, ny, km)) {
        set_error("licom_restart_read: bad argument");
        return LICOM_EINVAL;
    }
    File F;
    errno = 0;
    F.f = std::fopen(path, "rb");
    if (!F.f) return io_error("licom_restart_read: cannot open", path);
    const size_t n2 = (size_t)nx * ny, b2 = n2 * 8;
    const Order o{r->little_endian == 0};
    int idx = 0;
    auto rec2 = [&](double *dst) {
        const int rc = read_record(F.f, path, dst, b2, idx++, o);
        if (rc == LICOM_OK && dst) o.doubles(dst, n2);
        return rc;
    };
    auto rec3 = [&](double *dst) {
        for (int k = 0; k < km; ++k)
            if (int rc = rec2(dst ? dst + (size_t)k * n2 : nullptr)) return rc;
        return (int)LICOM_OK;
    };
    if (int rc = rec2(r->h0)) return rc;
    if (int rc = rec3(r->u)) return rc;
    if (int rc = rec3(r->v)) return rc;
    if (int rc = rec3(r->at)) return rc;
    if (int rc = rec3(r->at ? r->at + (size_t)km * n2 : nullptr)) return rc;
    if (int rc = rec3(r->ws)) return rc;
    double *surf[6] = {r->su, r->sv, r->swv, r->sshf, r->lthf, r->fresh};
    for (double *p : surf)
        if (int rc = rec2(p)) return rc;
    int32_t date[2] = {0, 0};
    if (int rc = read_record(F.f, path, date, sizeof date, idx++, o)) return rc;
    o.ints(date, 2);
    r->number_month = date[0];
    r->number_day = date[1];
    return LICOM_OK;
}

int licom_restart_write(const char *path, int nx, int ny, int km, const int32_t *kmt,
                        const int32_t *kmu, const licom_restart *r) {
    using namespace licom;
    if (!path || !r || !kmt || !kmu || bad_dims(nx, ny, km) || !r->h0 || !r->u || !r->v || !r->at ||
        !r->ws || !r->su || !r->sv || !r->swv || !r->sshf || !r->lthf || !r->fresh) {
        set_error("licom_restart_write: bad argument (every field is required)");
        return LICOM_EINVAL;
    }
    File F;
    errno = 0;
    F.f = std::fopen(path, "wb");
    if (!F.f) return io_error("licom_restart_write: cannot open", path);
    const size_t n2 = (size_t)nx * ny, b2 = n2 * 8;
    const Order o{r->little_endian == 0};
    std::vector<double> buf(n2);
    // land is written as 0: h0 and the T-grid fields by vit, u and v by viv
    // (ssave-cdf-instant.F90:173-191); the caller's arrays are not modified
    auto masked = [&](const double *src, const int32_t *kz, int k) {
        for (size_t q = 0; q < n2; ++q) buf[q] = k < kz[q] ? src[q] : 0.0;
        o.doubles(buf.data(), n2);
        return write_record(F.f, path, buf.data(), b2, o);
    };
    if (int rc = masked(r->h0, kmt, 0)) return rc;
    const double *f3[5] = {r->u, r->v, r->at, r->at + (size_t)km * n2, r->ws};
    const int32_t *m3[5] = {kmu, kmu, kmt, kmt, kmt};
    for (int a = 0; a < 5; ++a)
        for (int k = 0; k < km; ++k)
            if (int rc = masked(f3[a] + (size_t)k * n2, m3[a], k)) return rc;
    const double *surf[6] = {r->su, r->sv, r->swv, r->sshf, r->lthf, r->fresh};
    for (const double *p : surf) {
        std::memcpy(buf.data(), p, b2);
        o.doubles(buf.data(), n2);
        if (int rc = write_record(F.f, path, buf.data(), b2, o)) return rc;
    }
    int32_t date[2] = {r->number_month, r->number_day};
    o.ints(date, 2);
    if (int rc = write_record(F.f, path, date, sizeof date, o)) return rc;
    if (std::fflush(F.f) != 0) return io_error("licom_restart_write: flush", path);
    return LICOM_OK;
}

int licom_eos_read(const char *path, int km, double *to, double *so, double *po, double *eosc) {
    using namespace licom;
    if (!path || km <= 0 || km > LICOM_MAX_KM || !to || !so || !eosc) {
        set_error("licom_eos_read: bad argument");
        return LICOM_EINVAL;
    }
    File F;
    errno = 0;
    F.f = std::fopen(path, "r");
    if (!F.f) return io_error("licom_eos_read: cannot open", path);
    ListReader L{F.f};
    std::vector<double> pbuf(km), row(9);
    auto fail = [&](const char *what) {
        set_error("licom_eos_read %s: %s", path, what);
        return (int)LICOM_EINVAL;
    };
    if (!L.skip_record() || !L.values(km, to)) return fail("to");
    if (!L.skip_record() || !L.values(km, so)) return fail("so");
    if (!L.skip_record() || !L.values(km, pbuf.data())) return fail("po");
    if (po) std::memcpy(po, pbuf.data(), km * sizeof(double));
    for (int k = 0; k < km; ++k) {
        if (!L.skip_record() || !L.values(9, row.data())) return fail("C(k,1:9)");
        // C(K,M) = C(K,M) * 1000 (constant_mod.F90:357-363): g/cm^3 -> kg/m^3;
        // stored as eosc[m][k], the Fortran C(km,9) layout
        for (int m = 0; m < 9; ++m) eosc[(size_t)m * km + k] = row[m] * 1000.0;
    }
    return LICOM_OK;
}

}  // extern "C"
