// Restart records: the unformatted sequential file that the reference writes
// in src/write_restart_files.f90:22-47 and reads back in src/readfiles.f90:19-43.
//
// Every Fortran `write(3) list` is one record: a 4-byte little-endian length,
// the bytes of the list, and the same length again.  The records, in order:
//   (itime: int32, time: float64)
//   gradpcmf: float64                      -- only with constant mass flow
//   flmass (numInnerFaces), u, v, w, p, te, ed, t, vis (numTotal each),
//   uu, vv, ww, uv, uw, vw (numCells each), uo, vo, wo, teo, edo (numTotal each)
// The field records are plain float64 arrays whose lengths this code takes from
// the caller (writing) or from the file's markers (reading).  Records of 2 GiB
// or more would need the multi-marker form and are refused.

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fcap.h"

namespace {

struct RestartError {
    fc_status code;
    std::string msg;
};

constexpr long long kMaxRecord = 0x7fffffffLL;   // bytes in one record with 4-byte markers

void fail(fc_status code, const std::string &msg) { throw RestartError{code, msg}; }

struct File {
    std::FILE *f = nullptr;
    File(const char *path, const char *mode) : f(std::fopen(path, mode)) {}
    ~File() {
        if (f) std::fclose(f);
    }
};

void put_record(std::FILE *f, const void *data, long long bytes) {
    if (bytes < 0 || bytes > kMaxRecord) fail(FC_ERR_ARGUMENT, "record too large for 4-byte markers");
    const int32_t m = (int32_t)bytes;
    if (std::fwrite(&m, 4, 1, f) != 1 || (bytes && std::fwrite(data, 1, (size_t)bytes, f) != (size_t)bytes) ||
        std::fwrite(&m, 4, 1, f) != 1)
        fail(FC_ERR_ARGUMENT, "write failed");
}

// next record's length in bytes (the file is left at its data)
long long record_length(std::FILE *f) {
    int32_t m = 0;
    if (std::fread(&m, 4, 1, f) != 1) fail(FC_ERR_STRUCTURE, "unexpected end of file");
    if (m < 0) fail(FC_ERR_STRUCTURE, "multi-part record (2 GiB or more) is not supported");
    return m;
}

void get_record(std::FILE *f, void *data, long long bytes) {
    if (bytes && std::fread(data, 1, (size_t)bytes, f) != (size_t)bytes) fail(FC_ERR_STRUCTURE, "truncated record");
    int32_t m = 0;
    if (std::fread(&m, 4, 1, f) != 1 || m != (int32_t)bytes) fail(FC_ERR_STRUCTURE, "record markers disagree");
}

void skip_record(std::FILE *f, long long bytes) {
    if (std::fseek(f, (long)bytes, SEEK_CUR) != 0) fail(FC_ERR_STRUCTURE, "truncated record");
    int32_t m = 0;
    if (std::fread(&m, 4, 1, f) != 1 || m != (int32_t)bytes) fail(FC_ERR_STRUCTURE, "record markers disagree");
}

void read_header(std::FILE *f, int has_gradpcmf, int *itime, double *time, double *gradpcmf) {
    if (record_length(f) != 12) fail(FC_ERR_STRUCTURE, "first record is not (itime, time)");
    unsigned char buf[12];
    get_record(f, buf, 12);
    int32_t it;
    double t;
    std::memcpy(&it, buf, 4);
    std::memcpy(&t, buf + 4, 8);
    if (itime) *itime = it;
    if (time) *time = t;
    if (has_gradpcmf) {
        if (record_length(f) != 8) fail(FC_ERR_STRUCTURE, "gradpcmf record is not one float64");
        double g;
        get_record(f, &g, 8);
        if (gradpcmf) *gradpcmf = g;
    }
}

template <class F>
fc_status guarded(char *err, int errlen, F &&fn) {
    if (err && errlen > 0) err[0] = 0;
    try {
        fn();
    } catch (const RestartError &e) {
        if (err && errlen > 0) std::snprintf(err, (size_t)errlen, "%s", e.msg.c_str());
        return e.code;
    }
    return FC_OK;
}

}  // namespace

extern "C" fc_status fc_restart_write(const char *path, int itime, double time, int has_gradpcmf, double gradpcmf,
                                      int nrec, const long long *len, const double *const *rec, char *err,
                                      int errlen) {
    return guarded(err, errlen, [&]() {
        if (!path || nrec < 0 || (nrec > 0 && (!len || !rec))) fail(FC_ERR_ARGUMENT, "arguments");
        File out(path, "wb");
        if (!out.f) fail(FC_ERR_ARGUMENT, std::string("cannot open ") + path);
        unsigned char head[12];
        const int32_t it = itime;
        std::memcpy(head, &it, 4);
        std::memcpy(head + 4, &time, 8);
        put_record(out.f, head, 12);
        if (has_gradpcmf) put_record(out.f, &gradpcmf, 8);
        for (int k = 0; k < nrec; ++k) {
            if (len[k] < 0 || (len[k] > 0 && !rec[k])) fail(FC_ERR_ARGUMENT, "record " + std::to_string(k));
            put_record(out.f, rec[k], 8 * len[k]);
        }
        if (std::fflush(out.f) != 0) fail(FC_ERR_ARGUMENT, "write failed");
    });
}

extern "C" fc_status fc_restart_probe(const char *path, int has_gradpcmf, int *itime, double *time, double *gradpcmf,
                                      int max_rec, int *nrec, long long *len, char *err, int errlen) {
    return guarded(err, errlen, [&]() {
        if (!path || !nrec || max_rec < 0 || (max_rec > 0 && !len)) fail(FC_ERR_ARGUMENT, "arguments");
        File in(path, "rb");
        if (!in.f) fail(FC_ERR_ARGUMENT, std::string("cannot open ") + path);
        read_header(in.f, has_gradpcmf, itime, time, gradpcmf);
        int k = 0;
        for (;;) {
            int32_t m = 0;
            if (std::fread(&m, 4, 1, in.f) != 1) break;   // clean end of file
            if (m < 0) fail(FC_ERR_STRUCTURE, "multi-part record (2 GiB or more) is not supported");
            if (m % 8) fail(FC_ERR_STRUCTURE, "field record " + std::to_string(k) + " is not float64 data");
            if (k < max_rec) len[k] = m / 8;
            skip_record(in.f, m);
            ++k;
        }
        *nrec = k;
    });
}

extern "C" fc_status fc_restart_read(const char *path, int has_gradpcmf, int nrec, const long long *len,
                                     double *const *rec, char *err, int errlen) {
    return guarded(err, errlen, [&]() {
        if (!path || nrec < 0 || (nrec > 0 && (!len || !rec))) fail(FC_ERR_ARGUMENT, "arguments");
        File in(path, "rb");
        if (!in.f) fail(FC_ERR_ARGUMENT, std::string("cannot open ") + path);
        read_header(in.f, has_gradpcmf, nullptr, nullptr, nullptr);
        for (int k = 0; k < nrec; ++k) {
            const long long bytes = record_length(in.f);
            if (bytes != 8 * len[k])
                fail(FC_ERR_STRUCTURE, "record " + std::to_string(k) + " holds " + std::to_string(bytes / 8) +
                                           " values, expected " + std::to_string(len[k]));
            if (rec[k]) get_record(in.f, rec[k], bytes);
            else skip_record(in.f, bytes);
        }
    });
}
