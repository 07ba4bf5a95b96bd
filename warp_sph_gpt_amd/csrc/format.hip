// format.hip -- host-side ASCII formatting of particle fields for the VTK
// XML writer (output/output.py:14-24 writes "%.6g" floats six per line and
// "%d" integers twenty per line, joined by single spaces and newlines).
// C's %.6g and %d give the same text as Python's for finite values, which
// are the only ones the frame writer lets through.
#include "common.hpp"

#include <cstdio>
#include <cstring>

extern "C" {

// Format n doubles as "%.6g", per_line values on a line (space separated,
// lines joined by '\n', no trailing newline) into out[cap].  Returns the
// number of bytes written (not counting the terminating NUL), or -1 if cap
// is too small.
long sphdem_format_g6(const double *v, long n, int per_line, char *out, long cap) {
    if (!out || cap < 1 || per_line < 1 || n < 0 || (n > 0 && !v)) return -1;
    long pos = 0;
    char tmp[64];
    for (long i = 0; i < n; ++i) {
        const int len = std::snprintf(tmp, sizeof(tmp), "%.6g", v[i]);
        const char sep = (i + 1 == n) ? '\0' : ((i + 1) % per_line == 0 ? '\n' : ' ');
        if (pos + len + 2 > cap) return -1;
        std::memcpy(out + pos, tmp, (size_t)len);
        pos += len;
        if (sep) out[pos++] = sep;
    }
    out[pos] = '\0';
    return pos;
}

// The same for int32 values with "%d".
long sphdem_format_int(const int *v, long n, int per_line, char *out, long cap) {
    if (!out || cap < 1 || per_line < 1 || n < 0 || (n > 0 && !v)) return -1;
    long pos = 0;
    char tmp[32];
    for (long i = 0; i < n; ++i) {
        const int len = std::snprintf(tmp, sizeof(tmp), "%d", v[i]);
        const char sep = (i + 1 == n) ? '\0' : ((i + 1) % per_line == 0 ? '\n' : ' ');
        if (pos + len + 2 > cap) return -1;
        std::memcpy(out + pos, tmp, (size_t)len);
        pos += len;
        if (sep) out[pos++] = sep;
    }
    out[pos] = '\0';
    return pos;
}

}  // extern "C"
