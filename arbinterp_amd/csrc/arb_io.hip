// Field ingestion: numeric CSV files as the reference's examples read them,
// np.genfromtxt(path, delimiter=',') (examples/3D_ARBInterpExample.py:14,
// examples/4D_ARBInterpExample.py:14), parsed on host threads.
//
// Grammar (the files the examples load, and what genfromtxt makes of them):
//   * lines end at '\n' ('\r' before it is dropped); text from '#' on is a
//     comment; lines that are blank after that are skipped;
//   * fields are separated by ','; spaces and tabs around a field are
//     ignored; an empty field is NaN;
//   * a field is a decimal floating-point number (C strtod forms without hex)
//     or nan / inf / infinity in any case, with an optional sign; anything
//     else is NaN;
//   * every kept line must have the column count of the first one (else
//     ARB_ESHAPE, as genfromtxt raises).
// Host code only: no GPU is needed.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "arb_internal.h"

namespace {

struct Line {
    const char* b;
    const char* e;
};

int read_all(const char* path, std::string& buf) {
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return ARB_EINVAL;
    std::fseek(fp, 0, SEEK_END);
    const long sz = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    if (sz < 0) {
        std::fclose(fp);
        return ARB_EINVAL;
    }
    buf.resize((size_t)sz);
    const size_t got = sz ? std::fread(&buf[0], 1, (size_t)sz, fp) : 0;
    std::fclose(fp);
    return got == (size_t)sz ? ARB_OK : ARB_EINVAL;
}

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

// Kept lines (comment removed, trimmed, non-empty).
void split_lines(const std::string& buf, std::vector<Line>& lines) {
    const char* p = buf.data();
    const char* end = p + buf.size();
    while (p < end) {
        const char* nl = static_cast<const char*>(std::memchr(p, '\n', (size_t)(end - p)));
        const char* le = nl ? nl : end;
        const char* hash = static_cast<const char*>(std::memchr(p, '#', (size_t)(le - p)));
        const char* e = hash ? hash : le;
        const char* b = p;
        while (b < e && is_space(*b)) ++b;
        while (e > b && is_space(e[-1])) --e;
        if (e > b) lines.push_back({b, e});
        p = nl ? nl + 1 : end;
    }
}

int count_fields(const Line& l) {
    int n = 1;
    for (const char* c = l.b; c < l.e; ++c) n += (*c == ',');
    return n;
}

bool ieq(const char* a, size_t n, const char* lit) {
    if (std::strlen(lit) != n) return false;
    for (size_t i = 0; i < n; ++i)
        if (std::tolower((unsigned char)a[i]) != lit[i]) return false;
    return true;
}

double parse_field(const char* b, const char* e) {
    while (b < e && is_space(*b)) ++b;
    while (e > b && is_space(e[-1])) --e;
    if (b == e) return NAN;
    char tmp[64];
    const size_t n = (size_t)(e - b);
    if (n >= sizeof tmp) return NAN;
    std::memcpy(tmp, b, n);
    tmp[n] = '\0';
    const char* s = tmp;
    double sign = 1.0;
    if (*s == '+' || *s == '-') {
        if (*s == '-') sign = -1.0;
        ++s;
    }
    const size_t m = std::strlen(s);
    if (ieq(s, m, "nan")) return NAN;
    if (ieq(s, m, "inf") || ieq(s, m, "infinity")) return sign * INFINITY;
    for (const char* c = s; *c; ++c)
        if (!(std::isdigit((unsigned char)*c) || *c == '.' || *c == 'e' || *c == 'E' || *c == '+' ||
              *c == '-'))
            return NAN;                               // hex, words, stray characters
    char* stop = nullptr;
    const double v = std::strtod(tmp, &stop);
    return (stop == tmp + n) ? v : NAN;
}

int parse_lines(const std::vector<Line>& lines, double* out, int32_t ncols) {
    const int64_t nl = (int64_t)lines.size();
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nt = (int)std::min<int64_t>(std::min<unsigned>(hw, 16u), std::max<int64_t>(1, nl / 4096));
    std::vector<int> bad(nt, 0);
    auto work = [&](int t) {
        const int64_t r0 = nl * t / nt, r1 = nl * (t + 1) / nt;
        for (int64_t r = r0; r < r1; ++r) {
            const Line& l = lines[(size_t)r];
            const char* f = l.b;
            int c = 0;
            for (const char* p = l.b; p <= l.e; ++p) {
                if (p == l.e || *p == ',') {
                    if (c < ncols) out[r * ncols + c] = parse_field(f, p);
                    ++c;
                    f = p + 1;
                }
            }
            if (c != ncols) bad[t] = 1;
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(work, t);
    work(0);
    for (auto& x : th) x.join();
    for (int b : bad)
        if (b) return ARB_ESHAPE;
    return ARB_OK;
}

}  // namespace

int arb_csv_shape(const char* path, int64_t* nrows, int32_t* ncols) {
    if (!path || !nrows || !ncols) return ARB_EINVAL;
    std::string buf;
    if (read_all(path, buf) != ARB_OK) return ARB_EINVAL;
    std::vector<Line> lines;
    split_lines(buf, lines);
    *nrows = (int64_t)lines.size();
    *ncols = lines.empty() ? 0 : count_fields(lines[0]);
    return ARB_OK;
}

int arb_csv_read(const char* path, double* out, int64_t nrows, int32_t ncols) {
    if (!path || (!out && nrows > 0) || nrows < 0 || ncols < 0) return ARB_EINVAL;
    std::string buf;
    if (read_all(path, buf) != ARB_OK) return ARB_EINVAL;
    std::vector<Line> lines;
    split_lines(buf, lines);
    if ((int64_t)lines.size() != nrows) return ARB_ESHAPE;
    return parse_lines(lines, out, ncols);
}
