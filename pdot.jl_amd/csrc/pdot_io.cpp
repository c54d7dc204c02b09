// pdot_io.cpp -- host-side text output of solution vectors in the format of
// write_vector_to_file (scripts/solve.jl:5-11): one element per line, in
// column-major order, each printed as Julia's println prints a Float64.
//
// Julia prints the shortest decimal string that reads back to the same
// double (Ryu).  With digits d1..dk and decimal exponent E (value
// d1.d2..dk x 10^E), the layout is positional when -5 < E < 6
// ("0.0001", "123456.0", "0.5") and otherwise "d1.d2..dke<E>" with at least
// one fraction digit ("1.0e-5", "1.234567e6", "5.0e-324"); zero is "0.0" or
// "-0.0", non-finite values "NaN", "Inf", "-Inf".  std::to_chars gives the
// same shortest digits (the closest such string, as Ryu does).

#include <algorithm>
#include <cerrno>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "pdot.h"

namespace {

// Writes the Julia text of x (no newline) at p; returns the end.  p must
// have room for 32 bytes.
char* julia_float(double x, char* p) {
    if (std::isnan(x)) {
        std::memcpy(p, "NaN", 3);
        return p + 3;
    }
    if (std::isinf(x)) {
        if (x < 0) *p++ = '-';
        std::memcpy(p, "Inf", 3);
        return p + 3;
    }
    if (std::signbit(x)) *p++ = '-';
    x = std::fabs(x);
    if (x == 0.0) {
        std::memcpy(p, "0.0", 3);
        return p + 3;
    }
    char sci[40];
    auto res = std::to_chars(sci, sci + sizeof(sci), x, std::chars_format::scientific);
    // sci = "d[.ddd]e[+-]XX"
    char digits[24];
    int k = 0;
    const char* q = sci;
    for (; q < res.ptr && *q != 'e'; ++q)
        if (*q != '.') digits[k++] = *q;
    int E = 0;
    std::from_chars(q[1] == '+' ? q + 2 : q + 1, res.ptr, E);
    while (k > 1 && digits[k - 1] == '0') --k;
    const int pt = E + 1;  // digits before the decimal point
    if (pt > -4 && pt <= 6) {
        if (pt <= 0) {
            *p++ = '0';
            *p++ = '.';
            for (int z = 0; z < -pt; ++z) *p++ = '0';
            std::memcpy(p, digits, k);
            p += k;
        } else if (pt < k) {
            std::memcpy(p, digits, pt);
            p += pt;
            *p++ = '.';
            std::memcpy(p, digits + pt, k - pt);
            p += k - pt;
        } else {
            std::memcpy(p, digits, k);
            p += k;
            for (int z = 0; z < pt - k; ++z) *p++ = '0';
            *p++ = '.';
            *p++ = '0';
        }
        return p;
    }
    *p++ = digits[0];
    *p++ = '.';
    if (k > 1) {
        std::memcpy(p, digits + 1, k - 1);
        p += k - 1;
    } else {
        *p++ = '0';
    }
    *p++ = 'e';
    auto r2 = std::to_chars(p, p + 8, E);
    return r2.ptr;
}

void set_err(char* errbuf, int64_t errlen, const std::string& msg) {
    if (!errbuf || errlen <= 0) return;
    const size_t k = std::min(msg.size(), (size_t)(errlen - 1));
    msg.copy(errbuf, k);
    errbuf[k] = '\0';
}

}  // namespace

extern "C" int pdot_format_julia_float(double x, char* out, int64_t outlen) {
    if (!out || outlen < 33) return PDOT_E_ARG;
    char* e = julia_float(x, out);
    *e = '\0';
    return (int)(e - out);
}

extern "C" int pdot_write_julia_vector(const char* path, const double* x, int64_t count, char* errbuf,
                                       int64_t errlen) {
    if (!path || (count > 0 && !x) || count < 0) {
        set_err(errbuf, errlen, "pdot_write_julia_vector: invalid arguments");
        return PDOT_E_ARG;
    }
    FILE* f = std::fopen(path, "wb");
    if (!f) {
        set_err(errbuf, errlen, std::string("pdot_write_julia_vector: cannot open ") + path + ": " + std::strerror(errno));
        return PDOT_E_ARG;
    }
    // Chunks are formatted in parallel and written in order.
    const int64_t chunk = 1 << 18;  // values per thread and round (about 8.6 MB of text)
    unsigned hw = std::thread::hardware_concurrency();
    const int nthreads = (int)std::max(1u, std::min(hw == 0 ? 1u : hw, 16u));
    std::vector<std::vector<char>> bufs(nthreads, std::vector<char>((size_t)chunk * 33));
    std::vector<size_t> lens(nthreads);
    int rc = PDOT_OK;
    for (int64_t base = 0; base < count && rc == PDOT_OK; base += chunk * nthreads) {
        std::vector<std::thread> pool;
        for (int t = 0; t < nthreads; ++t) {
            pool.emplace_back([&, t]() {
                const int64_t lo = base + t * chunk;
                const int64_t hi = std::min(count, lo + chunk);
                char* p = bufs[t].data();
                for (int64_t i = lo; i < hi; ++i) {
                    p = julia_float(x[i], p);
                    *p++ = '\n';
                }
                lens[t] = (size_t)(p - bufs[t].data());
            });
        }
        for (auto& th : pool) th.join();
        for (int t = 0; t < nthreads; ++t) {
            if (lens[t] && std::fwrite(bufs[t].data(), 1, lens[t], f) != lens[t]) {
                set_err(errbuf, errlen, std::string("pdot_write_julia_vector: write failed: ") + std::strerror(errno));
                rc = PDOT_E_ARG;
                break;
            }
        }
    }
    if (std::fclose(f) != 0 && rc == PDOT_OK) {
        set_err(errbuf, errlen, std::string("pdot_write_julia_vector: close failed: ") + std::strerror(errno));
        rc = PDOT_E_ARG;
    }
    return rc;
}
