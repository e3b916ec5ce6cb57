// Number formats of the reference's writers and a threaded row formatter.
// Built without FMA contraction: frd_e12 follows frd.rs's arithmetic.

#include "axg_writers.h"

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

namespace axg {

// 10^e the way f64::powi computes it on this target: LLVM lowers powi with a
// variable exponent to the runtime's repeated squaring (binary
// exponentiation, then one reciprocal for a negative exponent).
static double powi_rt(double a, int b) {
    const bool recip = b < 0;
    double r = 1.0;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return recip ? 1.0 / r : r;
}

void frd_e12(double v, char out[12]) {
    if (!std::isfinite(v) || v == 0.0) {
        std::memcpy(out, " 0.00000E+00", 12);
        return;
    }
    const char sign = std::signbit(v) ? '-' : ' ';
    const double av = std::fabs(v);
    int ex = (int)std::floor(std::log10(av));
    double mant = av / powi_rt(10.0, ex);
    if (mant >= 10.0) {
        mant /= 10.0;
        ex += 1;
    } else if (mant < 1.0) {
        mant *= 10.0;
        ex -= 1;
    }
    double scaled = std::round(mant * 1e5);  // halfway cases away from zero, as f64::round
    if (scaled >= 1e6) {
        scaled /= 10.0;
        ex += 1;
    }
    // "%c%.5fE%+03d" of (sign, scaled / 1e5, ex), cut to 12 characters: scaled is
    // an integer in [1e5, 1e6), so its six digits are the printed ones (the
    // division's rounding error is far below the fifth decimal)
    if (!(scaled < 1e6)) {
        // a subnormal v: powi's 10^ex overflowed, the mantissa is infinite and
        // the reference prints it as such ("    infE-307")
        char buf[64];
        const int len = std::snprintf(buf, sizeof buf, "%c%.5fE%+03d", sign, scaled / 1e5, ex);
        const int pad = len < 12 ? 12 - len : 0;
        std::memset(out, ' ', (size_t)pad);
        std::memcpy(out + pad, buf, (size_t)(12 - pad));
        return;
    }
    int64_t k = (int64_t)scaled;
    out[0] = sign;
    for (int i = 7; i >= 3; i--) out[i] = (char)('0' + k % 10), k /= 10;
    out[1] = (char)('0' + k);
    out[2] = '.';
    out[8] = 'E';
    out[9] = ex < 0 ? '-' : '+';
    const int ae = ex < 0 ? -ex : ex;
    if (ae < 100) {
        out[10] = (char)('0' + ae / 10);
        out[11] = (char)('0' + ae % 10);
    } else {  // a three-digit exponent loses its last digit, as the reference's cut
        out[10] = (char)('0' + ae / 100);
        out[11] = (char)('0' + ae / 10 % 10);
    }
}

void rust_exp(std::string &s, double v, int prec, int width) {
    char body[64];
    int len;
    if (std::isnan(v)) {
        len = std::snprintf(body, sizeof body, "NaN");
    } else if (std::isinf(v)) {
        len = std::snprintf(body, sizeof body, v > 0 ? "inf" : "-inf");
    } else {
        // to_chars with a precision rounds exactly, as printf's %.*e; Rust prints
        // the exponent as a plain integer ("1.500000E-3", "1.000000E0")
        char c[64];
        const auto r = std::to_chars(c, c + sizeof c, v, std::chars_format::scientific, prec);
        char *e = std::find(c, r.ptr, 'e');
        const bool neg = e[1] == '-';
        int ev = 0;
        std::from_chars(e + 2, r.ptr, ev);
        len = (int)(e - c);
        std::memcpy(body, c, (size_t)len);
        body[len++] = 'E';
        if (neg && ev != 0) body[len++] = '-';
        len = (int)(std::to_chars(body + len, body + sizeof body, ev).ptr - body);
    }
    if (len < width) s.append((size_t)(width - len), ' ');
    s.append(body, (size_t)len);
}

void rust_int(std::string &s, int64_t v, int width) {
    char b[32];
    const int len = (int)(std::to_chars(b, b + sizeof b, v).ptr - b);
    if (len < width) s.append((size_t)(width - len), ' ');
    s.append(b, (size_t)len);
}

void format_rows(std::string &o, int64_t n,
                 const std::function<void(int64_t, std::string &)> &row, int threads) {
    if (threads <= 0) threads = (int)std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
    const int64_t min_chunk = 4096;
    const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(threads, n / min_chunk));
    if (nt == 1) {
        for (int64_t i = 0; i < n; i++) row(i, o);
        return;
    }
    std::vector<std::string> part(nt);
    auto work = [&](int t) {
        const int64_t a = n * t / nt, b = n * (t + 1) / nt;
        std::string &p = part[t];
        p.reserve((size_t)(b - a) * 96);
        for (int64_t i = a; i < b; i++) row(i, p);
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; t++) pool.emplace_back(work, t);
    work(0);
    for (auto &th : pool) th.join();
    size_t total = o.size();
    for (auto &p : part) total += p.size();
    o.reserve(total);
    for (auto &p : part) o += p;
}

}  // namespace axg
