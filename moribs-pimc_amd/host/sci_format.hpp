// Text of a double as printf's "%14.6e" writes it, for the 1.5-million-line 3D density file that
// SaveDensities3D (mc_estim.cc:1728-1820) rewrites after every block: about ten times faster
// than snprintf and the same bytes.
//
// A finite nonzero x = m * 10^(e-6) with seven significant digits m in [10^6, 10^7): the scaled
// value y = |x| * 10^(6-e) is formed with one correctly rounded multiplication or division by an
// exact power of ten (|6-e| <= 22), so y carries a relative error of at most 2^-53, under 1.2e-9
// absolute below 10^7.  Rounding y to the nearest integer therefore gives printf's digits unless
// y lies within that distance of a half-integer; values within 1e-6 of one, and zeros' signs,
// non-finite values and exponents outside the table take snprintf itself.
#pragma once

#include <cmath>
#include <cstdio>

namespace sci_format {

inline int snprintf_e14_6(double x, char* out) { return std::snprintf(out, 32, "%14.6e", x); }

// Writes 14 characters (no terminator) and returns 14, or, through snprintf, what that writes
// (14 characters for every finite x with a two-digit exponent; more for a three-digit one).
inline int e14_6(double x, char* out) {
    static const double p10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                                   1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
    const double a = std::fabs(x);
    if (!(a >= 1e-16 && a < 1e16)) {  // zero, tiny, huge, inf, nan
        if (x == 0.0 && !std::signbit(x)) {
            const char z[] = "  0.000000e+00";
            for (int i = 0; i < 14; ++i) out[i] = z[i];
            return 14;
        }
        return snprintf_e14_6(x, out);
    }
    int e = (int)std::floor(std::log10(a));
    double y = 0.0;
    for (int pass = 0; pass < 3; ++pass) {  // log10 may be off by one at a power of ten
        const int k = 6 - e;
        if (k > 22 || k < -22) return snprintf_e14_6(x, out);
        y = k >= 0 ? a * p10[k] : a / p10[-k];
        if (y < 1e6) {
            --e;
        } else if (y >= 1e7) {
            ++e;
        } else {
            break;
        }
    }
    if (!(y >= 1e6 && y < 1e7)) return snprintf_e14_6(x, out);
    const double fl = std::floor(y);
    const double f = y - fl;
    if (std::fabs(f - 0.5) < 1e-6) return snprintf_e14_6(x, out);  // too near a tie to decide here
    long m = (long)fl + (f > 0.5 ? 1 : 0);
    if (m == 10000000) {  // 9.9999995... rounds up to the next decade
        m = 1000000;
        ++e;
    }
    if (e > 99 || e < -99) return snprintf_e14_6(x, out);
    char* p = out;
    *p++ = ' ';
    *p++ = x < 0 ? '-' : ' ';
    char dig[7];
    for (int i = 6; i >= 0; --i) {
        dig[i] = (char)('0' + m % 10);
        m /= 10;
    }
    *p++ = dig[0];
    *p++ = '.';
    for (int i = 1; i < 7; ++i) *p++ = dig[i];
    *p++ = 'e';
    *p++ = e < 0 ? '-' : '+';
    const int ae = e < 0 ? -e : e;
    *p++ = (char)('0' + ae / 10);
    *p++ = (char)('0' + ae % 10);
    return 14;
}

}  // namespace sci_format
