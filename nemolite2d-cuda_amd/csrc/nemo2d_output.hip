// Snapshot files in the serial model's format (original/nemolite2d.f90:773-799):
// per T point of each row, 'x, y, depth, ssh, u, v' written (6e20.7e4),
// u and v the averages 0.5*(un(i-1,j) + un(i,j)), 0.5*(vn(i,j-1) + vn(i,j)),
// and an empty record after each row.  Host code only; rows are formatted
// by several threads into a buffer whose layout is fixed (121 bytes a line).
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "nemo2d.h"
#include "nemo2d_internal.h"

namespace nemo2d {

namespace {

constexpr int FW = 20;             // field width of e20.7e4
constexpr int LINE = 6 * FW + 1;   // one point's record with its newline

// x in Fortran's E20.7E4 form: right-justified '0.ddddddd' mantissa and a
// signed 4-digit exponent, with the processor's rounding of the binary value
// to 7 significant digits (round half to even, as the C library's %e),
// signed zero printed with its sign and an exponent of 0.
void fmt_e20_7e4(double x, char* out) {
    char t[40];
    if (!isfinite(x)) {
        const char* s = isnan(x) ? "NaN" : (x > 0 ? "Infinity" : "-Infinity");
        const int n = (int)strlen(s);
        memset(out, ' ', FW);
        memcpy(out + FW - n, s, n);
        return;
    }
    const bool neg = signbit(x);
    int e10 = 0;
    char dig[7];
    if (x == 0.0) {
        memset(dig, '0', 7);
    } else {
        // d.dddddde+XX: 7 significant digits
        snprintf(t, sizeof t, "%.6e", fabs(x));
        dig[0] = t[0];
        memcpy(dig + 1, t + 2, 6);
        e10 = atoi(t + 9) + 1;  // 0.ddddddd x 10^(e10)
    }
    char* p = out;
    memset(p, ' ', FW);
    // '-0.ddddddd' or ' 0.ddddddd' then 'E+eeee' at the end of the field
    char* q = out + FW - 15;
    if (neg) q[-1] = '-';
    q[0] = '0';
    q[1] = '.';
    memcpy(q + 2, dig, 7);
    q[9] = 'E';
    q[10] = e10 < 0 ? '-' : '+';
    int a = e10 < 0 ? -e10 : e10;
    for (int k = 14; k >= 11; --k) {
        q[k] = (char)('0' + a % 10);
        a /= 10;
    }
}

}  // namespace

long go2d_bytes(int jpi, int jpj) { return (long)jpj * ((long)jpi * LINE + 1); }

void format_go2d_rows(const Arr2& xt, const Arr2& yt, const Arr2& ht, const Arr2& ss, const Arr2& un,
                      const Arr2& vn, int jpi, int j0, int j1, char* buf) {
    for (int j = j0; j < j1; ++j) {
        char* p = buf + (long)(j - j0) * ((long)jpi * LINE + 1);
        for (int i = 1; i <= jpi; ++i) {
            const double u = 0.5 * (un.at(i - 1, j) + un.at(i, j));
            const double v = 0.5 * (vn.at(i, j - 1) + vn.at(i, j));
            fmt_e20_7e4(xt.at(i, j), p);
            fmt_e20_7e4(yt.at(i, j), p + FW);
            fmt_e20_7e4(ht.at(i, j), p + 2 * FW);
            fmt_e20_7e4(ss.at(i, j), p + 3 * FW);
            fmt_e20_7e4(u, p + 4 * FW);
            fmt_e20_7e4(v, p + 5 * FW);
            p[6 * FW] = '\n';
            p += LINE;
        }
        *p = '\n';
    }
}

int write_go2d(const char* path, const Arr2& xt, const Arr2& yt, const Arr2& ht, const Arr2& ss,
               const Arr2& un, const Arr2& vn, int jpi, int jpj, std::string& err) {
    FILE* f = fopen(path, "wb");
    if (!f) {
        err = std::string("cannot open ") + path;
        return NEMO2D_ERR_ARG;
    }
    const long row = (long)jpi * LINE + 1;
    const int chunk = (int)std::max(1L, std::min((long)jpj, (64L << 20) / row));  // ~64 MB
    const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<char> buf((size_t)chunk * row);
    int rc = NEMO2D_OK;
    for (int j0 = 1; j0 <= jpj && rc == NEMO2D_OK; j0 += chunk) {
        const int j1 = std::min(jpj + 1, j0 + chunk);
        const int n = j1 - j0;
        const int per = (n + nth - 1) / nth;
        std::vector<std::thread> th;
        for (int t = 0; t < nth; ++t) {
            const int a = j0 + t * per, b = std::min(j1, a + per);
            if (a >= b) break;
            th.emplace_back([&, a, b] {
                format_go2d_rows(xt, yt, ht, ss, un, vn, jpi, a, b, buf.data() + (long)(a - j0) * row);
            });
        }
        for (auto& x : th) x.join();
        if (fwrite(buf.data(), 1, (size_t)n * row, f) != (size_t)n * row) {
            err = std::string("write failed: ") + path;
            rc = NEMO2D_ERR_ARG;
        }
    }
    if (fclose(f) != 0 && rc == NEMO2D_OK) {
        err = std::string("close failed: ") + path;
        rc = NEMO2D_ERR_ARG;
    }
    return rc;
}

}  // namespace nemo2d

extern "C" {

int nemo2d_format_go2d(int jpi, int jpj, const double* xt, const double* yt, const double* ht,
                       const double* sshn, const double* un, const double* vn, char* buf, long cap,
                       long* len) {
    using namespace nemo2d;
    if (jpi < 1 || jpj < 1 || !xt || !yt || !ht || !sshn || !un || !vn || !len)
        return set_error(NEMO2D_ERR_ARG, "nemo2d_format_go2d: bad argument");
    const long need = go2d_bytes(jpi, jpj);
    *len = need;
    if (!buf) return NEMO2D_OK;  // size query
    if (cap < need) return set_error(NEMO2D_ERR_ARG, "nemo2d_format_go2d: buffer too small");
    // Fortran layouts: T (1:jpi,1:jpj), U (0:jpi,1:jpj), V (1:jpi,0:jpj)
    const Arr2 T_x{xt, -1L - jpi, jpi}, T_y{yt, -1L - jpi, jpi}, T_h{ht, -1L - jpi, jpi};
    const Arr2 T_s{sshn, -1L - jpi, jpi}, U{un, -(long)jpi - 1, (long)jpi + 1}, V{vn, -1L, jpi};
    format_go2d_rows(T_x, T_y, T_h, T_s, U, V, jpi, 1, jpj + 1, buf);
    return NEMO2D_OK;
}

}  // extern "C"
