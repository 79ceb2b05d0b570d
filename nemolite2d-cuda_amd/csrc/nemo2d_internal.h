// Host-side helpers shared by the library's translation units.
#pragma once

#include <math.h>

#include <string>
#include <vector>

#include "nemo2d_core.h"

namespace nemo2d {

struct Classes {
    std::vector<uint16_t> fl;        // per-point class words
    std::vector<double> flu, flv;    // Flather sqrt(g/h) at Flather points
    std::vector<int> uptr, vptr;     // Flather lines: rows (u) / columns (v)
    std::vector<long> upts, vpts;    // Flather point offsets in loop order
    bool chained = false;            // a Flather point's inner point is one too
};

// Tide of the clamped-ssh open boundary, original/nemolite2d.f90:629-635:
// amp * sin(omega * rtime) with omega = 2*3.14159/(period*3600) (the serial
// model's literal 3.14159, period 12.42 h, amp 0.2 m), on the host C library.
inline double tide_omega_of(double period_hours) { return 2.0 * 3.14159 / (period_hours * 3600.0); }
inline double tide_eval(double amp, double omega, double rtime) { return amp * sin(omega * rtime); }

// Reciprocals of a uniform grid's divisors (see nemo2d_core.h div_rcp).
void set_reciprocals(Scal& s, double fric);

// Record the message returned by nemo2d_last_error(); returns code.
int set_error(int code, const std::string& msg);

int classify(const Layout& L, const int* pt, const double* hu, const double* hv, Classes& c,
             std::string& err);

// A 2D host array read as p[off + j*stride + i] for point (i, j).
struct Arr2 {
    const double* p;
    long off, stride;
    double at(int i, int j) const { return p[off + (long)j * stride + i]; }
};

// go2d snapshot files (nemo2d_output.hip): bytes of a jpi x jpj file, rows
// j0..j1-1 formatted into buf, and a whole file written with threads.
long go2d_bytes(int jpi, int jpj);
void format_go2d_rows(const Arr2& xt, const Arr2& yt, const Arr2& ht, const Arr2& ss, const Arr2& un,
                      const Arr2& vn, int jpi, int j0, int j1, char* buf);
int write_go2d(const char* path, const Arr2& xt, const Arr2& yt, const Arr2& ht, const Arr2& ss,
               const Arr2& un, const Arr2& vn, int jpi, int jpj, std::string& err);

}  // namespace nemo2d
