// Shared parts of the command-line drivers: argument parsing of the data and
// access types, wall-clock timing, run statistics and CSV output.  The
// drivers take the arguments, in the same order, of the reference's example
// programs (example/testfloat_compress_fastmode1.cc,
// example/testfloat_decompress_fastmode1.cc) and print the same metrics.
#pragma once

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include <vector>

#include "szp.h"
#include "szp_rw.h"

namespace cli {

inline double now_seconds()
{
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    return tv.tv_sec + tv.tv_usec * 1e-6;
}

// "float" -> SZ_FLOAT, "double" -> SZ_DOUBLE, else -1.
inline int parse_data_type(const char *s)
{
    if (!strcmp(s, "float")) return SZ_FLOAT;
    if (!strcmp(s, "double")) return SZ_DOUBLE;
    return -1;
}

// "random" -> SZP_RANDOMACCESS, "nonrandom" -> SZP_NONRANDOMACCESS, else -1.
inline int parse_access(const char *s)
{
    if (!strcmp(s, "random")) return SZP_RANDOMACCESS;
    if (!strcmp(s, "nonrandom")) return SZP_NONRANDOMACCESS;
    return -1;
}

struct Stat {
    double mn, mx, avg, sd;
};

inline Stat stats(const std::vector<double> &v)
{
    Stat s{0, 0, 0, 0};
    if (v.empty()) return s;
    s.mn = s.mx = v[0];
    double sum = 0;
    for (double x : v) {
        s.mn = x < s.mn ? x : s.mn;
        s.mx = x > s.mx ? x : s.mx;
        sum += x;
    }
    s.avg = sum / v.size();
    double var = 0;
    for (double x : v) var += (x - s.avg) * (x - s.avg);
    s.sd = sqrt(var / v.size());
    return s;
}

inline void print_stat(const char *name, const std::vector<double> &v)
{
    const Stat s = stats(v);
    printf("%s: Min=%.6f, Max=%.6f, Avg=%.6f, StdDev=%.6f\n", name, s.mn, s.mx, s.avg, s.sd);
}

}  // namespace cli
