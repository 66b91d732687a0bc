// Built-in analytic test problems: u, f = Laplacian(u), du/dx, du/dy, du/dz.
// Problem 0 is u = sin(2 pi x) sin(2 pi y) sin(2 pi z); problems 1-3 are the
// reference's compile-time PROBLEM1..3 (src/problem_definition.cpp:1-76).
#pragma once

#include <cmath>

namespace mgp {
namespace problems {
inline int& current() {
    thread_local int p = 0;
    return p;
}
inline double u(double x, double y, double z) {
    switch (current()) {
    case 0: { const double p = 2.0 * M_PI; return sin(p * x) * sin(p * y) * sin(p * z); }
    case 1: return sin(3.0 * x) * sin(5.3 * y) * sin(-7.8 * z);
    case 2: return x * x * x * y * y * z;
    default: return cos(x * z * z) * sin(y * y * y);
    }
}
inline double f(double x, double y, double z) {
    switch (current()) {
    case 0: { const double p = 2.0 * M_PI; return -3.0 * p * p * sin(p * x) * sin(p * y) * sin(p * z); }
    case 1: return -(3.0 * 3.0 + 5.3 * 5.3 + -7.8 * -7.8) * sin(3.0 * x) * sin(5.3 * y) * sin(-7.8 * z);
    case 2: return 2.0 * x * x * x * z + 6.0 * x * y * y * z;
    default: {
        // sin and cos of one argument through sincos(): GCC merges such pairs
        // into sincos (whose last bit can differ from sin's) and the reference
        // was built with GCC, so the explicit call keeps every compiler on the
        // reference's values
        double sy, cy, sxz, cxz;
        sincos(y * y * y, &sy, &cy);
        sincos(x * z * z, &sxz, &cxz);
        return ((-4.0 * x * x * z * z - 9.0 * y * y * y * y - z * z * z * z) * sy + 6.0 * y * cy) * cxz -
               2.0 * x * sxz * sy;
    }
    }
}
inline double dudx(double x, double y, double z) {
    switch (current()) {
    case 0: { const double p = 2.0 * M_PI; return p * cos(p * x) * sin(p * y) * sin(p * z); }
    case 1: return 3.0 * cos(3.0 * x) * sin(5.3 * y) * sin(-7.8 * z);
    case 2: return 3.0 * x * x * y * y * z;
    default: return -z * z * sin(x * z * z) * sin(y * y * y);
    }
}
inline double dudy(double x, double y, double z) {
    switch (current()) {
    case 0: { const double p = 2.0 * M_PI; return p * sin(p * x) * cos(p * y) * sin(p * z); }
    case 1: return 5.3 * sin(3.0 * x) * cos(5.3 * y) * sin(-7.8 * z);
    case 2: return 2.0 * x * x * x * y * z;
    default: return 3.0 * cos(x * z * z) * y * y * cos(y * y * y);
    }
}
inline double dudz(double x, double y, double z) {
    switch (current()) {
    case 0: { const double p = 2.0 * M_PI; return p * sin(p * x) * sin(p * y) * cos(p * z); }
    case 1: return -7.8 * sin(3.0 * x) * sin(5.3 * y) * cos(-7.8 * z);
    case 2: return x * x * x * y * y;
    default: return -2 * x * z * sin(x * z * z) * sin(y * y * y);
    }
}
}  // namespace problems
}  // namespace mgp
