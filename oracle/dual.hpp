// dual.hpp - forward-mode dual numbers for the tangent build of the oracle.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).  A Dual carries a double value and ND directional
// derivatives; every operation computes the value exactly as the double build does (same
// operation, same rounding, no contraction), so the value part of a tangent run is bitwise the
// FP64 oracle, and each derivative component is the derivative of that FP64 computation along
// one seeded direction (the chain rule applied operation by operation).
#pragma once

#include <cmath>

namespace orc {

constexpr int ND = 4;

struct Dual {
    double v;
    double d[ND];
    Dual() : v(0) {
        for (int k = 0; k < ND; ++k) d[k] = 0;
    }
    Dual(double x) : v(x) {  // NOLINT: implicit, constants promote as in the double build
        for (int k = 0; k < ND; ++k) d[k] = 0;
    }
    explicit operator double() const { return v; }
    explicit operator float() const { return (float)v; }
    Dual &operator+=(const Dual &b) {
        v = v + b.v;
        for (int k = 0; k < ND; ++k) d[k] = d[k] + b.d[k];
        return *this;
    }
    Dual &operator-=(const Dual &b) {
        v = v - b.v;
        for (int k = 0; k < ND; ++k) d[k] = d[k] - b.d[k];
        return *this;
    }
    Dual &operator*=(const Dual &b);
    Dual &operator/=(const Dual &b);
};

inline Dual operator+(const Dual &a, const Dual &b) {
    Dual r = a;
    r += b;
    return r;
}
inline Dual operator-(const Dual &a, const Dual &b) {
    Dual r = a;
    r -= b;
    return r;
}
inline Dual operator-(const Dual &a) {
    Dual r;
    r.v = -a.v;
    for (int k = 0; k < ND; ++k) r.d[k] = -a.d[k];
    return r;
}
inline Dual operator*(const Dual &a, const Dual &b) {
    Dual r;
    r.v = a.v * b.v;
    for (int k = 0; k < ND; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
inline Dual operator/(const Dual &a, const Dual &b) {
    Dual r;
    r.v = a.v / b.v;
    // (a/b)' = (a' - (a/b) b') / b
    for (int k = 0; k < ND; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
    return r;
}
inline Dual &Dual::operator*=(const Dual &b) { return *this = *this * b; }
inline Dual &Dual::operator/=(const Dual &b) { return *this = *this / b; }

inline bool operator<(const Dual &a, const Dual &b) { return a.v < b.v; }
inline bool operator>(const Dual &a, const Dual &b) { return a.v > b.v; }
inline bool operator<=(const Dual &a, const Dual &b) { return a.v <= b.v; }
inline bool operator>=(const Dual &a, const Dual &b) { return a.v >= b.v; }
inline bool operator==(const Dual &a, const Dual &b) { return a.v == b.v; }
inline bool operator!=(const Dual &a, const Dual &b) { return a.v != b.v; }

// f(a) with f'(a.v) = df
inline Dual chain(const Dual &a, double fv, double df) {
    Dual r;
    r.v = fv;
    for (int k = 0; k < ND; ++k) r.d[k] = df * a.d[k];
    return r;
}
inline Dual exp(const Dual &a) {
    const double e = std::exp(a.v);
    return chain(a, e, e);
}
inline Dual log(const Dual &a) { return chain(a, std::log(a.v), 1.0 / a.v); }
inline Dual sqrt(const Dual &a) {
    const double s = std::sqrt(a.v);
    return chain(a, s, 0.5 / s);
}
inline Dual acos(const Dual &a) { return chain(a, std::acos(a.v), -1.0 / std::sqrt(1.0 - a.v * a.v)); }
inline Dual ceil(const Dual &a) { return Dual(std::ceil(a.v)); }
// x^y: d = y x^(y-1) dx + x^y ln(x) dy; the second term only where y carries a derivative (the
// exponents of the model are constants, and x may be negative under an integer exponent).
inline Dual pow(const Dual &x, const Dual &y) {
    Dual r;
    r.v = std::pow(x.v, y.v);
    bool dy = false;
    for (int k = 0; k < ND; ++k) dy = dy || y.d[k] != 0.0;
    const double gx = y.v * std::pow(x.v, y.v - 1.0);
    const double gy = dy ? r.v * std::log(x.v) : 0.0;
    for (int k = 0; k < ND; ++k) r.d[k] = gx * x.d[k] + gy * y.d[k];
    return r;
}

}  // namespace orc
