// SPDX-License-Identifier: GPL-3.0-or-later
// Forward-mode dual numbers: a value and N first derivatives, for the
// vector-Jacobian products of the face-flux adjoint.  Comparisons look at
// the value only, so branches (min/max of wave speeds) follow the forward
// evaluation and the derivative is that of the branch taken.
#ifndef REACTIVE_AMD_DUAL_H
#define REACTIVE_AMD_DUAL_H
#include <hip/hip_runtime.h>

namespace ramd {

template <int N>
struct Dual {
    double v;
    double d[N];
    __device__ Dual() : v(0) {
        for (int i = 0; i < N; ++i) d[i] = 0;
    }
    __device__ Dual(double x) : v(x) {  // a constant
        for (int i = 0; i < N; ++i) d[i] = 0;
    }
    // the independent variable number `slot`
    __device__ static Dual variable(double x, int slot) {
        Dual r(x);
        r.d[slot] = 1;
        return r;
    }
};

#define RAMD_DUAL template <int N> __device__ inline
RAMD_DUAL Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v + b.v);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
RAMD_DUAL Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v - b.v);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
RAMD_DUAL Dual<N> operator-(const Dual<N>& a) {
    Dual<N> r(-a.v);
    for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
    return r;
}
RAMD_DUAL Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v * b.v);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
RAMD_DUAL Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    const double inv = 1 / b.v, q = a.v * inv;
    Dual<N> r(q);
    for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - q * b.d[i]) * inv;
    return r;
}
RAMD_DUAL Dual<N> operator+(const Dual<N>& a, double b) { return a + Dual<N>(b); }
RAMD_DUAL Dual<N> operator+(double a, const Dual<N>& b) { return Dual<N>(a) + b; }
RAMD_DUAL Dual<N> operator-(const Dual<N>& a, double b) { return a - Dual<N>(b); }
RAMD_DUAL Dual<N> operator-(double a, const Dual<N>& b) { return Dual<N>(a) - b; }
RAMD_DUAL Dual<N> operator*(const Dual<N>& a, double b) {
    Dual<N> r(a.v * b);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b;
    return r;
}
RAMD_DUAL Dual<N> operator*(double a, const Dual<N>& b) { return b * a; }
RAMD_DUAL Dual<N> operator/(const Dual<N>& a, double b) { return a / Dual<N>(b); }
RAMD_DUAL Dual<N> operator/(double a, const Dual<N>& b) { return Dual<N>(a) / b; }
RAMD_DUAL bool operator<(const Dual<N>& a, const Dual<N>& b) { return a.v < b.v; }
RAMD_DUAL bool operator>(const Dual<N>& a, const Dual<N>& b) { return a.v > b.v; }
RAMD_DUAL bool operator<(const Dual<N>& a, double b) { return a.v < b; }
RAMD_DUAL bool operator>(const Dual<N>& a, double b) { return a.v > b; }
RAMD_DUAL bool operator<(double a, const Dual<N>& b) { return a < b.v; }
RAMD_DUAL bool operator>(double a, const Dual<N>& b) { return a > b.v; }
#undef RAMD_DUAL

__device__ inline double value(double x) { return x; }
template <int N> __device__ inline double value(const Dual<N>& x) { return x.v; }

}  // namespace ramd
#endif
