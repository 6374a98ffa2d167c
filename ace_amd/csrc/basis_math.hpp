// basis_math.hpp -- scalar pieces of the symmetry functions, templated on the real type T (float or
// double) used by the basis kernels.
#pragma once
#include "ace_common.hpp"

template <class T>
__device__ inline T t_exp(T x);
template <>
__device__ inline float t_exp<float>(float x) { return fast_exp(x); }
template <>
__device__ inline double t_exp<double>(double x) { return exp(x); }

// cutoff value and derivative (cutoff.cpp:62-117); COS is the reference's 0.5 (cos(pi r/rc) + 1)
template <class T>
__device__ inline void t_cutoff(int name, T rc, T r, T& f, T& g) {
    if (r >= rc) { f = T(0); g = T(0); return; }
    switch (name) {
        case ACE_CUT_COS: {
            const T pirc = T(3.141592653589793) / rc;
            T s, c;
            if constexpr (sizeof(T) == 4) __sincosf(r * pirc, &s, &c);
            else sincos(r * pirc, &s, &c);
            f = T(0.5) * (c + T(1));
            g = T(-0.5) * pirc * s;
        } break;
        case ACE_CUT_STEP: f = T(1); g = T(0); break;
        case ACE_CUT_TANH: {
            const T t = tanh(T(1) - r / rc);
            f = t * t * t;
            g = T(-3) * t * t * (T(1) - t * t) / rc;
        } break;
        case ACE_CUT_POLY3: {
            const T x = r / rc;
            f = (T(2) * x - T(3)) * x * x + T(1);
            g = T(6) * x * (x - T(1)) / rc;
        } break;
        default: f = T(0); g = T(0);
    }
}

// one radial function value and r-derivative (basis_radial.cpp:175-257)
template <class T>
__device__ inline void t_radial(int form, T eta, T rs, T r, T fc, T gc, T& v, T& dv) {
    switch (form) {
        case ACE_RAD_GAUSSIAN: {
            const T arg = eta * (r - rs);
            const T e = t_exp<T>(-arg * arg);
            v = e * fc;
            dv = e * (T(-2) * eta * arg * fc + gc);
        } break;
        case ACE_RAD_SECH: {
            const T e = t_exp<T>(-eta * (r - rs));
            const T e2 = e * e;
            const T den = T(1) / (T(1) + e2);
            v = T(2) * e * den * fc;
            dv = T(2) * e * den * (-eta * (T(1) - e2) * den * fc + gc);
        } break;
        case ACE_RAD_LOGISTIC: {
            const T e = t_exp<T>(T(-2) * eta * (r - rs));
            const T den = T(1) / (T(1) + e);
            const T t = (T(1) - e) * den;
            const T s2 = T(4) * e * den * den;
            v = s2 * fc;
            dv = s2 * (T(-2) * eta * t * fc + gc);
        } break;
        case ACE_RAD_TANH: {
            const T t = tanh(-eta * (r - rs));
            v = T(0.5) * (t + T(1)) * fc;
            dv = T(0.5) * (-eta * (T(1) - t * t) * fc + (T(1) + t) * gc);
        } break;
        case ACE_RAD_LOGCOSH: {
            const T e = t_exp<T>(T(-2) * eta * (r - rs));
            const T l = T(0.5) * log1p(e);
            v = l * fc;
            dv = -eta * e / (T(1) + e) * fc + l * gc;
        } break;
        case ACE_RAD_LOGCOSH2: {
            const T rad2 = T(1.414213562373095);
            const T arg = -eta * (r - rs);
            if (arg == T(0)) {
                v = T(0.5) * fc / rad2;
                dv = T(0.5) * (-eta * gc + fc / rad2);
            } else {
                const T e = t_exp<T>(T(2) * rad2 * arg);
                const T den = T(1) / (e - T(1));
                v = arg * e * den * fc;
                dv = e * (eta * (T(1) + T(2) * rad2 * arg - e) * den * fc + arg * gc) * den;
            }
        } break;
        default: v = T(0); dv = T(0);
    }
}

// x^z for the angular window x in [0,1] (z > 0): binary powering for integer z, else exp/log
template <class T>
__device__ inline T t_pow_window(T x, T z, int zint) {
    if (zint >= 0) {
        T y = T(1), w = x;
        int n = zint;
        while (n) {
            if (n & 1) y *= w;
            w *= w;
            n >>= 1;
        }
        return y;
    }
    return (x > T(0)) ? pow(x, z) : T(0);
}

template <class T>
__device__ inline T t_wave_sum(T v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}
