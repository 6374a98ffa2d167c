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
            // the accurate sincosf: the hardware sine/cosine's absolute error (~1e-6) enters every
            // neighbour's cutoff and, summed over the pairs of an atom, moved FP32 forces by ~1e-5 eV/A
            if constexpr (sizeof(T) == 4) sincosf(r * pirc, &s, &c);
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

// Cross-lane moves through DPP (data-parallel primitives: a VALU operand modifier, no LDS round trip).
// CTRL: 0xB1 quad_perm(1,0,3,2), 0x4E quad_perm(2,3,0,1), 0x141 row_half_mirror, 0x140 row_mirror,
// 0x142 row_bcast:15, 0x143 row_bcast:31.  Rows outside ROWMASK receive 0.
template <int CTRL, int ROWMASK>
__device__ inline float dpp_mov(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, ROWMASK, 0xF, false));
}
template <int CTRL, int ROWMASK>
__device__ inline double dpp_mov(double v) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, ROWMASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWMASK, 0xF, false);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}

// sum over the 16 lanes of each row, returned in every lane of the row
template <class T>
__device__ inline T t_row_sum(T v) {
    v += dpp_mov<0xB1, 0xF>(v);
    v += dpp_mov<0x4E, 0xF>(v);
    v += dpp_mov<0x141, 0xF>(v);
    v += dpp_mov<0x140, 0xF>(v);
    return v;
}

// sum over the wave, valid in lane 63 (row sums, then row_bcast:15 into rows 1 and 3 and
// row_bcast:31 into rows 2 and 3)
template <class T>
__device__ inline T t_wave_sum63(T v) {
    v = t_row_sum<T>(v);
    v += dpp_mov<0x142, 0xA>(v);
    v += dpp_mov<0x143, 0xC>(v);
    return v;
}

// sum over aligned groups of n lanes (n a power of two, wave-uniform), returned in every lane of the group
template <class T>
__device__ inline T t_group_sum(T v, int n) {
    if (n >= 2) v += dpp_mov<0xB1, 0xF>(v);
    if (n >= 4) v += dpp_mov<0x4E, 0xF>(v);
    if (n >= 8) v += dpp_mov<0x141, 0xF>(v);
    if (n >= 16) v += dpp_mov<0x140, 0xF>(v);
    if (n >= 32) v += __shfl_xor(v, 16, 64);
    if (n >= 64) v += __shfl_xor(v, 32, 64);
    return v;
}

// lane exchange across lane bit B: bits 5 and 4 by permutes, bit 3 by row_mirror (lane ^ 15 in a row),
// bit 2 by row_half_mirror (lane ^ 7), bits 1 and 0 by quad permutes.  Each is an involution whose
// partner differs in bit B, which is what the reductions below need.
template <int B, class T>
__device__ inline T lane_swap(T v) {
    if constexpr (B == 5) return __shfl_xor(v, 32, 64);
    else if constexpr (B == 4) return __shfl_xor(v, 16, 64);
    else if constexpr (B == 3) return dpp_mov<0x140, 0xF>(v);
    else if constexpr (B == 2) return dpp_mov<0x141, 0xF>(v);
    else if constexpr (B == 1) return dpp_mov<0x4E, 0xF>(v);
    else return dpp_mov<0xB1, 0xF>(v);
}

// halving steps from bit B down: lanes with bit B set keep the upper half of v, the others the lower half,
// and add the partner's copy of the half they keep
template <int B, int H, class T>
__device__ inline void rs_steps(T* v, int lane) {
    if constexpr (H >= 1) {
        const bool hi = (lane >> B) & 1;
#pragma unroll
        for (int j = 0; j < H; ++j) {
            const T send = hi ? v[j] : v[j + H];
            const T keep = hi ? v[j + H] : v[j];
            v[j] = keep + lane_swap<B>(send);
        }
        rs_steps<B - 1, H / 2>(v, lane);
    }
}

template <int B, class T>
__device__ inline T sum_steps(T v) {
    if constexpr (B >= 0) return sum_steps<B - 1>(v + lane_swap<B>(v));
    else return v;
}

// Reduce-scatter of M = 2^LM values per lane over the wave: returns the wave total of value
// lane >> (6 - LM), held by every lane of its group of 2^(6 - LM) lanes.  M - 1 + 6 - LM exchanges
// instead of 6 per value.
template <int LM, class T>
__device__ inline T wave_reduce_scatter(T* v, int lane) {
    rs_steps<5, (1 << LM) / 2>(v, lane);
    return sum_steps<5 - LM>(v[0]);
}

template <class T>
__device__ inline T t_wave_sum(T v) {
    v = t_wave_sum63<T>(v);
    if constexpr (sizeof(T) == 8) {
        const long long b = __builtin_bit_cast(long long, v);
        const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), 63);
        const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
        return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
    } else {
        return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
    }
}
