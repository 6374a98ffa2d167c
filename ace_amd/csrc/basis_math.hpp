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

// 0.5 (cos(pi r / rc) + 1) and its r-derivative in FP64, for 0 <= r < rc: with t = pi r / rc - pi/2 in
// [-pi/2, pi/2], cos(pi r / rc) = -sin t and sin(pi r / rc) = cos t, and sin t, cos t are their Taylor
// series through t^17 and t^18 (truncation below 5e-14 on that interval) in Horner form, about twenty FP64
// FMAs.  Not inlined: its eighteen FP64 constants would otherwise hold registers across the whole kernel.
__device__ __attribute__((noinline)) void cos_cutoff64(double r, double rc, double& f, double& g) {
    const double k = 3.141592653589793 / rc;
    const double t = r * k - 1.5707963267948966;
    const double u = t * t;
    double p = 1.0 / 355687428096000.0;                 // 1/17!
    p = p * u - 1.0 / 1307674368000.0;                  // 1/15!
    p = p * u + 1.0 / 6227020800.0;                     // 1/13!
    p = p * u - 1.0 / 39916800.0;                       // 1/11!
    p = p * u + 1.0 / 362880.0;                         // 1/9!
    p = p * u - 1.0 / 5040.0;                           // 1/7!
    p = p * u + 1.0 / 120.0;                            // 1/5!
    p = p * u - 1.0 / 6.0;                              // 1/3!
    p = p * u + 1.0;
    const double sin_t = t * p;
    double q = -1.0 / 6402373705728000.0;               // 1/18!
    q = q * u + 1.0 / 20922789888000.0;                 // 1/16!
    q = q * u - 1.0 / 87178291200.0;                    // 1/14!
    q = q * u + 1.0 / 479001600.0;                      // 1/12!
    q = q * u - 1.0 / 3628800.0;                        // 1/10!
    q = q * u + 1.0 / 40320.0;                          // 1/8!
    q = q * u - 1.0 / 720.0;                            // 1/6!
    q = q * u + 1.0 / 24.0;                             // 1/4!
    q = q * u - 0.5;                                    // 1/2!
    const double cos_t = q * u + 1.0;
    f = 0.5 * (1.0 - sin_t);
    g = -0.5 * k * cos_t;
}

// Cutoff value and r-derivative for |d| known to FP64 precision (r) and the cutoff radius rc, in T
// (cutoff.cpp:62-117).  Every neighbour's cutoff multiplies all of its symmetry-function terms, so an FP32
// cosine's absolute error (~1e-7, from rounding pi r / rc) alone puts ~1e-5 eV/A into the forces: COS runs
// in FP64 and is rounded once.  TANH and POLY3 run in T on y = (rc - r) / rc with rc - r formed in FP64,
// so their rounding stays relative where they vanish at rc:
//   TANH:  tanh^3 y;                                     f' = -3 t^2 (1 - t^2) / rc
//   POLY3: (2x - 3) x^2 + 1 = y^2 (3 - 2y), x = r / rc;  f' = -6 (1 - y) y / rc
template <class T>
__device__ inline void t_cutoff_r64(int name, double rc, double r, T& f, T& g) {
    if (r >= rc) { f = T(0); g = T(0); return; }
    const T rcT = (T)rc;
    const T y = (T)(rc - r) / rcT;
    switch (name) {
        case ACE_CUT_COS: {
            double f64, g64;
            cos_cutoff64(r, rc, f64, g64);
            f = (T)f64;
            g = (T)g64;
        } break;
        case ACE_CUT_STEP: f = T(1); g = T(0); break;
        case ACE_CUT_TANH: {
            const T t = tanh(y);
            f = t * t * t;
            g = T(-3) * t * t * (T(1) - t * t) / rcT;
        } break;
        case ACE_CUT_POLY3: {
            f = y * y * (T(3) - T(2) * y);
            g = T(-6) * (T(1) - y) * y / rcT;
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
// 0x142 row_bcast:15, 0x143 row_bcast:31, 0x120 + r row_ror:r (rotation by r within the 16-lane row).  Rows outside ROWMASK receive 0.
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

// gfx950 lane swaps between two registers: B = 5 exchanges the upper 32 lanes of a with the lower 32 lanes
// of b (v_permlane32_swap), B = 4 the odd 16-lane rows of a with the even rows of b (v_permlane16_swap).
// Afterwards a + b holds a(l) + a(l ^ 2^B) in the lanes l whose bit B is clear and b(l) + b(l ^ 2^B) in the
// others.
template <int B>
__device__ inline void swap_rows(float& a, float& b) {
    const unsigned x = __builtin_bit_cast(unsigned, a), y = __builtin_bit_cast(unsigned, b);
    unsigned p, q;
    if constexpr (B == 5) {
        const auto r = __builtin_amdgcn_permlane32_swap(x, y, false, false);
        p = r[0]; q = r[1];
    } else {
        static_assert(B == 4, "lane swaps exchange 32- or 16-lane halves");
        const auto r = __builtin_amdgcn_permlane16_swap(x, y, false, false);
        p = r[0]; q = r[1];
    }
    a = __builtin_bit_cast(float, p);
    b = __builtin_bit_cast(float, q);
}
template <int B>
__device__ inline void swap_rows(double& a, double& b) {
    const long long x = __builtin_bit_cast(long long, a), y = __builtin_bit_cast(long long, b);
    float xl = __builtin_bit_cast(float, (unsigned)(x & 0xffffffffll)), xh = __builtin_bit_cast(float, (unsigned)(x >> 32));
    float yl = __builtin_bit_cast(float, (unsigned)(y & 0xffffffffll)), yh = __builtin_bit_cast(float, (unsigned)(y >> 32));
    swap_rows<B>(xl, yl);
    swap_rows<B>(xh, yh);
    a = __builtin_bit_cast(double, ((long long)__builtin_bit_cast(unsigned, xh) << 32) | (long long)__builtin_bit_cast(unsigned, xl));
    b = __builtin_bit_cast(double, ((long long)__builtin_bit_cast(unsigned, yh) << 32) | (long long)__builtin_bit_cast(unsigned, yl));
}

// minimum over the wave, in every lane: row rotations, then lane swaps across rows
__device__ inline int wave_min_i(int v) {
    v = min(v, __builtin_amdgcn_update_dpp(0, v, 0x121, 0xF, 0xF, false));
    v = min(v, __builtin_amdgcn_update_dpp(0, v, 0x122, 0xF, 0xF, false));
    v = min(v, __builtin_amdgcn_update_dpp(0, v, 0x124, 0xF, 0xF, false));
    v = min(v, __builtin_amdgcn_update_dpp(0, v, 0x128, 0xF, 0xF, false));
    float a = __builtin_bit_cast(float, v), b = a;
    swap_rows<4>(a, b);
    v = min(__builtin_bit_cast(int, a), __builtin_bit_cast(int, b));
    a = __builtin_bit_cast(float, v);
    b = a;
    swap_rows<5>(a, b);
    return min(__builtin_bit_cast(int, a), __builtin_bit_cast(int, b));
}

// sum over aligned groups of n lanes (n a power of two, wave-uniform), returned in every lane of the group
template <class T>
__device__ inline T t_group_sum(T v, int n) {
    if (n >= 2) v += dpp_mov<0xB1, 0xF>(v);
    if (n >= 4) v += dpp_mov<0x4E, 0xF>(v);
    if (n >= 8) v += dpp_mov<0x141, 0xF>(v);
    if (n >= 16) v += dpp_mov<0x140, 0xF>(v);
    if (n >= 32) {
        T a = v, b = v;
        swap_rows<4>(a, b);
        v = a + b;
    }
    if (n >= 64) {
        T a = v, b = v;
        swap_rows<5>(a, b);
        v = a + b;
    }
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
// and add the partner's copy of the half they keep (bits 5 and 4: one lane swap of the pair of halves)
template <int B, int H, class T>
__device__ inline void rs_steps(T* v, int lane) {
    if constexpr (H >= 1) {
        if constexpr (B >= 4) {
#pragma unroll
            for (int j = 0; j < H; ++j) {
                T a = v[j], b = v[j + H];
                swap_rows<B>(a, b);
                v[j] = a + b;
            }
        } else {
            const bool hi = (lane >> B) & 1;
#pragma unroll
            for (int j = 0; j < H; ++j) {
                const T send = hi ? v[j] : v[j + H];
                const T keep = hi ? v[j + H] : v[j];
                v[j] = keep + lane_swap<B>(send);
            }
        }
        rs_steps<B - 1, H / 2>(v, lane);
    }
}

template <int B, class T>
__device__ inline T sum_steps(T v) {
    if constexpr (B >= 4) {
        T a = v, b = v;
        swap_rows<B>(a, b);
        return sum_steps<B - 1>(a + b);
    } else if constexpr (B >= 0) {
        return sum_steps<B - 1>(v + lane_swap<B>(v));
    } else {
        return v;
    }
}

// sum over the lanes l' with l' = l mod n (n a power of two up to 32, wave-uniform), in every lane:
// row rotations by n, 2n, .. 8 (DPP row_ror) sum within the 16-lane row, lane swaps across rows
template <class T>
__device__ inline T t_cross_sum(T v, int n) {
    if (n <= 1) v += dpp_mov<0x121, 0xF>(v);
    if (n <= 2) v += dpp_mov<0x122, 0xF>(v);
    if (n <= 4) v += dpp_mov<0x124, 0xF>(v);
    if (n <= 8) v += dpp_mov<0x128, 0xF>(v);
    if (n <= 16) {
        T a = v, b = v;
        swap_rows<4>(a, b);
        v = a + b;
    }
    if (n <= 32) {
        T a = v, b = v;
        swap_rows<5>(a, b);
        v = a + b;
    }
    return v;
}

// Reduce-scatter of M = 2^LM values per lane over the wave: returns the wave total of value
// lane >> (6 - LM), held by every lane of its group of 2^(6 - LM) lanes.  M - 1 + 6 - LM exchanges
// instead of 6 per value.
template <int LM, class T>
__device__ inline T wave_reduce_scatter(T* v, int lane) {
    rs_steps<5, (1 << LM) / 2>(v, lane);
    return sum_steps<5 - LM>(v[0]);
}

// Reduce-scatter of any M <= 64 values per lane (no padding to a power of two): halving steps from lane bit B
// down.  A step splits the list of m values into P = floor(m / 2) pairs (j, j + m - P) -- lanes with the bit
// set keep the second of each pair, the others the first, and add their partner's copy -- and, for odd m, one
// single value (index P) that both halves keep, summed with the partner's copy.  The list shrinks to m - P.
template <int B, int M, class T>
__device__ inline void rs_any(T* v, int lane) {
    if constexpr (B >= 0 && M >= 1) {
        constexpr int P = M / 2, H = M - P;
        if constexpr (B >= 4) {
#pragma unroll
            for (int j = 0; j < P; ++j) {
                T a = v[j], b = v[j + H];
                swap_rows<B>(a, b);
                v[j] = a + b;
            }
            if constexpr (M & 1) {
                T a = v[P], b = v[P];
                swap_rows<B>(a, b);
                v[P] = a + b;
            }
        } else {
            const bool hi = (lane >> B) & 1;
#pragma unroll
            for (int j = 0; j < P; ++j) {
                const T send = hi ? v[j] : v[j + H];
                const T keep = hi ? v[j + H] : v[j];
                v[j] = keep + lane_swap<B>(send);
            }
            if constexpr (M & 1) v[P] = v[P] + lane_swap<B>(v[P]);
        }
        rs_any<B - 1, H>(v, lane);
    }
}

// The value index whose wave total rs_any<5, M> leaves in this lane, and whether the lane is the one of its
// group that writes it (lanes that differ only in the bits at which that value was a single hold copies).
template <int M>
__device__ inline void rs_any_owner(int lane, int& f, bool& writer) {
    int sz[6];
    int m = M;
#pragma unroll
    for (int B = 5; B >= 0; --B) {
        sz[B] = m;
        m -= m / 2;
    }
    int pos = 0;
    writer = true;
#pragma unroll
    for (int B = 0; B <= 5; ++B) {   // back from the last step
        const int P = sz[B] / 2, H = sz[B] - P;
        const bool hi = (lane >> B) & 1;
        if (pos < P) pos = hi ? pos + H : pos;
        else writer = writer && !hi;
    }
    f = pos;
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

// e^y for an exponent formed in FP64: y log2(e) split into FP32 high and low parts, 2^yh (1 + yl ln 2) with
// the hardware exp2 on yh (relative error a few FP32 ulps whatever the size of y).
__device__ inline float exp_r64(double y) {
    const double z = y * 1.4426950408889634;
    const float zh = (float)z;
    const float zl = (float)(z - (double)zh);
    const float e = __builtin_amdgcn_exp2f(zh);
    return fmaf(e, zl * 0.6931471805599453f, e);
}

// Distance factor F(s) of an angular form, s = r_ij^2 + r_ik^2, and its logarithmic derivative q = d ln F / ds
// (basis_angular.cpp:211-249 values, 261-343 derivatives), in FP64: GAUSS exp(-s/eta^2); GAUSS2 2e/(1+e) and
// SECH 2e/(1+e^2), e = exp(-pi/2 s/eta^2); STUDENT3/4/5 (1 + s/eta^2)^(-p/2), p = 4, 5, 6.  ie = 1/eta^2.
__device__ inline void angular_factor(int form, double ie, double s, double& F, double& q) {
    const double hp = 1.5707963267948966;   // pi / 2
    switch (form) {
        case ACE_ANG_GAUSS2: {
            const double e = exp(-hp * ie * s), den = 1.0 / (1.0 + e);
            F = 2.0 * e * den;
            q = -hp * ie * den;
        } break;
        case ACE_ANG_SECH: {
            const double e = exp(-hp * ie * s), e2 = e * e, den = 1.0 / (1.0 + e2);
            F = 2.0 * e * den;
            q = -hp * ie * (1.0 - e2) * den;
        } break;
        case ACE_ANG_STUDENT3: case ACE_ANG_STUDENT4: case ACE_ANG_STUDENT5: {
            const int p = form == ACE_ANG_STUDENT3 ? 4 : (form == ACE_ANG_STUDENT4 ? 5 : 6);
            const double d2 = 1.0 / (1.0 + ie * s);
            F = p == 4 ? d2 * d2 : (p == 5 ? d2 * d2 * sqrt(d2) : d2 * d2 * d2);
            q = -0.5 * p * ie * d2;
        } break;
        default: {
            F = exp(-ie * s);
            q = -ie;
        }
    }
}

// One radial function value and r-derivative with |d| known to FP64 precision (r) and the cutoff factors
// in T (basis_radial.cpp:175-257).  T = float: each form's exponent is formed in FP64 from r - rs and
// evaluated by exp_r64, the rest in FP32 (an exponent formed in FP32 carries its relative rounding times
// its own size, up to ~25 for a steep Gaussian, into the function); TANH as e / (1 + e) with e <= 1.
// T = double: t_radial in FP64.
template <class T>
__device__ inline void t_radial_r64(int form, double eta, double rs, double r, T fc, T gc, T& v, T& dv) {
    if constexpr (sizeof(T) == 8) {
        t_radial<double>(form, eta, rs, r, fc, gc, v, dv);
    } else {
        const double x = r - rs;
        const float ef = (float)eta;
        switch (form) {
            case ACE_RAD_GAUSSIAN: {
                const double a = eta * x;
                const float e = exp_r64(-(a * a));
                v = e * fc;
                dv = e * (-2.f * ef * (float)a * fc + gc);
            } break;
            case ACE_RAD_SECH: {
                const float e = exp_r64(-eta * x);
                const float e2 = e * e;
                const float den = 1.f / (1.f + e2);
                v = 2.f * e * den * fc;
                dv = 2.f * e * den * (-ef * (1.f - e2) * den * fc + gc);
            } break;
            case ACE_RAD_LOGISTIC: {
                const float e = exp_r64(-2.0 * eta * x);
                const float den = 1.f / (1.f + e);
                const float t = (1.f - e) * den;
                const float s2 = 4.f * e * den * den;
                v = s2 * fc;
                dv = s2 * (-2.f * ef * t * fc + gc);
            } break;
            case ACE_RAD_TANH: {
                // 0.5 (tanh(-eta x) + 1) = e / (1 + e) (x >= 0) or 1 / (1 + e) (x < 0), e = exp(-2 eta |x|)
                const float e = exp_r64(-2.0 * eta * fabs(x));
                const float den = 1.f / (1.f + e);
                const float h = (x >= 0.0 ? e : 1.f) * den;
                v = h * fc;
                dv = -2.f * ef * e * den * den * fc + h * gc;
            } break;
            case ACE_RAD_LOGCOSH: {
                const float e = exp_r64(-2.0 * eta * x);
                const float l = 0.5f * log1pf(e);
                v = l * fc;
                dv = -ef * e / (1.f + e) * fc + l * gc;
            } break;
            case ACE_RAD_LOGCOSH2: {
                const float rad2 = 1.414213562373095f;
                const double a = -eta * x;
                if (a == 0.0) {
                    v = 0.5f * fc / rad2;
                    dv = 0.5f * (-ef * gc + fc / rad2);
                } else {
                    // e = exp(y), y = 2 sqrt(2) a; near y = 0, e - 1 = expm1(y) and 1 + y - e = -(expm1(y) - y)
                    // by its series y^2 (1/2 + y/6 + ...), where the differences would cancel
                    const double y64 = 2.0 * 1.414213562373095 * a;
                    const float y = (float)y64, e = exp_r64(y64);
                    const bool small = fabsf(y) < 0.5f;
                    const float em1 = small ? expm1f(y) : e - 1.f;
                    float q;
                    if (small) {
                        float p = 1.f / 3628800.f;                    // 1/10!
                        p = p * y + 1.f / 362880.f;
                        p = p * y + 1.f / 40320.f;
                        p = p * y + 1.f / 5040.f;
                        p = p * y + 1.f / 720.f;
                        p = p * y + 1.f / 120.f;
                        p = p * y + 1.f / 24.f;
                        p = p * y + 1.f / 6.f;
                        p = p * y + 0.5f;
                        q = y * y * p;
                    } else {
                        q = em1 - y;
                    }
                    const float den = 1.f / em1, af = (float)a;
                    v = af * e * den * fc;
                    dv = e * (-ef * q * den * fc + af * gc) * den;
                }
            } break;
            default: v = 0.f; dv = 0.f;
        }
    }
}
