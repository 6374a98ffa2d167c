// ace_common.hpp -- device-side model tables and helpers shared by the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/ace_hip.h"

#define ACE_MAXT ACE_MAX_TYPES
#define ACE_MAXB (ACE_MAX_TYPES * (ACE_MAX_TYPES + 1) / 2)
#define ACE_MAXL ACE_MAX_LAYERS
#define ACE_MAXG 4          // distinct angular eta values per species (exp groups)
#define ACE_WAVE 64

// One species (NNH): offsets follow NNH::init_input (nnh.cpp:183-201).
struct DevSpecies {
    int K;                         // input width
    int nR;                        // radial inputs
    int ldp;                       // parameter count
    // radial blocks, per neighbour species J
    int r_form[ACE_MAXT], r_cut[ACE_MAXT], r_n[ACE_MAXT], r_goff[ACE_MAXT], r_poff[ACE_MAXT];
    float r_rc[ACE_MAXT];
    // angular blocks, per TensorTri index b; one cutoff and one set of exp groups per species
    int a_cut;
    float a_rc;
    int ng;
    float ietar2[ACE_MAXG];        // 1/eta^2 of each distinct eta (BasisA::init run grouping)
    int a_form[ACE_MAXB], a_n[ACE_MAXB], a_goff[ACE_MAXB], a_poff[ACE_MAXB];
    int a_allint[ACE_MAXB];        // 1 if every zeta of the block is an integer in [1,64]
    float rc_max;                  // largest basis cutoff of this species
    // network
    int L;
    int width[ACE_MAXL + 1];
    int act;
    float c;
    int boff[ACE_MAXL], woff[ACE_MAXL];   // offsets in the species' parameter block
    int poff;                      // offset of the species' parameter block in the global array
    int in_off;                    // offset of inpw/inpb
    float outw, outb;
    double E0;
};

struct DevModel {
    int ntypes;
    int kmax;
    int ng_max;
    DevSpecies sp[ACE_MAXT];
};

// Per-function parameter arrays (float), indexed by r_poff / a_poff.
struct DevBasisParams {
    const float* r_rs;
    const float* r_eta;
    const float* a_lamf;   // 0.5*lambda
    const float* a_zeta;
    const int*   a_zint;   // integer zeta or -1
    const int*   a_grp;    // exp group of the function
};

// Per-structure geometry for the neighbour build.
struct StrucInfo {
    double R[9];      // column-major, columns = lattice vectors
    double Rinv[9];   // column-major inverse
    int atom_off, natoms;
    int nc[3];        // cells per dimension
    int st[3];        // stencil half-width per dimension
    int cf[3];        // reference image range ceil(3 rc / |a_d|)
    int cell_off;
    int pad;
};

// host-side device buffer with capacity
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t cap = 0;
};

__host__ __device__ inline int tri_index(int j, int k) {
    return (j > k) ? j * (j + 1) / 2 + k : k * (k + 1) / 2 + j;
}

// ---------------------------------------------------------------------------------------------
// cutoff functions (cutoff.cpp:62-117), FP32.  coscut/sincut are the reference's Taylor forms of
// cos/sin; in FP32 the hardware cos/sin of the same argument agree to rounding.
// ---------------------------------------------------------------------------------------------
__device__ inline void cutoff_fg(int name, float rc, float r, float& f, float& g) {
    if (r >= rc) { f = 0.f; g = 0.f; return; }
    switch (name) {
        case ACE_CUT_COS: {
            const float pirc = 3.14159265358979f / rc;
            float s, c;
            __sincosf(r * pirc, &s, &c);
            f = 0.5f * (c + 1.f);
            g = -0.5f * pirc * s;
        } break;
        case ACE_CUT_STEP: f = 1.f; g = 0.f; break;
        case ACE_CUT_TANH: {
            const float t = tanhf(1.f - r / rc);
            f = t * t * t;
            g = -3.f * t * t * (1.f - t * t) / rc;
        } break;
        case ACE_CUT_POLY3: {
            const float x = r / rc;
            f = (2.f * x - 3.f) * x * x + 1.f;
            g = 6.f * x * (x - 1.f) / rc;
        } break;
        default: f = 0.f; g = 0.f;
    }
}

__device__ inline float fast_exp(float x) {   // e^x through the hardware exp2
    return __builtin_amdgcn_exp2f(x * 1.4426950408889634f);
}

// ---------------------------------------------------------------------------------------------
// radial function value and derivative (basis_radial.cpp:175-257), FP32.
// ---------------------------------------------------------------------------------------------
__device__ inline void radial_fd(int form, float eta, float rs, float r, float fc, float gc,
                                 float& v, float& dv) {
    switch (form) {
        case ACE_RAD_GAUSSIAN: {
            const float arg = eta * (r - rs);
            const float e = fast_exp(-arg * arg);
            v = e * fc;
            dv = e * (-2.f * eta * arg * fc + gc);
        } break;
        case ACE_RAD_SECH: {
            const float e = fast_exp(-eta * (r - rs));
            const float e2 = e * e;
            const float den = 1.f / (1.f + e2);
            v = 2.f * e * den * fc;
            dv = 2.f * e * den * (-eta * (1.f - e2) * den * fc + gc);
        } break;
        case ACE_RAD_LOGISTIC: {
            const float e = fast_exp(-2.f * eta * (r - rs));
            const float den = 1.f / (1.f + e);
            const float t = (1.f - e) * den;
            const float s2 = 4.f * e * den * den;
            v = s2 * fc;
            dv = s2 * (-2.f * eta * t * fc + gc);
        } break;
        case ACE_RAD_TANH: {
            const float t = tanhf(-eta * (r - rs));
            v = 0.5f * (t + 1.f) * fc;
            dv = 0.5f * (-eta * (1.f - t * t) * fc + (1.f + t) * gc);
        } break;
        case ACE_RAD_LOGCOSH: {
            const float e = fast_exp(-2.f * eta * (r - rs));
            const float l = 0.5f * log1pf(e);
            v = l * fc;
            dv = -eta * e / (1.f + e) * fc + l * gc;
        } break;
        case ACE_RAD_LOGCOSH2: {
            const float rad2 = 1.41421356237310f;
            const float arg = -eta * (r - rs);
            if (arg == 0.f) {
                v = 0.5f * fc / rad2;
                dv = 0.5f * (-eta * gc + fc / rad2);
            } else {
                const float e = fast_exp(2.f * rad2 * arg);
                const float den = 1.f / (e - 1.f);
                v = arg * e * den * fc;
                dv = e * (eta * (1.f + 2.f * rad2 * arg - e) * den * fc + arg * gc) * den;
            }
        } break;
        default: v = 0.f; dv = 0.f;
    }
}

// orders LDS traffic between the lanes of one wave (no cross-wave barrier)
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// wave-wide sum (all lanes receive the result)
__device__ inline float wave_sum(float v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}
__device__ inline int wave_sum_i(int v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

__device__ inline unsigned long long lane_mask_lt() {
    const int lane = threadIdx.x & 63;
    return (lane == 0) ? 0ull : (~0ull >> (64 - lane));
}

// integer power by binary exponentiation (special.cpp:139-148), n >= 0
__device__ inline float powi_f(float x, int n) {
    float y = 1.f, w = x;
    while (n) {
        if (n & 1) y *= w;
        w *= w;
        n >>= 1;
    }
    return y;
}

// x^z for x in [0,1], z > 0 (angular window), through log2/exp2 when z is not an integer
__device__ inline float pow_window(float x, float z, int zint) {
    if (zint >= 0) return powi_f(x, zint);
    return (x > 0.f) ? __builtin_amdgcn_exp2f(z * __builtin_amdgcn_logf(x)) : 0.f;
}

// packed image vector: 3 x 10-bit signed components
__host__ __device__ inline int img_pack(int a, int b, int c) {
    return ((a & 1023)) | ((b & 1023) << 10) | ((c & 1023) << 20);
}
__host__ __device__ inline void img_unpack(int p, int& a, int& b, int& c) {
    a = (p << 22) >> 22;
    b = (p << 12) >> 22;
    c = (p << 2) >> 22;
}
