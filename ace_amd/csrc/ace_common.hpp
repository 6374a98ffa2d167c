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
struct __attribute__((aligned(16))) DevSpecies {
    int K;                         // input width
    int nR;                        // radial inputs
    int ldp;                       // parameter count
    // radial blocks, per neighbour species J
    int r_form[ACE_MAXT], r_cut[ACE_MAXT], r_n[ACE_MAXT], r_goff[ACE_MAXT], r_poff[ACE_MAXT];
    float r_rc[ACE_MAXT];
    double r_rc64[ACE_MAXT];
    // angular blocks, per TensorTri index b; one cutoff and one set of exp groups per species
    int a_cut;
    float a_rc;
    double a_rc64;
    int ng;
    float ietar2[ACE_MAXG];        // 1/eta^2 of each distinct eta (BasisA::init run grouping)
    double ietar2_64[ACE_MAXG];
    int a_form[ACE_MAXB], a_n[ACE_MAXB], a_goff[ACE_MAXB], a_poff[ACE_MAXB];
    int a_canon[ACE_MAXB];         // 1: every function of block b has lambda = +-1, zeta = 2^k, distinct (s, k, g)
    int nz;                        // canonical power levels in use (largest k + 1)
    float rc_max;                  // largest basis cutoff of this species
    // network
    int L;
    int width[ACE_MAXL + 1];
    int act;
    float c;
    int boff[ACE_MAXL], woff[ACE_MAXL];   // offsets in the species' parameter block
    int poff;                      // offset of the species' parameter block in the global array
    // matrix-core layout of the same parameters (16-byte aligned block at wp_off, wp_n floats, a multiple
    // of 4): W_l as [k][o] at wl_off[l], round16(w_l) rows of stride round16(w_{l+1}) + 1, zero padded;
    // biases b_l at bl_off[l]
    int wp_off, wp_n;
    int wl_off[ACE_MAXL], bl_off[ACE_MAXL];
    // BF16 copies of W_0 (offset in BF16 values): high parts, low parts of [o][k] (round16(w_1) rows of
    // round32(K)), then high and low parts of [k][o] (round16(K) rows of round32(w_1)), zero padded
    int bf_off;
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

// Per-function parameter arrays (FP32 and FP64 copies), indexed by r_poff / a_poff.
struct DevBasisParams {
    const float* r_rs;
    const float* r_eta;
    const float* a_lamf;   // 0.5*lambda
    const float* a_zeta;
    const double* r_rs64;
    const double* r_eta64;
    const double* a_lamf64;
    const double* a_zeta64;
    const int*   a_zint;   // integer zeta or -1
    const int*   a_grp;    // exp group of the function
    const int*   a_sel;    // window-power cache index log2(zeta) + 7*(lambda<0), or -1
    const int*   a_cidx;   // canonical slot (7*(lambda<0) + log2 zeta)*NG + group, canonical blocks only
    const float* inpw;     // network input scaling, applied when G rows are written
    const float* inpb;
    const double* inpw64;
    const double* inpb64;
    int chunk_w;           // angular triplet loops: a chunk's set-up and flush cost in eighths of one pair
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

__device__ inline float fast_exp(float x) {   // e^x through the hardware exp2
    return __builtin_amdgcn_exp2f(x * 1.4426950408889634f);
}

// row stride of a network activation tile and of a padded weight block: width rounded up to 16 (four
// matrix-core k-steps) plus 1 (odd, so rows fall in distinct LDS banks)
__host__ __device__ inline int mlp_ld(int w) { return ((w + 15) & ~15) + 1; }

// activation value and derivative of the hidden layers (NN::ANN af_* of ml/nn.cpp:149-240), FP32
__device__ inline void act_fd(int act, float c, float z, float& a, float& d) {
    switch (act) {
        case ACE_ACT_TANH: {
            const float t = tanhf(c * z);
            a = t;
            d = c * (1.f - t * t);
        } break;
        case ACE_ACT_SIGMOID: {
            const float e = fast_exp(-c * fabsf(z));
            const float f = 1.f / (1.f + e);
            a = (z >= 0.f) ? f : e * f;
            d = c * e * f * f;
        } break;
        case ACE_ACT_SWISH: {
            const float zz = c * z;
            const float e = fast_exp(-fabsf(zz));
            const float den = 1.f / (1.f + e);
            if (zz > 0.f) { a = z * den; d = (1.f + e * (1.f + zz)) * den * den; }
            else { a = z * e * den; d = e * (1.f + zz + e) * den * den; }
        } break;
        case ACE_ACT_SOFTPLUS: {
            const float e = fast_exp(-c * fabsf(z));
            const float den = 1.f / (e + 1.f);
            const float l = (log1pf(e) - 0.693147180559945f) / c;
            a = (z >= 0.f) ? z + l : l;
            d = (z >= 0.f) ? den : e * den;
        } break;
        default: a = z; d = 1.f;
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

// packed image vector: 3 x 10-bit signed components
__host__ __device__ inline int img_pack(int a, int b, int c) {
    return ((a & 1023)) | ((b & 1023) << 10) | ((c & 1023) << 20);
}
__host__ __device__ inline void img_unpack(int p, int& a, int& b, int& c) {
    a = (p << 22) >> 22;
    b = (p << 12) >> 22;
    c = (p << 2) >> 22;
}
