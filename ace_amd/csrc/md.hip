// md.hip -- velocity-Verlet molecular dynamics on the device, metal units (A, ps, amu, eV).
//
// One step follows the integrator template of src/sim/integrator.cpp:548-571: half kick with F(t),
// drift, forces at the new positions, half kick with F(t + dt).  ace has no NNP molecular dynamics and its
// Verlet template carries no unit conversion; here a = F / (m * MD_MVV2E).  Forces are read directly from
// the fixed-point accumulators of the last energy/force launch (scale 2^34), so no conversion launch sits
// between the force kernel and the kicks.  Positions and velocities are FP64.
#include "ace_common.hpp"

#define MD_THREADS 256
#define MD_MVV2E 1.0364269e-4   // eV per amu (A/ps)^2
#define MD_FIX_INV (1.0 / 17179869184.0)

// v += dt/2 F(t) / (m mvv2e); r += dt v.  ctr[0] counts steps (read by k_md_kick of the same step).
__global__ __launch_bounds__(MD_THREADS) void k_md_kick_drift(int N, double hdt, double dt,
                                                              const int* __restrict__ type,
                                                              const double* __restrict__ minv,
                                                              const unsigned long long* __restrict__ F_fix,
                                                              double* __restrict__ pos, double* __restrict__ vel,
                                                              unsigned int* __restrict__ ctr) {
    const int i = blockIdx.x * MD_THREADS + threadIdx.x;
    if (i == 0) atomicAdd(ctr, 1u);
    if (i >= N) return;
    const double w = hdt * minv[type[i]] * MD_FIX_INV;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double v = vel[3 * i + d] + w * (double)(long long)F_fix[3 * i + d];
        vel[3 * i + d] = v;
        pos[3 * i + d] += dt * v;
    }
}

// v += dt/2 F(t + dt) / (m mvv2e); the step's kinetic energy sum_i m_i mvv2e |v_i|^2 / 2 and potential
// energy sum_s E_s are added to rec[2 step + 1] and written to rec[2 step] (fixed point, scale 2^34).
__global__ __launch_bounds__(MD_THREADS) void k_md_kick(int N, int S, double hdt, const int* __restrict__ type,
                                                        const double* __restrict__ minv,
                                                        const double* __restrict__ mm,
                                                        const unsigned long long* __restrict__ F_fix,
                                                        const unsigned long long* __restrict__ E_fix,
                                                        double* __restrict__ vel, unsigned long long* __restrict__ rec,
                                                        int rec_cap, const unsigned int* __restrict__ ctr) {
    __shared__ double s_ke[MD_THREADS / 64];
    const int i = blockIdx.x * MD_THREADS + threadIdx.x;
    double ke = 0.0;
    if (i < N) {
        const int t = type[i];
        const double w = hdt * minv[t] * MD_FIX_INV;
        double v2 = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double v = vel[3 * i + d] + w * (double)(long long)F_fix[3 * i + d];
            vel[3 * i + d] = v;
            v2 += v * v;
        }
        ke = 0.5 * mm[t] * v2;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) ke += __shfl_xor(ke, o, 64);
    if ((threadIdx.x & 63) == 0) s_ke[threadIdx.x >> 6] = ke;
    __syncthreads();
    const unsigned int step = ctr[0] - 1u;
    if (threadIdx.x == 0 && step < (unsigned int)rec_cap) {
        double b = 0.0;
#pragma unroll
        for (int w = 0; w < MD_THREADS / 64; ++w) b += s_ke[w];
        atomicAdd(&rec[2 * step + 1], (unsigned long long)__double2ll_rn(b * 17179869184.0));
        if (blockIdx.x == 0) {
            unsigned long long pe = 0;
            for (int s = 0; s < S; ++s) pe += E_fix[s];
            rec[2 * step] = pe;
        }
    }
}

// Returns every atom to its cell, r -> R (f - floor(f)) with f = R^-1 r (Cell::returnToCell, cell.cpp:73-83,
// which the reference's Verlet applies after every drift, integrator.cpp:553).  Run just before each
// neighbour-list rebuild: between rebuilds the list's image vectors refer to the positions as they were
// at the build, and the build's image range (ceil(3 rc / |a_d|), neighbor.cpp:49-53) must cover them.
__global__ __launch_bounds__(MD_THREADS) void k_md_wrap(int N, const int* __restrict__ struc_of,
                                                        const double* __restrict__ cell, double* __restrict__ pos) {
    const int i = blockIdx.x * MD_THREADS + threadIdx.x;
    if (i >= N) return;
    const double* R = cell + 9 * struc_of[i];   // column-major: R[3 c + r] = component r of lattice vector c
    // R^-1 by cofactors
    const double c00 = R[4] * R[8] - R[7] * R[5], c01 = R[7] * R[2] - R[1] * R[8], c02 = R[1] * R[5] - R[4] * R[2];
    const double c10 = R[6] * R[5] - R[3] * R[8], c11 = R[0] * R[8] - R[6] * R[2], c12 = R[3] * R[2] - R[0] * R[5];
    const double c20 = R[3] * R[7] - R[6] * R[4], c21 = R[6] * R[1] - R[0] * R[7], c22 = R[0] * R[4] - R[3] * R[1];
    const double idet = 1.0 / (R[0] * c00 + R[3] * c01 + R[6] * c02);
    const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
    double f0 = (c00 * x + c10 * y + c20 * z) * idet;
    double f1 = (c01 * x + c11 * y + c21 * z) * idet;
    double f2 = (c02 * x + c12 * y + c22 * z) * idet;
    f0 -= floor(f0);
    f1 -= floor(f1);
    f2 -= floor(f2);
    pos[3 * i] = R[0] * f0 + R[3] * f1 + R[6] * f2;
    pos[3 * i + 1] = R[1] * f0 + R[4] * f1 + R[7] * f2;
    pos[3 * i + 2] = R[2] * f0 + R[5] * f1 + R[8] * f2;
}

// The neighbour build clears the list-overflow flag (scal[over_idx]) and the basis kernels' refusal flag
// (scal[3]) for the next list; across the rebuilds of a run they are collected here (sticky[0] != 0).
__global__ void k_md_flags(const int* __restrict__ scal, int over_idx, unsigned int* __restrict__ sticky) {
    if (threadIdx.x == 0 && (scal[3] | scal[over_idx])) atomicOr(sticky, 1u);
}
