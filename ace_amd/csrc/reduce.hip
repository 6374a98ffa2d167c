// reduce.hip -- fixed-point accumulator readout and the per-structure energy sum of the energy-only loss.
//
// The basis kernels add, for each pair p = (i -> j, n) with g_p = dE_i/dd_p and d_p = r_i - r_j - R n,
// +g_p to F_j and -g_p to F_i (the scatter of NNP::compute_force, src/nnp/nnp.cpp:442-471) as 64-bit
// fixed-point integers, so the sums are exact and independent of the order of the additions.
#include "ace_common.hpp"
#include "basis_math.hpp"

#define RED_THREADS 256

// fixed-point accumulators (units of 2^-34) to FP64: forces [3N], energies [S], virials [9S]
__global__ __launch_bounds__(256) void k_finalize(int n3, int ns, const unsigned long long* __restrict__ F_fix,
                                                  const unsigned long long* __restrict__ E_fix,
                                                  const unsigned long long* __restrict__ W_fix, double* __restrict__ F,
                                                  double* __restrict__ E, double* __restrict__ W) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    constexpr double inv = 1.0 / 17179869184.0;
    if (t < n3) F[t] = (double)(long long)F_fix[t] * inv;
    if (t < ns) E[t] = (double)(long long)E_fix[t] * inv;
    if (t < 9 * ns) W[t] = (double)(long long)W_fix[t] * inv;
}

// n 64-bit words set to zero (a kernel, not hipMemsetAsync, so that callers may capture it in a HIP graph)
__global__ __launch_bounds__(256) void k_zero_u64(size_t n, unsigned long long* __restrict__ p) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) p[t] = 0ull;
}

// block per structure: E_s = sum e_atom (fixed-order FP64 tree); the training loss without forces
__global__ __launch_bounds__(RED_THREADS) void k_struct_reduce(const int* __restrict__ atom_off,
                                                                const int* __restrict__ natoms_of,
                                                                const double* __restrict__ e_atom, double* __restrict__ E) {
    __shared__ double s[RED_THREADS];
    const int st = blockIdx.x, t = threadIdx.x;
    const int a0 = atom_off[st], n = natoms_of[st];
    double acc = 0.0;
    for (int k = t; k < n; k += RED_THREADS) acc += e_atom[a0 + k];
    s[t] = acc;
    __syncthreads();
    for (int h = RED_THREADS / 2; h > 0; h >>= 1) {
        if (t < h) s[t] += s[t + h];
        __syncthreads();
    }
    if (t == 0) E[st] = s[0];
}

// network-input rows x = inpw*(G + inpb) (species-sorted, FP32) -> G = x/inpw - inpb in atom order
__global__ void k_export_rows(int natoms, const int* __restrict__ row_of, const int* __restrict__ type,
                              const DevModel* __restrict__ M, const double* __restrict__ inpw,
                              const double* __restrict__ inpb, int ldg, const float* __restrict__ G, int ldo,
                              double* __restrict__ out) {
    const int i = blockIdx.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= natoms || k >= ldo) return;
    const DevSpecies& sp = M->sp[type[i]];
    double g = 0.0;
    if (k < sp.K) {
        const int o = sp.in_off + k;
        g = (double)G[(long long)row_of[i] * ldg + k] / inpw[o] - inpb[o];
    }
    out[(long long)i * ldo + k] = g;
}
