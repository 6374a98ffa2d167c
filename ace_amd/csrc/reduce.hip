// reduce.hip -- fixed-point accumulator readout and the per-structure energy sum of the energy-only loss.
//
// For each pair p = (i -> j, n) with g_p = dE_i/dd_p and d_p = r_i - r_j - R n, F_j gets +g_p and F_i -g_p
// (the scatter of NNP::compute_force, src/nnp/nnp.cpp:442-471): the fused and FP64-basis kernels add them
// as 64-bit fixed-point integers, so the sums are exact and independent of the order of the additions;
// the FP32 split pass leaves g_p at the reverse entry of j's key row and k_force_gather sums each row in a
// fixed order and stores the fixed-point force.
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

// Forces from per-pair gradients, the gathered form of the same scatter: the gradient launch wrote g of
// every pair (i -> a, n) at the entry of its reverse (a -> i, -n) in a's row and -sum g of a's own pairs to
// fown[a], so F_a = fown[a] + sum over a's row; wave per atom, each lane summing its entries in row order
// and the wave in a fixed butterfly (lane 0's total), so the forces depend on neither the batch nor the
// launch; stored once as fixed point.  A row longer than cap belongs to a refused build.
#define FG_WAVES 4
__global__ __launch_bounds__(FG_WAVES * 64) void k_force_gather(int natoms, int cap, const int* __restrict__ counts,
                                                               const double* __restrict__ gin,
                                                               const double* __restrict__ fown,
                                                               unsigned long long* __restrict__ F_fix) {
    const int lane = threadIdx.x & 63;
    const int a = blockIdx.x * FG_WAVES + (threadIdx.x >> 6);
    if (a >= natoms) return;
    const int n = min(counts[a], cap);
    const double* g = gin + 3 * (long long)a * cap;
    double f[3] = {0.0, 0.0, 0.0};
    for (int e = lane; e < n; e += 64) {
#pragma unroll
        for (int c = 0; c < 3; ++c) f[c] += g[3 * e + c];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) f[c] += __shfl_xor(f[c], o, 64);
    }
    if (lane < 3) F_fix[3 * a + lane] = fix_of(fown[3 * a + lane] + f[lane]);
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
