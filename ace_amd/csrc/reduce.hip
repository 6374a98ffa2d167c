// reduce.hip -- forces from per-pair gradients, FP64 per-structure energy/virial reductions, and
// displacement refresh for moved atoms.
//
// With g_p = dE_i/dd_p for pair p = (i -> j, n), d_p = r_i - r_j - R n, and the reverse pair
// rev(p) = (j -> i, -n):  F_i = -dE/dr_i = sum_{p in list(i)} (g_rev(p) - g_p).
// This is the scatter F_i +=, F_j -= of NNP::compute_force (src/nnp/nnp.cpp:442-471) done as a
// deterministic gather in FP64, without atomics.
#include "ace_common.hpp"
#include "basis_math.hpp"

// GL lanes per atom: force (FP64) and the atom's virial share -sum_p g_p (x) d_p (FP64, 9 values);
// each lane sums a strided subset of the atom's pairs, the GL partials are combined by shuffles in a
// fixed order.
#define GL 8
__global__ __launch_bounds__(256) void k_force_gather(int natoms, const int* __restrict__ nb_off,
                                                      const int* __restrict__ nb_rev, const float4* __restrict__ nb_d,
                                                      const float4* __restrict__ pair_g, double* __restrict__ F,
                                                      double* __restrict__ vir_atom) {
    const int gt = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = gt / GL, sub = gt % GL;
    const bool on = i < natoms;
    double f[3] = {0, 0, 0};
    double w[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (on) {
        for (int p = nb_off[i] + sub; p < nb_off[i + 1]; p += GL) {
            const float4 g = pair_g[p];
            const float4 gr = pair_g[nb_rev[p]];
            const float4 d = nb_d[p];
            f[0] += (double)gr.x - (double)g.x;
            f[1] += (double)gr.y - (double)g.y;
            f[2] += (double)gr.z - (double)g.z;
            const double ga[3] = {g.x, g.y, g.z}, da[3] = {d.x, d.y, d.z};
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) w[3 * a + b] -= ga[a] * da[b];
        }
    }
    // sums over the GL = 8 lanes of the atom (DPP quad and half-row moves)
#pragma unroll
    for (int q = 0; q < 3; ++q) f[q] = t_group_sum<double>(f[q], GL);
    if (vir_atom) {
#pragma unroll
        for (int q = 0; q < 9; ++q) w[q] = t_group_sum<double>(w[q], GL);
    }
    if (on && sub == 0) {
        F[3 * i] = f[0];
        F[3 * i + 1] = f[1];
        F[3 * i + 2] = f[2];
        if (vir_atom) {
#pragma unroll
            for (int q = 0; q < 9; ++q) vir_atom[9 * i + q] = w[q];
        }
    }
}

// block per structure: E_s = sum e_atom, W_s = sum vir_atom (fixed-order FP64 tree)
#define RED_THREADS 256
__global__ __launch_bounds__(RED_THREADS) void k_struct_reduce(
    const int* __restrict__ atom_off, const int* __restrict__ natoms_of, const double* __restrict__ e_atom,
    const double* __restrict__ vir_atom, double* __restrict__ E, double* __restrict__ vir) {
    __shared__ double s[RED_THREADS][10];
    const int st = blockIdx.x, t = threadIdx.x;
    const int a0 = atom_off[st], n = natoms_of[st];
    double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = t; k < n; k += RED_THREADS) {
        acc[0] += e_atom[a0 + k];
        if (vir_atom)
            for (int q = 0; q < 9; ++q) acc[1 + q] += vir_atom[9 * (a0 + k) + q];
    }
    for (int q = 0; q < 10; ++q) s[t][q] = acc[q];
    __syncthreads();
    for (int h = RED_THREADS / 2; h > 0; h >>= 1) {
        if (t < h)
            for (int q = 0; q < 10; ++q) s[t][q] += s[t + h][q];
        __syncthreads();
    }
    if (t == 0) {
        E[st] = s[0][0];
        if (vir)
            for (int q = 0; q < 9; ++q) vir[9 * st + q] = s[0][1 + q];
    }
}

// thread per pair: d_p = r_i - (r_j + R n) in FP64 (nnp.cpp:363), stored as FP32 with |d|
__global__ void k_nb_disp(const int* __restrict__ npairs, const int* __restrict__ nb_src, const int* __restrict__ nb_j,
                          const int* __restrict__ nb_img, const int* __restrict__ struc_of,
                          const StrucInfo* __restrict__ sinfo, const double* __restrict__ pos,
                          float4* __restrict__ nb_d) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= *npairs) return;
    const int i = nb_src[p], j = nb_j[p];
    const StrucInfo& S = sinfo[struc_of[i]];
    int c0, c1, c2;
    img_unpack(nb_img[p], c0, c1, c2);
    double rv[3];
    for (int d = 0; d < 3; ++d) {
        const double sh = S.R[d] * c0 + S.R[d + 3] * c1 + S.R[d + 6] * c2;
        rv[d] = pos[3 * i + d] - (pos[3 * j + d] + sh);
    }
    const double r = sqrt(rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]);
    nb_d[p] = make_float4((float)rv[0], (float)rv[1], (float)rv[2], (float)r);
}

// G rows (species-sorted, FP32) -> FP64 rows in atom order with stride ldo
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
