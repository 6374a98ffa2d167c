// reduce.hip -- forces from per-pair gradients, FP64 per-structure energy/virial reductions, and
// displacement refresh for moved atoms.
//
// With g_p = dE_i/dd_p for pair p = (i -> j, n), d_p = r_i - r_j - R n, and the reverse pair
// rev(p) = (j -> i, -n):  F_i = -dE/dr_i = sum_{p in list(i)} (g_rev(p) - g_p).
// This is the scatter F_i +=, F_j -= of NNP::compute_force (src/nnp/nnp.cpp:442-471) done as a
// deterministic gather in FP64, without atomics.
#include "ace_common.hpp"
#include "basis_math.hpp"

// GL lanes per atom, 16 atoms of one structure per workgroup.  Each lane sums a strided subset of the
// atom's pairs (four pairs per round: the reverse indices, then the gradients, are loaded together) into
// the force (FP64) and the atom's virial share -sum_p g_p (x) d_p; DPP row sums combine the GL lanes.
// The workgroup's 16 (energy, virial) values are summed in a fixed order into one partial per
// workgroup; the last workgroup of a structure (per-structure ticket) sums the structure's partials in
// workgroup order.  Every result is independent of block timing.
#define GL 16
#define GB 4
#define RED_THREADS 256
#define GATHER_ATOMS (RED_THREADS / GL)

// gtab[b] = (structure, first atom) of workgroup b; sblk[s] = (first workgroup, workgroups) of structure s
__global__ __launch_bounds__(RED_THREADS) void k_force_gather(
    const int2* __restrict__ gtab, const int2* __restrict__ sblk, const int* __restrict__ atom_off,
    const int* __restrict__ natoms_of, const int* __restrict__ nb_off, const int* __restrict__ nb_rev,
    const float4* __restrict__ nb_d, const float4* __restrict__ pair_g, const double* __restrict__ e_atom,
    double* __restrict__ F, double* __restrict__ part, int* __restrict__ ticket, double* __restrict__ E,
    double* __restrict__ vir) {
    __shared__ double s_v[GATHER_ATOMS][10];
    __shared__ int s_last;
    const int2 gb = gtab[blockIdx.x];
    const int st = gb.x;
    const int aend = atom_off[st] + natoms_of[st];
    const int i = gb.y + threadIdx.x / GL, sub = threadIdx.x % GL;
    const bool on = i < aend;
    double f[3] = {0, 0, 0};
    double w[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (on) {
        const int p1 = nb_off[i + 1];
        for (int p0 = nb_off[i] + sub; p0 < p1; p0 += GL * GB) {
            int rv[GB];
            float4 g[GB], d[GB], gr[GB];
#pragma unroll
            for (int u = 0; u < GB; ++u) {
                const int p = p0 + GL * u;
                rv[u] = (p < p1) ? nb_rev[p] : 0;
                g[u] = (p < p1) ? pair_g[p] : make_float4(0.f, 0.f, 0.f, 0.f);
                d[u] = (p < p1) ? nb_d[p] : make_float4(0.f, 0.f, 0.f, 0.f);
            }
#pragma unroll
            for (int u = 0; u < GB; ++u) gr[u] = (p0 + GL * u < p1) ? pair_g[rv[u]] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int u = 0; u < GB; ++u) {
                f[0] += (double)gr[u].x - (double)g[u].x;
                f[1] += (double)gr[u].y - (double)g[u].y;
                f[2] += (double)gr[u].z - (double)g[u].z;
                const double ga[3] = {g[u].x, g[u].y, g[u].z}, da[3] = {d[u].x, d[u].y, d[u].z};
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int b = 0; b < 3; ++b) w[3 * a + b] -= ga[a] * da[b];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) f[q] = t_group_sum<double>(f[q], GL);
#pragma unroll
    for (int q = 0; q < 9; ++q) w[q] = t_group_sum<double>(w[q], GL);
    const int la = threadIdx.x / GL;
    if (sub == 0) {
        if (on) {
            F[3 * i] = f[0];
            F[3 * i + 1] = f[1];
            F[3 * i + 2] = f[2];
        }
        s_v[la][0] = on ? e_atom[i] : 0.0;
#pragma unroll
        for (int q = 0; q < 9; ++q) s_v[la][1 + q] = on ? w[q] : 0.0;
    }
    __syncthreads();
    if (threadIdx.x < 10) {
        double acc = 0.0;
        for (int a = 0; a < GATHER_ATOMS; ++a) acc += s_v[a][threadIdx.x];
        part[10 * (long long)blockIdx.x + threadIdx.x] = acc;
    }
    __threadfence();
    __syncthreads();
    const int2 sb = sblk[st];
    if (threadIdx.x == 0) {
        const int last = atomicAdd(&ticket[st], 1) == sb.y - 1;
        if (last) ticket[st] = 0;
        s_last = last;
    }
    __syncthreads();
    if (s_last) {
        // thread t sums partials t, t + 256, ... in order; then fixed-order wave (DPP) and block sums
        __threadfence();
        double acc[10];
#pragma unroll
        for (int q = 0; q < 10; ++q) acc[q] = 0.0;
        for (int b = threadIdx.x; b < sb.y; b += RED_THREADS) {
            const double* pb = part + 10 * (long long)(sb.x + b);
#pragma unroll
            for (int q = 0; q < 10; ++q) acc[q] += pb[q];
        }
        const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
        for (int q = 0; q < 10; ++q) {
            const double v = t_wave_sum63<double>(acc[q]);
            if (lane == 63) s_v[wv][q] = v;
        }
        __syncthreads();
        if (threadIdx.x < 10) {
            double v = 0.0;
            for (int w = 0; w < RED_THREADS / 64; ++w) v += s_v[w][threadIdx.x];
            if (threadIdx.x == 0) E[st] = v;
            else vir[9 * st + threadIdx.x - 1] = v;
        }
    }
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
