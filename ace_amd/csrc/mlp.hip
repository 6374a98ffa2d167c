// mlp.hip -- per-species atomic networks over species-sorted atom tiles.
//
// Forward: ANN::fpbp (src/ml/nn.cpp:1402-1420): ins = inpw*(G+inpb) (applied by the basis kernel when it
// writes the rows), z_l = b_l + W_l a_{l-1},
// a_l = f(c z_l) (hidden), last layer linear, out = outb + outw*a_last.
// Backward: DODZ::grad row 0 (src/ml/nn.cpp:1836-1862): dE/dG = dO/dz_0 W_0 diag(inpw).
// A workgroup holds a tile of TM atoms of one species; activations and derivatives live in LDS.
#include "ace_common.hpp"

#define MLP_THREADS 256

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

// dynamic LDS: x [TM][K], then per layer l: a_l [TM][w_{l+1}], d_l [TM][w_{l+1}]
__host__ __device__ inline int mlp_lds_floats(const DevSpecies& sp, int tm) {
    int n = tm * sp.K;
    for (int l = 0; l < sp.L; ++l) n += 2 * tm * sp.width[l + 1];
    return n;
}

// tiles: (species, first row, rows, 0)
__global__ __launch_bounds__(MLP_THREADS) void k_mlp_energy_grad(
    const DevModel* __restrict__ M, const float* __restrict__ params, const float* __restrict__ inpw,
    const float* __restrict__ inpb, const int4* __restrict__ tiles, int ldg, const float* __restrict__ G,
    float* __restrict__ dEdG, const int* __restrict__ atom_of_row, double* __restrict__ e_atom, int tm) {
    extern __shared__ float smem[];
    const int4 tile = tiles[blockIdx.x];
    const int I = tile.x, r0 = tile.y, nr = tile.z;
    const DevSpecies& sp = M->sp[I];
    const int K = sp.K, L = sp.L;
    float* X = smem;
    float* A[ACE_MAXL];
    float* D[ACE_MAXL];
    {
        float* q = smem + tm * K;
        for (int l = 0; l < L; ++l) {
            A[l] = q; q += tm * sp.width[l + 1];
            D[l] = q; q += tm * sp.width[l + 1];
        }
    }
    const float* p = params + sp.poff;
    const float* iw = inpw + sp.in_off;
    const float* ib = inpb + sp.in_off;
    const int tid = threadIdx.x;
    (void)ib;
    for (int idx = tid; idx < tm * K; idx += MLP_THREADS) {   // rows already hold inpw*(G + inpb)
        const int r = idx / K, k = idx - r * K;
        X[r * K + k] = (r < nr) ? G[(long long)(r0 + r) * ldg + k] : 0.f;
    }
    __syncthreads();
    for (int l = 0; l < L; ++l) {
        const int no = sp.width[l + 1], ni = sp.width[l];
        const float* W = p + sp.woff[l];
        const float* b = p + sp.boff[l];
        const int act = (l == L - 1) ? ACE_ACT_LINEAR : sp.act;
        for (int idx = tid; idx < tm * no; idx += MLP_THREADS) {
            const int r = idx / no, o = idx - r * no;
            float z = b[o];
            if (l == 0) {
                for (int k = 0; k < ni; ++k) z += W[o + k * no] * X[r * ni + k];
            } else {
                for (int k = 0; k < ni; ++k) z += W[o + k * no] * A[l - 1][r * ni + k];
            }
            float a, d;
            act_fd(act, sp.c, z, a, d);
            A[l][r * no + o] = a;
            D[l][r * no + o] = d;
        }
        __syncthreads();
    }
    // energies
    for (int r = tid; r < nr; r += MLP_THREADS) {
        const float out = sp.outb + sp.outw * A[L - 1][r];
        e_atom[atom_of_row[r0 + r]] = (double)out + sp.E0;
    }
    __syncthreads();
    // backward: delta of the output layer, then hidden layers; deltas overwrite a[]
    for (int r = tid; r < tm; r += MLP_THREADS) A[L - 1][r] = sp.outw * D[L - 1][r];
    __syncthreads();
    for (int l = L - 1; l > 0; --l) {
        const int no = sp.width[l + 1], ni = sp.width[l];
        const float* W = p + sp.woff[l];
        for (int idx = tid; idx < tm * ni; idx += MLP_THREADS) {
            const int r = idx / ni, k = idx - r * ni;
            float s = 0.f;
            for (int o = 0; o < no; ++o) s += W[o + k * no] * A[l][r * no + o];
            A[l - 1][r * ni + k] = s * D[l - 1][r * ni + k];
        }
        __syncthreads();
    }
    {
        const int no = sp.width[1];
        const float* W = p + sp.woff[0];
        for (int idx = tid; idx < tm * K; idx += MLP_THREADS) {
            const int r = idx / K, k = idx - r * K;
            if (r >= nr) continue;
            float s = 0.f;
            for (int o = 0; o < no; ++o) s += W[o + k * no] * A[0][r * no + o];
            dEdG[(long long)(r0 + r) * ldg + k] = s * iw[k];
        }
    }
}
