// train.hip -- kernels of the batched training step.
//
// Energy term: NNPTEFR::error_cost_o1 (src/nnp/nnptefr.cpp:1336-1382): dU = (E - E_ref) * norm(n),
// loss(dU), gfac = w_E * loss'(dU) * norm, gradient gfac * sum_i dE_i/dW (NN::Cost::grad with dcdo=1,
// src/ml/nn.cpp:1747-1808).
// Force term (no reference counterpart; its network factor is NN::D2ODZDI, src/ml/nn.cpp:2047-2140):
// L_F = w_F * sum_a 0.5 |F_a - F_ref_a|^2, eps_a = dL/dF_a, v_i = sum_a eps_a . dG_i/dr_a (basis
// tangent mode, dG/dr recomputed), and dL_F/dW = -sum_i d/dW (dE_i/dG . v_i) by forward-over-reverse
// through the atomic network.  Per atom the network sweep differentiates J_i = gfac E_i - Q_i,
// Q_i = dE_i/dG . v_i, so both terms share one reverse pass.
#include <vector>
#include "ace_common.hpp"

struct TrainState {
    bool have_targets = false;
    bool have_forces = false;
    bool g_valid = false;            // G rows of the current positions are in HBM
    int tm = 16;                     // rows per training tile
    std::vector<int4> tiles;
    std::vector<int> tile_first;     // first tile of each species (ntypes + 1)
    DBuf<int4> d_tiles;
    DBuf<double> d_Eref, d_Fref, d_gfac, d_eps, d_out, d_grad, d_gsum;
    DBuf<float> d_V, d_part;
    int ldpart = 0;
};

__device__ inline void act_fd2(int act, float c, float z, float& a, float& d, float& d2) {
    switch (act) {
        case ACE_ACT_TANH: {
            const float t = tanhf(c * z);
            a = t;
            d = c * (1.f - t * t);
            d2 = -2.f * c * c * t * (1.f - t * t);
        } break;
        case ACE_ACT_SIGMOID: {
            const float e = fast_exp(-c * fabsf(z));
            const float f = 1.f / (1.f + e);
            a = (z >= 0.f) ? f : e * f;
            d = c * e * f * f;
            d2 = (z >= 0.f) ? -c * c * e * (1.f - e) * f * f * f : -c * c * e * (e - 1.f) * f * f * f;
        } break;
        case ACE_ACT_SWISH: {
            const float zz = c * z;
            const float e = fast_exp(-fabsf(zz));
            const float den = 1.f / (1.f + e);
            if (zz > 0.f) {
                a = z * den;
                d = (1.f + e * (1.f + zz)) * den * den;
                d2 = c * e * (2.f - zz + e * (2.f + zz)) * den * den * den;
            } else {
                a = z * e * den;
                d = e * (1.f + zz + e) * den * den;
                d2 = c * e * (2.f + zz + e * (2.f - zz)) * den * den * den;
            }
        } break;
        case ACE_ACT_SOFTPLUS: {
            const float e = fast_exp(-c * fabsf(z));
            const float den = 1.f / (e + 1.f);
            const float l = (log1pf(e) - 0.693147180559945f) / c;
            a = (z >= 0.f) ? z + l : l;
            d = (z >= 0.f) ? den : e * den;
            d2 = c * e * den * den;
        } break;
        default: a = z; d = 1.f; d2 = 0.f;
    }
}

// block per structure: energy loss terms, gfac; force residual eps and force loss (FP64)
#define LOSS_THREADS 256
__global__ __launch_bounds__(LOSS_THREADS) void k_loss(
    const int* __restrict__ atom_off, const int* __restrict__ natoms_of, const double* __restrict__ E,
    const double* __restrict__ E_ref, const double* __restrict__ F, const double* __restrict__ F_ref,
    ace_loss_desc Ld, double* __restrict__ gfac, double* __restrict__ eps, double* __restrict__ out /*S x 3*/) {
    __shared__ double s[LOSS_THREADS];
    const int st = blockIdx.x, t = threadIdx.x;
    const int a0 = atom_off[st], n = natoms_of[st];
    double lf = 0;
    if (F_ref) {
        for (int q = t; q < 3 * n; q += LOSS_THREADS) {
            const double r = F[3 * a0 + q] - F_ref[3 * a0 + q];
            lf += 0.5 * r * r;
            eps[3 * a0 + q] = Ld.w_force * r;
        }
    }
    s[t] = lf;
    __syncthreads();
    for (int h = LOSS_THREADS / 2; h > 0; h >>= 1) {
        if (t < h) s[t] += s[t + h];
        __syncthreads();
    }
    if (t == 0) {
        const double nA = n;
        double nf = 1.0;
        switch (Ld.norm) {                       // nnptefr.cpp:609-632
            case ACE_NORM_LINEAR: nf = 1.0 / nA; break;
            case ACE_NORM_SQRT: nf = 1.0 / sqrt(nA); break;
            case ACE_NORM_CBRT: nf = 1.0 / cbrt(nA); break;
            case ACE_NORM_LOG: nf = 1.0 / (log(nA) + 1.0); break;
            default: nf = 1.0;
        }
        const double dE = E[st] - E_ref[st];
        const double dU = dE * nf;
        double g = nf, le = 0;
        switch (Ld.loss) {                       // nnptefr.cpp:1353-1378
            case ACE_LOSS_MAE: le = fabs(dU); g *= (dU > 0) - (dU < 0); break;
            case ACE_LOSS_HUBER: {
                const double arg = dU / Ld.delta, sq = sqrt(1.0 + arg * arg);
                le = Ld.delta * Ld.delta * (sq - 1.0);
                g *= dU / sq;
            } break;
            case ACE_LOSS_ASINH: {
                const double arg = dU / Ld.delta, sq = sqrt(1.0 + arg * arg), lg = log(arg + sq);
                le = Ld.delta * Ld.delta * (1.0 - sq + arg * lg);
                g *= lg * Ld.delta;
            } break;
            default: le = 0.5 * dU * dU; g *= dU;
        }
        gfac[st] = Ld.w_energy * g;
        out[3 * st] = Ld.w_energy * le;
        out[3 * st + 1] = Ld.w_force * s[0];
        out[3 * st + 2] = dE * dE / (nA * nA);
    }
}

// dynamic LDS per tile: X, Xt [tm][K]; per layer A, D, D2, At, Zt [tm][w_{l+1}]
__host__ __device__ inline int train_lds_floats(const DevSpecies& sp, int tm) {
    int n = 2 * tm * sp.K;
    for (int l = 0; l < sp.L; ++l) n += 5 * tm * sp.width[l + 1];
    return n;
}

// tile of tm atoms of one species: per-tile partial gradient of sum_i (gfac E_i - dE_i/dG . v_i)
__global__ __launch_bounds__(256) void k_mlp_train(
    const DevModel* __restrict__ M, const float* __restrict__ params, const float* __restrict__ inpw,
    const float* __restrict__ inpb, const int4* __restrict__ tiles, int ldg, const float* __restrict__ G,
    const float* __restrict__ V, const int* __restrict__ atom_of_row, const int* __restrict__ struc_of,
    const double* __restrict__ gfac, int tm, int ldpart, float* __restrict__ part) {
    extern __shared__ float smem[];
    const int4 tile = tiles[blockIdx.x];
    const int I = tile.x, r0 = tile.y, nr = tile.z;
    const DevSpecies& sp = M->sp[I];
    const int K = sp.K, L = sp.L;
    const bool tan = (V != nullptr);
    float* X = smem;
    float* Xt = X + tm * K;
    float *A[ACE_MAXL], *D[ACE_MAXL], *D2[ACE_MAXL], *At[ACE_MAXL], *Zt[ACE_MAXL];
    {
        float* q = Xt + tm * K;
        for (int l = 0; l < L; ++l) {
            const int w = tm * sp.width[l + 1];
            A[l] = q; q += w;
            D[l] = q; q += w;
            D2[l] = q; q += w;
            At[l] = q; q += w;
            Zt[l] = q; q += w;
        }
    }
    const float* p = params + sp.poff;
    const float* iw = inpw + sp.in_off;
    const float* ib = inpb + sp.in_off;
    const int tid = threadIdx.x;
    for (int idx = tid; idx < tm * K; idx += 256) {
        const int r = idx / K, k = idx - r * K;
        const long long gi = (long long)(r0 + r) * ldg + k;
        X[idx] = (r < nr) ? G[gi] : 0.f;          // rows already hold inpw*(G + inpb)
        Xt[idx] = (r < nr && tan) ? iw[k] * V[gi] : 0.f;
    }
    __syncthreads();
    // forward with tangent
    for (int l = 0; l < L; ++l) {
        const int no = sp.width[l + 1], ni = sp.width[l];
        const float* W = p + sp.woff[l];
        const float* b = p + sp.boff[l];
        const float* in = (l == 0) ? X : A[l - 1];
        const float* int_ = (l == 0) ? Xt : At[l - 1];
        const int act = (l == L - 1) ? ACE_ACT_LINEAR : sp.act;
        for (int idx = tid; idx < tm * no; idx += 256) {
            const int r = idx / no, o = idx - r * no;
            float z = b[o], zt = 0.f;
            for (int k = 0; k < ni; ++k) {
                const float w = W[o + k * no];
                z += w * in[r * ni + k];
                zt += w * int_[r * ni + k];
            }
            float a, d, d2;
            act_fd2(act, sp.c, z, a, d, d2);
            A[l][idx] = a;
            D[l][idx] = d;
            D2[l][idx] = d2;
            Zt[l][idx] = zt;
            At[l][idx] = d * zt;
        }
        __syncthreads();
    }
    // seeds at the linear output layer: primal adjoint gfac*outw, tangent adjoint -outw
    for (int r = tid; r < tm; r += 256) {
        float g = 0.f;
        if (r < nr) g = (float)gfac[struc_of[atom_of_row[r0 + r]]];
        A[L - 1][r] = g * sp.outw;
        At[L - 1][r] = (tan && r < nr) ? -sp.outw : 0.f;
    }
    __syncthreads();
    float* out = part + (long long)blockIdx.x * ldpart;
    for (int l = L - 1; l >= 0; --l) {
        const int no = sp.width[l + 1], ni = sp.width[l];
        const float* W = p + sp.woff[l];
        // zb (into D2) and ztb (into Zt), elementwise
        for (int idx = tid; idx < tm * no; idx += 256) {
            const float ab = A[l][idx], atb = At[l][idx], d = D[l][idx];
            const float zb = D2[l][idx] * Zt[l][idx] * atb + d * ab;
            D2[l][idx] = zb;
            Zt[l][idx] = d * atb;
        }
        __syncthreads();
        const float* in = (l == 0) ? X : A[l - 1];
        const float* int_ = (l == 0) ? Xt : At[l - 1];
        for (int o = tid; o < no; o += 256) {
            float s = 0.f;
            for (int r = 0; r < tm; ++r) s += D2[l][r * no + o];
            out[sp.boff[l] + o] = s;
        }
        for (int idx = tid; idx < no * ni; idx += 256) {
            const int o = idx % no, k = idx / no;
            float s = 0.f;
            for (int r = 0; r < tm; ++r) s += D2[l][r * no + o] * in[r * ni + k] + Zt[l][r * no + o] * int_[r * ni + k];
            out[sp.woff[l] + idx] = s;
        }
        __syncthreads();
        if (l > 0) {
            for (int idx = tid; idx < tm * ni; idx += 256) {
                const int r = idx / ni, k = idx - r * ni;
                float s = 0.f, st = 0.f;
                for (int o = 0; o < no; ++o) {
                    const float w = W[o + k * no];
                    s += w * D2[l][r * no + o];
                    st += w * Zt[l][r * no + o];
                }
                A[l - 1][idx] = s;
                At[l - 1][idx] = st;
            }
            __syncthreads();
        }
    }
}

// Gradient of the batch from the per-tile partials (tile rows of ldpart floats), in two passes whose sums
// run in a fixed order: pass 1 adds the partials of GRED_GROUP consecutive tiles of one species (FP64),
// workgroups spread over (parameter chunk, tile group, species); pass 2 adds the groups of each parameter.
#define GRED_GROUP 16
struct GradRed {
    int nt;
    int np[ACE_MAXT];      // parameters of each species
    int off[ACE_MAXT];     // offset of the species in the gradient
    int t0[ACE_MAXT];      // its tiles [t0, t1)
    int t1[ACE_MAXT];
    long long gofs[ACE_MAXT];   // offset of its group sums [groups][np] in the pass-1 buffer
};

__global__ __launch_bounds__(256) void k_grad_reduce1(GradRed R, int ldpart, const float* __restrict__ part,
                                                      double* __restrict__ gsum) {
    const int I = blockIdx.z, q = blockIdx.x * 256 + threadIdx.x, gi = blockIdx.y;
    const int ta = R.t0[I] + gi * GRED_GROUP;
    if (q >= R.np[I] || ta >= R.t1[I]) return;
    const int n = min(GRED_GROUP, R.t1[I] - ta);
    float v[GRED_GROUP];
#pragma unroll
    for (int u = 0; u < GRED_GROUP; ++u) v[u] = (u < n) ? part[(long long)(ta + u) * ldpart + q] : 0.f;
    double s = 0;
#pragma unroll
    for (int u = 0; u < GRED_GROUP; ++u) s += (double)v[u];
    gsum[R.gofs[I] + (long long)gi * R.np[I] + q] = s;
}

__global__ __launch_bounds__(256) void k_grad_reduce2(GradRed R, const double* __restrict__ gsum,
                                                      double* __restrict__ grad) {
    const int I = blockIdx.z, q = blockIdx.x * 256 + threadIdx.x;
    if (q >= R.np[I]) return;
    const int ng = (R.t1[I] - R.t0[I] + GRED_GROUP - 1) / GRED_GROUP;
    double s = 0;
    for (int g = 0; g < ng; ++g) s += gsum[R.gofs[I] + (long long)g * R.np[I] + q];
    grad[R.off[I] + q] = s;
}
