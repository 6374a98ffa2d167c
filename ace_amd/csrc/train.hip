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
    bool fwd_ok = false;             // ace_forward ran; valid while ctx->seq == fwd_seq
    unsigned long long fwd_seq = 0;
    int tm = 16;                     // rows per training tile
    std::vector<int4> tiles;
    std::vector<int> tile_first;     // first tile of each species (ntypes + 1)
    DBuf<int4> d_tiles;
    DBuf<double> d_Eref, d_Fref, d_gfac, d_eps, d_out, d_grad, d_gsum;
    DBuf<float> d_V, d_part;
    int ldpart = 0;
    bool use_graph = false;           // replay the device part of a step from a captured HIP graph
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    long long graph_sig[2][12] = {};
    // device optimiser: FP64 master parameters (species concatenated), ADAM moments, (beta1^t, beta2^t)
    bool opt_on = false;
    bool opt_host_stale = false;      // the device parameters are newer than the host copies
    bool opt_dev_stale = false;       // the host parameters were set after the device ones
    double opt_gamma = 0, opt_eps = 1e-16;
    int opt_n = 0;
    DBuf<double> d_p, d_m, d_v, d_bt;
    DBuf<int> d_pmap;
    DBuf<int> d_adone;               // completion counter of the ADAM launch (left at zero)
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

// the per-structure loss terms from the energy error and the summed force term: gfac = dL/dE, out = (energy
// loss, force loss, (dE / N)^2)
__device__ inline void loss_terms(int st, int n, double dE, double lf, const ace_loss_desc& Ld, double* gfac,
                                  double* out) {
    const double nA = n;
    double nf = 1.0;
    switch (Ld.norm) {                       // nnptefr.cpp:609-632
        case ACE_NORM_LINEAR: nf = 1.0 / nA; break;
        case ACE_NORM_SQRT: nf = 1.0 / sqrt(nA); break;
        case ACE_NORM_CBRT: nf = 1.0 / cbrt(nA); break;
        case ACE_NORM_LOG: nf = 1.0 / (log(nA) + 1.0); break;
        default: nf = 1.0;
    }
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
    out[3 * st + 1] = Ld.w_force * lf;
    out[3 * st + 2] = dE * dE / (nA * nA);
}

// fixed-order FP64 tree over the block
#define LOSS_THREADS 256
__device__ inline double loss_block_sum(double* s, double v) {
    const int t = threadIdx.x;
    s[t] = v;
    __syncthreads();
    for (int h = LOSS_THREADS / 2; h > 0; h >>= 1) {
        if (t < h) s[t] += s[t + h];
        __syncthreads();
    }
    return s[0];
}

// Block per structure, for the step with forces: the structure's fixed-point force, energy and virial sums
// converted to FP64 (as k_finalize), then the loss on them -- force residual eps = w_force (F - F_ref) and the
// force loss in the same pass over the forces, the energy terms and gfac -- in one launch.
__global__ __launch_bounds__(LOSS_THREADS) void k_finalize_loss(
    const int* __restrict__ atom_off, const int* __restrict__ natoms_of, const unsigned long long* __restrict__ F_fix,
    const unsigned long long* __restrict__ E_fix, const unsigned long long* __restrict__ W_fix, double* __restrict__ F,
    double* __restrict__ E, double* __restrict__ W, const double* __restrict__ E_ref, const double* __restrict__ F_ref,
    ace_loss_desc Ld, double* __restrict__ gfac, double* __restrict__ eps, double* __restrict__ out /*S x 3*/) {
    __shared__ double s[LOSS_THREADS];
    constexpr double inv = 1.0 / 17179869184.0;
    const int st = blockIdx.x, t = threadIdx.x;
    const int a0 = atom_off[st], n = natoms_of[st];
    double lf = 0;
    for (int q = t; q < 3 * n; q += LOSS_THREADS) {
        const double f = (double)(long long)F_fix[3 * a0 + q] * inv;
        F[3 * a0 + q] = f;
        const double r = f - F_ref[3 * a0 + q];
        lf += 0.5 * r * r;
        eps[3 * a0 + q] = Ld.w_force * r;
    }
    if (t < 9) W[9 * st + t] = (double)(long long)W_fix[9 * st + t] * inv;
    const double tot = loss_block_sum(s, lf);
    if (t == 0) {
        const double e = (double)(long long)E_fix[st] * inv;
        E[st] = e;
        loss_terms(st, n, e - E_ref[st], tot, Ld, gfac, out);
    }
}

// Block per structure, for the step without forces: E_s = sum of the atomic energies (the fixed-order tree of
// k_struct_reduce), then the energy loss terms and gfac.
__global__ __launch_bounds__(LOSS_THREADS) void k_reduce_loss(const int* __restrict__ atom_off,
                                                              const int* __restrict__ natoms_of,
                                                              const double* __restrict__ e_atom, double* __restrict__ E,
                                                              const double* __restrict__ E_ref, ace_loss_desc Ld,
                                                              double* __restrict__ gfac, double* __restrict__ out) {
    __shared__ double s[LOSS_THREADS];
    const int st = blockIdx.x, t = threadIdx.x;
    const int a0 = atom_off[st], n = natoms_of[st];
    double acc = 0.0;
    for (int k = t; k < n; k += LOSS_THREADS) acc += e_atom[a0 + k];
    const double e = loss_block_sum(s, acc);
    if (t == 0) {
        E[st] = e;
        loss_terms(st, n, e - E_ref[st], 0.0, Ld, gfac, out);
    }
}

// ---- matrix-core training pass -------------------------------------------------------------------
// One workgroup of four waves per tile of 16 atoms of one species (the v_mfma_f32_16x16x4f32 helpers and the
// zero-padded weight layout of the inference network kernel); the 16-column tiles of every product are spread
// over the waves, with one barrier per phase.  Per layer l (z = W a + b, a' = f(z)):
//   forward, primal and tangent:  Z = A_{l-1} W^T, Zt = At_{l-1} W^T; A = f(Z), D = f'(Z), D2 = f''(Z),
//                                  At = D Zt
//   backward from the seeds (primal adjoint gfac outw, tangent adjoint -outw at the linear output):
//     zb = D2 Zt atb + D ab,  ztb = D atb                        (elementwise, into D2 and Zt)
//     d bias = sum_rows zb,  d W = zb^T A_{l-1} + ztb^T At_{l-1}  (16 x 16 tiles, K = the 16 rows)
//     ab_{l-1} = zb W,  atb_{l-1} = ztb W                         (into A_{l-1}, At_{l-1})
// LDS rows have odd strides (mlp_ld) and start zeroed, so padding rows and columns contribute nothing; rows
// past the tile's atoms get zero seeds.  The tile's gradient goes to part (reference parameter order).
__host__ __device__ inline int train_mc_lds_floats(const DevSpecies& sp) {
    int n = 2 * MLP_ROWS * mlp_ld(sp.K);
    for (int l = 0; l < sp.L; ++l) n += 5 * MLP_ROWS * mlp_ld(sp.width[l + 1]);
    return n;
}

// D[o][k] (16 x 16 tile at (o0, k0)) = sum_{r < 16} P[r][o] Q[r][k] + Pt[r][o] Qt[r][k]; P, Pt rows of stride
// ldo, Q, Qt rows of stride ldi, all in LDS.  A operand = P^T: lane l supplies A[o0 + l % 16][4 u + l / 16].
__device__ inline mfma_f4 mfma_wgrad(int P, int Pt, int ldo, int Q, int Qt, int ldi, int o0, int k0, int lane) {
    const int c = lane & 15, q = lane >> 4;
    mfma_f4 a0 = mfma_f4{0.f, 0.f, 0.f, 0.f}, a1 = mfma_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int r = 4 * u + q;
        a0 = __builtin_amdgcn_mfma_f32_16x16x4f32(mlp_smem[P + r * ldo + o0 + c], mlp_smem[Q + r * ldi + k0 + c], a0, 0, 0, 0);
        a1 = __builtin_amdgcn_mfma_f32_16x16x4f32(mlp_smem[Pt + r * ldo + o0 + c], mlp_smem[Qt + r * ldi + k0 + c], a1, 0, 0, 0);
    }
    return a0 + a1;
}

#define TRAIN_WAVES 4

// the 16-column output tiles of a product, tile t on wave t % TRAIN_WAVES (one matrix-core chain per tile)
template <class EPI>
__device__ inline void mfma_tiles(int a, int lda, int k, const float* __restrict__ Bg, int so, int sk, int n, int lane,
                                  int wv, EPI epi) {
    for (int o0 = 16 * wv; o0 < n; o0 += 16 * TRAIN_WAVES) {
        mfma_f4 acc[1];
        mfma_group<false, 1>(a, lda, k, 0, Bg, so, sk, n, o0, lane, acc);
        epi(acc[0], o0);
    }
}

__global__ __launch_bounds__(64 * TRAIN_WAVES) void k_mlp_train_mc(
    const DevModel* __restrict__ M, const float* __restrict__ wpad, const float* __restrict__ inpw,
    const int4* __restrict__ tiles, int ldg, const float* __restrict__ G, const float* __restrict__ V,
    const int* __restrict__ atom_of_row, const int* __restrict__ struc_of, const double* __restrict__ gfac,
    int ldpart, float* __restrict__ part) {
    constexpr int NTH = 64 * TRAIN_WAVES;
    __shared__ DevSpecies s_sp;
    __shared__ int s_off[5 * ACE_MAXL + 2];
    const int4 tile = tiles[blockIdx.x];
    const int I = tile.x, r0 = tile.y, nr = tile.z;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (tid < 64) {
        constexpr int n16 = (int)(sizeof(DevSpecies) / 16);
        static_assert(n16 >= 64 && n16 <= 128, "DevSpecies copy covers 64 to 128 words");
        const int4* src = (const int4*)&M->sp[I];
        int4* dst = (int4*)&s_sp;
        const int4 v0 = src[lane];
        const int4 v1 = (64 + lane < n16) ? src[64 + lane] : v0;
        dst[lane] = v0;
        if (64 + lane < n16) dst[64 + lane] = v1;
    }
    __syncthreads();
    const DevSpecies& sp = s_sp;
    const int K = sp.K, L = sp.L, ldx = mlp_ld(K);
    const bool tan = (V != nullptr);
    if (tid == 0) {
        int q = 2 * MLP_ROWS * ldx;
        for (int l = 0; l < L; ++l) {
            const int w = MLP_ROWS * mlp_ld(sp.width[l + 1]);
            for (int j = 0; j < 5; ++j) s_off[5 * l + j] = q + j * w;   // A, D, D2, At, Zt
            q += 5 * w;
        }
        s_off[5 * ACE_MAXL] = q;
    }
    __syncthreads();
    const int total = s_off[5 * ACE_MAXL];
    for (int e = tid; e < total; e += NTH) mlp_smem[e] = 0.f;
    __syncthreads();
    const int X = 0, Xt = MLP_ROWS * ldx;
    const float* iw = inpw + sp.in_off;
    for (int e = tid; e < MLP_ROWS * K; e += NTH) {
        const int r = e / K, k = e - r * K;
        if (r < nr) {
            const long long gi = (long long)(r0 + r) * ldg + k;
            mlp_smem[X + r * ldx + k] = G[gi];                    // rows hold inpw (G + inpb)
            if (tan) mlp_smem[Xt + r * ldx + k] = iw[k] * V[gi];
        }
    }
    __syncthreads();
    const float* Wg = wpad + sp.wp_off;
    const int cr = lane & 15, rq = 4 * (lane >> 4);
    // forward, primal and tangent (the primal epilogue of a tile reads back the tangent values the same
    // lanes wrote for that tile)
    int ain = X, atin = Xt, lda = ldx;
    for (int l = 0; l < L; ++l) {
        const int no = sp.width[l + 1], ni = sp.width[l], ldo = mlp_ld(no);
        const int Ao = s_off[5 * l], Do = s_off[5 * l + 1], D2o = s_off[5 * l + 2], Ato = s_off[5 * l + 3],
                  Zto = s_off[5 * l + 4];
        const int act = (l == L - 1) ? ACE_ACT_LINEAR : sp.act;
        const float* bl = Wg + sp.bl_off[l];
        mfma_tiles(atin, lda, ni, Wg + sp.wl_off[l], 1, ldo, no, lane, wv, [&](const mfma_f4& acc, int o0) {
            const int o = o0 + cr;
            if (o < no)
#pragma unroll
                for (int r = 0; r < 4; ++r) mlp_smem[Zto + (rq + r) * ldo + o] = acc[r];
        });
        mfma_tiles(ain, lda, ni, Wg + sp.wl_off[l], 1, ldo, no, lane, wv, [&](const mfma_f4& acc, int o0) {
            const int o = o0 + cr;
            if (o < no) {
                const float bo = bl[o];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int e = (rq + r) * ldo + o;
                    float a, d, d2;
                    act_fd2(act, sp.c, acc[r] + bo, a, d, d2);
                    mlp_smem[Ao + e] = a;
                    mlp_smem[Do + e] = d;
                    mlp_smem[D2o + e] = d2;
                    mlp_smem[Ato + e] = d * mlp_smem[Zto + e];
                }
            }
        });
        __syncthreads();
        ain = Ao;
        atin = Ato;
        lda = ldo;
    }
    // seeds at the linear output layer: primal adjoint gfac outw, tangent adjoint -outw
    {
        const int ldL = mlp_ld(1), AL = s_off[5 * (L - 1)], AtL = s_off[5 * (L - 1) + 3];
        if (tid < MLP_ROWS) {
            const bool on = tid < nr;
            const float g = on ? (float)gfac[struc_of[atom_of_row[r0 + tid]]] : 0.f;
            mlp_smem[AL + tid * ldL] = g * sp.outw;
            mlp_smem[AtL + tid * ldL] = (tan && on) ? -sp.outw : 0.f;
        }
    }
    __syncthreads();
    float* out = part + (long long)blockIdx.x * ldpart;
    for (int l = L - 1; l >= 0; --l) {
        const int no = sp.width[l + 1], ni = sp.width[l], ldo = mlp_ld(no);
        const int Ao = s_off[5 * l], Do = s_off[5 * l + 1], D2o = s_off[5 * l + 2], Ato = s_off[5 * l + 3],
                  Zto = s_off[5 * l + 4];
        // zb (into D2), ztb (into Zt)
        for (int e = tid; e < MLP_ROWS * no; e += NTH) {
            const int r = e / no, o = e - r * no, q = r * ldo + o;
            const float ab = mlp_smem[Ao + q], atb = mlp_smem[Ato + q], d = mlp_smem[Do + q];
            mlp_smem[D2o + q] = mlp_smem[D2o + q] * mlp_smem[Zto + q] * atb + d * ab;
            mlp_smem[Zto + q] = d * atb;
        }
        __syncthreads();
        // bias gradient: column sums of zb
        for (int o = tid; o < no; o += NTH) {
            float sb = 0.f;
#pragma unroll
            for (int r = 0; r < MLP_ROWS; ++r) sb += mlp_smem[D2o + r * ldo + o];
            out[sp.boff[l] + o] = sb;
        }
        // weight gradient zb^T A_{l-1} + ztb^T At_{l-1}, written as W (no x ni) column-major; 16 x 16 tiles
        // over the waves
        const int in = (l == 0) ? X : s_off[5 * (l - 1)], int_ = (l == 0) ? Xt : s_off[5 * (l - 1) + 3];
        const int ldi = (l == 0) ? ldx : mlp_ld(ni);
        const int nob = (no + 15) >> 4, nkb = (ni + 15) >> 4;
        for (int t = wv; t < nob * nkb; t += TRAIN_WAVES) {
            const int o0 = 16 * (t % nob), k0 = 16 * (t / nob);
            const mfma_f4 acc = mfma_wgrad(D2o, Zto, ldo, in, int_, ldi, o0, k0, lane);
            const int k = k0 + cr;
            if (k < ni)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int o = o0 + rq + r;
                    if (o < no) out[sp.woff[l] + o + k * no] = acc[r];
                }
        }
        __syncthreads();
        if (l > 0) {
            // ab_{l-1} = zb W, atb_{l-1} = ztb W: B[o][k] = W_l[o][k] (wpad rows k, stride ldo)
            const int ldm = mlp_ld(ni), Am = s_off[5 * (l - 1)], Atm = s_off[5 * (l - 1) + 3];
            mfma_tiles(D2o, ldo, no, Wg + sp.wl_off[l], ldo, 1, ni, lane, wv, [&](const mfma_f4& acc, int k0) {
                const int k = k0 + cr;
                if (k < ni)
#pragma unroll
                    for (int r = 0; r < 4; ++r) mlp_smem[Am + (rq + r) * ldm + k] = acc[r];
            });
            mfma_tiles(Zto, ldo, no, Wg + sp.wl_off[l], ldo, 1, ni, lane, wv, [&](const mfma_f4& acc, int k0) {
                const int k = k0 + cr;
                if (k < ni)
#pragma unroll
                    for (int r = 0; r < 4; ++r) mlp_smem[Atm + (rq + r) * ldm + k] = acc[r];
            });
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
