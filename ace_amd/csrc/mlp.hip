// mlp.hip -- per-species atomic networks over species-sorted atom tiles.
//
// Forward: ANN::fpbp (src/ml/nn.cpp:1402-1420): ins = inpw*(G+inpb) (applied by the basis kernel when it
// writes the rows), z_l = b_l + W_l a_{l-1},
// a_l = f(c z_l) (hidden), last layer linear, out = outb + outw*a_last.
// Backward: DODZ::grad row 0 (src/ml/nn.cpp:1836-1862): dE/dG = dO/dz_0 W_0 diag(inpw).
// A workgroup holds a tile of TM atoms of one species; activations and derivatives live in LDS.
#include "ace_common.hpp"

#define MLP_THREADS 256

// ---- matrix-core network pass -------------------------------------------------------------------
// One wave per tile of 16 atoms of one species.  Every layer, forward and backward, is a chain of
// v_mfma_f32_16x16x4f32 (D[16x16] += A[16x4] B[4x16], FP32 products and sums):
//   lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16], and receives D[4 (l / 16) + r][l % 16] in
//   register r.  A (activations, deltas) comes from LDS; B (weights) from LDS when the species' weights
//   fit next to the activations (one batched copy), else from global memory.
// LDS rows have an odd stride (width rounded up to 16, plus 1) so the 16 rows of one A read fall in
// distinct banks; padding columns and the padded weight rows and columns hold zeros, so every operand
// load is unconditional and a k-step or column past the width adds nothing.  LDS is
// addressed through integer offsets into mlp_smem so every access compiles to a ds_ instruction.
typedef float mfma_f4 __attribute__((ext_vector_type(4)));

#define MLP_ROWS 16

extern __shared__ float mlp_smem[];


// floats of LDS for one tile: X [16][ld(K)], then per layer A_l, D_l [16][ld(w_{l+1})]
__host__ __device__ inline int mlp_lds_floats(const DevSpecies& sp) {
    int n = MLP_ROWS * mlp_ld(sp.K);
    for (int l = 0; l < sp.L; ++l) n += 2 * MLP_ROWS * mlp_ld(sp.width[l + 1]);
    return n;
}

// C[16 x n] = A[16 x k] * B for NT 16-wide output tiles starting at o0, o0 + ts, o0 + 2 ts, ... (ts = 16:
// adjacent tiles); A at LDS offset a (row stride lda), B[kk][o] at LDS offset b (BLDS) or global Bg, element
// o * so + kk * sk.  Two accumulators per tile (even and odd k-steps) give 2 NT independent matrix-core
// chains; the A operand of a k-step is shared by the NT tiles.  acc[t] receives the lane's 4 values of tile t.
template <bool BLDS, int NT>
__device__ inline void mfma_group(int a, int lda, int k, int b, const float* __restrict__ Bg, int so, int sk, int n,
                                  int o0, int lane, mfma_f4* acc, int ts = 16) {
    mfma_f4 c0[NT], c1[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        c0[t] = mfma_f4{0.f, 0.f, 0.f, 0.f};
        c1[t] = mfma_f4{0.f, 0.f, 0.f, 0.f};
    }
    const int r = lane & 15, q = lane >> 4;
    const int k16 = (k + 15) & ~15;   // rows [k, k16) of A and B are zero padding
    (void)n;
    for (int kb = 0; kb < k16; kb += 16) {
        float av[4], bv[4][NT];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int kk = kb + 4 * u + q;
            av[u] = mlp_smem[a + r * lda + kk];
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const int o = o0 + ts * t + r;
                if (BLDS) bv[u][t] = mlp_smem[b + o * so + kk * sk];
                else bv[u][t] = Bg[o * so + kk * sk];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; u += 2)
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                c0[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[u], bv[u][t], c0[t], 0, 0, 0);
                c1[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[u + 1], bv[u + 1][t], c1[t], 0, 0, 0);
            }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = c0[t] + c1[t];
}

// ---- BF16 first layer ---------------------------------------------------------------------------
// v_mfma_f32_16x16x32_bf16 (D[16x16] += A[16x32] B[32x16], FP32 sums): lane l supplies the 8 values
// A[l % 16][8 (l / 16) + i] and B[8 (l / 16) + i][l % 16], i = 0..7, and receives D as above.
// Each FP32 operand x is split into x_hi = bf16(x) and x_lo = bf16(x - x_hi), so x = x_hi + x_lo to
// about 2^-17 |x|, and the product is taken as A_hi B_hi + A_hi B_lo + A_lo B_hi (the dropped A_lo B_lo is
// ~2^-16 of it).  A single BF16 product (2^-9 relative per operand) leaves errors of ~1e-3 eV/A in the
// forces of a 128-wide layer; the three products keep the layer within ~1e-5 of FP32 at the BF16 rate.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ inline void bf_split8(const float* v, bf16x8& hi, bf16x8& lo) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const __bf16 h = (__bf16)v[i];
        hi[i] = h;
        lo[i] = (__bf16)(v[i] - (float)h);
    }
}

// C[16 x n] = A[16 x k] B for NT 16-wide output tiles from o0.  A: FP32 in LDS at offset a (row stride lda),
// columns >= k read as zero.  B: BF16 pairs in global memory, B[kk][o] = Bh[o * sb + kk] (high parts) and
// Bh[lo + o * sb + kk] (low parts), zero padded to sb = round32(k) values per row and to whole 16-row tiles,
// so each lane's 8 values are one 16-byte load.
template <int NT>
__device__ inline void bf3_group(int a, int lda, int k, const __bf16* __restrict__ Bh, int lo, int sb, int o0,
                                 int lane, mfma_f4* acc, int ts) {
    mfma_f4 c0[NT], c1[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        c0[t] = mfma_f4{0.f, 0.f, 0.f, 0.f};
        c1[t] = mfma_f4{0.f, 0.f, 0.f, 0.f};
    }
    const int r = lane & 15, q = lane >> 4;
    for (int kb = 0; kb < sb; kb += 32) {
        float av[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int kk = kb + 8 * q + i;
            const float v = mlp_smem[a + r * lda + min(kk, k - 1)];
            av[i] = kk < k ? v : 0.f;
        }
        bf16x8 ah, al;
        bf_split8(av, ah, al);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int e = (o0 + ts * t + r) * sb + kb + 8 * q;
            const bf16x8 bh = *(const bf16x8*)(Bh + e);
            const bf16x8 bl = *(const bf16x8*)(Bh + lo + e);
            c0[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, c0[t], 0, 0, 0);
            c1[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, c1[t], 0, 0, 0);
            c1[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, c1[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = c0[t] + c1[t];
}

// The 16-wide output tiles of a layer dealt to the W waves of a tile's workgroup: wave wv takes tiles wv,
// wv + W, wv + 2W, ... in groups of 4, 2 or 1; epi(acc, tile count, first column, column stride).
template <class EPI>
__device__ inline void bf3_layer(int a, int lda, int k, const __bf16* __restrict__ Bh, int lo, int n, int lane, EPI epi,
                                 int wv = 0, int W = 1) {
    const int sb = (k + 31) & ~31;
    const int ts = 16 * W;
    for (int o0 = 16 * wv; o0 < n;) {
        const int rem = (n - o0 + ts - 1) / ts;   // this wave's tiles from o0 on
        if (rem >= 4) {
            mfma_f4 acc[4];
            bf3_group<4>(a, lda, k, Bh, lo, sb, o0, lane, acc, ts);
            epi(acc, 4, o0, ts);
            o0 += 4 * ts;
        } else if (rem >= 2) {
            mfma_f4 acc[2];
            bf3_group<2>(a, lda, k, Bh, lo, sb, o0, lane, acc, ts);
            epi(acc, 2, o0, ts);
            o0 += 2 * ts;
        } else {
            mfma_f4 acc[1];
            bf3_group<1>(a, lda, k, Bh, lo, sb, o0, lane, acc, ts);
            epi(acc, 1, o0, ts);
            o0 += ts;
        }
    }
}

template <bool BLDS, class EPI>
__device__ inline void mfma_layer(int a, int lda, int k, int b, const float* __restrict__ Bg, int so, int sk, int n,
                                  int lane, EPI epi, int wv = 0, int W = 1) {
    const int ts = 16 * W;
    for (int o0 = 16 * wv; o0 < n;) {
        const int rem = (n - o0 + ts - 1) / ts;
        if (rem >= 4) {
            mfma_f4 acc[4];
            mfma_group<BLDS, 4>(a, lda, k, b, Bg, so, sk, n, o0, lane, acc, ts);
            epi(acc, 4, o0, ts);
            o0 += 4 * ts;
        } else if (rem >= 2) {
            mfma_f4 acc[2];
            mfma_group<BLDS, 2>(a, lda, k, b, Bg, so, sk, n, o0, lane, acc, ts);
            epi(acc, 2, o0, ts);
            o0 += 2 * ts;
        } else {
            mfma_f4 acc[1];
            mfma_group<BLDS, 1>(a, lda, k, b, Bg, so, sk, n, o0, lane, acc, ts);
            epi(acc, 1, o0, ts);
            o0 += ts;
        }
    }
}

// tiles: (species, first row, rows, 0), rows <= 16.  wpad: the parameters in the matrix-core layout
// (DevSpecies::wl_off / bl_off).  STAGE: the species' block is first copied to LDS with batched 16-byte
// loads (one memory round trip instead of one per k-step); otherwise it is read from global memory.
// BF: the first layer's forward product and the dE/dG product run on BF16 matrix cores (three-product
// split above) with the weights from w0bf (DevSpecies::bf_off); the other layers stay FP32.
// W (1, 2 or 4) waves share a tile: they stage its rows together and split every layer's 16-wide output
// tiles (wave w takes tiles w, w + W, ...), with a workgroup barrier between layers; more waves per tile
// keep more matrix-core chains and memory loads in flight per SIMD for the same LDS.
template <bool STAGE, bool BF, int W>
__global__ __launch_bounds__(64 * W) void k_mlp_energy_grad(
    const DevModel* __restrict__ M, const float* __restrict__ wpad, const __bf16* __restrict__ w0bf,
    const float* __restrict__ inpw, const int4* __restrict__ tiles, int ldg, const float* __restrict__ G,
    float* __restrict__ dEdG, const int* __restrict__ atom_of_row, double* __restrict__ e_atom) {
    static_assert(W == 1 || W == 2 || W == 4, "one, two or four waves per tile");
    constexpr int BT = 64 * W;
    __shared__ int s_off[2 * ACE_MAXL + 2];   // A_l, D_l offsets; weights offset
    __shared__ DevSpecies s_sp;                 // the species' tables, read from LDS in the layer loops
    const int4 tile = tiles[blockIdx.x];
    const int I = tile.x, r0 = tile.y, nr = tile.z;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = W > 1 ? __builtin_amdgcn_readfirstlane(tid >> 6) : 0;
    auto sync = [] {
        if constexpr (W > 1) __syncthreads();
        else wave_sync();
    };
    if (wv == 0) {
        constexpr int n16 = (int)(sizeof(DevSpecies) / 16);
        static_assert(n16 >= 64 && n16 <= 128, "DevSpecies copy covers 64 to 128 words");
        const int4* src = (const int4*)&M->sp[I];
        int4* dst = (int4*)&s_sp;
        const int4 v0 = src[lane];
        const int4 v1 = (64 + lane < n16) ? src[64 + lane] : v0;
        dst[lane] = v0;
        if (64 + lane < n16) dst[64 + lane] = v1;
    }
    sync();
    const DevSpecies& sp = s_sp;
    const int K = sp.K, L = sp.L;
    const int ldx = mlp_ld(K);
    if (tid == 0) {
        int q = MLP_ROWS * ldx;
        for (int l = 0; l < L; ++l) {
            const int w = MLP_ROWS * mlp_ld(sp.width[l + 1]);
            s_off[2 * l] = q;
            s_off[2 * l + 1] = q + w;
            q += 2 * w;
        }
        s_off[2 * ACE_MAXL] = q;   // 16-byte aligned: every ld is odd and the row count 16
    }
    sync();
    const int wofs = s_off[2 * ACE_MAXL];
    const float* Wg = wpad + sp.wp_off;
    if (STAGE) {
        const float4* src = (const float4*)Wg;
        float4* dst = (float4*)(mlp_smem + wofs);
        const int n4 = sp.wp_n >> 2;
        for (int b0 = 0; b0 < n4; b0 += BT * 8) {
            float4 v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int e = b0 + BT * u + tid;
                v[u] = (e < n4) ? src[e] : make_float4(0.f, 0.f, 0.f, 0.f);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int e = b0 + BT * u + tid;
                if (e < n4) dst[e] = v[u];
            }
        }
    }
    // input rows (already inpw (G + inpb)): 16-byte loads of the row prefix [0, round4(K)); the row tail
    // past K holds zeros in G; rows past the tile are zero
    {
        const int k4 = (K + 3) >> 2;
        const int n = MLP_ROWS * k4;
        for (int b0 = 0; b0 < n; b0 += BT * 8) {
            float4 v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int e = b0 + BT * u + tid;
                const int r = e / k4, c = e - r * k4;
                v[u] = (e < n && r < nr) ? *(const float4*)(G + (long long)(r0 + r) * ldg + 4 * c)
                                         : make_float4(0.f, 0.f, 0.f, 0.f);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int e = b0 + BT * u + tid;
                if (e < n) {
                    const int r = e / k4, c = e - r * k4;
                    const int x = r * ldx + 4 * c;
                    mlp_smem[x] = v[u].x; mlp_smem[x + 1] = v[u].y; mlp_smem[x + 2] = v[u].z; mlp_smem[x + 3] = v[u].w;
                }
            }
        }
        const int pad = ldx - 4 * k4;   // columns [4 k4, ldx) of every row
        for (int e = tid; e < MLP_ROWS * pad; e += BT) mlp_smem[(e / pad) * ldx + 4 * k4 + e % pad] = 0.f;
    }
    for (int idx = MLP_ROWS * ldx + tid; idx < wofs; idx += BT) mlp_smem[idx] = 0.f;
    sync();
    const int cr = lane & 15, rq = 4 * (lane >> 4);
    // forward: z_l = b_l + W_l a_{l-1}; W_l[o][k] at wl_off + k mlp_ld(no) + o
    int ain = 0, lda = ldx;
    for (int l = 0; l < L; ++l) {
        const int no = sp.width[l + 1], ni = sp.width[l];
        const int ao = s_off[2 * l], dofs = s_off[2 * l + 1], ldo = mlp_ld(no);
        const int act = (l == L - 1) ? ACE_ACT_LINEAR : sp.act;
        const float* bias = STAGE ? nullptr : Wg + sp.bl_off[l];
        const int boff = wofs + sp.bl_off[l];
        const float cc = sp.c;
        auto epi = [&](const mfma_f4* acc, int nt, int o0, int ts) {
            for (int t = 0; t < nt; ++t) {
                const int o = o0 + ts * t + cr;
                if (o < no) {
                    const float bo = STAGE ? mlp_smem[boff + o] : bias[o];
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        float a, d;
                        act_fd(act, cc, acc[t][r] + bo, a, d);
                        mlp_smem[ao + (rq + r) * ldo + o] = a;
                        mlp_smem[dofs + (rq + r) * ldo + o] = d;
                    }
                }
            }
        };
        if (BF && l == 0)   // forward block: B[k][o] = W_0[o][k] at [o][k], round16(w_1) x round32(K) per part
            bf3_layer(ain, lda, ni, w0bf + sp.bf_off, ((no + 15) & ~15) * ((ni + 31) & ~31), no, lane, epi, wv, W);
        else
            mfma_layer<STAGE>(ain, lda, ni, wofs + sp.wl_off[l], Wg + sp.wl_off[l], 1, mlp_ld(no), no, lane, epi, wv, W);
        ain = ao;
        lda = ldo;
        sync();
    }
    // energies: out = outb + outw a_last (the last layer has width 1)
    const int aL = s_off[2 * (L - 1)], dL = s_off[2 * (L - 1) + 1], ldL = mlp_ld(1);
    if (wv == 0 && lane < nr) e_atom[atom_of_row[r0 + lane]] = (double)(sp.outb + sp.outw * mlp_smem[aL + lane * ldL]) + sp.E0;
    sync();
    // backward: delta_l overwrites A_l; delta_{l-1} = (delta_l W_l) * f'(z_{l-1})
    if (wv == 0 && lane < MLP_ROWS) mlp_smem[aL + lane * ldL] = sp.outw * mlp_smem[dL + lane * ldL];
    sync();
    for (int l = L - 1; l > 0; --l) {
        const int no = sp.width[l + 1], ni = sp.width[l];
        const int al = s_off[2 * l], am = s_off[2 * (l - 1)], dm = s_off[2 * (l - 1) + 1];
        const int ldl = mlp_ld(no), ldm = mlp_ld(ni);
        // B[o][k] = W_l[o][k]: output column k, reduction index o
        mfma_layer<STAGE>(al, ldl, no, wofs + sp.wl_off[l], Wg + sp.wl_off[l], mlp_ld(no), 1, ni, lane,
                          [&](const mfma_f4* acc, int nt, int k0, int ts) {
                              for (int t = 0; t < nt; ++t) {
                                  const int k = k0 + ts * t + cr;
                                  if (k < ni) {
#pragma unroll
                                      for (int r = 0; r < 4; ++r) {
                                          const int e = (rq + r) * ldm + k;
                                          mlp_smem[am + e] = acc[t][r] * mlp_smem[dm + e];
                                      }
                                  }
                              }
                          }, wv, W);
        sync();
    }
    // dE/dG = (delta_0 W_0) diag(inpw)
    {
        const int no = sp.width[1];
        const float* iw = inpw + sp.in_off;
        const int a0 = s_off[0], ld0 = mlp_ld(no);
        auto epi = [&](const mfma_f4* acc, int nt, int k0, int ts) {
            for (int t = 0; t < nt; ++t) {
                const int k = k0 + ts * t + cr;
                if (k < K) {
                    const float w = iw[k];
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        if (rq + r < nr) dEdG[(long long)(r0 + rq + r) * ldg + k] = acc[t][r] * w;
                }
            }
        };
        if (BF) {   // backward block after the forward one: B[o][k] = W_0[o][k] at [k][o], round16(K) x round32(w_1)
            const int fwd = ((no + 15) & ~15) * ((K + 31) & ~31);
            bf3_layer(a0, ld0, no, w0bf + sp.bf_off + 2 * fwd, ((K + 15) & ~15) * ((no + 31) & ~31), K, lane, epi, wv, W);
        } else {
            mfma_layer<STAGE>(a0, ld0, no, wofs + sp.wl_off[0], Wg + sp.wl_off[0], mlp_ld(no), 1, K, lane, epi, wv, W);
        }
    }
}
