// basis.hip -- symmetry functions G (NNP::compute_symm, src/nnp/nnp.cpp:350-392) and their adjoint
// dE/dd for every neighbour pair (the geometric half of NNP::compute_force, src/nnp/nnp.cpp:412-476).
//
// One wavefront per centre atom.  The atom's neighbours inside the basis cutoff are bucketed by
// species into LDS (unit vector, distance, and per distinct angular eta the separable factor
// E_s = fc(r_s) exp(-r_s^2/eta^2) with its r-derivative), so a GAUSS triplet term
//   fc(rij) fc(rik) |1/2 + lambda/2 cos|^zeta exp(-(rij^2+rik^2)/eta^2)      (basis_angular.cpp:371-382)
// costs one product E_j E_k instead of an exp.  Triplets of an angular block (J,K) -- the rectangle
// J x K, or the triangle j<k of J x J -- are flattened over the 64 lanes; each lane keeps the block's
// functions in registers and the wave reduces them at the end of the block.
#include "ace_common.hpp"

#define BW 4            // waves per workgroup

// Modes: 0 = G rows; 1 = adjoint (per-pair dE/dd from dE/dG rows); 2 = tangent (directional
// derivative of G along per-atom displacements eps: v_i = sum_a eps_a . dG_i/dr_a).
// Dynamic LDS, per wave: u[nbm] (float4: unit vector, r), E[ng][nbm], row[kmax]; modes 1 and 2 add
// Ed[ng][nbm]; mode 1 adds g[3 nbm], p[nbm]; mode 2 adds te[nbm] (float4: eps_i - eps_j, projection).
// nbm = neighbour slots per atom (largest list, rounded up).
__host__ __device__ inline int basis_lds_wave_bytes(int nbm, int ng, int kmax, int mode) {
    int b = 16 * nbm + 4 * ng * nbm + 4 * ((kmax + 3) & ~3);
    if (mode >= 1) b += 4 * ng * nbm;
    if (mode == 1) b += 12 * nbm + 4 * nbm;
    if (mode == 2) b += 16 * nbm;
    return (b + 15) & ~15;
}

struct BasisLds {
    float4* u;
    float4* te;
    float* E;
    float* Ed;
    float* g;
    int* p;
    float* row;
    int nbm;
};

__device__ inline BasisLds basis_lds(void* base, int wv, int nbm, int ng, int kmax, int mode) {
    char* c = (char*)base + (size_t)wv * basis_lds_wave_bytes(nbm, ng, kmax, mode);
    BasisLds L;
    L.nbm = nbm;
    L.u = (float4*)c; c += 16 * nbm;
    L.te = (float4*)c; if (mode == 2) c += 16 * nbm;
    L.E = (float*)c; c += 4 * ng * nbm;
    L.row = (float*)c; c += 4 * ((kmax + 3) & ~3);
    L.Ed = (float*)c; if (mode >= 1) c += 4 * ng * nbm;
    L.g = (float*)c; if (mode == 1) c += 12 * nbm;
    L.p = (int*)c;
    return L;
}

// (a, c) of flattened triplet t in block (J,K): rectangle a in J, c in K; triangle a<c in J.
__device__ inline void tri_decode(int t, bool same, int nJ, int nK, int& a, int& c) {
    if (!same) {
        a = t / nK;
        c = t - a * nK;
    } else {
        const float n2 = 2.f * nJ - 1.f;
        int r = (int)floorf(0.5f * (n2 - sqrtf(n2 * n2 - 8.f * (float)t)));
        if (r < 0) r = 0;
        // S(r) = r (2 nJ - r - 1) / 2 pairs precede row r
        while (r > 0 && (r * (2 * nJ - r - 1)) / 2 > t) --r;
        while (((r + 1) * (2 * nJ - r - 2)) / 2 <= t) ++r;
        a = r;
        c = t - (r * (2 * nJ - r - 1)) / 2 + r + 1;
    }
}

// angular damping product and its partials for group g (GAUSS: separable)
template <int ADJ>
__device__ inline void load_groups(const BasisLds& L, int ng, int sa, int sc, float* Da, float* Dc,
                                   float* Dda, float* Ddc) {
#pragma unroll
    for (int g = 0; g < ACE_MAXG; ++g) {
        if (g < ng) {
            Da[g] = L.E[g * L.nbm + sa];
            Dc[g] = L.E[g * L.nbm + sc];
            if (ADJ) {
                Dda[g] = L.Ed[g * L.nbm + sa];
                Ddc[g] = L.Ed[g * L.nbm + sc];
            }
        } else {
            Da[g] = 0.f; Dc[g] = 0.f;
            if (ADJ) { Dda[g] = 0.f; Ddc[g] = 0.f; }
        }
    }
}

__device__ inline float pick4(const float* v, int g) {
    float r = v[0];
    r = (g == 1) ? v[1] : r;
    r = (g == 2) ? v[2] : r;
    r = (g == 3) ? v[3] : r;
    return r;
}

template <int MAXA, int MAXR, int MODE>
__global__ __launch_bounds__(BW * 64) void k_basis(
    int natoms, const DevModel* __restrict__ M, DevBasisParams P, const int* __restrict__ type,
    const int* __restrict__ nb_off, const int* __restrict__ nb_j, const float4* __restrict__ nb_d,
    const int* __restrict__ row_of, int ldg, float* __restrict__ G, const float* __restrict__ dEdG,
    float4* __restrict__ pair_g, int* __restrict__ err, int nbm, const double* __restrict__ eps) {
    constexpr int ADJ = (MODE == 1);
    constexpr int TAN = (MODE == 2);
    extern __shared__ float4 basis_smem[];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * BW + wv;
    if (i >= natoms) return;
    const int I = type[i];
    const DevSpecies& sp = M->sp[I];
    const BasisLds L = basis_lds(basis_smem, wv, nbm, M->ng_max, M->kmax, MODE);
    const int nt = M->ntypes;
    const int o0 = nb_off[i], o1 = nb_off[i + 1];
    const float rcut = sp.rc_max;
    const int row = row_of[i];

    // --- bucket the neighbours by species ---------------------------------------------------
    int cnt[ACE_MAXT];
#pragma unroll
    for (int J = 0; J < ACE_MAXT; ++J) cnt[J] = 0;
    for (int cb = o0; cb < o1; cb += 64) {
        const int p = cb + lane;
        int Jl = -1;
        if (p < o1 && nb_d[p].w < rcut) Jl = type[nb_j[p]];
#pragma unroll
        for (int J = 0; J < ACE_MAXT; ++J)
            if (J < nt) cnt[J] += __popcll(__ballot(Jl == J));
    }
    int bstart[ACE_MAXT + 1];
    bstart[0] = 0;
#pragma unroll
    for (int J = 0; J < ACE_MAXT; ++J) bstart[J + 1] = bstart[J] + ((J < nt) ? cnt[J] : 0);
    const int ntot = bstart[ACE_MAXT];
    if (ntot > nbm) {
        if (lane == 0) atomicAdd(err, 1);
        return;
    }
    int fill[ACE_MAXT];
#pragma unroll
    for (int J = 0; J < ACE_MAXT; ++J) fill[J] = bstart[J];
    for (int cb = o0; cb < o1; cb += 64) {
        const int p = cb + lane;
        int Jl = -1;
        float4 d = make_float4(0.f, 0.f, 0.f, 0.f);
        if (p < o1) {
            d = nb_d[p];
            if (d.w < rcut) Jl = type[nb_j[p]];
            else if (ADJ) pair_g[p] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        int slot = -1;
#pragma unroll
        for (int J = 0; J < ACE_MAXT; ++J) {
            if (J < nt) {
                const unsigned long long m = __ballot(Jl == J);
                if (Jl == J) slot = fill[J] + __popcll(m & lane_mask_lt());
                fill[J] += __popcll(m);
            }
        }
        if (slot >= 0) {
            const float r = d.w, ir = 1.f / r;
            L.u[slot] = make_float4(d.x * ir, d.y * ir, d.z * ir, r);
            if (ADJ) L.p[slot] = p;
            float fc, gc;
            cutoff_fg(sp.a_cut, sp.a_rc, r, fc, gc);
#pragma unroll
            for (int g = 0; g < ACE_MAXG; ++g) {
                if (g < sp.ng) {
                    const float e = fast_exp(-r * r * sp.ietar2[g]);
                    L.E[g * nbm + slot] = fc * e;
                    if (ADJ || TAN) L.Ed[g * nbm + slot] = (gc - 2.f * r * sp.ietar2[g] * fc) * e;
                }
            }
            if (ADJ) {
                L.g[3 * slot] = 0.f;
                L.g[3 * slot + 1] = 0.f;
                L.g[3 * slot + 2] = 0.f;
            }
            if (TAN) {
                const int jj = nb_j[p];
                const float ex = (float)(eps[3 * i] - eps[3 * jj]);
                const float ey = (float)(eps[3 * i + 1] - eps[3 * jj + 1]);
                const float ez = (float)(eps[3 * i + 2] - eps[3 * jj + 2]);
                L.te[slot] = make_float4(ex, ey, ez, (d.x * ex + d.y * ey + d.z * ez) * ir);
            }
        }
    }
    // row buffer: zero (forward) or dE/dG (adjoint)
    for (int k = lane; k < sp.K; k += 64) L.row[k] = ADJ ? dEdG[(long long)row * ldg + k] : 0.f;
    wave_sync();

    // --- radial blocks -------------------------------------------------------------------------
    for (int J = 0; J < nt; ++J) {
        const int nJ = cnt[J];
        const int nR = sp.r_n[J];
        if (nJ == 0 || nR == 0) continue;
        const int form = sp.r_form[J], cut = sp.r_cut[J];
        const float rc = sp.r_rc[J];
        const int goff = sp.r_goff[J], poff = sp.r_poff[J];
        for (int n0 = 0; n0 < nR; n0 += MAXR) {
            float acc[MAXR];
#pragma unroll
            for (int n = 0; n < MAXR; ++n) acc[n] = 0.f;
            for (int q = lane; q < nJ; q += 64) {
                const int slot = bstart[J] + q;
                const float4 u = L.u[slot];
                float fc, gc;
                cutoff_fg(cut, rc, u.w, fc, gc);
                float dEdr = 0.f;
#pragma unroll
                for (int n = 0; n < MAXR; ++n) {
                    if (n0 + n < nR) {
                        float v, dv;
                        radial_fd(form, P.r_eta[poff + n0 + n], P.r_rs[poff + n0 + n], u.w, fc, gc, v, dv);
                        if (ADJ) dEdr += L.row[goff + n0 + n] * dv;
                        else if (TAN) acc[n] += dv * L.te[slot].w;
                        else acc[n] += v;
                    }
                }
                if (ADJ) {
                    L.g[3 * slot] += dEdr * u.x;
                    L.g[3 * slot + 1] += dEdr * u.y;
                    L.g[3 * slot + 2] += dEdr * u.z;
                }
            }
            if (!ADJ) {
#pragma unroll
                for (int n = 0; n < MAXR; ++n) {
                    if (n0 + n < nR) {
                        const float t = wave_sum(acc[n]);
                        if (lane == 0) L.row[goff + n0 + n] = t;
                    }
                }
            }
        }
    }
    wave_sync();

    // --- angular blocks (J <= K) ------------------------------------------------------------------
    const int ng = sp.ng;
    for (int K = 0; K < nt; ++K) {
        for (int J = 0; J <= K; ++J) {
            const int b = K * (K + 1) / 2 + J;
            const int nA = sp.a_n[b];
            const bool same = (J == K);
            const int nJ = cnt[J], nK = cnt[K];
            const int T = same ? nJ * (nJ - 1) / 2 : nJ * nK;
            if (nA == 0 || T == 0) continue;
            const int goff = sp.a_goff[b], poff = sp.a_poff[b];
            const int bJ = bstart[J], bK = bstart[K];
            for (int m0 = 0; m0 < nA; m0 += MAXA) {
                float acc[MAXA];
                float w[MAXA];
#pragma unroll
                for (int m = 0; m < MAXA; ++m) {
                    acc[m] = 0.f;
                    w[m] = (ADJ && m0 + m < nA) ? L.row[goff + m0 + m] : 0.f;
                }
                for (int t = lane; t < T; t += 64) {
                    int a, c;
                    tri_decode(t, same, nJ, nK, a, c);
                    const int sa = bJ + a, sc = bK + c;
                    const float4 ua = L.u[sa], uc = L.u[sc];
                    const float cs = ua.x * uc.x + ua.y * uc.y + ua.z * uc.z;
                    float Da[ACE_MAXG], Dc[ACE_MAXG], Dda[ACE_MAXG], Ddc[ACE_MAXG];
                    load_groups<(MODE >= 1)>(L, ng, sa, sc, Da, Dc, Dda, Ddc);
                    if (MODE == 0) {
                        float D[ACE_MAXG];
#pragma unroll
                        for (int g = 0; g < ACE_MAXG; ++g) D[g] = Da[g] * Dc[g];
#pragma unroll
                        for (int m = 0; m < MAXA; ++m) {
                            if (m0 + m < nA) {
                                const int fi = poff + m0 + m;
                                const float base = fabsf(0.5f + P.a_lamf[fi] * cs);
                                const float pw = pow_window(base, P.a_zeta[fi], P.a_zint[fi]);
                                acc[m] += pw * pick4(D, P.a_grp[fi]);
                            }
                        }
                    } else if (TAN) {
                        const float4 ta = L.te[sa], tc = L.te[sc];
                        const float dcs = (uc.x * ta.x + uc.y * ta.y + uc.z * ta.z - cs * ta.w) / ua.w +
                                          (ua.x * tc.x + ua.y * tc.y + ua.z * tc.z - cs * tc.w) / uc.w;
                        float X[ACE_MAXG], Y[ACE_MAXG];
#pragma unroll
                        for (int g = 0; g < ACE_MAXG; ++g) {
                            X[g] = Da[g] * Dc[g] * dcs;
                            Y[g] = Dda[g] * Dc[g] * ta.w + Da[g] * Ddc[g] * tc.w;
                        }
#pragma unroll
                        for (int m = 0; m < MAXA; ++m) {
                            if (m0 + m < nA) {
                                const int fi = poff + m0 + m;
                                const float lamf = P.a_lamf[fi], zeta = P.a_zeta[fi];
                                const int zint = P.a_zint[fi];
                                const float base = fabsf(0.5f + lamf * cs);
                                const float p1 = (zint >= 0) ? powi_f(base, zint - 1)
                                               : pow_window(base, zeta - 1.f, -1);
                                const int g = P.a_grp[fi];
                                acc[m] += zeta * lamf * p1 * pick4(X, g) + p1 * base * pick4(Y, g);
                            }
                        }
                    } else {
                        float S[ACE_MAXG], Sp[ACE_MAXG];
#pragma unroll
                        for (int g = 0; g < ACE_MAXG; ++g) { S[g] = 0.f; Sp[g] = 0.f; }
#pragma unroll
                        for (int m = 0; m < MAXA; ++m) {
                            if (m0 + m < nA) {
                                const int fi = poff + m0 + m;
                                const float lamf = P.a_lamf[fi], zeta = P.a_zeta[fi];
                                const int zint = P.a_zint[fi];
                                const float base = fabsf(0.5f + lamf * cs);
                                const float p1 = (zint >= 0) ? powi_f(base, zint - 1)
                                               : pow_window(base, zeta - 1.f, -1);
                                const float s = w[m] * p1 * base;
                                const float sp_ = w[m] * zeta * lamf * p1;
                                const int g = P.a_grp[fi];
#pragma unroll
                                for (int gg = 0; gg < ACE_MAXG; ++gg) {
                                    S[gg] += (g == gg) ? s : 0.f;
                                    Sp[gg] += (g == gg) ? sp_ : 0.f;
                                }
                            }
                        }
                        float dEdc = 0.f, dEda = 0.f, dEdcc = 0.f;
#pragma unroll
                        for (int g = 0; g < ACE_MAXG; ++g) {
                            dEdc += Sp[g] * Da[g] * Dc[g];
                            dEda += S[g] * Dda[g] * Dc[g];
                            dEdcc += S[g] * Da[g] * Ddc[g];
                        }
                        const float ia = 1.f / ua.w, ic = 1.f / uc.w;
                        const float ka = dEda - dEdc * cs * ia, kc = dEdcc - dEdc * cs * ic;
                        const float xa = dEdc * ia, xc = dEdc * ic;
                        atomicAdd(&L.g[3 * sa], ka * ua.x + xa * uc.x);
                        atomicAdd(&L.g[3 * sa + 1], ka * ua.y + xa * uc.y);
                        atomicAdd(&L.g[3 * sa + 2], ka * ua.z + xa * uc.z);
                        atomicAdd(&L.g[3 * sc], kc * uc.x + xc * ua.x);
                        atomicAdd(&L.g[3 * sc + 1], kc * uc.y + xc * ua.y);
                        atomicAdd(&L.g[3 * sc + 2], kc * uc.z + xc * ua.z);
                    }
                }
                if (!ADJ) {
#pragma unroll
                    for (int m = 0; m < MAXA; ++m) {
                        if (m0 + m < nA) {
                            const float t = wave_sum(acc[m]);
                            if (lane == 0) L.row[goff + m0 + m] = t;
                        }
                    }
                }
            }
        }
    }
    wave_sync();
    if (!ADJ) {
        for (int k = lane; k < sp.K; k += 64) G[(long long)row * ldg + k] = L.row[k];   // G or v
    } else {
        for (int q = lane; q < ntot; q += 64)
            pair_g[L.p[q]] = make_float4(L.g[3 * q], L.g[3 * q + 1], L.g[3 * q + 2], 0.f);
    }
}

