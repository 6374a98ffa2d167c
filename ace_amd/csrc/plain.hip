// plain.hip -- unbatched FP64 GPU port of the reference loops: one structure per launch, one
// thread per atom, FP64 throughout, force scatter with FP64 atomics.  It follows NNP::compute_symm
// (src/nnp/nnp.cpp:350-392) and NNP::compute_force (src/nnp/nnp.cpp:412-476) literally, with the
// reference's fmexp (src/math/special.cpp:226-276) and pow.  It is the denominator of the batched
// path's speed-up, not a product path of its own.  Forms: GAUSSIAN radial, GAUSS angular, COS
// cutoff, TANH/LINEAR activations.
#include "ace_common.hpp"

#define PL_KMAX 256   // G row stride of the port's scratch
#define PL_WMAX 128   // widest input or layer of the port
#define PL_LMAX 4     // deepest network of the port

struct PlainSpecies {
    int K, nR, L, act;
    int r_n[ACE_MAXT], r_goff[ACE_MAXT], r_poff[ACE_MAXT];
    double r_rc[ACE_MAXT];
    int a_n[ACE_MAXB], a_goff[ACE_MAXB], a_poff[ACE_MAXB];
    double a_rc[ACE_MAXB];
    int width[ACE_MAXL + 1];
    int boff[ACE_MAXL], woff[ACE_MAXL];
    int poff, in_off;
    double c, outw, outb, E0;
};
struct PlainModel {
    int ntypes;
    PlainSpecies sp[ACE_MAXT];
};
struct PlainParams {
    const double* r_rs;
    const double* r_eta;
    const double* a_eta;
    const double* a_zeta;
    const double* a_lamf;
    const double* nn;      // network parameters
    const double* inpw;
    const double* inpb;
};

__device__ inline double pl_fmexp(double x) {
    const double l2e = 1.4426950408889634074;
    if (x < -1022.0 / l2e) return 0;
    if (x > 1023.0 / l2e) return __longlong_as_double(0x7ff0000000000000LL);
    x *= l2e;
    const double ipart = floor(x + 0.5);
    const double fpart = x - ipart;
    const double epart = __longlong_as_double((long long)((((int)ipart) + 1023)) << 52);
    x = fpart * fpart;
    double px = 2.30933477057345225087e-2;
    px = px * x + 2.02020656693165307700e1;
    double qx = x + 2.33184211722314911771e2;
    px = px * x + 1.51390680115615096133e3;
    qx = qx * x + 4.36821166879210612817e3;
    px = px * fpart;
    return epart * (1.0 + 2.0 * (px / (qx - px)));
}
__device__ inline double pl_coscut(double x) {
    x -= 0.5 * 3.141592653589793;
    const double x2 = x * x;
    return x * (-1.0 + x2 * (1.0 / 6.0 + x2 * (-1.0 / 120.0 + x2 * (1.0 / 5040.0 + x2 * (-1.0 / 362880.0
           + x2 * (1.0 / 39916800.0 + x2 * (-1.0 / 6227020800.0 + x2 * 1.0 / 1307674368000.0)))))));
}
__device__ inline double pl_sincut(double x) {
    x -= 0.5 * 3.141592653589793;
    const double x2 = x * x;
    return 1.0 + x2 * (-1.0 / 2.0 + x2 * (1.0 / 24.0 + x2 * (-1.0 / 720.0 + x2 * (1.0 / 40320.0
           + x2 * (-1.0 / 3628800.0 + x2 * (1.0 / 479001600.0 + x2 * (-1.0 / 87178291200.0
           + x2 * 1.0 / 20922789888000.0)))))));
}
__device__ inline double pl_cutf(double rc, double r) {
    return (r < rc) ? 0.5 * (pl_coscut(r * 3.141592653589793 / rc) + 1.0) : 0.0;
}
__device__ inline double pl_cutg(double rc, double r) {
    const double pirc = 3.141592653589793 / rc;
    return (r < rc) ? -0.5 * pirc * pl_sincut(r * pirc) : 0.0;
}

__device__ inline void pl_disp(const double* R, const double* pos, int i, int j, int im, double* out) {
    int a, b, c;
    img_unpack(im, a, b, c);
    for (int d = 0; d < 3; ++d) out[d] = pos[3 * i + d] - (pos[3 * j + d] + (R[d] * a + R[d + 3] * b + R[d + 6] * c));
}

// thread per atom of one structure
__global__ void k_plain_symm(int n, int a0, const PlainModel* __restrict__ M, PlainParams P,
                             const double* __restrict__ pos, const int* __restrict__ type, const double* __restrict__ R,
                             const int* __restrict__ nb_off, const int* __restrict__ nb_j, const int* __restrict__ nb_img,
                             double* __restrict__ G) {
    const int li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= n) return;
    const int i = a0 + li;
    const int II = type[i];
    const PlainSpecies& sp = M->sp[II];
    double* Gi = G + (long long)li * PL_KMAX;
    for (int k = 0; k < sp.K; ++k) Gi[k] = 0;
    for (int j = nb_off[i]; j < nb_off[i + 1]; ++j) {
        const int jj = nb_j[j], JJ = type[jj];
        double rIJ[3];
        pl_disp(R, pos, i, jj, nb_img[j], rIJ);
        const double drIJ = sqrt(rIJ[0] * rIJ[0] + rIJ[1] * rIJ[1] + rIJ[2] * rIJ[2]);
        {
            const double cf = pl_cutf(sp.r_rc[JJ], drIJ);
            for (int q = 0; q < sp.r_n[JJ]; ++q) {
                const double arg = P.r_eta[sp.r_poff[JJ] + q] * (drIJ - P.r_rs[sp.r_poff[JJ] + q]);
                Gi[sp.r_goff[JJ] + q] += pl_fmexp(-arg * arg) * cf;
            }
        }
        for (int k = j + 1; k < nb_off[i + 1]; ++k) {
            const int kk = nb_j[k], KK = type[kk];
            double rIK[3];
            pl_disp(R, pos, i, kk, nb_img[k], rIK);
            const double drIK = sqrt(rIK[0] * rIK[0] + rIK[1] * rIK[1] + rIK[2] * rIK[2]);
            const double cosv = (rIJ[0] * rIK[0] + rIJ[1] * rIK[1] + rIJ[2] * rIK[2]) / (drIJ * drIK);
            const int b = tri_index(JJ, KK);
            const double cprod = pl_cutf(sp.a_rc[b], drIJ) * pl_cutf(sp.a_rc[b], drIK);
            const double r2s = drIJ * drIJ + drIK * drIK;
            for (int q = 0; q < sp.a_n[b]; ++q) {
                const int o = sp.a_poff[b] + q;
                const double eta = P.a_eta[o];
                Gi[sp.a_goff[b] + q] += cprod * pow(fabs(0.5 + P.a_lamf[o] * cosv), P.a_zeta[o]) *
                                        pl_fmexp(-r2s / (eta * eta));
            }
        }
    }
}

__global__ void k_plain_force(int n, int a0, const PlainModel* __restrict__ M, PlainParams P,
                              const double* __restrict__ pos, const int* __restrict__ type, const double* __restrict__ R,
                              const int* __restrict__ nb_off, const int* __restrict__ nb_j, const int* __restrict__ nb_img,
                              const double* __restrict__ G, double* __restrict__ F, double* __restrict__ E) {
    const int li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= n) return;
    const int i = a0 + li;
    const int II = type[i];
    const PlainSpecies& sp = M->sp[II];
    const double* Gi = G + (long long)li * PL_KMAX;
    const double* p = P.nn + sp.poff;
    // ANN::fpbp
    double x[PL_WMAX];
    double a[PL_LMAX][PL_WMAX], d[PL_LMAX][PL_WMAX];
    for (int k = 0; k < sp.K; ++k) x[k] = P.inpw[sp.in_off + k] * (Gi[k] + P.inpb[sp.in_off + k]);
    for (int l = 0; l < sp.L; ++l) {
        const int no = sp.width[l + 1], ni = sp.width[l];
        const double* W = p + sp.woff[l];
        const double* in = (l == 0) ? x : a[l - 1];
        for (int o = 0; o < no; ++o) {
            double z = 0;
            for (int k = 0; k < ni; ++k) z += W[o + k * no] * in[k];
            z += p[sp.boff[l] + o];
            if (l == sp.L - 1 || sp.act == ACE_ACT_LINEAR) { a[l][o] = z; d[l][o] = 1.0; }
            else { const double t = tanh(sp.c * z); a[l][o] = t; d[l][o] = sp.c * (1.0 - t * t); }
        }
    }
    atomicAdd(&E[0], sp.outb + sp.outw * a[sp.L - 1][0] + sp.E0);
    // DODZ::grad row 0
    double cur[PL_WMAX], nxt[PL_WMAX];
    cur[0] = sp.outw * d[sp.L - 1][0];
    for (int l = sp.L - 1; l > 0; --l) {
        const int no = sp.width[l + 1], ni = sp.width[l];
        const double* W = p + sp.woff[l];
        for (int k = 0; k < ni; ++k) {
            double s = 0;
            for (int o = 0; o < no; ++o) s += cur[o] * W[o + k * no];
            nxt[k] = s * d[l - 1][k];
        }
        for (int k = 0; k < ni; ++k) cur[k] = nxt[k];
    }
    double* dEdG = x;   // inputs are no longer needed
    {
        const int no = sp.width[1];
        const double* W = p + sp.woff[0];
        for (int k = 0; k < sp.K; ++k) {
            double s = 0;
            for (int o = 0; o < no; ++o) s += cur[o] * W[o + k * no];
            dEdG[k] = s * P.inpw[sp.in_off + k];
        }
    }
    double fi[3] = {0, 0, 0};
    for (int j = nb_off[i]; j < nb_off[i + 1]; ++j) {
        const int jj = nb_j[j], JJ = type[jj];
        double rIJ[3];
        pl_disp(R, pos, i, jj, nb_img[j], rIJ);
        const double drIJ = sqrt(rIJ[0] * rIJ[0] + rIJ[1] * rIJ[1] + rIJ[2] * rIJ[2]);
        const double dIJi = 1.0 / drIJ;
        double amp = 0;
        {
            const double cf = pl_cutf(sp.r_rc[JJ], drIJ), cg = pl_cutg(sp.r_rc[JJ], drIJ);
            for (int q = 0; q < sp.r_n[JJ]; ++q) {
                const double eta = P.r_eta[sp.r_poff[JJ] + q];
                const double arg = eta * (drIJ - P.r_rs[sp.r_poff[JJ] + q]);
                amp -= dEdG[sp.r_goff[JJ] + q] * pl_fmexp(-arg * arg) * (-2.0 * eta * arg * cf + cg);
            }
        }
        amp *= dIJi;
        for (int dd = 0; dd < 3; ++dd) {
            fi[dd] += amp * rIJ[dd];
            atomicAdd(&F[3 * jj + dd], -amp * rIJ[dd]);
        }
        for (int k = j + 1; k < nb_off[i + 1]; ++k) {
            const int kk = nb_j[k], KK = type[kk];
            double rIK[3];
            pl_disp(R, pos, i, kk, nb_img[k], rIK);
            const double drIK = sqrt(rIK[0] * rIK[0] + rIK[1] * rIK[1] + rIK[2] * rIK[2]);
            const double dIKi = 1.0 / drIK;
            const double cosv = (rIJ[0] * rIK[0] + rIJ[1] * rIK[1] + rIJ[2] * rIK[2]) * dIJi * dIKi;
            const int b = tri_index(JJ, KK);
            const double c0 = pl_cutf(sp.a_rc[b], drIJ), c1 = pl_cutf(sp.a_rc[b], drIK);
            const double g0 = pl_cutg(sp.a_rc[b], drIJ), g1 = pl_cutg(sp.a_rc[b], drIK);
            const double r2s = drIJ * drIJ + drIK * drIK;
            const double dij = 2.0 * drIJ * c0, dik = 2.0 * drIK * c1;
            double phi = 0, e0 = 0, e1 = 0;
            for (int q = 0; q < sp.a_n[b]; ++q) {
                const int o = sp.a_poff[b] + q;
                const double eta = P.a_eta[o], ieta2 = 1.0 / (eta * eta);
                const double lamf = P.a_lamf[o], zeta = P.a_zeta[o];
                const double cw = fabs(0.5 + lamf * cosv);
                const double gangle = dEdG[sp.a_goff[b] + q] * pow(cw, zeta - 1.0) * pl_fmexp(-ieta2 * r2s);
                const double fangle = cw * gangle;
                phi -= zeta * lamf * gangle;
                e0 -= fangle * (-dij * ieta2 + g0);
                e1 -= fangle * (-dik * ieta2 + g1);
            }
            phi *= c0 * c1;
            e0 *= c1;
            e1 *= c0;
            for (int dd = 0; dd < 3; ++dd) {
                fi[dd] += (phi * (dIKi - cosv * dIJi) + e0) * rIJ[dd] * dIJi;
                fi[dd] += (phi * (dIJi - cosv * dIKi) + e1) * rIK[dd] * dIKi;
                atomicAdd(&F[3 * jj + dd], -((-phi * cosv * dIJi + e0) * rIJ[dd] * dIJi + phi * dIJi * rIK[dd] * dIKi));
                atomicAdd(&F[3 * kk + dd], -((-phi * cosv * dIKi + e1) * rIK[dd] * dIKi + phi * dIKi * rIJ[dd] * dIJi));
            }
        }
    }
    for (int dd = 0; dd < 3; ++dd) atomicAdd(&F[3 * i + dd], fi[dd]);
}
